"""Benchmark of the CCSD(T) triples body on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--slices S]

A step is one pass of the body over the whole triple list of the production
case: nocc=30, nvir=300, k blocks of 16, all 4960 triples i >= j >= k, with the
job's arrays already resident in HBM.  Both forms are timed in the same run:

* the plain FP64 form (FP64 MFMA GEMMs with the fused energy epilogue), and
* the split form (the contractions as Ozaki INT8 split products with S slices
  and FP64 recombination -- for S <= 3 both halves, K=nvir and K=nocc, each with
  its own row scales; for larger S the K=nocc half in FP64 -- and the FP64
  energy epilogue).

Each form runs the configuration the library's autotuner (ccsdt_job_autotune)
picks on the full list -- work-item order, tile variant, plain or persistent
grid, and for the split form the slice count (the smallest whose E[T] is within
1e-9 of the FP64 form's) -- captured in a hipGraph and replayed once per step.
--fp64-opts / --split-opts pin a configuration instead (profiling runs).

`value` is the split form's FP64-equivalent TFLOP/s (16*nvir^2*(nvir+nocc) flops
per triple, 2 per multiply-add) when its E[T] = emp4+emp5 agrees with the FP64
form's to 1e-8 relative, else the FP64 form's.  The JSON line also carries:

* the roofline of the dominant kernel (HIP events on the library's stream);
* per-step p50/p99 latency (step_p50_ms, step_p99_ms) and the harness's own
  OPS/PERF figures (omp/test.c:216-226) for both forms;
* the separate energy pass's HBM rate and the synchronous drop-in's
  PCIe-inclusive rate;
* a serving sample (48 jobs of the serving workload, Poisson at 40 jobs/s,
  after one unmeasured warm pass): serving_job_p50_ms / serving_job_p99_ms;
* the same list as a full-tensor job (distinct slices per occupied index);
* the CPU oracle timed on the host's cores over the whole list after a
  warm-up call (cpu_baseline), and from the same run the oracle's E[T], each
  form's rel_dE_vs_oracle and the FP64 form's largest per-term error.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "(T) FP64-eq TFLOP/s nocc=30 nvir=300, split vs FP64 @rel dE<1e-8; job p99"
FP64_MFMA_PEAK = 78.6   # TFLOP/s, MI355X dense FP64 matrix
INT8_MFMA_PEAK = 5000.0  # TOP/s, MI355X dense INT8 matrix
HBM_PEAK = 8000.0        # GB/s
READ_REF_BYTES = 4 << 30  # size of the streaming-read and copy reference passes
ET_TOL = 1e-8


def load_pkg():
    import importlib.util
    name = "nwchem_ccsd_trpdrv_amd"
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(ROOT, "nwchem-ccsd-trpdrv_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def traffic_from_profile(form, opts):
    """HBM-side bytes per launch of the form's fused kernel from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated):
    the entry taken with the same run options, or None when there is none."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_nocc30_nvir300.json" % form)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    keys = ("order", "variant", "persistent", "slices", "occ_digits")
    for e in d.get("entries", []):
        prof = e.get("opts", {})
        if all(prof.get(k, 0) == opts.get(k, 0) for k in keys):
            return e.get("bytes_per_launch")
    return None


def kernel_name(opts):
    """Name of the fused kernel a configuration launches."""
    p = "true" if opts.get("persistent") else "false"
    if opts["mode"] == "ozaki_int8":
        # S = 3 with occupied digit planes is template code 3 + 16 (kSlicesOccDigits)
        kern = "k_split_rows" if opts.get("variant", 0) == 2 else "k_fused_split"
        return "%s<%d,true,false,%s>" % (kern, opts["slices"] + 16 * bool(opts.get("occ_digits")), p)
    return ("k_fused" if opts.get("variant", 0) == 0 else "k_fused2") + "<true,false,%s>" % p


def time_form(job, trip, steps, warmup, **kw):
    for _ in range(warmup):
        job.run(trip, **kw)
    step_ms, kern_ms, res = [], [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        res = job.run(trip, terms=True, **kw)
        step_ms.append((time.perf_counter() - s0) * 1e3)
        kern_ms.append(res["ms_main"])
    wall = time.perf_counter() - t0
    return {"ms_per_step": wall * 1e3 / steps, "kernel_ms": float(np.mean(kern_ms)),
            "p50_ms": float(np.percentile(step_ms, 50)),
            "p99_ms": float(np.percentile(step_ms, 99)), "res": res}


def job_alone(cc, jb):
    if "tensors" in jb:
        return cc.Job.from_tensors(jb["tensors"], jb["nocc"], jb["nvir"])
    return cc.Job(jb["arrays"], jb["nocc"], jb["nvir"], jb["kchunk"])


def run_serving(cc, jobs, arrivals, mode, slices, items_per_tick=0, budget=0, occ_digits=0,
                variant=-1):
    """Feeds the jobs to a server at their arrival times (seconds after start)
    and returns per-job results, latency (ms), the span (s) and the server's
    options (tile and tick size as resolved: -1 / 0 take the tuning cache's
    server pick, else tile 0 and 131072 items)."""
    return cc.serve(jobs, arrivals, budget_bytes=budget, mode=mode, slices=slices,
                    items_per_tick=items_per_tick, occ_digits=occ_digits, variant=variant)


def serving_report(cc, rate, njobs, mode, slices, seed=0xC0DE0004, check=True, tensors=False,
                   budget=0, oracle_gflop=20.0, jobs=None, occ_digits=0, variant=-1,
                   items_per_tick=0):
    """One Poisson serving run.  check: every job's E[T] against the same job
    run alone in the FP64 form on the GPU, and the jobs of at most
    oracle_gflop GEMM work against the CPU oracle (on gathered arrays for
    tensor jobs)."""
    if jobs is None:
        jobs = cc.serving_jobs(njobs, seed, tensors=tensors)
    arrivals = cc.poisson_arrivals(njobs, rate, seed)
    occ_digits = int(occ_digits) if mode == "ozaki_int8" else 0
    res, lat, span, sopts = run_serving(cc, jobs, arrivals, mode, slices, budget=budget,
                                        occ_digits=occ_digits, variant=variant,
                                        items_per_tick=items_per_tick)
    flops = sum(len(jb["triples"]) * cc.gemm_flops_per_triple(jb["nocc"], jb["nvir"]) for jb in jobs)
    if mode == "fp64":
        label = mode
    elif slices == cc.SLICES_PER_JOB:
        label = "%s per-job slices (%d jobs at s=3 occ_digits, %d at s=4)" % (
            mode, sum(r["slices"] == 3 for r in res), sum(r["slices"] == 4 for r in res))
    else:
        label = "%s s=%d%s" % (mode, slices, " occ_digits" if occ_digits else "")
    out = {"rate_jobs_per_s": rate, "jobs": njobs, "mode": label, "variant": sopts["variant"],
           "items_per_tick": sopts["items_per_tick"],
           "job_format": "tensors" if tensors else "harness",
           "jobs_per_s": njobs / span, "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
           "fp64_eq_tflops_over_span": flops / span / 1e12,
           "footprint_gb_max": max(r["footprint"] for r in res) / 1e9,
           "budget_gb": (budget or 200e9) / 1e9}
    if check:
        import oracle
        worst, worst_o, nor = 0.0, 0.0, 0
        per_job = []
        for jb, r in zip(jobs, res):
            with job_alone(cc, jb) as job:
                a = job.run(jb["triples"], mode="fp64")
            E = a["emp4"] + a["emp5"]
            dE = abs(r["emp4"] + r["emp5"] - E) / abs(E)
            worst = max(worst, dE)
            per_job.append({"nocc": jb["nocc"], "nvir": jb["nvir"], "triples": len(jb["triples"]),
                            "slices": r["slices"], "occ_digits": r["occ_digits"],
                            "latency_ms": (r["t_finish"] - r["t_submit"]) * 1e3, "rel_dE_vs_fp64": dE})
            if len(jb["triples"]) * cc.gemm_flops_per_triple(jb["nocc"], jb["nvir"]) <= oracle_gflop * 1e9:
                no, nv = jb["nocc"], jb["nvir"]
                if "tensors" in jb:
                    Eo = 0.0
                    for q in jb["triples"]:
                        arr = cc.gather_triple(jb["tensors"], no, nv, int(q[1]), int(q[2]), int(q[3]))
                        t = oracle.run_list(arr, 0, no, nv, np.array([[q[0], q[1], q[2], q[3], q[3]]], np.int32))
                        Eo += t["emp4"] + t["emp5"]
                else:
                    t = oracle.run_list(jb["arrays"], 0, no, nv, jb["triples"])
                    Eo = t["emp4"] + t["emp5"]
                worst_o = max(worst_o, abs(r["emp4"] + r["emp5"] - Eo) / abs(Eo))
                nor += 1
        out["max_rel_dE_vs_fp64_alone"] = worst
        out["per_job"] = per_job
        out["oracle_checked_jobs"] = nor
        out["max_rel_dE_vs_oracle"] = worst_o
    return out


def cpu_baseline(arrays, trip, flops, seconds):
    """The oracle body (the reference's operations, one call per triple, MKL
    DGEMM and the energy loop threaded over all host cores) over the same list,
    after one untimed warm-up call; per-call wall time summed as in
    omp/test.c:166-179.  seconds > 0 stops after that much CPU time (a bounded
    sample), 0 runs the whole list.  Also returns the oracle's terms, summand
    scales and compensated totals of the triples it ran."""
    import oracle
    oracle.blas_select(0)
    cores = oracle.max_threads()
    oracle.run_list(arrays, 0, 30, 300, trip[:1])  # warm-up: MKL and OpenMP start-up
    t0 = time.perf_counter()
    out = oracle.run_list(arrays, 0, 30, 300, trip, with_abs=True, max_seconds=seconds)
    wall = time.perf_counter() - t0
    done = int(out["done"])
    secs = float(np.sum(out["secs"]))
    what = "all %d triples" % done if done == len(trip) else "first %d of the %d triples" % (done, len(trip))
    return {"value": done * flops / secs / 1e12, "unit": "TFLOP/s", "cores": cores,
            "kind": "port",
            "sample": "%s through the oracle body (%s DGEMM, %d OpenMP threads) after one "
                      "warm-up call, per-call wall time summed as in omp/test.c:166-179"
                      % (what, oracle.blas_name(), cores),
            "triples_per_s": done / secs, "seconds": secs, "wall_s": wall}, out


def reference_perf(nocc, nvir, ms_per_triple):
    """The harness's own OPS/PERF figures (omp/test.c:216-226) for an average
    time per triple: dgemm_flops = 8 nvir^2 (nvir+nocc) (one flop per
    multiply-add), dgemm_mops = 8 (4 nvir^2 + 2 nvir nocc), tengy_ops = 94 nvir^2;
    GF/s = (dgemm_flops + tengy_ops) / tavg, GB/s = 8 (dgemm_mops + tengy_ops) / tavg."""
    dgemm_flops = 8.0 * nvir * nvir * (nvir + nocc)
    dgemm_mops = 8.0 * (4.0 * nvir * nvir + 2.0 * nvir * nocc)
    tengy_ops = 94.0 * nvir * nvir
    tavg = ms_per_triple * 1e-3
    return {"tavg_s": tavg, "dgemm_flops": dgemm_flops, "dgemm_mops": dgemm_mops,
            "tengy_ops": tengy_ops, "GF_per_s": 1e-9 * (dgemm_flops + tengy_ops) / tavg,
            "GB_per_s": 8e-9 * (dgemm_mops + tengy_ops) / tavg}


def batched_report(cc, steps, warmup, slices):
    """Config of the batched persistent launch: nocc=40, nvir=400, uneven k blocks
    of 8..48, every (i>=j>=k) triple in one graph-captured launch sequence.  The
    autotuner picks the tile and schedule for each form on the full list, then
    the chosen configuration is timed."""
    nocc, nvir = 40, 400
    arrays, blocks, trip, kchunk = cc.config_case(3)
    n = len(trip)
    flops = n * cc.gemm_flops_per_triple(nocc, nvir)
    out = {"metric": "batched (T) FP64-eq TFLOP/s nocc=40 nvir=400 uneven k blocks",
           "unit": "TFLOP/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
           "higher_is_better": True,
           "config": {"workload": "triples body, nocc=40 nvir=400, uneven k blocks %s, "
                                  "%d triples (i>=j>=k)" % (blocks, n),
                      "nocc": nocc, "nvir": nvir, "kchunk": kchunk, "ntriples": n},
           "forms": {}}
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        # slices 0: the autotuner also picks the split form's slice count
        for mode, S in (("fp64", 0), ("ozaki_int8", slices)):
            t0 = time.perf_counter()
            best, table = job.autotune(trip, mode=mode, slices=S, graph=True, reps=2)
            tune_s = time.perf_counter() - t0
            kw = dict(best)
            r = time_form(job, trip, steps, warmup, **kw)
            print("bench: batched %s %.2f ms/step (%s)" % (mode, r["ms_per_step"], best),
                  file=sys.stderr, flush=True)
            out["forms"][mode] = {
                "value": flops / (r["ms_per_step"] * 1e-3) / 1e12,
                "ms_per_step": r["ms_per_step"], "device_ms": r["kernel_ms"],
                "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
                "emp4": r["res"]["emp4"], "emp5": r["res"]["emp5"],
                "chosen": best, "autotune_seconds": tune_s,
                "candidates": [{k: v for k, v in row.items() if k != "mode"} for row in table]}
    f, sp = out["forms"]["fp64"], out["forms"]["ozaki_int8"]
    e_f = f["emp4"] + f["emp5"]
    dev = abs(sp["emp4"] + sp["emp5"] - e_f) / abs(e_f)
    sp["rel_dE_vs_fp64"] = dev
    sp["within_tol"] = dev <= ET_TOL
    head = sp if dev <= ET_TOL else f
    out["value"] = head["value"]
    out["ms_per_step"] = head["ms_per_step"]
    out["dtype"] = ("i8x%d+f64" % sp["chosen"]["slices"]) if dev <= ET_TOL else "f64"
    out["frac_fp64_peak_fp64_form"] = f["value"] / FP64_MFMA_PEAK
    return out


def dropin_report(cc, arrays, trip, nocc, nvir, flops, ncalls=24):
    """The synchronous drop-in body (ccsd_trpdrv_omp_cbody_) on host buffers:
    every call copies the slices it reads over PCIe and, with f write-back on
    (the reference leaves f1n..f4t filled), copies 8 nvir^2 doubles back."""
    out = {}
    f = [np.zeros(nvir * nvir) for _ in range(8)]
    # pointer arguments built once, as a compiled caller has them
    body = cc.BodyCall(f, arrays["eorb"], 0, nocc, nvir, arrays)
    for wb in (1, 0):
        cc.lib().ccsdt_set_f_writeback(wb)
        for t in trip[:2]:
            body.call(*t)
        t0 = time.perf_counter()
        for t in trip[:ncalls]:
            body.call(*t)
        dt = (time.perf_counter() - t0) / ncalls
        out["f_writeback" if wb else "no_f_writeback"] = {
            "ms_per_call": dt * 1e3, "triples_per_s": 1.0 / dt,
            "fp64_eq_tflops": flops / dt / 1e12}
    cc.lib().ccsdt_set_f_writeback(1)
    out["calls"] = ncalls
    return out


def tensor_job_report(cc, steps, warmup, slices, pinned=None):
    """The nocc=30/nvir=300 list (4960 triples) as a full-tensor job: every
    (i,j) and (k,j) slice distinct, as real amplitudes and integrals are, so
    operand panels are shared between triples only through common occupied
    indices (the harness-shaped job shares its i-indexed panels across the whole
    list).  Both forms autotuned (or run with pinned[mode], JSON run options)
    and timed as graph replays."""
    nocc, nvir = 30, 300
    t = cc.make_tensor_inputs(nocc, nvir, cc.config_seed(1))
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, 16))
    flops = len(trip) * cc.gemm_flops_per_triple(nocc, nvir)
    out = {"metric": "full-tensor job FP64-eq TFLOP/s nocc=30 nvir=300", "unit": "TFLOP/s",
           "n_gpus": 1, "steps": steps, "warmup": warmup,
           "footprint_gb": cc.tensor_footprint(nocc, nvir, len(trip)) / 1e9, "forms": {}}
    with cc.Job.from_tensors(t, nocc, nvir) as job:
        for mode, S in (("fp64", 0), ("ozaki_int8", slices)):
            if pinned and pinned.get(mode):
                best = dict(json.loads(pinned[mode]), mode=mode, graph=True)
                best.setdefault("slices", S or (4 if mode == "ozaki_int8" else 0))
            else:
                best, table = job.autotune(trip, mode=mode, slices=S, graph=True, reps=2)
            r = time_form(job, trip, steps, warmup, **best)
            kern = [job.run(trip, **dict(best, graph=False))["ms_main"] for _ in range(3)]
            out["forms"][mode] = {"value": flops / (r["ms_per_step"] * 1e-3) / 1e12,
                                  "kernel_tflops": flops / (float(np.median(kern)) * 1e-3) / 1e12,
                                  "ms_per_step": r["ms_per_step"], "chosen": best,
                                  "emp4": r["res"]["emp4"], "emp5": r["res"]["emp5"]}
            print("bench: tensor job %s %.1f TFLOP/s (kernel %.1f) %s" % (
                mode, out["forms"][mode]["value"], out["forms"][mode]["kernel_tflops"], best),
                file=sys.stderr, flush=True)
    f, sp = out["forms"]["fp64"], out["forms"]["ozaki_int8"]
    E = f["emp4"] + f["emp5"]
    sp["rel_dE_vs_fp64"] = abs(sp["emp4"] + sp["emp5"] - E) / abs(E)
    return out


def split_sweep_report(cc, steps, warmup):
    """The split-precision workload at nocc=30, nvir=300: INT8 slice counts 3..7
    (and 3 with the K=nocc half as digit planes too),
    the BF16x3 form and, as a precision comparison point, the FP32 form, on the
    well-conditioned inputs and on the ill-conditioned
    variant (J/K entries scaled by 10^u, u ~ U(-6,0)); per case the FP64-eq
    TFLOP/s and E[T] against the FP64 form on the same inputs (itself checked
    against the oracle in the tests)."""
    nocc, nvir = 30, 300
    out = {"metric": "split form sweep nocc=30 nvir=300: FP64-eq TFLOP/s and rel dE in E[T]",
           "unit": "TFLOP/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "cases": []}
    for ill in (False, True):
        arrays, blocks, trip, kchunk = cc.config_case(2, illcond=ill)
        flops = len(trip) * cc.gemm_flops_per_triple(nocc, nvir)
        with cc.Job(arrays, nocc, nvir, kchunk) as job:
            f = time_form(job, trip, steps, warmup, mode="fp64", order=1, persistent=True, graph=True)
            E = f["res"]["emp4"] + f["res"]["emp5"]
            case = {"inputs": "ill-conditioned" if ill else "well-conditioned",
                    "fp64": {"value": flops / (f["ms_per_step"] * 1e-3) / 1e12, "E": E}, "split": []}
            levels = [(3, 1)] + [(S, 0) for S in range(3, 8)]  # (slices, occ_digits)
            for S, occ in levels:
                # the block-pair tile for every S; the row-pair tile where it runs
                # (S <= 4), whose E[T] is bitwise the same
                tiles = (0, 2) if S <= 4 else (0,)
                row = {"slices": S, "occ_digits": occ}
                for v in tiles:
                    r = time_form(job, trip, steps, warmup, mode="ozaki_int8", slices=S, order=1,
                                  graph=True, occ_digits=occ, variant=v)
                    e = r["res"]["emp4"] + r["res"]["emp5"]
                    val = flops / (r["ms_per_step"] * 1e-3) / 1e12
                    if v == 0:
                        row.update({"value": val, "ms_per_step": r["ms_per_step"],
                                    "rel_dE": abs(e - E) / abs(E)})
                    else:
                        row["row_pair"] = {"value": val, "ms_per_step": r["ms_per_step"],
                                           "rel_dE": abs(e - E) / abs(E)}
                case["split"].append(row)
                print("bench: split sweep %s S=%d%s %.1f TFLOP/s%s rel dE %.2e" % (
                    case["inputs"], S, " occ_digits" if occ else "", row["value"],
                    " (row-pair tile %.1f)" % row["row_pair"]["value"] if "row_pair" in row else "",
                    row["rel_dE"]), file=sys.stderr, flush=True)
            for mode in ("bf16x3", "fp32"):
                r = time_form(job, trip, steps, warmup, mode=mode, order=1, graph=True)
                e = r["res"]["emp4"] + r["res"]["emp5"]
                case[mode] = {"value": flops / (r["ms_per_step"] * 1e-3) / 1e12,
                              "ms_per_step": r["ms_per_step"], "rel_dE": abs(e - E) / abs(E)}
                print("bench: split sweep %s %s %.1f TFLOP/s rel dE %.2e" % (
                    case["inputs"], mode, case[mode]["value"], case[mode]["rel_dE"]),
                    file=sys.stderr, flush=True)
        out["cases"].append(case)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--slices", type=int, default=0,
                    help="Ozaki slices of the split form (0: chosen by the autotuner; with "
                         "--serving: 0 = 4 slices, -1 = each job's form from its size)")
    ap.add_argument("--fp64-opts", default="", metavar="JSON",
                    help="run the FP64 form with these run options instead of autotuning, "
                         "e.g. '{\"order\": 1, \"variant\": 0, \"persistent\": true, "
                         "\"wg_per_cu\": 0}'")
    ap.add_argument("--split-opts", default="", metavar="JSON",
                    help="the same for the split form (may hold \"slices\")")
    ap.add_argument("--tune-cache", default=os.environ.get(
                        "CCSDT_TUNE_CACHE", os.path.join(tempfile.gettempdir(), "ccsdt_tune_cache.txt")),
                    help="file of the autotuner's kept picks (read, then appended to); a run "
                         "after one that tuned the same lists takes no tuning time")
    ap.add_argument("--no-extras", action="store_true",
                    help="time the two forms only (no drop-in, energy-pass, serving, full-tensor "
                         "or CPU legs): for profiling runs")
    ap.add_argument("--cpu-seconds", type=float, default=0.0,
                    help="bound the CPU baseline to this many seconds (0: the whole list)")
    ap.add_argument("--serving-jobs-sample", type=int, default=48,
                    help="jobs of the default line's serving sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving", default="", metavar="RATES",
                    help="only run the serving loop at these Poisson rates (jobs/s, comma list)")
    ap.add_argument("--serving-jobs", type=int, default=24)
    ap.add_argument("--serving-variant", type=int, default=-1,
                    help="server tile (CCSDT_TILE_*) for --serving; -1: the tuning cache's "
                         "server pick, else tile 0")
    ap.add_argument("--items-per-tick", type=int, default=0,
                    help="work items per serving tick for --serving; 0: the tuning cache's "
                         "server pick, else 131072")
    ap.add_argument("--serving-tune", action="store_true",
                    help="with --serving: tune the server (tile x items per tick, lowest p99 "
                         "at the first rate) for each mode first and keep the picks in the "
                         "tuning cache")
    ap.add_argument("--serving-no-check", action="store_true",
                    help="--serving without the per-job accuracy checks (timing sweeps)")
    ap.add_argument("--occ-digits", action="store_true",
                    help="serving in the split form with the K=nocc half as INT8 digit planes "
                         "too (needs --slices 3)")
    ap.add_argument("--tensors", action="store_true",
                    help="serving jobs as full-tensor jobs (every occupied slice resident)")
    ap.add_argument("--budget-gb", type=float, default=0.0,
                    help="serving HBM budget in GB (0: the server's 200 GB default)")
    ap.add_argument("--batched", action="store_true",
                    help="only run the batched configuration (nocc=40, nvir=400, uneven k "
                         "blocks) with the autotuner")
    ap.add_argument("--tensor-job", action="store_true",
                    help="only run the nocc=30/nvir=300 list as a full-tensor job (distinct "
                         "slices for every occupied index)")
    ap.add_argument("--split-sweep", action="store_true",
                    help="only run the split form at 3..7 slices on well- and ill-conditioned "
                         "inputs (nocc=30, nvir=300)")
    ap.add_argument("--batched-slices", type=int, default=0,
                    help="slices of the batched split form (0: chosen by the autotuner)")
    ap.add_argument("--tengy-pass", action="store_true",
                    help="only run the separate energy pass (f written by the FP64 GEMM, then "
                         "k_tengy) over the nocc=30/nvir=300 list, with a plain device copy of "
                         "the same bytes as the HBM reference point")
    ap.add_argument("--load-rate", action="store_true",
                    help="only measure the per-instruction cost of wave-wide loads of 4, 8 "
                         "and 16 bytes per lane (and 16-byte global->LDS DMA) on L1-resident "
                         "data")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (reference point for the "
                         "FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the triples path has no multi-GPU form in this "
              "build (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    cc = load_pkg()
    cc.init(0)
    kept = cc.tune_cache_file(args.tune_cache or None)
    print("bench: tuning cache %s (%d picks kept)" % (args.tune_cache, kept), file=sys.stderr, flush=True)
    if args.serving:
        rates = [float(x) for x in args.serving.split(",")]
        S = args.slices or 4  # -1: per-job slice counts (CCSDT_SLICES_PER_JOB)
        jobs = cc.serving_jobs(args.serving_jobs, 0xC0DE0004, tensors=args.tensors)
        kw = dict(tensors=args.tensors, budget=int(args.budget_gb * 1e9), jobs=jobs,
                  items_per_tick=args.items_per_tick)
        tuning = []
        if args.serving_tune:
            arr = cc.poisson_arrivals(len(jobs), rates[0], 0xC0DE0004)
            for mode, sl, oc in (("ozaki_int8", S, int(args.occ_digits)), ("fp64", 0, 0)):
                best, table = cc.tune_server(jobs, arr, mode=mode, slices=sl, occ_digits=oc,
                                             budget_bytes=kw["budget"])
                tuning.append({"mode": mode, "slices": sl, "rate_jobs_per_s": rates[0],
                               "pick": best, "table": table})
                print("bench: server tuned for %s slices %d: tile %d, %d items per tick (p99 %.1f ms)" % (
                    mode, sl, best["variant"], best["items_per_tick"], best["p99_ms"]),
                    file=sys.stderr, flush=True)

        occ = dict(occ_digits=int(args.occ_digits))
        # warm-up: one unmeasured pass of the same jobs, so the measured runs see
        # a server in steady state (kernels loaded, pinned staging and device
        # memory pools populated)
        serving_report(cc, rates[0], args.serving_jobs, "ozaki_int8", S, check=False,
                       variant=args.serving_variant, **kw, **occ)
        reps = []
        for rate, mode, sl in [(r, "ozaki_int8", S) for r in rates] + [(rates[-1], "fp64", 0)]:
            r = serving_report(cc, rate, args.serving_jobs, mode, sl, variant=args.serving_variant,
                               check=not args.serving_no_check, **kw, **occ)
            reps.append(r)
            print("bench: serving %s %s variant %d items/tick %d rate %.0f: %.1f jobs/s p50 %.1f ms "
                  "p99 %.1f ms" % (r["job_format"], r["mode"], r["variant"], r["items_per_tick"],
                                   r["rate_jobs_per_s"], r["jobs_per_s"], r["p50_ms"], r["p99_ms"]),
                  file=sys.stderr, flush=True)
        print(json.dumps({"metric": "serving: Poisson stream of triples jobs, nocc log-U[5,40], "
                                    "nvir log-U[50,400], one GPU", "unit": "jobs/s", "runs": reps,
                          "server_tuning": tuning}),
              flush=True)
        return
    if args.tensor_job:
        rep = tensor_job_report(cc, args.steps, args.warmup, args.slices,
                                pinned={"fp64": args.fp64_opts, "ozaki_int8": args.split_opts})
        if args.copy_calibration > 0:
            nbytes = int(args.copy_calibration * 2 ** 20)
            ms = cc.lib().ccsdt_debug_copy(nbytes, 3)
            rep["copy_calibration"] = {"bytes": nbytes, "ms": ms, "GB_per_s": 2 * nbytes / ms / 1e6}
        print(json.dumps(rep), flush=True)
        return
    if args.split_sweep:
        print(json.dumps(split_sweep_report(cc, args.steps, args.warmup)), flush=True)
        return
    if args.batched:
        print(json.dumps(batched_report(cc, args.steps, args.warmup, args.batched_slices)),
              flush=True)
        return
    if args.tengy_pass:
        nocc, nvir = 30, 300
        arrays, blocks, trip, kchunk = cc.config_case(1)
        nb = cc.tengy_bytes_per_triple(nvir)
        with cc.Job(arrays, nocc, nvir, kchunk) as job:
            job.run(trip, mode="fp64_split")
            runs = [job.run(trip, mode="fp64_split") for _ in range(max(1, args.steps))]
        ms = float(np.median([r["ms_tengy"] for r in runs]))
        copy_bytes = READ_REF_BYTES
        copy_ms = cc.lib().ccsdt_debug_copy(copy_bytes, 5)
        read_ms = cc.lib().ccsdt_debug_read(copy_bytes, 5)
        line = {"metric": "separate energy pass (k_tengy) HBM GB/s, nocc=30 nvir=300, 4960 triples",
                "unit": "GB/s", "kernel": "k_tengy", "ms": ms, "bytes": len(trip) * nb,
                "GB_per_s": len(trip) * nb / (ms * 1e-3) / 1e9,
                "hbm_frac": len(trip) * nb / (ms * 1e-3) / 1e9 / HBM_PEAK,
                "copy_reference": {"bytes_read": copy_bytes, "bytes_written": copy_bytes, "ms": copy_ms,
                                   "GB_per_s": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9},
                "read_reference": {"bytes_read": copy_bytes, "ms": read_ms,
                                   "GB_per_s": copy_bytes / (read_ms * 1e-3) / 1e9},
                "emp4": runs[-1]["emp4"], "emp5": runs[-1]["emp5"]}
        print(json.dumps(line), flush=True)
        return
    if args.load_rate:
        clock_ghz = 2.4  # MI355X peak engine clock
        rows = []
        for nbytes, lds in ((4, 0), (8, 0), (16, 0), (16, 1)):
            ns = cc.lib().ccsdt_debug_load_rate(nbytes, lds)
            if ns < 0:
                raise RuntimeError("load-rate probe failed: %s" % cc.lib().ccsdt_last_error_string())
            rows.append({"bytes_per_lane": nbytes, "to_lds": bool(lds), "ns_per_wave_load_per_cu": ns,
                         "cycles_at_2.4GHz": ns * clock_ghz,
                         "bytes_per_cycle_per_cu": 64 * nbytes / (ns * clock_ghz)})
        print(json.dumps({"metric": "vector-memory cost per wave-wide load (L1-resident, every CU)",
                          "runs": rows}), flush=True)
        return
    nocc, nvir = 30, 300
    arrays, blocks, trip, kchunk = cc.config_case(1)
    n = len(trip)
    flops = cc.gemm_flops_per_triple(nocc, nvir)
    total_flops = n * flops
    job = cc.Job(arrays, nocc, nvir, kchunk)

    # Each form: the autotuner picks item order, tile and schedule (and, for the
    # split form with --slices 0, the slice count) on the full list; the pick
    # then runs as a captured graph per step.  The fused kernel's own time comes
    # from eager runs of the same configuration (HIP events around the launch).
    forms = {}
    for mode, slices, pinned in (("fp64", 0, args.fp64_opts), ("ozaki_int8", args.slices, args.split_opts)):
        if pinned:
            best = dict(json.loads(pinned), mode=mode, graph=True)
            best.setdefault("slices", slices or (4 if mode == "ozaki_int8" else 0))
            table = []
        else:
            t0 = time.perf_counter()
            best, table = job.autotune(trip, mode=mode, slices=slices, graph=True, reps=2)
            best = dict(best)
        r = time_form(job, trip, args.steps, args.warmup, **best)
        r["tune_s"] = 0.0 if pinned else time.perf_counter() - t0
        r["tune_cached"] = bool(not pinned and not table)
        eager = dict(best, graph=False)
        kern = [job.run(trip, **eager)["ms_main"] for _ in range(3)]
        r["kernel_ms"] = float(np.median(kern))
        r["opts"], r["table"] = best, table
        forms[mode] = r
        print("bench: %s form %.2f ms/step, kernel %.2f ms (%s)" % (mode, r["ms_per_step"],
              r["kernel_ms"], best), file=sys.stderr, flush=True)
    fp, sp = forms["fp64"], forms["ozaki_int8"]
    S = sp["opts"]["slices"]
    if args.no_extras:
        sep = {"ms_tengy": float("nan"), "ms_main": float("nan")}
        dropin = None
    else:
        sep = job.run(trip, mode="fp64_split")  # GEMM writes f to HBM, separate energy pass
        dropin = dropin_report(cc, arrays, trip, nocc, nvir, flops)
        read_ms = cc.lib().ccsdt_debug_read(READ_REF_BYTES, 5)
        sep["read_GB_per_s"] = READ_REF_BYTES / (read_ms * 1e-3) / 1e9 if read_ms > 0 else None
    job.close()

    e_fp = fp["res"]["emp4"] + fp["res"]["emp5"]
    e_sp = sp["res"]["emp4"] + sp["res"]["emp5"]
    dev = abs(e_sp - e_fp) / abs(e_fp)
    split_ok = dev <= ET_TOL
    head = sp if split_ok else fp

    def tf(ms):
        return total_flops / (ms * 1e-3) / 1e12

    fp_kernel = kernel_name(fp["opts"])
    sp_kernel = kernel_name(sp["opts"])
    # roofline of the dominant kernel of the headline form
    if split_ok:
        # Ozaki digit products over K=nvir, and over K=nocc when that half has
        # digit planes too; otherwise the K=nocc products run in FP64
        tail8 = bool(sp["opts"].get("occ_digits"))
        int8_ops = n * 2.0 * 8 * nvir ** 2 * (nvir + (nocc if tail8 else 0)) * S * (S + 1) / 2
        fp64_part = 0.0 if tail8 else n * 2.0 * 8 * nvir ** 2 * nocc
        t = sp["kernel_ms"] * 1e-3
        ideal = int8_ops / (INT8_MFMA_PEAK * 1e12) + fp64_part / (FP64_MFMA_PEAK * 1e12)
        roof = {"bound": "mfma", "kernel": sp_kernel,
                "achieved": int8_ops / t / 1e12, "peak": INT8_MFMA_PEAK, "unit": "TOP/s",
                "frac": int8_ops / t / 1e12 / INT8_MFMA_PEAK,
                "traffic": traffic_from_profile("split", sp["opts"]),
                "kernel_ms": sp["kernel_ms"], "int8_ops_per_launch": int8_ops,
                "fp64_flops_per_launch": fp64_part,
                "ideal_time_frac": ideal / t,
                "fp64_eq_tflops": tf(sp["kernel_ms"])}
    else:
        roof = {"bound": "mfma", "kernel": fp_kernel,
                "achieved": tf(fp["kernel_ms"]), "peak": FP64_MFMA_PEAK, "unit": "TFLOP/s",
                "frac": tf(fp["kernel_ms"]) / FP64_MFMA_PEAK,
                "traffic": traffic_from_profile("fp64", fp["opts"]), "kernel_ms": fp["kernel_ms"],
                "flops_per_launch": total_flops}

    if split_ok and not args.no_extras:
        # the rate the INT8 matrix cores sustain on register operands on this
        # device: the datasheet figure assumes operands whose bits do not matter
        rates = {}
        for shape, name in ((0, "16x16x64"), (1, "32x32x32")):
            for rnd in (1, 0):
                rates["%s_%s" % (name, "random" if rnd else "constant")] = \
                    cc.lib().ccsdt_debug_mfma_rate(shape, rnd)
        roof["measured_int8_mfma_TOPs"] = rates
        if rates["16x16x64_random"] > 0:
            roof["frac_of_measured_16x16x64_random"] = roof["achieved"] / rates["16x16x64_random"]

    def cand(table):
        return [{k: v for k, v in row.items() if k not in ("mode", "graph", "status")}
                for row in table]

    tengy_gbs = n * cc.tengy_bytes_per_triple(nvir) / (sep["ms_tengy"] * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": tf(head["ms_per_step"]), "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("i8x%d+f64" % S) if split_ok else "f64",
        "data": "synthetic: seeded PCG inputs at the harness's array shapes",
        "config": {"workload": "triples body, nocc=30 nvir=300, k blocks of 16, 4960 triples "
                               "(i>=j>=k), a = first virtual; headline = %s form"
                               % ("split" if split_ok else "FP64"),
                   "nocc": nocc, "nvir": nvir, "kchunk": kchunk, "ntriples": n,
                   "mode": ("ozaki_int8 s=%d" % S) if split_ok else "fp64"},
        "roofline": roof,
        "step_p50_ms": head["p50_ms"], "step_p99_ms": head["p99_ms"],
        "reference_perf_convention": {
            "fp64_form": reference_perf(nocc, nvir, fp["ms_per_step"] / n),
            "split_form": reference_perf(nocc, nvir, sp["ms_per_step"] / n),
            "note": "omp/test.c:216-226 figures for the average time per triple of one "
                    "step (GEMM flops counted once per multiply-add)"},
        "fp64_form": {"value": tf(fp["ms_per_step"]), "ms_per_step": fp["ms_per_step"],
                      "kernel": fp_kernel, "kernel_ms": fp["kernel_ms"],
                      "kernel_tflops": tf(fp["kernel_ms"]),
                      "frac_fp64_peak": tf(fp["kernel_ms"]) / FP64_MFMA_PEAK,
                      "traffic": traffic_from_profile("fp64", fp["opts"]),
                      "p50_ms": fp["p50_ms"], "p99_ms": fp["p99_ms"],
                      "emp4": fp["res"]["emp4"], "emp5": fp["res"]["emp5"],
                      "autotuned": fp["opts"], "candidates": cand(fp["table"]),
                      "tune_seconds": fp["tune_s"], "tune_cached": fp["tune_cached"]},
        "split_form": {"slices": S, "kernel": sp_kernel, "autotuned": sp["opts"],
                       "candidates": cand(sp["table"]), "value": tf(sp["ms_per_step"]),
                       "tune_seconds": sp["tune_s"], "tune_cached": sp["tune_cached"],
                       "ms_per_step": sp["ms_per_step"], "kernel_ms": sp["kernel_ms"],
                       "kernel_tflops": tf(sp["kernel_ms"]),
                       "p50_ms": sp["p50_ms"], "p99_ms": sp["p99_ms"],
                       "emp4": sp["res"]["emp4"], "emp5": sp["res"]["emp5"],
                       "rel_dE_vs_fp64": dev, "within_tol": split_ok,
                       "speedup_vs_fp64": fp["ms_per_step"] / sp["ms_per_step"]},
        "dropin_pcie_inclusive": dropin,
        "tengy_pass": {"kernel": "k_tengy", "ms": sep["ms_tengy"], "GB_per_s": tengy_gbs,
                       "hbm_frac": tengy_gbs / HBM_PEAK,
                       "bytes_per_triple": cc.tengy_bytes_per_triple(nvir),
                       "gemm_only_ms": sep["ms_main"],
                       # a plain streaming read of 4 GiB on the same device: the
                       # practical ceiling for a kernel that only reads
                       "read_ceiling_GB_per_s": sep.get("read_GB_per_s"),
                       "frac_of_read_ceiling": (tengy_gbs / sep["read_GB_per_s"]
                                                if sep.get("read_GB_per_s") else None)},
    }
    if args.no_extras:
        line["cpu_baseline"] = None
        print(json.dumps(line), flush=True)
        return
    # serving sample: one unmeasured pass of the same jobs first (kernels loaded,
    # pinned and device pools populated), then the measured pass, every job
    # checked against its FP64 run alone and the small ones against the oracle.
    # Per-job split forms: three slices with occupied digit planes for jobs of at
    # least 1e8 energy summands, four below.
    sjobs = cc.serving_jobs(args.serving_jobs_sample, 0xC0DE0004)
    serving_report(cc, 40.0, len(sjobs), "ozaki_int8", cc.SLICES_PER_JOB, check=False, jobs=sjobs)
    srv = serving_report(cc, 40.0, len(sjobs), "ozaki_int8", cc.SLICES_PER_JOB, check=True, jobs=sjobs,
                         oracle_gflop=0.0 if args.no_cpu else 20.0)
    srv["within_tol"] = srv["max_rel_dE_vs_fp64_alone"] <= ET_TOL
    srv["warm"] = True
    line["serving"] = srv
    line["serving_job_p50_ms"] = srv["p50_ms"]
    line["serving_job_p99_ms"] = srv["p99_ms"]
    print("bench: serving %d jobs at 40/s: p50 %.1f ms p99 %.1f ms, %.1f jobs/s" % (
        len(sjobs), srv["p50_ms"], srv["p99_ms"], srv["jobs_per_s"]), file=sys.stderr, flush=True)
    line["tensor_job"] = tensor_job_report(cc, max(3, args.steps // 4), 1, args.slices)
    if not args.no_cpu:
        cb, ref = cpu_baseline(arrays, trip, flops, args.cpu_seconds)
        line["cpu_baseline"] = cb
        done = int(ref["done"])
        print("bench: cpu baseline %.4f TFLOP/s over %d triples (%.1f s)" % (
            cb["value"], done, cb["seconds"]), file=sys.stderr, flush=True)
        got = fp["res"]["terms"][:done]
        scale = np.maximum(np.abs(ref["terms"]), ref["abs4"])
        par = {"triples": done,
               "fp64_max_rel_err_terms": float(np.max(np.abs(got - ref["terms"]) / scale)),
               "term_tolerance": 1e-11}
        if done == n:
            e_or = ref["emp4"] + ref["emp5"]
            par.update({"oracle_emp4": ref["emp4"], "oracle_emp5": ref["emp5"],
                        "oracle_E": e_or,
                        "fp64_rel_dE_vs_oracle": abs(e_fp - e_or) / abs(e_or),
                        "split_rel_dE_vs_oracle": abs(e_sp - e_or) / abs(e_or),
                        "split_max_rel_err_terms": float(np.max(
                            np.abs(sp["res"]["terms"] - ref["terms"]) / scale))})
            line["split_form"]["rel_dE_vs_oracle"] = par["split_rel_dE_vs_oracle"]
            line["fp64_form"]["rel_dE_vs_oracle"] = par["fp64_rel_dE_vs_oracle"]
        line["parity_oracle"] = par
    else:
        line["cpu_baseline"] = None
    if args.copy_calibration > 0:
        nbytes = int(args.copy_calibration * 2 ** 20)
        ms = cc.lib().ccsdt_debug_copy(nbytes, 3)
        line["copy_calibration"] = {"bytes": nbytes, "ms": ms, "GB_per_s": 2 * nbytes / ms / 1e6}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
