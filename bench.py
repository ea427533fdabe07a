"""Benchmark of the CCSD(T) triples body on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode fp64]

A step is one pass of the body over the whole triple list of the production
case: nocc=30, nvir=300, k blocks of 16, all 4960 triples i >= j >= k, with the
job's arrays already resident in HBM.  It prints one JSON line with the FP64-
equivalent TFLOP/s (16*nvir^2*(nvir+nocc) flops per triple, 2 per multiply-add),
the roofline of the dominant kernel measured with HIP events on its own
stream, and the CPU oracle timed on the host's cores over a bounded sample of
the same list.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "(T) FP64-eq TFLOP/s nocc=30 nvir=300, split vs FP64 @rel dE<1e-8; job p99"
FP64_MFMA_PEAK = 78.6  # TFLOP/s, MI355X dense FP64 matrix
HBM_PEAK = 8000.0      # GB/s
KERNEL = {"fp64": "k_fused<true,false>", "fp64_split": "k_fused<false,true>",
          "ozaki_int8": "k_fused_split<S,true,false>"}


def load_pkg():
    import importlib.util
    name = "nwchem_ccsd_trpdrv_amd"
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(ROOT, "nwchem-ccsd-trpdrv_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def traffic_from_profile(mode):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_nocc30_nvir300.json" % mode)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("bytes_per_launch")


def cpu_baseline(cc, arrays, trip, flops, seconds):
    import oracle
    provider = oracle.blas_select(0)
    cores = oracle.max_threads()
    out = oracle.run_list(arrays, 0, 30, 300, trip, max_seconds=seconds)
    done = int(out["done"])
    secs = float(np.sum(out["secs"]))
    return {"value": done * flops / secs / 1e12, "unit": "TFLOP/s", "cores": cores,
            "kind": "port",
            "sample": "first %d of the %d triples, oracle body (%s DGEMM, %d OpenMP threads), "
                      "per-call wall time summed as in omp/test.c:166-179"
                      % (done, len(trip), oracle.blas_name(), cores),
            "triples_per_s": done / secs, "provider": oracle.blas_name(), "_terms": out["terms"],
            "_done": done}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="fp64", choices=["fp64", "fp64_split", "ozaki_int8"])
    ap.add_argument("--slices", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
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
    nocc, nvir = 30, 300
    arrays, blocks, trip, kchunk = cc.config_case(1)
    n = len(trip)
    flops = cc.gemm_flops_per_triple(nocc, nvir)
    job = cc.Job(arrays, nocc, nvir, kchunk)
    for _ in range(args.warmup):
        job.run(trip, mode=args.mode, slices=args.slices)
    print("bench: warmup done (%d steps), timing %d steps" % (args.warmup, args.steps),
          file=sys.stderr, flush=True)
    step_ms, main_ms, res = [], [], None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        res = job.run(trip, mode=args.mode, slices=args.slices, terms=True)
        step_ms.append((time.perf_counter() - s0) * 1e3)
        main_ms.append(res["ms_main"])
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = n * flops / (ms_per_step * 1e-3) / 1e12
    kern_ms = float(np.mean(main_ms))
    achieved = n * flops / (kern_ms * 1e-3) / 1e12

    # separate energy pass (GEMM writes f to HBM, tengy reads it back)
    sp = job.run(trip, mode="fp64_split")
    tengy_gbs = n * cc.tengy_bytes_per_triple(nvir) / (sp["ms_tengy"] * 1e-3) / 1e9

    line = {
        "metric": METRIC, "value": value, "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded PCG inputs at the harness's array shapes",
        "config": {"workload": "plain FP64 triples body, nocc=30 nvir=300, k blocks of 16, "
                               "4960 triples (i>=j>=k), a = first virtual",
                   "nocc": nocc, "nvir": nvir, "kchunk": kchunk, "ntriples": n,
                   "mode": args.mode},
        "roofline": {"bound": "mfma", "kernel": KERNEL[args.mode], "achieved": achieved,
                     "peak": FP64_MFMA_PEAK, "unit": "TFLOP/s",
                     "frac": achieved / FP64_MFMA_PEAK,
                     "traffic": traffic_from_profile(args.mode),
                     "kernel_ms": kern_ms, "flops_per_launch": n * flops},
        "job_p50_ms": float(np.percentile(step_ms, 50)),
        "job_p99_ms": float(np.percentile(step_ms, 99)),
        "tengy_pass": {"kernel": "k_tengy", "ms": sp["ms_tengy"], "GB_per_s": tengy_gbs,
                       "hbm_frac": tengy_gbs / HBM_PEAK,
                       "bytes_per_triple": cc.tengy_bytes_per_triple(nvir),
                       "gemm_ms": sp["ms_main"]},
        "emp4": res["emp4"], "emp5": res["emp5"],
    }
    if not args.no_cpu:
        cb = cpu_baseline(cc, arrays, trip, flops, args.cpu_seconds)
        done = cb.pop("_done")
        ref_terms = cb.pop("_terms")
        got = res["terms"][:done]
        scale = np.maximum(np.abs(ref_terms), 1e-300)
        line["cpu_baseline"] = cb
        line["parity_sample"] = {"triples": done,
                                 "max_rel_err_terms": float(np.max(np.abs(got - ref_terms) / scale))}
    else:
        line["cpu_baseline"] = None
    job.close()
    if args.copy_calibration > 0:
        nbytes = int(args.copy_calibration * 2 ** 20)
        ms = cc.lib().ccsdt_debug_copy(nbytes, 3)
        line["copy_calibration"] = {"bytes": nbytes, "ms": ms, "GB_per_s": 2 * nbytes / ms / 1e6}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
