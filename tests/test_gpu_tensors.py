"""Full-tensor jobs (ccsdt_job_create_tensors) against the oracle (needs a GPU).

Each triple of a tensor job must give the terms the oracle gives on the
harness-shaped arrays gathered for that triple (FP64 form: 1e-11 per term; split
form: 1e-8 in E[T]); a tensor job whose slices are all equal must reproduce a
harness job bitwise; schedules and serving keep their bitwise guarantees."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device(cc):
    cc.init(0)


def gathered_oracle(cc, orc, t, nocc, nvir, trip):
    """Per-triple oracle terms on gathered arrays, and the list totals."""
    terms, abs4 = [], []
    for a, i, j, k, klo in trip:
        arr = cc.gather_triple(t, nocc, nvir, int(i), int(j), int(k))
        r = orc.run_list(arr, 0, nocc, nvir, np.array([[a, i, j, k, k]], np.int32), with_abs=True)
        terms.append(r["terms"][0])
        abs4.append(r["abs4"][0])
    terms, abs4 = np.array(terms), np.array(abs4)
    ik = trip[:, 1] != trip[:, 3]
    E = terms[:, 0].sum() + terms[:, 1].sum() + terms[ik, 2].sum() + terms[ik, 3].sum()
    return terms, abs4, E


@pytest.mark.parametrize("nocc,nvir", [(4, 37), (5, 20)])
def test_tensor_job_against_gathered_oracle(cc, orc, nocc, nvir):
    t = cc.make_tensor_inputs(nocc, nvir, 515)
    trip = cc.triple_list(nocc, [(1, nocc)])          # every i >= j >= k
    terms, abs4, E = gathered_oracle(cc, orc, t, nocc, nvir, trip)
    with cc.Job.from_tensors(t, nocc, nvir) as job:
        f = job.run(trip, terms=True)
        s = job.run(trip, mode="ozaki_int8", slices=4, terms=True)
    scale = np.maximum(np.abs(terms), abs4)
    assert np.max(np.abs(f["terms"] - terms) / scale) < 1e-11
    assert abs(f["emp4"] + f["emp5"] - E) <= 1e-11 * abs(E)
    assert abs(s["emp4"] + s["emp5"] - E) <= 1e-8 * abs(E)


def test_tensor_job_any_k_order(cc, orc):
    """Triples outside i >= j >= k (k > i, j > i) read the right slices too."""
    nocc, nvir = 5, 24
    t = cc.make_tensor_inputs(nocc, nvir, 77)
    trip = np.array([[nocc + 1, 1, 4, 5, 1], [nocc + 2, 2, 5, 3, 2], [nocc + 3, 5, 1, 2, 1],
                     [nocc + 1, 3, 3, 3, 3]], np.int32)
    terms, abs4, E = gathered_oracle(cc, orc, t, nocc, nvir, trip)
    with cc.Job.from_tensors(t, nocc, nvir) as job:
        f = job.run(trip, terms=True, fout=False)
    assert np.max(np.abs(f["terms"] - terms) / np.maximum(np.abs(terms), abs4)) < 1e-11


def test_constant_tensors_equal_harness_job(cc):
    """Slice gather = harness path: with every occupied slice equal to the
    harness arrays' single slice, the two jobs compute identical panels and
    give bitwise identical terms."""
    nocc, nvir, kchunk = 4, 33, 4
    h = cc.make_inputs(nocc, nvir, kchunk, 31)
    vv, ov = nvir * nvir, nocc * nvir
    # harness arrays with all k slices equal to the i/ij slice
    for a, b, n in (("tkj", "tij", vv), ("jka", "jia", vv), ("kka", "kia", vv),
                    ("tka", "tia", ov), ("xka", "xia", ov), ("jkj", "jij", ov), ("kkj", "kij", ov)):
        h[a] = np.tile(h[b][:n], kchunk)
    t = {"eorb": h["eorb"],
         "t2": np.tile(h["tij"][:vv], nocc * nocc), "jv": np.tile(h["jia"][:vv], nocc),
         "kv": np.tile(h["kia"][:vv], nocc), "t1o": np.tile(h["tia"][:ov], nocc),
         "xo": np.tile(h["xia"][:ov], nocc), "jo": np.tile(h["jij"][:ov], nocc * nocc),
         "ko": np.tile(h["kij"][:ov], nocc * nocc)}
    for name in cc.VECTOR_NAMES:
        t[name] = h[name][:nvir]
    trip = cc.triple_list(nocc, [(1, nocc)])
    with cc.Job(h, nocc, nvir, kchunk) as a, cc.Job.from_tensors(t, nocc, nvir) as b:
        ra = a.run(trip, terms=True)
        rb = b.run(trip, terms=True)
    assert np.array_equal(ra["terms"], rb["terms"])


@pytest.mark.parametrize("driver", [False, True])
@pytest.mark.parametrize("mode,slices,occ", [("fp64", 0, 0), ("ozaki_int8", 3, 0), ("ozaki_int8", 3, 1)])
def test_tensor_job_schedules_bitwise(cc, mode, slices, occ, driver):
    nocc, nvir = 6, 40
    t = cc.make_tensor_inputs(nocc, nvir, 5, driver=driver)
    trip = cc.triple_list(nocc, [(1, 3), (4, 6)])
    with cc.Job.from_tensors(t, nocc, nvir) as job:
        ref = job.run(trip, mode=mode, slices=slices, occ_digits=occ, terms=True)
        for kw in (dict(order=1), dict(order=1, persistent=True, graph=True),
                   dict(persistent=True)):
            got = job.run(trip, mode=mode, slices=slices, occ_digits=occ, terms=True, **kw)
            assert np.array_equal(got["terms"], ref["terms"]), kw
        if mode == "fp64":
            two = job.run(trip, terms=True, variant=1, order=1)
            assert np.max(np.abs(two["terms"] - ref["terms"])) <= 1e-12 * np.max(np.abs(ref["terms"]))


def test_tensor_and_harness_jobs_serve_together(cc):
    """Tensor and harness-shaped jobs interleave in one server; each result is
    bitwise its plain run."""
    jobs = cc.serving_jobs(3, 99, nocc_range=(3, 7), nvir_range=(20, 50))
    tens = [(5, 30, cc.make_tensor_inputs(5, 30, 1)), (4, 45, cc.make_tensor_inputs(4, 45, 2))]
    alone = []
    for jb in jobs:
        with cc.Job(jb["arrays"], jb["nocc"], jb["nvir"], jb["kchunk"]) as job:
            alone.append(job.run(jb["triples"]))
    ttrip = [cc.triple_list(no, [(1, no)]) for no, nv, t in tens]
    for (no, nv, t), tr in zip(tens, ttrip):
        with cc.Job.from_tensors(t, no, nv) as job:
            alone.append(job.run(tr))
    with cc.Server(items_per_tick=300) as srv:
        ids = []
        for q, jb in enumerate(jobs):
            ids.append(srv.submit(jb["arrays"], jb["nocc"], jb["nvir"], jb["kchunk"], jb["triples"]))
            if q < len(tens):
                no, nv, t = tens[q]
                ids.append(srv.submit_tensors(t, no, nv, ttrip[q]))
        srv.drain()
        res = {i: srv.result(i) for i in ids}
    order = [ids[0], ids[2], ids[4], ids[1], ids[3]]  # harness jobs, then tensor jobs
    for jid, a in zip(order, alone):
        m = res[jid]
        assert m["status"] == 0 and (m["emp4"], m["emp5"]) == (a["emp4"], a["emp5"])


# ------------------------------------------------------------------ driver tensors
@pytest.mark.parametrize("mode,slices,variant,occ", [("fp64", 0, 0, 0), ("fp64", 0, 1, 0), ("fp64_split", 0, 0, 0),
                                                     ("ozaki_int8", 4, 0, 0), ("ozaki_int8", 3, 0, 1),
                                                     ("bf16x3", 0, 0, 0)])
def test_driver_tensor_job_against_gathered_oracle(cc, orc, mode, slices, variant, occ):
    """Per-triple vectors read in place from t1, dja, djx: each triple's terms
    match the oracle on the host-gathered arrays (rows by the same slicing)."""
    nocc, nvir = 5, 37
    t = cc.make_tensor_inputs(nocc, nvir, 808, driver=True)
    trip = cc.triple_list(nocc, [(1, nocc)])
    terms, abs4, E = gathered_oracle(cc, orc, t, nocc, nvir, trip)
    with cc.Job.from_tensors(t, nocc, nvir) as job:
        r = job.run(trip, mode=mode, slices=slices, variant=variant, occ_digits=occ, terms=True)
    if mode in ("fp64", "fp64_split"):
        assert np.max(np.abs(r["terms"] - terms) / np.maximum(np.abs(terms), abs4)) < 1e-11
        assert abs(r["emp4"] + r["emp5"] - E) <= 1e-11 * abs(E)
    elif slices == 3:
        # S = 3 on a list this small: 1e-6 of the summed term scale (about 21 bits
        # per operand row; E[T] at three slices has measured up to 2.1e-8 on lists
        # of a few hundred triples, so the 1e-8 E[T] bound applies to the lists
        # that three slices are chosen for: 1e8 summands and up)
        assert np.max(np.abs(r["terms"] - terms) / abs4) < 1e-6
        assert abs(r["emp4"] + r["emp5"] - E) <= 1e-6 * np.sum(abs4)
    else:
        assert abs(r["emp4"] + r["emp5"] - E) <= 1e-8 * abs(E)


def test_driver_rows_equal_to_fixed_vectors_bitwise(cc):
    """Driver tensors whose rows all equal one set of vectors give the fixed-
    vector tensor job's terms bitwise (the pool rows only change addresses)."""
    nocc, nvir = 4, 29
    t = cc.make_tensor_inputs(nocc, nvir, 12)
    t["t1v2"] = t["t1v1"]
    t["dintc2"] = t["dintc1"]
    t["dintx2"] = t["dintx1"]
    d = {k: v for k, v in t.items() if k not in cc.VECTOR_NAMES}
    d["t1"] = np.tile(t["t1v1"], nocc)
    d["dja"] = np.tile(t["dintc1"], nocc * nocc)
    d["djx"] = np.tile(t["dintx1"], nocc * nocc)
    trip = cc.triple_list(nocc, [(1, 2), (3, 4)])
    for mode, slices in (("fp64", 0), ("ozaki_int8", 3), ("fp32", 0)):
        with cc.Job.from_tensors(t, nocc, nvir) as a, cc.Job.from_tensors(d, nocc, nvir) as b:
            ra = a.run(trip, mode=mode, slices=slices, terms=True)
            rb = b.run(trip, mode=mode, slices=slices, terms=True)
        assert np.array_equal(ra["terms"], rb["terms"]), mode


def test_driver_tensor_job_missing_tensor_rejected(cc):
    nocc, nvir = 3, 20
    t = cc.make_tensor_inputs(nocc, nvir, 4, driver=True)
    t["djx"] = None
    with pytest.raises(cc.CcsdtError) as e:
        cc.Job.from_tensors(t, nocc, nvir)
    assert e.value.code == -1


def test_driver_tensor_jobs_serve_bitwise(cc):
    """Driver-tensor jobs in a server mix give their plain runs' totals bitwise."""
    specs = [(5, 30, 1), (4, 45, 2), (6, 24, 3)]
    tens = [(no, nv, cc.make_tensor_inputs(no, nv, sd, driver=True)) for no, nv, sd in specs]
    trips = [cc.triple_list(no, [(1, no)]) for no, nv, t in tens]
    alone = []
    for (no, nv, t), tr in zip(tens, trips):
        with cc.Job.from_tensors(t, no, nv) as job:
            alone.append(job.run(tr))
    with cc.Server(items_per_tick=257) as srv:
        ids = [srv.submit_tensors(t, no, nv, tr) for (no, nv, t), tr in zip(tens, trips)]
        srv.drain()
        res = [srv.result(i) for i in ids]
    for a, m in zip(alone, res):
        assert m["status"] == 0 and (m["emp4"], m["emp5"]) == (a["emp4"], a["emp5"])
