"""Parity of the HIP path with the oracle and the golden fixtures (needs a GPU).

Tolerance (FP64 form): every per-triple term emp4i, emp5i, emp4k, emp5k within
1e-11 relative, where the denominator is max(|oracle term|, sum of the absolute
summands of that term); totals within 1e-11 relative.  f matrices within 1e-13
relative to their largest element.
"""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-11
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def device(cc):
    cc.init(0)
    cc.lib().ccsdt_set_error_policy(1)


def rel_terms(got, ref, abs4):
    return np.max(np.abs(got - ref) / np.maximum(np.abs(ref), abs4))


def run_case(cc, orc, nocc, nvir, kchunk, seed, order="ijk", mode="fp64", ntrip=None):
    arrays = cc.make_inputs(nocc, nvir, kchunk, seed)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk), order)
    if ntrip is not None:
        trip = trip[:ntrip]
    ref = orc.run_list(arrays, 0, nocc, nvir, trip, with_abs=True)
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        got = job.run(trip, mode=mode, terms=True)
    return arrays, trip, ref, got


def test_mfma_lane_layout(cc):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    d = cc.selftest_mfma(A.ravel(), B.ravel())
    D = A @ B
    want = np.array([[D[l // 16 + 4 * r, l % 16] for r in range(4)] for l in range(64)]).ravel()
    assert np.allclose(d, want, rtol=1e-14, atol=1e-14)


def test_config0_terms_and_totals(cc, orc):
    arrays, trip, ref, got = run_case(cc, orc, 8, 40, 8, cc.config_seed(0))
    assert rel_terms(got["terms"], ref["terms"], ref["abs4"]) < TOL
    assert abs(got["emp4"] - ref["emp4"]) <= TOL * abs(ref["emp4"])
    assert abs(got["emp5"] - ref["emp5"]) <= TOL * abs(ref["emp5"])


def test_config0_golden(cc, orc):
    g = np.load(os.path.join(GOLD, "config0.npz"), allow_pickle=False)
    arrays, blocks, trip, kchunk = cc.config_case(0)
    ref = orc.run_list(arrays, 0, 8, 40, trip, with_abs=True)
    with cc.Job(arrays, 8, 40, kchunk) as job:
        got = job.run(g["triples"], terms=True)
        one = job.run(g["triples"][7:8], fout=True)
    assert rel_terms(got["terms"], g["terms"], ref["abs4"]) < TOL
    assert np.allclose([got["emp4"], got["emp5"]], g["emp"], rtol=TOL, atol=0)
    f = one["f"]
    assert np.max(np.abs(f - g["f7"])) <= 1e-13 * np.max(np.abs(g["f7"]))


def test_reference_loop_order(cc, orc):
    # the omp/test.c:160-165 enumeration (k <= i, j free), two k blocks
    arrays, trip, ref, got = run_case(cc, orc, 6, 40, 3, 5, order="ref")
    assert len(trip) == 6 * 21
    assert rel_terms(got["terms"], ref["terms"], ref["abs4"]) < TOL


@pytest.mark.parametrize("nocc,nvir,kchunk", [(1, 1, 1), (3, 17, 2), (5, 37, 5), (2, 16, 1),
                                              (13, 50, 4), (4, 64, 3), (7, 33, 7)])
def test_ragged_shapes(cc, orc, nocc, nvir, kchunk):
    arrays, trip, ref, got = run_case(cc, orc, nocc, nvir, kchunk, 100 + nvir)
    assert rel_terms(got["terms"], ref["terms"], ref["abs4"]) < TOL


def test_split_mode_matches_fused(cc, orc):
    arrays, trip, ref, got = run_case(cc, orc, 8, 40, 8, 3, mode="fp64_split")
    assert rel_terms(got["terms"], ref["terms"], ref["abs4"]) < TOL


@pytest.mark.parametrize("nocc,nvir,kchunk", [(3, 1, 1), (4, 33, 2), (5, 45, 3), (4, 47, 4)])
def test_separate_pass_odd_nvir(cc, orc, nocc, nvir, kchunk):
    """The separate energy pass (f written to HBM laid out by tile pair, then
    k_tengy) at odd nvir, where a thread's row pair can start at the last row and
    is read one row early: nvir = 1 (single loads), 33 (a last tile one row wide),
    45 and 47 (odd last tiles), against the oracle per term."""
    arrays, trip, ref, got = run_case(cc, orc, nocc, nvir, kchunk, 11 + nvir, mode="fp64_split")
    assert rel_terms(got["terms"], ref["terms"], ref["abs4"]) < TOL


def test_deterministic(cc):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    with cc.Job(arrays, 8, 40, kchunk) as job:
        a = job.run(trip, terms=True)
        b = job.run(trip, terms=True)
    assert np.array_equal(a["terms"], b["terms"]) and a["emp4"] == b["emp4"]


def test_fout_matches_oracle(cc, orc):
    nocc, nvir, kchunk = 5, 37, 3
    arrays = cc.make_inputs(nocc, nvir, kchunk, 17)
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        for k, klo in [(3, 1), (5, 4)]:
            trip = np.array([[nocc + 2, 5, 4, k, klo]], dtype=np.int32)
            f = job.run(trip, fout=True)["f"]
            want = orc.fmats(arrays, nocc, nvir, k, klo)
            assert np.max(np.abs(f - want)) <= 1e-13 * np.max(np.abs(want))


def test_dropin_body(cc, orc):
    nocc, nvir, kchunk = 6, 45, 3
    arrays = cc.make_inputs(nocc, nvir, kchunk, 23)
    trips = [(nocc + 1, 4, 2, 2, 1), (nocc + 1, 3, 3, 3, 1), (nocc + 5, 6, 5, 5, 4)]
    e4 = e5 = 0.0
    r4 = r5 = 0.0
    for a, i, j, k, klo in trips:
        f = [np.full(nvir * nvir, np.nan) for _ in range(8)]
        e4, e5 = cc.cbody(f, arrays["eorb"], 0, nocc, nvir, e4, e5, a, i, j, k, klo, arrays)
        want = orc.fmats(arrays, nocc, nvir, k, klo)
        for m in range(8):
            got = f[m].reshape(nvir, nvir).T
            assert np.max(np.abs(got - want[m])) <= 1e-13 * np.max(np.abs(want[m]))
        t = orc.run_list(arrays, 0, nocc, nvir, np.array([[a, i, j, k, klo]], np.int32))["terms"][0]
        r4 += t[0] + (t[2] if i != k else 0.0)
        r5 += t[1] + (t[3] if i != k else 0.0)
    assert abs(e4 - r4) <= TOL * abs(r4) and abs(e5 - r5) <= TOL * abs(r5)


def test_prepared_body_call_equals_cbody(cc):
    """BodyCall (pointer arguments built once, as bench.py times the drop-in)
    gives bitwise the f matrices and running totals of per-call cbody, over a
    sequence of triples at the 30/300 shapes with uploads of new contents."""
    arrays, blocks, trip, kchunk = cc.config_case(1)
    trip = trip[::613][:8]
    fa = [np.zeros(300 * 300) for _ in range(8)]
    fb = [np.zeros(300 * 300) for _ in range(8)]
    body = cc.BodyCall(fb, arrays["eorb"], 0, 30, 300, arrays)
    e4 = e5 = 0.0
    for t in trip:
        e4, e5 = cc.cbody(fa, arrays["eorb"], 0, 30, 300, e4, e5, *t, arrays)
        body.call(*t)
        assert (body.emp4, body.emp5) == (e4, e5)
        assert all(np.array_equal(x, y) for x, y in zip(fa, fb))


@pytest.mark.parametrize("symbol", ["ccsd_trpdrv_cbody_", "ccsd_trpdrv_omp_fbody_"])
def test_dropin_aliases(cc, symbol):
    nocc, nvir = 4, 20
    arrays = cc.make_inputs(nocc, nvir, 1, 2)
    fa = [np.zeros(nvir * nvir) for _ in range(8)]
    fb = [np.zeros(nvir * nvir) for _ in range(8)]
    a = cc.cbody(fa, arrays["eorb"], 0, nocc, nvir, 1.0, 2.0, nocc + 1, 3, 2, 1, 1, arrays)
    b = cc.cbody(fb, arrays["eorb"], 0, nocc, nvir, 1.0, 2.0, nocc + 1, 3, 2, 1, 1, arrays,
                 symbol=symbol)
    assert a == b and all(np.array_equal(x, y) for x, y in zip(fa, fb))


@pytest.mark.parametrize("symbol", ["ccsd_tengy_omp", "ccsd_tengy", "ccsd_tengy_omp_"])
def test_dropin_tengy(cc, orc, symbol):
    for nocc, nvir in [(5, 37), (8, 40), (2, 1)]:
        arrays = cc.make_inputs(nocc, nvir, 1, 31)
        rng = np.random.default_rng(nvir)
        f = [rng.uniform(-0.05, 0.05, nvir * nvir) for _ in range(8)]
        vec = {n: arrays[n] for n in ("dintc1", "dintx1", "t1v1", "dintc2", "dintx2", "t1v2")}
        ea = orc.eaijk(arrays["eorb"], 0, nocc + 1, nocc, 1, 1)
        got = np.array(cc.tengy(f, vec, arrays["eorb"], ea, 0, nocc, nvir, symbol=symbol))
        ref = np.array(orc.tengy(f, arrays, ea, 0, nocc, nvir))
        assert np.all(np.abs(got - ref) <= 1e-12 * np.maximum(np.abs(ref), 1e-10))


def test_dropin_bad_args_policy0(cc):
    L = cc.lib()
    L.ccsdt_set_error_policy(0)
    try:
        nocc, nvir = 3, 8
        arrays = cc.make_inputs(nocc, nvir, 1, 1)
        f = [np.zeros(nvir * nvir) for _ in range(8)]
        e4, e5 = cc.cbody(f, arrays["eorb"], 0, nocc, nvir, 0.0, 0.0, nocc + 1, 1, 1, 5, 1, arrays)
        assert np.isnan(e4) and L.ccsdt_last_error() == -1
    finally:
        L.ccsdt_set_error_policy(1)


def test_injected_failure_recovers(cc):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    L = cc.lib()
    L.ccsdt_debug_inject_failure(1)
    with pytest.raises(cc.CcsdtError) as e:
        cc.Job(arrays, 8, 40, kchunk)
    assert e.value.code == -6
    with cc.Job(arrays, 8, 40, kchunk) as job:
        L.ccsdt_debug_inject_failure(1)
        with pytest.raises(cc.CcsdtError):
            job.run(trip)
        assert np.isfinite(job.run(trip)["emp4"])


def test_config1_head_golden(cc, orc):
    g = np.load(os.path.join(GOLD, "config1_head.npz"), allow_pickle=False)
    arrays, blocks, trip, kchunk = cc.config_case(1)
    ref = orc.run_list(arrays, 0, 30, 300, g["triples"], with_abs=True)
    with cc.Job(arrays, 30, 300, kchunk) as job:
        got = job.run(g["triples"], terms=True)
    assert rel_terms(got["terms"], g["terms"], ref["abs4"]) < TOL


def test_config1_strided_sample_vs_oracle(cc, orc):
    """Every 10th triple of the 30/300 list (496 triples: every k block, i, j and
    k pattern of the list) against the oracle replaying the reference body: the
    FP64 form per term within 1e-11 of max(|term|, sum of |summands|) and in
    E[T] within 1e-11; the split form the bench reports (INT8, 3 slices, the
    K = nocc half as digit planes) within 1e-8 in E[T] and 1e-6 of the summand
    scale per term."""
    arrays, blocks, trip, kchunk = cc.config_case(1)
    sample = trip[::10]
    ref = orc.run_list(arrays, 0, 30, 300, sample, with_abs=True)
    E = ref["emp4"] + ref["emp5"]
    with cc.Job(arrays, 30, 300, kchunk) as job:
        got = job.run(sample, terms=True)
        spl = job.run(sample, mode="ozaki_int8", slices=3, occ_digits=1, terms=True)
    assert rel_terms(got["terms"], ref["terms"], ref["abs4"]) < TOL
    assert abs(got["emp4"] + got["emp5"] - E) <= TOL * abs(E)
    assert abs(spl["emp4"] + spl["emp5"] - E) <= 1e-8 * abs(E)
    assert np.max(np.abs(spl["terms"] - ref["terms"]) / ref["abs4"]) < 1e-6


def test_config1_full_list_properties(cc):
    """All 4960 triples at 30/300: totals equal the compensated sum of the
    per-triple terms with the i != k rule, the run is bitwise repeatable, and
    scaling every GEMM operand by 2 scales the terms by exactly 16 and 4."""
    arrays, blocks, trip, kchunk = cc.config_case(1)
    with cc.Job(arrays, 30, 300, kchunk) as job:
        a = job.run(trip, terms=True)
        b = job.run(trip, terms=True)
    assert np.array_equal(a["terms"], b["terms"])
    t = a["terms"]
    ik = trip[:, 1] != trip[:, 3]
    s4 = np.sum(np.concatenate([t[:, 0], t[ik, 2]]))
    s5 = np.sum(np.concatenate([t[:, 1], t[ik, 3]]))
    assert abs(a["emp4"] - s4) <= 1e-12 * np.sum(np.abs(t[:, [0, 2]]))
    assert abs(a["emp5"] - s5) <= 1e-12 * np.sum(np.abs(t[:, [1, 3]]))
    sc = dict(arrays)
    for n in ["tij", "tkj", "tia", "tka", "xia", "xka", "jia", "jka", "kia", "kka", "jij",
              "jkj", "kij", "kkj"]:
        sc[n] = arrays[n] * 2.0
    with cc.Job(sc, 30, 300, kchunk) as job:
        c = job.run(trip[::37], terms=True)
    assert np.array_equal(c["terms"], a["terms"][::37] * np.array([16.0, 4.0, 16.0, 4.0]))


def wrap1000(x):
    """The harness's guard against overflow, omp/test.c:188-191."""
    if x > 1000.0:
        x -= 1000.0
    if x < -1000.0:
        x += 1000.0
    return x


def harness_replay(orc, arrays, nocc, nvir, trip):
    """(emp4, emp5, scale) after each call of omp/test.c's loop: the oracle's
    terms accumulated with the body's i != k rule
    (omp/ccsd_trpdrv_omp_cbody.c:133-142) and the harness's wrap; scale is the
    running sum of the absolute summands of every term."""
    r = orc.run_list(arrays, 0, nocc, nvir, trip, with_abs=True)
    e4 = e5 = s = 0.0
    out = []
    for q, (a, i, j, k, klo) in enumerate(trip):
        t = r["terms"][q]
        e4 += t[0]
        e5 += t[1]
        if i != k:
            e4 += t[2]
            e5 += t[3]
        e4, e5 = wrap1000(e4), wrap1000(e5)
        s += float(np.sum(r["abs4"][q]))
        out.append((e4, e5, s))
    return out


@pytest.mark.parametrize("nocc,nvir,maxiter,nkpass", [(8, 40, 100, 1), (8, 40, -1, 2), (5, 33, -1, 1)])
def test_harness_loop_through_dropin(cc, orc, nocc, nvir, maxiter, nkpass):
    """omp/test.c's main loop with its own inputs (make_array, a = 1, its k
    blocks, maxiter, the +-1000 wrap), calling the drop-in body: the running
    totals after every call match the oracle's replay of the same loop."""
    arrays, kchunk = cc.make_harness_inputs(nocc, nvir, nkpass)
    trip = cc.harness_triples(nocc, kchunk, maxiter)
    want = harness_replay(orc, arrays, nocc, nvir, trip)
    f = [np.zeros(nvir * nvir) for _ in range(8)]
    e4 = e5 = 0.0
    for q, (a, i, j, k, klo) in enumerate(trip):
        e4, e5 = cc.cbody(f, arrays["eorb"], 0, nocc, nvir, e4, e5, int(a), int(i), int(j),
                          int(k), int(klo), arrays)
        e4, e5 = wrap1000(e4), wrap1000(e5)
        w4, w5, scale = want[q]
        assert abs(e4 - w4) <= TOL * scale and abs(e5 - w5) <= TOL * scale, q


@pytest.mark.parametrize("binary", ["test_c99_dropin", "test_f77_dropin"])
def test_reference_harness_binary(cc, orc, binary):
    """The reference harness itself (omp/test.c, built by oracle/Makefile against
    this library; test_f77_dropin calls the Fortran body's symbol) runs and
    prints the totals of the oracle's replay of its loop, to its %f precision."""
    exe = os.path.join(ROOT, "oracle", "_ref", binary)
    if not os.path.exists(exe):
        pytest.skip("reference harness not built (reference sources absent at build time)")
    for nocc, nvir, maxiter, nkpass in ((8, 40, -1, 2), (30, 300, 100, 1)):
        out = subprocess.run([exe, str(nocc), str(nvir), str(maxiter), str(nkpass)],
                             capture_output=True, text=True, timeout=180)
        assert out.returncode == 0 and "SUCCESS" in out.stdout, out.stdout + out.stderr
        m = re.search(r"emp4=(\S+) emp5=(\S+)", out.stdout)
        got4, got5 = float(m.group(1)), float(m.group(2))
        arrays, kchunk = cc.make_harness_inputs(nocc, nvir, nkpass)
        w4, w5, scale = harness_replay(orc, arrays, nocc, nvir,
                                       cc.harness_triples(nocc, kchunk, maxiter))[-1]
        assert abs(got4 - w4) <= 6e-7 + TOL * scale and abs(got5 - w5) <= 6e-7 + TOL * scale

