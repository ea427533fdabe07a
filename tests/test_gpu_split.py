"""Split forms against the FP64 oracle (needs a GPU): Ozaki INT8 digits and
the BF16x3 planes.

Tolerance: E[T] = emp4 + emp5 over the triple list within 1e-8 relative of the
oracle for INT8 slices >= 4 and for BF16x3 (the forms' stated bound); every
per-triple term within 1e-6 of its absolute-summand scale; more INT8 slices
must converge toward FP64.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ET_TOL = 1e-8
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module", autouse=True)
def device(cc):
    cc.init(0)


def et(e4, e5):
    return e4 + e5


def run(cc, arrays, nocc, nvir, kchunk, trip, slices, **kw):
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        return job.run(trip, mode="ozaki_int8", slices=slices, **kw)


def run_mode(cc, arrays, nocc, nvir, kchunk, trip, mode, **kw):
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        return job.run(trip, mode=mode, **kw)


def test_config0_slices_converge(cc, orc):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    ref = orc.run_list(arrays, 0, 8, 40, trip, with_abs=True)
    E = et(ref["emp4"], ref["emp5"])
    errs = []
    for s in (3, 4, 5, 6, 7):
        got = run(cc, arrays, 8, 40, kchunk, trip, s, terms=True)
        err = abs(et(got["emp4"], got["emp5"]) - E) / abs(E)
        errs.append(err)
        assert np.max(np.abs(got["terms"] - ref["terms"]) / ref["abs4"]) < 1e-6
        if s >= 4:
            assert err < ET_TOL, (s, err)
    assert errs[-1] < 1e-12 and errs[-1] < errs[0]


def test_illconditioned(cc, orc):
    nocc, nvir, kchunk = 8, 40, 8
    arrays = cc.make_inputs(nocc, nvir, kchunk, cc.config_seed(2), illcond=True)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    ref = orc.run_list(arrays, 0, nocc, nvir, trip)
    E = et(ref["emp4"], ref["emp5"])
    for s in (4, 5):
        got = run(cc, arrays, nocc, nvir, kchunk, trip, s)
        assert abs(et(got["emp4"], got["emp5"]) - E) / abs(E) < ET_TOL


@pytest.mark.parametrize("nocc,nvir,kchunk", [(1, 1, 1), (3, 17, 2), (5, 70, 5), (4, 65, 2)])
def test_ragged_shapes(cc, orc, nocc, nvir, kchunk):
    arrays = cc.make_inputs(nocc, nvir, kchunk, 300 + nvir)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    ref = orc.run_list(arrays, 0, nocc, nvir, trip, with_abs=True)
    got = run(cc, arrays, nocc, nvir, kchunk, trip, 7, terms=True)
    assert np.max(np.abs(got["terms"] - ref["terms"]) / np.maximum(np.abs(ref["terms"]), ref["abs4"])) < 1e-11


@pytest.mark.parametrize("nocc,nvir,kchunk", [(1, 20, 1), (13, 33, 4), (64, 24, 8), (70, 50, 5)])
def test_s3_occupied_digit_planes(cc, orc, nocc, nvir, kchunk):
    """occ_digits at S = 3: the K = nocc half also runs as INT8 digit planes, with
    row scales of its own, one 64-wide K step per 64 occupied orbitals: nocc = 1, a
    partial step, exactly one step and two steps.  Every term within 1e-6 of its
    summand scale (a dropped or mis-scaled K = nocc half is off by O(1)), and the
    same list with the FP64 tail agrees to the same bound."""
    arrays = cc.make_inputs(nocc, nvir, kchunk, 500 + nocc)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    trip = trip[:: max(1, len(trip) // 60)]
    ref = orc.run_list(arrays, 0, nocc, nvir, trip, with_abs=True)
    got = run(cc, arrays, nocc, nvir, kchunk, trip, 3, terms=True, occ_digits=1)
    assert np.max(np.abs(got["terms"] - ref["terms"]) / ref["abs4"]) < 1e-6
    E = et(ref["emp4"], ref["emp5"])
    assert abs(et(got["emp4"], got["emp5"]) - E) <= 1e-6 * np.sum(ref["abs4"])
    tail = run(cc, arrays, nocc, nvir, kchunk, trip, 3, terms=True)
    assert np.max(np.abs(got["terms"] - tail["terms"]) / ref["abs4"]) < 1e-6


@pytest.mark.parametrize("nocc,nvir,kchunk", [(1, 20, 1), (13, 33, 4), (30, 64, 16), (32, 40, 8)])
def test_folded_occupied_planes_bitwise(cc, nocc, nvir, kchunk):
    """nocc <= 32: two occupied digits share each 64-wide K step (fewer loads and
    MFMAs).  The integer digit products are the same, so every term equals the
    one-digit-per-step layout's bitwise; switching the layout inside one job
    re-packs, and a graph-replayed persistent run agrees too."""
    arrays = cc.make_inputs(nocc, nvir, kchunk, 900 + nocc)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    trip = trip[:: max(1, len(trip) // 80)]
    L = cc.lib()
    kw = dict(mode="ozaki_int8", slices=3, occ_digits=1, terms=True)
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        folded = job.run(trip, **kw)
        try:
            L.ccsdt_debug_occ_fold(0)
            plain = job.run(trip, **kw)
            plain_g = job.run(trip, order=1, graph=True, persistent=True, **kw)
        finally:
            L.ccsdt_debug_occ_fold(1)
        folded_g = job.run(trip, order=1, graph=True, persistent=True, **kw)
    assert np.array_equal(folded["terms"], plain["terms"])
    assert np.array_equal(folded_g["terms"], plain["terms"])
    assert np.array_equal(plain_g["terms"], plain["terms"])
    assert (folded["emp4"], folded["emp5"]) == (plain["emp4"], plain["emp5"])


def test_occ_digits_needs_three_slices(cc):
    nocc, nvir, kchunk = 4, 20, 2
    arrays = cc.make_inputs(nocc, nvir, kchunk, 5)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        for S in (4, 7):
            with pytest.raises(cc.CcsdtError) as e:
                job.run(trip, mode="ozaki_int8", slices=S, occ_digits=1)
            assert e.value.code == -1
        with pytest.raises(cc.CcsdtError):
            job.run(trip, mode="ozaki_int8", slices=3, occ_digits=2)
        # other modes ignore it
        a = job.run(trip, mode="fp64")
        b = job.run(trip, mode="fp64", occ_digits=1)
        assert (a["emp4"], a["emp5"]) == (b["emp4"], b["emp5"])


def test_split_f_matrices(cc, orc):
    nocc, nvir, kchunk = 5, 37, 3
    arrays = cc.make_inputs(nocc, nvir, kchunk, 17)
    trip = np.array([[nocc + 2, 5, 4, 3, 1]], dtype=np.int32)
    f = run(cc, arrays, nocc, nvir, kchunk, trip, 7, fout=True)["f"]
    want = orc.fmats(arrays, nocc, nvir, 3, 1)
    assert np.max(np.abs(f - want)) <= 1e-12 * np.max(np.abs(want))


def test_production_shape_golden(cc, orc):
    g = np.load(os.path.join(GOLD, "config1_head.npz"), allow_pickle=False)
    arrays, blocks, trip, kchunk = cc.config_case(1)
    got = run(cc, arrays, 30, 300, kchunk, g["triples"], 4, terms=True)
    t = g["terms"]
    ik = g["triples"][:, 1] != g["triples"][:, 3]
    E = t[:, 0].sum() + t[:, 1].sum() + t[ik, 2].sum() + t[ik, 3].sum()
    assert abs(et(got["emp4"], got["emp5"]) - E) / abs(E) < ET_TOL


def test_full_list_vs_fp64_form(cc):
    """All 4960 triples at 30/300: the split form's E[T] within 1e-8 of the FP64
    form's at 4 slices, at 3, and at 3 with occupied digit planes (the level the
    autotuner picks at this shape), and bitwise repeatable."""
    arrays, blocks, trip, kchunk = cc.config_case(1)
    with cc.Job(arrays, 30, 300, kchunk) as job:
        a = job.run(trip, mode="fp64")
        b = job.run(trip, mode="ozaki_int8", slices=4, terms=True)
        c = job.run(trip, mode="ozaki_int8", slices=4, terms=True)
        d = job.run(trip, mode="ozaki_int8", slices=3, order=1)
        e = job.run(trip, mode="ozaki_int8", slices=3, order=1, occ_digits=1, terms=True)
        e2 = job.run(trip, mode="ozaki_int8", slices=3, order=0, occ_digits=1, terms=True,
                     persistent=True, graph=True)
    E = et(a["emp4"], a["emp5"])
    assert abs(et(b["emp4"], b["emp5"]) - E) / abs(E) < ET_TOL
    assert abs(et(d["emp4"], d["emp5"]) - E) / abs(E) < ET_TOL
    assert abs(et(e["emp4"], e["emp5"]) - E) / abs(E) < ET_TOL
    assert np.array_equal(b["terms"], c["terms"])
    assert np.array_equal(e["terms"], e2["terms"])


@pytest.mark.parametrize("illcond", [False, True])
def test_split_precision_workload(cc, illcond):
    """The split-precision workload at nocc=30, nvir=300 (its own seeds), on the
    well-conditioned inputs and on the variant whose J/K entries are scaled by
    10^u, u ~ U(-6,0): E[T] within 1e-8 of the FP64 form's at 3 slices, with the
    FP64 occupied tail and with occupied digit planes, and the deviation shrinks
    by at least 100x from 3 to 5 slices."""
    arrays, blocks, trip, kchunk = cc.config_case(2, illcond=illcond)
    with cc.Job(arrays, 30, 300, kchunk) as job:
        a = job.run(trip, mode="fp64", order=1)
        E = et(a["emp4"], a["emp5"])
        dev = {}
        for S, occ in ((3, 1), (3, 0), (5, 0)):
            r = job.run(trip, mode="ozaki_int8", slices=S, occ_digits=occ, order=1)
            dev[S, occ] = abs(et(r["emp4"], r["emp5"]) - E) / abs(E)
    assert dev[3, 1] < ET_TOL and dev[3, 0] < ET_TOL, dev
    assert dev[5, 0] * 100 < dev[3, 0], dev


def test_bad_slice_count(cc):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    with cc.Job(arrays, 8, 40, kchunk) as job:
        with pytest.raises(cc.CcsdtError):
            job.run(trip, mode="ozaki_int8", slices=9)


def test_int8_and_bf16_mfma_layouts(cc):
    """The operand and result lane layouts the split kernels assume: lane l holds
    16 bytes of row l%16 (A) / column l%16 (B) at K offset (bytes per lane)*(l/16);
    result element r of lane l is D[4*(l/16) + r][l%16]."""
    rng = np.random.default_rng(3)
    # INT8 16x16x64
    A = rng.integers(-100, 101, (16, 64)).astype(np.int8)
    B = rng.integers(-100, 101, (64, 16)).astype(np.int8)
    a = np.array([A[l % 16, 16 * (l // 16):16 * (l // 16) + 16] for l in range(64)])
    b = np.array([B[16 * (l // 16):16 * (l // 16) + 16, l % 16] for l in range(64)])
    d = cc.selftest_mfma_i8(a, b, 0)
    ref = A.astype(np.int64) @ B.astype(np.int64)
    got = np.array([[d[16 * (r // 4) + c, r % 4] for c in range(16)] for r in range(16)])
    assert np.array_equal(got, ref)
    # BF16 16x16x32 (small integers are exact in BF16)
    A = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    B = rng.integers(-4, 5, (32, 16)).astype(np.float32)

    def bits(x):
        return (np.asarray(x, np.float32).view(np.uint32) >> 16).astype(np.uint16)
    a = np.array([bits(A[l % 16, 8 * (l // 16):8 * (l // 16) + 8]) for l in range(64)])
    b = np.array([bits(B[8 * (l // 16):8 * (l // 16) + 8, l % 16]) for l in range(64)])
    d = cc.selftest_mfma_i8(a.view(np.int8).reshape(64, 16), b.view(np.int8).reshape(64, 16), 2)
    D = d.view(np.float32)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    got = np.array([[D[16 * (r // 4) + c, r % 4] for c in range(16)] for r in range(16)])
    assert np.array_equal(got, ref)
    # INT8 32x32x32: lane l holds row l%32 (A) / column l%32 (B) at K bytes
    # 16*(l/32); result element e of lane l is D[8*(e/4) + 4*(l/32) + e%4][l%32]
    A = rng.integers(-100, 101, (32, 32)).astype(np.int8)
    B = rng.integers(-100, 101, (32, 32)).astype(np.int8)
    a = np.array([A[l % 32, 16 * (l // 32):16 * (l // 32) + 16] for l in range(64)])
    b = np.array([B[16 * (l // 32):16 * (l // 32) + 16, l % 32] for l in range(64)])
    d = cc.selftest_mfma_i8(a, b, 1)
    got = np.zeros((32, 32), np.int64)
    for l in range(64):
        for e in range(16):
            got[8 * (e // 4) + 4 * (l // 32) + e % 4, l % 32] = d[l, e]
    assert np.array_equal(got, A.astype(np.int64) @ B.astype(np.int64))


def test_mfma_rate_probe(cc):
    """ccsdt_debug_mfma_rate runs both INT8 instructions on constant and random
    register operands and reports a rate in TOP/s below the 5.0 POP/s datasheet
    figure (the values themselves are measurements, reported by bench.py)."""
    for shape in (0, 1):
        for rnd in (0, 1):
            r = cc.lib().ccsdt_debug_mfma_rate(shape, rnd)
            assert 100.0 < r < 5200.0, (shape, rnd, r)
    assert cc.lib().ccsdt_debug_mfma_rate(2, 0) == -1.0


@pytest.mark.parametrize("illcond", [False, True])
def test_bf16x3_config0(cc, orc, illcond):
    arrays, blocks, trip, kchunk = cc.config_case(0, illcond=illcond)
    ref = orc.run_list(arrays, 0, 8, 40, trip, with_abs=True)
    got = run_mode(cc, arrays, 8, 40, kchunk, trip, "bf16x3", terms=True)
    E = et(ref["emp4"], ref["emp5"])
    assert abs(et(got["emp4"], got["emp5"]) - E) / abs(E) < ET_TOL
    assert np.max(np.abs(got["terms"] - ref["terms"]) / ref["abs4"]) < 1e-6


@pytest.mark.parametrize("nocc,nvir,kchunk", [(1, 1, 1), (3, 17, 2), (5, 70, 5), (4, 33, 2)])
def test_bf16x3_ragged_shapes(cc, orc, nocc, nvir, kchunk):
    """nvir not a multiple of the BF16 K step (32) nor of 16."""
    arrays = cc.make_inputs(nocc, nvir, kchunk, 400 + nvir)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    ref = orc.run_list(arrays, 0, nocc, nvir, trip, with_abs=True)
    got = run_mode(cc, arrays, nocc, nvir, kchunk, trip, "bf16x3", terms=True)
    scale = np.maximum(np.maximum(np.abs(ref["terms"]), ref["abs4"]), 1e-300)
    assert np.max(np.abs(got["terms"] - ref["terms"]) / scale) < 1e-6
    E = et(ref["emp4"], ref["emp5"])
    assert abs(et(got["emp4"], got["emp5"]) - E) / abs(E) < ET_TOL


def test_bf16x3_full_list_and_schedules(cc):
    """All 4960 triples at 30/300: BF16x3 E[T] within 1e-8 of the FP64 form's;
    pair-major / persistent / graph runs bitwise equal to the plain run."""
    arrays, blocks, trip, kchunk = cc.config_case(1)
    with cc.Job(arrays, 30, 300, kchunk) as job:
        a = job.run(trip, mode="fp64")
        b = job.run(trip, mode="bf16x3", terms=True)
        c = job.run(trip, mode="bf16x3", terms=True, order=1, persistent=True, graph=True)
    E = et(a["emp4"], a["emp5"])
    assert abs(et(b["emp4"], b["emp5"]) - E) / abs(E) < ET_TOL
    assert np.array_equal(b["terms"], c["terms"])


def test_load_rate_probe(cc):
    """The vector-memory probe behind the split kernel's design runs, rejects
    widths it has no kernel for, and finds a 16-byte load no cheaper per
    instruction than a 4-byte one."""
    L = cc.lib()
    ns = {(b, l): L.ccsdt_debug_load_rate(b, l) for b, l in ((4, 0), (8, 0), (16, 0), (16, 1))}
    assert all(0 < v < 1e4 for v in ns.values()), ns
    assert ns[(16, 0)] >= ns[(4, 0)]
    assert L.ccsdt_debug_load_rate(2, 0) == -1.0 and L.ccsdt_debug_load_rate(8, 1) == -1.0
