"""Pins the CPU oracle before it is trusted as the checker.

* PCG32 against the generator's published known-answer outputs.
* tengy and the body glue (eaijk + i!=k accumulation) bit for bit against the
  reference's own C sources compiled into oracle/_ref (single thread), and the
  Fortran twin within rounding.
* The 16 GEMMs against the textbook DGEMM restated with numpy from the
  reference's cblas argument lists, against an exact rational computation on
  tiny shapes, and between the two DGEMM providers.
* Invariants the mathematics fixes: n<->t swap symmetry, exact power-of-two
  homogeneity, f = 0 => 0, and the i == k rule.
* tengy against its constant-denominator closed form (Frobenius products and
  bilinear forms of the f matrices, numpy matrix algebra), which also rejects a
  variant with one operand left untransposed.
"""
import ctypes
import ctypes.util
import os
from fractions import Fraction

import numpy as np
import pytest

# Known answers of PCG32 XSH-RR 64/32 seeded with (initstate=42, initseq=54):
# the first six outputs printed by the generator's reference demo program.
PCG_KAT = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def _omp_threads(n):
    """Sets the thread count in every OpenMP runtime mapped into the process
    (MKL can bring a second libgomp that other libraries then bind to)."""
    paths = {ctypes.util.find_library("gomp") or "libgomp.so.1"}
    with open("/proc/self/maps") as fh:
        for line in fh:
            p = line.split()[-1]
            if "/libgomp" in p or "/libomp" in p or "/libiomp" in p:
                paths.add(p)
    for p in paths:
        ctypes.CDLL(p).omp_set_num_threads(n)


@pytest.fixture
def one_thread(orc):
    _omp_threads(1)
    yield
    _omp_threads(orc.max_threads())


def rand_case(cc, nocc, nvir, kchunk=1, seed=7):
    return cc.make_inputs(nocc, nvir, kchunk, seed)


def rand_f(nvir, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-0.05, 0.05, nvir * nvir) for _ in range(8)]


def test_pcg_known_answers(orc, cc):
    assert list(orc.pcg_raw(42, 54, 6)) == PCG_KAT
    assert list(cc.pcg_raw(42, 54, 6)) == PCG_KAT


def test_pcg_uniform_construction(orc, cc):
    raw = [int(x) for x in orc.pcg_raw(5, 9, 8)]
    want = [((raw[2 * q] >> 5) * 2.0 ** 26 + (raw[2 * q + 1] >> 6)) * 2.0 ** -53
            for q in range(4)]
    got = orc.pcg_uniform(5, 9, 0.0, 1.0, 4)
    assert list(got) == want
    assert np.array_equal(cc.pcg_uniform(11, 33, -0.05, 0.05, 1000),
                          orc.pcg_uniform(11, 33, -0.05, 0.05, 1000))


@pytest.mark.parametrize("nocc,nvir", [(8, 40), (5, 37), (3, 1), (2, 33)])
def test_tengy_bitwise_vs_reference(orc, cc, one_thread, nocc, nvir):
    L = orc.ref_tengy_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    arrays = rand_case(cc, nocc, nvir)
    f = rand_f(nvir, nvir)
    for a, i, j, k in [(nocc + 1, 1, 1, 1), (nocc + nvir, nocc, 1, nocc)]:
        ea = orc.eaijk(arrays["eorb"], 0, a, i, j, k)
        got = orc.tengy(f, arrays, ea, 0, nocc, nvir, nthreads=1)
        ref = orc.ref_tengy(L, f, arrays, ea, 0, nocc, nvir)
        assert got == ref


def test_tengy_fortran_twin(orc, cc):
    L = orc.ref_tengy_f77_lib()
    if L is None:
        pytest.skip("Fortran reference twin not built")
    nocc, nvir = 6, 45
    arrays = rand_case(cc, nocc, nvir)
    f = rand_f(nvir, 3)
    ea = orc.eaijk(arrays["eorb"], 0, nocc + 2, 3, 2, 1)
    got = np.array(orc.tengy(f, arrays, ea, 0, nocc, nvir, nthreads=1))
    ref = np.array(orc.ref_tengy_f77(L, f, arrays, ea, 0, nocc, nvir))
    assert np.allclose(got, ref, rtol=1e-13, atol=0)


def test_body_glue_bitwise_vs_reference(orc, cc, one_thread):
    L = orc.ref_cbody_skipdgemm_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    orc.blas_select(1)
    nocc, nvir, kchunk = 5, 24, 3
    arrays = cc.make_inputs(nocc, nvir, kchunk, 99)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk), "ref")
    out = orc.run_list(arrays, 0, nocc, nvir, trip, nthreads=1)
    e4 = e5 = 0.0
    for t, (a, i, j, k, klo) in enumerate(trip):
        f = orc.fmats(arrays, nocc, nvir, int(k), int(klo))
        fcm = [np.ascontiguousarray(f[m].T) for m in range(8)]
        e4, e5 = orc.ref_cbody_glue(L, fcm, arrays, 0, nocc, nvir, e4, e5, *map(int, (a, i, j, k, klo)))
    assert (out["emp4_plain"], out["emp5_plain"]) == (e4, e5)
    orc.blas_select(0)


def numpy_fmats(arrays, nocc, nvir, k, klo):
    """The 16 cblas_dgemm calls of omp/ccsd_trpdrv_omp_cbody.c:71-116, written
    as numpy products of the column-major buffers."""
    nv, no = nvir, nocc
    vv, ov = nv * nv, no * nv
    s = k - klo
    F = lambda x, r, c: np.asarray(x)[: r * c].reshape(r, c, order="F")
    sl = lambda x, n: np.asarray(x)[s * n:(s + 1) * n]
    Jia, Kia, Tij = F(arrays["jia"], nv, nv), F(arrays["kia"], nv, nv), F(arrays["tij"], nv, nv)
    Tia, Xia = F(arrays["tia"], nv, no), F(arrays["xia"], nv, no)
    Kij, Jij = F(arrays["kij"], no, nv), F(arrays["jij"], no, nv)
    Tkj, Jka, Kka = (F(sl(arrays[n], vv), nv, nv) for n in ("tkj", "jka", "kka"))
    Kkj, Jkj = (F(sl(arrays[n], ov), no, nv) for n in ("kkj", "jkj"))
    Tka, Xka = (F(sl(arrays[n], ov), nv, no) for n in ("tka", "xka"))
    f1n = Jia @ Tkj.T - Tia @ Kkj
    f2n = Kia @ Tkj.T - Xia @ Kkj
    f3n = Jia @ Tkj - Tia @ Jkj
    f4n = Kia @ Tkj - Xia @ Jkj
    f1t = Jka @ Tij.T - Tka @ Kij
    f2t = Kka @ Tij.T - Xka @ Kij
    f3t = Jka @ Tij - Tka @ Jij
    f4t = Kka @ Tij - Xka @ Jij
    return np.array([f1n, f1t, f2n, f2t, f3n, f3t, f4n, f4t])


@pytest.mark.parametrize("provider", [1, 0])
def test_fmats_vs_numpy(orc, cc, provider):
    orc.blas_select(provider)
    nocc, nvir, kchunk = 7, 29, 4
    arrays = cc.make_inputs(nocc, nvir, kchunk, 3)
    for k, klo in [(1, 1), (4, 2), (7, 4)]:
        got = orc.fmats(arrays, nocc, nvir, k, klo)
        want = numpy_fmats(arrays, nocc, nvir, k, klo)
        assert np.allclose(got, want, rtol=0, atol=1e-15)
    orc.blas_select(0)


def test_fmats_exact_rational(orc, cc):
    """Tiny dyadic inputs: every product and sum is exact in binary64, so the
    oracle must equal the exact rational GEMMs."""
    orc.blas_select(1)
    nocc, nvir, kchunk, k, klo = 2, 3, 2, 2, 1
    rng = np.random.default_rng(0)
    arrays = cc.make_inputs(nocc, nvir, kchunk, 1)
    for n in arrays:
        if n != "eorb":
            arrays[n] = rng.integers(-8, 9, arrays[n].size) / 8.0
    got = orc.fmats(arrays, nocc, nvir, k, klo)
    fr = {n: [Fraction(float(x)) for x in arrays[n]] for n in arrays}
    nv, no, s = nvir, nocc, k - klo
    vv, ov = nv * nv, no * nv

    def g(name, r, c, off=0):
        return lambda i, j: fr[name][off + i + j * r]
    Jia, Kia, Tij = g("jia", nv, nv), g("kia", nv, nv), g("tij", nv, nv)
    Tia, Xia = g("tia", nv, no), g("xia", nv, no)
    Kij, Jij = g("kij", no, nv), g("jij", no, nv)
    Tkj, Jka, Kka = g("tkj", nv, nv, s * vv), g("jka", nv, nv, s * vv), g("kka", nv, nv, s * vv)
    Kkj, Jkj = g("kkj", no, nv, s * ov), g("jkj", no, nv, s * ov)
    Tka, Xka = g("tka", nv, no, s * ov), g("xka", nv, no, s * ov)
    defs = [  # (A, B, transB, C, D) for f = A op(B) - C D, body order
        (Jia, Tkj, True, Tia, Kkj), (Jka, Tij, True, Tka, Kij),
        (Kia, Tkj, True, Xia, Kkj), (Kka, Tij, True, Xka, Kij),
        (Jia, Tkj, False, Tia, Jkj), (Jka, Tij, False, Tka, Jij),
        (Kia, Tkj, False, Xia, Jkj), (Kka, Tij, False, Xka, Jij)]
    for m, (A, B, tb, C, D) in enumerate(defs):
        for b in range(nv):
            for c in range(nv):
                v = sum(A(b, d) * (B(c, d) if tb else B(d, c)) for d in range(nv))
                v -= sum(C(b, l) * D(l, c) for l in range(no))
                assert got[m][b, c] == float(v)
    orc.blas_select(0)


def test_providers_agree(orc, cc):
    if orc.blas_select(2) != 2:
        pytest.skip("MKL not in this image")
    nocc, nvir = 6, 50
    arrays = cc.make_inputs(nocc, nvir, 2, 4)
    a = orc.fmats(arrays, nocc, nvir, 2, 1)
    orc.blas_select(1)
    b = orc.fmats(arrays, nocc, nvir, 2, 1)
    orc.blas_select(0)
    assert np.allclose(a, b, rtol=0, atol=1e-15)


def swap_nt(arrays, nocc, nvir, s):
    """Exchanges the n-side and t-side operands of slice s (and the t1v/dint
    vectors), which turns f?n into f?t and back."""
    vv, ov = nvir * nvir, nocc * nvir
    A = {n: np.array(v, copy=True) for n, v in arrays.items()}
    for x, y, nx, ny in [("jia", "jka", vv, vv), ("kia", "kka", vv, vv), ("tia", "tka", ov, ov),
                         ("xia", "xka", ov, ov), ("tij", "tkj", vv, vv), ("kij", "kkj", ov, ov),
                         ("jij", "jkj", ov, ov)]:
        a0 = arrays[x][:nx].copy()
        A[x][:nx] = arrays[y][s * ny:(s + 1) * ny]
        A[y][s * ny:(s + 1) * ny] = a0
    for x, y in [("t1v1", "t1v2"), ("dintx1", "dintx2"), ("dintc1", "dintc2")]:
        A[x], A[y] = arrays[y].copy(), arrays[x].copy()
    return A


def test_nt_swap_symmetry(orc, cc, one_thread):
    nocc, nvir, kchunk = 5, 30, 2
    arrays = cc.make_inputs(nocc, nvir, kchunk, 21)
    trip = np.array([[nocc + 1, 4, 3, 2, 1]], dtype=np.int32)
    t0 = orc.run_list(arrays, 0, nocc, nvir, trip, nthreads=1)["terms"][0]
    sw = swap_nt(arrays, nocc, nvir, 1)
    t1 = orc.run_list(sw, 0, nocc, nvir, trip, nthreads=1)["terms"][0]
    assert np.allclose(t1, t0[[2, 3, 0, 1]], rtol=1e-14, atol=0)


def test_driver_rows_keep_the_ik_exchange(orc, cc, one_thread):
    """The driver-tensor slicing of a full-tensor job (t1v1 = t1[i], dintc1 =
    dja[j][k], dintx1 = djx[j][k], and the same with i, k exchanged for the
    second set) makes triple (k,j,i) the i <-> k exchange of triple (i,j,k):
    its (emp4i, emp5i) are the other's (emp4k, emp5k).  A slicing that took a
    vector from the wrong occupied index breaks this."""
    nocc, nvir = 4, 14
    t = cc.make_tensor_inputs(nocc, nvir, 61, driver=True)
    for i, j, k in [(3, 2, 1), (4, 4, 2), (2, 1, 4)]:
        tr = np.array([[nocc + 2, i, j, k, k]], np.int32)
        ts = np.array([[nocc + 2, k, j, i, i]], np.int32)
        a = orc.run_list(cc.gather_triple(t, nocc, nvir, i, j, k), 0, nocc, nvir, tr, nthreads=1)
        b = orc.run_list(cc.gather_triple(t, nocc, nvir, k, j, i), 0, nocc, nvir, ts, nthreads=1)
        assert np.allclose(b["terms"][0], a["terms"][0][[2, 3, 0, 1]], rtol=1e-13, atol=0)
    # the mistaken slicing dintc1 = dja[j][i] fails the same check
    i, j, k = 3, 2, 1
    ga = cc.gather_triple(t, nocc, nvir, i, j, k)
    gb = cc.gather_triple(t, nocc, nvir, k, j, i)
    ga["dintc1"], gb["dintc1"] = ga["dintc2"].copy(), gb["dintc2"].copy()
    a = orc.run_list(ga, 0, nocc, nvir, np.array([[nocc + 2, i, j, k, k]], np.int32), nthreads=1)
    b = orc.run_list(gb, 0, nocc, nvir, np.array([[nocc + 2, k, j, i, i]], np.int32), nthreads=1)
    assert not np.allclose(b["terms"][0], a["terms"][0][[2, 3, 0, 1]], rtol=1e-6, atol=0)


def test_homogeneity_and_zero(orc, cc, one_thread):
    nocc, nvir = 4, 20
    arrays = cc.make_inputs(nocc, nvir, 1, 8)
    trip = np.array([[nocc + 3, 3, 2, 1, 1]], dtype=np.int32)
    base = orc.run_list(arrays, 0, nocc, nvir, trip, nthreads=1)["terms"][0]
    sc = dict(arrays)
    for n in ["tij", "tkj", "tia", "tka", "xia", "xka", "jia", "jka", "kia", "kka", "jij",
              "jkj", "kij", "kkj"]:
        sc[n] = arrays[n] * 2.0  # f scales by exactly 4
    two = orc.run_list(sc, 0, nocc, nvir, trip, nthreads=1)["terms"][0]
    assert np.array_equal(two, base * np.array([16.0, 4.0, 16.0, 4.0]))
    zero = dict(arrays)
    for n in ["jia", "kia", "jka", "kka", "tia", "xia", "tka", "xka"]:
        zero[n] = np.zeros_like(arrays[n])
    assert not np.any(orc.run_list(zero, 0, nocc, nvir, trip, nthreads=1)["terms"])


def test_ik_rule(orc, cc):
    nocc, nvir = 4, 16
    arrays = cc.make_inputs(nocc, nvir, nocc, 5)
    trip = cc.triple_list(nocc, [(1, nocc)], "ijk")
    out = orc.run_list(arrays, 0, nocc, nvir, trip)
    t = out["terms"]
    ik = trip[:, 1] != trip[:, 3]
    e4 = t[:, 0].sum() + t[ik, 2].sum()
    e5 = t[:, 1].sum() + t[ik, 3].sum()
    assert np.isclose(out["emp4"], e4, rtol=1e-13) and np.isclose(out["emp5"], e5, rtol=1e-13)
    assert np.isclose(out["emp4_plain"], out["emp4"], rtol=1e-12)


def _round32(arrays):
    return {n: np.asarray(v, dtype=np.float32).astype(np.float64) for n, v in arrays.items()}


def test_fp32_reference_body_tracks_fp64(orc, cc):
    """The FP32 reference body (gpu-omp, built with its SKIP_DGEMM switch) on
    FP32-rounded inputs agrees with the FP64 oracle on the same inputs to FP32
    accuracy, term by term; a dropped or mis-signed term would show at the
    percent level."""
    L = orc.ref_cbody_f32_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    nocc, nvir, kchunk = 5, 24, 5
    arrays = _round32(cc.make_inputs(nocc, nvir, kchunk, 7))
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    for a, i, j, k, klo in trip[::3]:
        f = orc.fmats(arrays, nocc, nvir, int(k), int(klo))
        fcm = [np.ascontiguousarray(f[m].T) for m in range(8)]
        e4, e5 = orc.ref_cbody_f32_glue(L, fcm, arrays, 0, nocc, nvir, *map(int, (a, i, j, k, klo)))
        r = orc.run_list(arrays, 0, nocc, nvir, np.array([[a, i, j, k, klo]], np.int32), with_abs=True)
        t, s = r["terms"][0], r["abs4"][0]
        ik = i != k
        E4, S4 = t[0] + ik * t[2], s[0] + ik * s[2]
        E5, S5 = t[1] + ik * t[3], s[1] + ik * s[3]
        assert abs(e4 - E4) <= 1e-5 * S4 and abs(e5 - E5) <= 1e-5 * S5


def _closed_form_tengy(F, v, delta, swap_one=False):
    """The four sums of omp/ccsd_tengy_c99.c:91-111 when every denominator is
    the same delta, as matrix expressions: a summand term X_bc Y_bc summed over
    all (b, c) is the Frobenius product <X, Y>, X_cb Y_bc is <X^T, Y>, and
    t1v[b] d[c] X_bc sums to t1v^T X d.  F[name] is the nvir x nvir matrix with
    F[b, c] = f(b + c*nvir).  swap_one uses F1n where F1n^T belongs (a plausible
    transcription error the check must catch)."""
    def ip(x, y):
        return float(np.sum(x * y))

    def e4(n, t):
        # n, t: dicts "1".."4" -> the n-side and t-side matrices of this sum
        T = (lambda m: m) if swap_one else (lambda m: m.T)
        a1 = t["1"] + T(n["1"]) + t["2"].T + n["3"] + n["4"].T
        b1 = t["1"] - 2 * t["2"] - 2 * t["3"] + t["4"]
        a2 = n["1"] + t["1"].T + n["2"].T + n["3"].T
        b2 = 2 * t["1"] - t["2"] - t["3"] + 2 * t["4"]
        c3 = ip(n["1"], n["1"] + n["3"].T + 2 * t["4"].T) + ip(n["2"], t["2"].T) + ip(n["3"], t["4"])
        return delta * (ip(a1, b1) - ip(a2, b2) + 3 * c3)

    def e5(n, t, t1v, dx, dc):
        mx = (t["1"] + n["2"] + n["4"].T - 2 * (t["3"] + n["4"] + n["2"].T + n["1"] + t["2"] + n["3"].T)
              + 4 * (n["3"] + t["4"] + n["1"].T))
        mc = n["1"] + n["4"] + t["1"].T - 2 * (n["2"] + n["3"] + t["2"].T)
        return delta * (t1v @ mx @ dx + t1v @ mc @ dc)

    fn = {q: F[q + "n"] for q in "1234"}
    ft = {q: F[q + "t"] for q in "1234"}
    return (e4(fn, ft), e5(fn, ft, v["t1v1"], v["dintx1"], v["dintc1"]),
            e4(ft, fn), e5(ft, fn, v["t1v2"], v["dintx2"], v["dintc2"]))


@pytest.mark.parametrize("nvir", [1, 7, 23])
def test_tengy_constant_denominator_closed_form(orc, cc, one_thread, nvir):
    """With every virtual orbital energy equal, denom = -1/(2 e_v + eaijk) is
    one constant, and emp4i/emp5i/emp4k/emp5k are Frobenius products, traces
    and bilinear forms of the f matrices (computed here with numpy matrix
    algebra, not the summation loop).  The oracle's reduction must match them;
    the same closed form with one operand left untransposed must not."""
    nocc = 3
    arrays = dict(rand_case(cc, nocc, nvir, nocc, seed=31))
    eorb = np.array(arrays["eorb"], dtype=np.float64)
    eorb[nocc:] = 0.75
    arrays["eorb"] = eorb
    ea = -1.9
    delta = -1.0 / (0.75 + 0.75 + ea)
    rng = np.random.default_rng(nvir)
    names = ["1n", "1t", "2n", "2t", "3n", "3t", "4n", "4t"]  # the tengy argument order
    F = {nm: rng.uniform(-1.0, 1.0, (nvir, nvir)) for nm in names}
    got = orc.tengy([F[nm].T for nm in names], arrays, ea, 0, nocc, nvir)  # [c][b] = column-major
    v = {nm: np.asarray(arrays[nm][:nvir], dtype=np.float64) for nm in
         ("t1v1", "t1v2", "dintx1", "dintx2", "dintc1", "dintc2")}
    want = _closed_form_tengy(F, v, delta)
    scale = delta * 64 * nvir * nvir  # the sums' size: entries of order one, a few dozen terms each
    for g, w in zip(got, want):
        assert abs(g - w) <= 1e-13 * abs(scale), (got, want)
    if nvir > 1:
        wrong = _closed_form_tengy(F, v, delta, swap_one=True)
        assert max(abs(g - w) for g, w in zip(got, wrong)) > 1e-3 * abs(scale)


def test_fp32_golden_matches_reference_build(cc, orc):
    """tests/golden/fp32_ref.npz (the FP32 reference body's per-triple emp4 and
    emp5, used by tests/test_gpu_fp32.py where oracle/_ref is absent) is
    reproduced bit for bit by the reference build, under the settings that wrote
    it (the oracle's first DGEMM provider, one thread)."""
    L = orc.ref_cbody_f32_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fp32_ref.npz"),
                allow_pickle=False)
    orc.blas_select(0)
    for case in (0, 1):
        nocc, nvir, kchunk, seed = (int(x) for x in g["case%d_shape" % case])
        arrays = cc.make_inputs(nocc, nvir, kchunk, seed)
        a32 = {n: np.asarray(v, dtype=np.float32).astype(np.float64) for n, v in arrays.items()}
        trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
        assert np.array_equal(trip, g["case%d_triples" % case])
        for t, (a, i, j, k, klo) in enumerate(trip.tolist()):
            f = orc.fmats(a32, nocc, nvir, k, klo)
            fcm = [np.ascontiguousarray(f[m].T).astype(np.float32) for m in range(8)]
            got = orc.ref_cbody_f32_glue(L, fcm, a32, 0, nocc, nvir, a, i, j, k, klo)
            assert tuple(got) == tuple(g["case%d_e45" % case][t]), (case, t)

