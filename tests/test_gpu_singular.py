"""A singular energy denominator (needs a GPU).

The reference divides without a guard: denom = -1/(e_b + e_c + eaijk)
(omp/ccsd_tengy_c99.c:65), so a (b,c) point where e_b + e_c + eaijk is exactly
zero makes that triple's terms non-finite, and nothing else.  The inputs below
put such a point into one triple on purpose, with dyadic orbital energies so
that every sum involved is exact: eaijk = e_a - ((e_i + e_j) + e_k) =
0.5 - ((1.25 + 1) + 1) = -2.75 and e_b + e_c = 1.375 + 1.375 = 2.75 at
b = c = 5 (0-based virtual index).  The drop-in body, the drop-in energy
reduction and the job API (FP64 and split forms) must all return non-finite
terms for that triple, as the oracle does, leave every other triple of the list
finite and equal to the oracle's, and stay usable afterwards.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-11
NOCC, NVIR, KCHUNK = 3, 20, 3
B_SING = 5  # 0-based virtual index of the singular point (b = c)


@pytest.fixture(scope="module", autouse=True)
def device(cc):
    cc.init(0)
    cc.lib().ccsdt_set_error_policy(1)


def singular_inputs(cc):
    arrays = cc.make_inputs(NOCC, NVIR, KCHUNK, 9090)
    eorb = np.array(arrays["eorb"])
    eorb[:NOCC] = [1.0, 1.25, 1.5]
    ev = 5.0 + np.arange(NVIR) * 0.1  # every other sum e_b + e_c is at least 5.6
    ev[0] = 0.5
    ev[B_SING] = 1.375
    eorb[NOCC:] = ev
    arrays["eorb"] = eorb
    return arrays


# (a, i, j, k, klo), 1-based: the second triple is singular (i=2, j=1, k=1)
TRIPLES = np.array([[NOCC + 1, 3, 2, 1, 1],
                    [NOCC + 1, 2, 1, 1, 1],
                    [NOCC + 1, 3, 3, 3, 1]], dtype=np.int32)


def test_inputs_are_singular_exactly_once(cc, orc):
    arrays = singular_inputs(cc)
    e = arrays["eorb"]
    for t, (a, i, j, k, _) in enumerate(TRIPLES):
        ea = orc.eaijk(e, 0, int(a), int(i), int(j), int(k))
        s = e[NOCC:, None] + e[None, NOCC:] + ea
        assert (np.count_nonzero(s == 0.0) == 1) == (t == 1)


def check_terms(cc, orc, arrays, got):
    ref = orc.run_list(arrays, 0, NOCC, NVIR, TRIPLES, with_abs=True)
    # the oracle's singular triple is non-finite in every term that the point enters
    assert not np.all(np.isfinite(ref["terms"][1]))
    bad = ~np.isfinite(ref["terms"][1])
    assert np.all(~np.isfinite(got["terms"][1][bad]))
    for t in (0, 2):
        assert np.all(np.isfinite(got["terms"][t]))
        rel = np.abs(got["terms"][t] - ref["terms"][t]) / np.maximum(np.abs(ref["terms"][t]),
                                                                     ref["abs4"][t])
        assert np.max(rel) < TOL
    assert not np.isfinite(got["emp4"] + got["emp5"])


@pytest.mark.parametrize("mode,slices,occ", [("fp64", 0, 0), ("ozaki_int8", 4, 0),
                                              ("ozaki_int8", 3, 1)])
def test_job_api_returns_non_finite_and_stays_usable(cc, orc, mode, slices, occ):
    arrays = singular_inputs(cc)
    with cc.Job(arrays, NOCC, NVIR, KCHUNK) as job:
        got = job.run(TRIPLES, mode=mode, slices=slices, occ_digits=occ, terms=True)
        if mode == "fp64":
            check_terms(cc, orc, arrays, got)
        else:
            assert not np.all(np.isfinite(got["terms"][1]))
            assert np.all(np.isfinite(got["terms"][[0, 2]]))
        # the same job afterwards, on the list without the singular triple
        ok = TRIPLES[[0, 2]]
        again = job.run(ok, mode="fp64", terms=True)
        ref = orc.run_list(arrays, 0, NOCC, NVIR, ok, with_abs=True)
        assert np.max(np.abs(again["terms"] - ref["terms"]) /
                      np.maximum(np.abs(ref["terms"]), ref["abs4"])) < TOL
        assert abs(again["emp4"] - ref["emp4"]) <= TOL * abs(ref["emp4"])


def test_dropins_return_non_finite(cc, orc):
    arrays = singular_inputs(cc)
    a, i, j, k, klo = (int(x) for x in TRIPLES[1])
    f = [np.zeros(NVIR * NVIR) for _ in range(8)]
    e4, e5 = cc.cbody(f, arrays["eorb"], 0, NOCC, NVIR, 0.0, 0.0, a, i, j, k, klo, arrays)
    assert not np.isfinite(e4 + e5)
    # the f matrices do not involve the denominators: they match the oracle's
    want = orc.fmats(arrays, NOCC, NVIR, k, klo)
    for m in range(8):
        got = f[m].reshape(NVIR, NVIR).T
        assert np.max(np.abs(got - want[m])) <= 1e-13 * np.max(np.abs(want[m]))
    ea = orc.eaijk(arrays["eorb"], 0, a, i, j, k)
    vec = {n: arrays[n][:NVIR] for n in ("dintc1", "dintx1", "t1v1", "dintc2", "dintx2", "t1v2")}
    for sym in ("ccsd_tengy_omp", "ccsd_tengy", "ccsd_tengy_omp_"):
        out = cc.tengy(f, vec, arrays["eorb"], ea, 0, NOCC, NVIR, symbol=sym)
        ref = orc.tengy(f, arrays, ea, 0, NOCC, NVIR)
        for g, r in zip(out, ref):
            assert np.isfinite(g) == np.isfinite(r)
        assert not all(np.isfinite(out))
    # the drop-in keeps working: a regular triple afterwards matches the oracle
    a, i, j, k, klo = (int(x) for x in TRIPLES[0])
    e4, e5 = cc.cbody(f, arrays["eorb"], 0, NOCC, NVIR, 0.0, 0.0, a, i, j, k, klo, arrays)
    ref = orc.run_list(arrays, 0, NOCC, NVIR, TRIPLES[:1])
    assert abs(e4 - ref["emp4"]) <= TOL * abs(ref["emp4"])
    assert abs(e5 - ref["emp5"]) <= TOL * abs(ref["emp5"])
