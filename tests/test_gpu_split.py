"""Ozaki INT8 split form against the FP64 oracle (needs a GPU).

Tolerance: E[T] = emp4 + emp5 over the triple list within 1e-8 relative of the
oracle for slices >= 4 (the form's stated bound); every per-triple term within
1e-6 of its absolute-summand scale; more slices must converge toward FP64.
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
    form's at 4 slices and at 3 (the count the autotuner picks at this shape),
    and bitwise repeatable."""
    arrays, blocks, trip, kchunk = cc.config_case(1)
    with cc.Job(arrays, 30, 300, kchunk) as job:
        a = job.run(trip, mode="fp64")
        b = job.run(trip, mode="ozaki_int8", slices=4, terms=True)
        c = job.run(trip, mode="ozaki_int8", slices=4, terms=True)
        d = job.run(trip, mode="ozaki_int8", slices=3, order=1)
    E = et(a["emp4"], a["emp5"])
    assert abs(et(b["emp4"], b["emp5"]) - E) / abs(E) < ET_TOL
    assert abs(et(d["emp4"], d["emp5"]) - E) / abs(E) < ET_TOL
    assert np.array_equal(b["terms"], c["terms"])


def test_bad_slice_count(cc):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    with cc.Job(arrays, 8, 40, kchunk) as job:
        with pytest.raises(cc.CcsdtError):
            job.run(trip, mode="ozaki_int8", slices=9)
