"""FP32 form (CCSDT_MODE_FP32) against the FP32 reference body (needs a GPU).

The FP32 form rounds the job's operands to FP32 and evaluates the GEMMs and the
energy summand in FP32, as gpu-omp/ccsd_trpdrv_omp_cbody.c does.  Per triple,
its emp4 and emp5 (i != k rule applied) must match the FP32 reference body run
on the same FP32 inputs and FP32-rounded f matrices within 1e-5 of the terms'
absolute-summand scale (two FP32 evaluations in different summation orders);
against the FP64 oracle E[T] agrees to FP32 accuracy.  The reference body's
results are also committed (tests/golden/fp32_ref.npz) for boxes without the
reference build."""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device(cc):
    cc.init(0)


def _round32(arrays):
    return {n: np.asarray(v, dtype=np.float32).astype(np.float64) for n, v in arrays.items()}


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fp32_ref.npz")


@pytest.mark.parametrize("case", [0, 1])
def test_fp32_form_against_fp32_reference(cc, orc, case):
    """Cases (8, 40, kchunk 8) and (5, 37, kchunk 2).  The FP32 reference body
    runs live when oracle/_ref is built; otherwise its per-triple results come
    from tests/golden/fp32_ref.npz (written by the same reference build, inputs
    checked by digest; tests/test_oracle.py checks the fixture against the
    reference build where it exists), so the pin holds on a box without it."""
    g = np.load(GOLD, allow_pickle=False)
    nocc, nvir, kchunk, seed = (int(x) for x in g["case%d_shape" % case])
    arrays = cc.make_inputs(nocc, nvir, kchunk, seed)
    names = list(g["case%d_names" % case])
    assert [hashlib.sha256(np.ascontiguousarray(arrays[n], dtype="<f8").tobytes()).hexdigest()
            for n in names] == list(g["case%d_digests" % case])
    a32 = _round32(arrays)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    assert np.array_equal(trip, g["case%d_triples" % case])
    L = orc.ref_cbody_f32_lib()
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        got = job.run(trip, mode="fp32", terms=True)["terms"]
    ref = orc.run_list(a32, 0, nocc, nvir, trip, with_abs=True)
    for t, (a, i, j, k, klo) in enumerate(trip):
        if L is not None:
            f = orc.fmats(a32, nocc, nvir, int(k), int(klo))
            fcm = [np.ascontiguousarray(f[m].T).astype(np.float32) for m in range(8)]
            e4, e5 = orc.ref_cbody_f32_glue(L, fcm, a32, 0, nocc, nvir, *map(int, (a, i, j, k, klo)))
        else:
            e4, e5 = g["case%d_e45" % case][t]
        ik = i != k
        g4 = got[t, 0] + ik * got[t, 2]
        g5 = got[t, 1] + ik * got[t, 3]
        s = ref["abs4"][t]
        assert abs(g4 - e4) <= 1e-5 * (s[0] + ik * s[2]), t
        assert abs(g5 - e5) <= 1e-5 * (s[1] + ik * s[3]), t


def test_fp32_form_totals_and_schedules(cc, orc):
    """E[T] of the FP32 form within 1e-5 of the FP64 oracle's; persistent,
    pair-major and graph runs bitwise equal to the plain run; f output of a
    single triple within FP32 rounding of the oracle's."""
    arrays, blocks, trip, kchunk = cc.config_case(0)
    ref = orc.run_list(arrays, 0, 8, 40, trip)
    E = ref["emp4"] + ref["emp5"]
    with cc.Job(arrays, 8, 40, kchunk) as job:
        a = job.run(trip, mode="fp32", terms=True)
        b = job.run(trip, mode="fp32", terms=True, order=1, persistent=True, graph=True)
        f = job.run(trip[7:8], mode="fp32", fout=True)["f"]
    assert abs(a["emp4"] + a["emp5"] - E) <= 1e-5 * abs(E)
    assert np.array_equal(a["terms"], b["terms"])
    want = orc.fmats(arrays, 8, 40, int(trip[7, 3]), int(trip[7, 4]))
    assert np.max(np.abs(f - want)) <= 1e-5 * np.max(np.abs(want))
