"""FP32 form (CCSDT_MODE_FP32) against the FP32 reference body (needs a GPU).

The FP32 form rounds the job's operands to FP32 and evaluates the GEMMs and the
energy summand in FP32, as gpu-omp/ccsd_trpdrv_omp_cbody.c does.  Per triple,
its emp4 and emp5 (i != k rule applied) must match the FP32 reference body run
on the same FP32 inputs and FP32-rounded f matrices within 1e-5 of the terms'
absolute-summand scale (two FP32 evaluations in different summation orders);
against the FP64 oracle E[T] agrees to FP32 accuracy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device(cc):
    cc.init(0)


def _round32(arrays):
    return {n: np.asarray(v, dtype=np.float32).astype(np.float64) for n, v in arrays.items()}


@pytest.mark.parametrize("nocc,nvir,kchunk", [(8, 40, 8), (5, 37, 2)])
def test_fp32_form_against_fp32_reference(cc, orc, nocc, nvir, kchunk):
    L = orc.ref_cbody_f32_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    arrays = cc.make_inputs(nocc, nvir, kchunk, 808)
    a32 = _round32(arrays)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        got = job.run(trip, mode="fp32", terms=True)["terms"]
    ref = orc.run_list(a32, 0, nocc, nvir, trip, with_abs=True)
    for t, (a, i, j, k, klo) in enumerate(trip):
        f = orc.fmats(a32, nocc, nvir, int(k), int(klo))
        fcm = [np.ascontiguousarray(f[m].T).astype(np.float32) for m in range(8)]
        e4, e5 = orc.ref_cbody_f32_glue(L, fcm, a32, 0, nocc, nvir, *map(int, (a, i, j, k, klo)))
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
