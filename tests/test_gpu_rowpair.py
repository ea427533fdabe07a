"""The row-pair tile of the Ozaki INT8 split form (needs a GPU).

CCSDT_TILE_ROW_PAIR runs two 16x16 block pairs stacked in one column block on
four waves, one product group each.  Its integer digit products are exact and
its recombination combines the same exact level sums with the same scales, so
the f matrices of a triple must be bitwise the two-wave tile's (which the
split-form tests pin against the oracle).  The energy summand is evaluated by
other lanes in another order, so per-triple terms and totals agree with the
two-wave tile's to rounding: within 1e-13 of the largest term of the triple
(measured differences are below 6e-14 of a term's own magnitude, 5e-13 for a
small term that cancels), far under the split form's 1e-8 bound.  Checked
on ragged shapes (odd block counts leave an item with one block; nvir = 1 and
17 leave a padded row block), with and without the occupied digit planes
(folded for nocc <= 32 and not), at 3 and 4 slices, on every schedule, where
each tile's own results are bitwise the same whatever the schedule, and at the
largest K the tile takes (nvir = 768), with operands that push its INT32 level
sums near their bound.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROWS = 2  # CCSDT_TILE_ROW_PAIR
RTOL = 1e-13  # the two tiles' energy sums run in different orders


def close(got, ref):
    """Terms of two tiles agree to RTOL of the largest term of their triple (a
    small term that cancels carries the rounding of the large summands it is
    made of), totals to RTOL of the larger total."""
    scale = np.max(np.abs(ref["terms"]), axis=1, keepdims=True)
    d = np.abs(got["terms"] - ref["terms"]) / np.maximum(scale, 1e-300)
    assert np.max(d) < RTOL, np.max(d)
    big = max(abs(ref["emp4"]), abs(ref["emp5"]))
    for key in ("emp4", "emp5"):
        assert abs(got[key] - ref[key]) <= RTOL * big


@pytest.fixture(scope="module", autouse=True)
def device(cc):
    cc.init(0)
    cc.lib().ccsdt_set_error_policy(1)


FORMS = [(3, 1), (3, 0), (4, 0)]  # (slices, occ_digits)


@pytest.mark.parametrize("nocc,nvir,kchunk", [(1, 1, 1), (3, 17, 2), (5, 37, 5), (8, 40, 8),
                                              (4, 64, 3), (13, 50, 4), (40, 33, 8), (7, 100, 7)])
@pytest.mark.parametrize("slices,occ", FORMS)
def test_bitwise_two_wave_tile(cc, nocc, nvir, kchunk, slices, occ):
    arrays = cc.make_inputs(nocc, nvir, kchunk, 700 + nvir + nocc)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    kw = dict(mode="ozaki_int8", slices=slices, occ_digits=occ, terms=True)
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        ref = job.run(trip, variant=0, **kw)
        got = job.run(trip, variant=ROWS, **kw)
    close(got, ref)


@pytest.mark.parametrize("slices,occ", FORMS)
def test_schedules_bitwise(cc, slices, occ):
    """Plain and persistent grids, graph replays with eager work between them,
    and both item orders: bitwise the tile's own plain run, and close to the
    two-wave tile's."""
    nocc, nvir, kchunk = 9, 90, 4
    arrays = cc.make_inputs(nocc, nvir, kchunk, 811)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    kw = dict(mode="ozaki_int8", slices=slices, occ_digits=occ, terms=True)
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        close(job.run(trip, variant=ROWS, **kw), job.run(trip, variant=0, **kw))
        ref = job.run(trip, variant=ROWS, **kw)
        for order in (0, 1):
            for persistent in (False, True):
                for wg in ((0, 1) if persistent else (0,)):
                    got = job.run(trip, variant=ROWS, order=order, persistent=persistent,
                                  wg_per_cu=wg, **kw)
                    assert np.array_equal(got["terms"], ref["terms"]), (order, persistent, wg)
                    for rep in range(3):
                        g = job.run(trip, variant=ROWS, order=order, persistent=persistent,
                                    wg_per_cu=wg, graph=True, **kw)
                        assert np.array_equal(g["terms"], ref["terms"]), (order, persistent, rep)
                        job.run(trip[:5], variant=0, **kw)  # eager work between replays


@pytest.mark.parametrize("slices,occ", FORMS)
def test_f_matrices_bitwise(cc, slices, occ):
    nocc, nvir, kchunk = 6, 45, 3
    arrays = cc.make_inputs(nocc, nvir, kchunk, 23)
    kw = dict(mode="ozaki_int8", slices=slices, occ_digits=occ, fout=True)
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        for t in ([nocc + 1, 4, 2, 2, 1], [nocc + 3, 6, 5, 5, 4]):
            trip = np.array([t], dtype=np.int32)
            ref = job.run(trip, variant=0, **kw)["f"]
            got = job.run(trip, variant=ROWS, **kw)["f"]
            assert np.array_equal(got, ref)


@pytest.mark.parametrize("slices,occ", FORMS)
def test_largest_k(cc, slices, occ):
    """nvir = 768, the largest K the tile takes: twelve 64-wide K steps, where the
    exact INT32 pairing of adjacent digit levels, (acc_L << 7) + acc_(L+1), still
    holds (its bound, (20352 * 128 + 24448) K < 2^31, allows K <= 816).  Random
    operands: f bitwise the two-wave tile's and terms close to it.  Constant
    operands put the top digit of every entry near 127, so the level-0 sums reach
    about 127^2 * 768 and their pairing about 1.6e9, near the 2^31 it can hold; an
    overflow there would be an O(1) error in f, against a split error of ~1e-6
    (three slices) or ~1e-8 (four) of the largest |f|, measured against the FP64
    form."""
    nocc, nvir, kchunk = 3, 768, 1
    arrays = cc.make_inputs(nocc, nvir, kchunk, 31)
    trips = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))[:2]
    kw = dict(mode="ozaki_int8", slices=slices, occ_digits=occ, fout=True)
    const = dict(arrays)
    for name in ("tij", "tkj", "tia", "tka", "xia", "xka") + tuple(cc.JK_NAMES):
        const[name] = np.full_like(arrays[name], 0.062)
    tol = 1e-5 if slices == 3 else 1e-7
    for t in trips:  # f output takes one triple per run
        trip = t[None]
        with cc.Job(arrays, nocc, nvir, kchunk) as job:
            ref = job.run(trip, variant=0, terms=True, **kw)
            got = job.run(trip, variant=ROWS, terms=True, **kw)
        assert np.array_equal(got["f"], ref["f"])
        close(got, ref)
        with cc.Job(const, nocc, nvir, kchunk) as job:
            f64 = job.run(trip, fout=True)["f"]
            got = job.run(trip, variant=ROWS, **kw)["f"]
        scale = np.max(np.abs(f64))
        assert scale > 0.0
        assert np.max(np.abs(got - f64)) <= tol * scale, np.max(np.abs(got - f64)) / scale


def test_production_list_bitwise(cc):
    """The 30/300 list (4960 triples) at S = 3 with occupied digit planes, the
    headline form: close to the two-wave tile, bitwise the same plain and
    persistent, and within 1e-8 in E[T] of the FP64 form."""
    arrays, blocks, trip, kchunk = cc.config_case(2)
    kw = dict(mode="ozaki_int8", slices=3, occ_digits=1, terms=True, order=1)
    with cc.Job(arrays, 30, 300, kchunk) as job:
        ref = job.run(trip, variant=0, **kw)
        got = job.run(trip, variant=ROWS, **kw)
        per = job.run(trip, variant=ROWS, persistent=True, graph=True, **kw)
        f64 = job.run(trip, mode="fp64")
    close(got, ref)
    assert np.array_equal(per["terms"], got["terms"])
    E = f64["emp4"] + f64["emp5"]
    assert abs(got["emp4"] + got["emp5"] - E) <= 1e-8 * abs(E)


def test_tensor_job_bitwise(cc):
    nocc, nvir = 5, 52
    t = cc.make_tensor_inputs(nocc, nvir, 31, driver=True)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, nocc))
    kw = dict(mode="ozaki_int8", slices=3, occ_digits=1, terms=True)
    with cc.Job.from_tensors(t, nocc, nvir) as job:
        ref = job.run(trip, variant=0, **kw)
        got = job.run(trip, variant=ROWS, **kw)
    close(got, ref)


@pytest.mark.parametrize("slices,occ", [(4, 0), (3, 1)])
def test_serving_on_row_pair_tile(cc, slices, occ):
    """A server on the row-pair tile: each job bitwise its plain run on that tile."""
    jobs = cc.serving_jobs(5, 4242, nocc_range=(3, 9), nvir_range=(20, 70))
    kw = dict(mode="ozaki_int8", slices=slices, occ_digits=occ)
    alone = []
    for jb in jobs:
        with cc.Job(jb["arrays"], jb["nocc"], jb["nvir"], jb["kchunk"]) as job:
            alone.append(job.run(jb["triples"], variant=ROWS, **kw))
    with cc.Server(variant=ROWS, items_per_tick=512, **kw) as srv:
        ids = [srv.submit(jb["arrays"], jb["nocc"], jb["nvir"], jb["kchunk"], jb["triples"])
               for jb in jobs]
        srv.drain()
        mixed = [srv.result(i) for i in ids]
    for a, m in zip(alone, mixed):
        assert m is not None and m["status"] == 0
        assert (a["emp4"], a["emp5"]) == (m["emp4"], m["emp5"])


def test_serving_per_job_forms_on_row_pair_tile(cc):
    """Per-job forms on the row-pair tile: a job of at least 1e8 energy summands
    runs at three slices with occupied digit planes, the small ones at four, in
    the same ticks (one launch per form); each job bitwise its plain run on the
    tile in its form."""
    big = dict(nocc=22, nvir=230, kchunk=22)
    big["triples"] = cc.triple_list(22, cc.blocks_uniform(22, 22))
    big["arrays"] = cc.make_inputs(22, 230, 22, 5151)
    assert len(big["triples"]) * 230 ** 2 >= 1e8
    jobs = [big] + cc.serving_jobs(4, 4242, nocc_range=(3, 9), nvir_range=(20, 70))
    with cc.Server(mode="ozaki_int8", slices=cc.SLICES_PER_JOB, variant=ROWS, items_per_tick=4096) as srv:
        ids = [srv.submit(jb["arrays"], jb["nocc"], jb["nvir"], jb["kchunk"], jb["triples"])
               for jb in jobs]
        srv.drain()
        mixed = [srv.result(i) for i in ids]
    assert [(m["slices"], m["occ_digits"]) for m in mixed] == [(3, 1)] + [(4, 0)] * 4
    for jb, m in zip(jobs, mixed):
        assert m["status"] == 0
        with cc.Job(jb["arrays"], jb["nocc"], jb["nvir"], jb["kchunk"]) as job:
            a = job.run(jb["triples"], mode="ozaki_int8", slices=m["slices"], occ_digits=m["occ_digits"],
                        variant=ROWS)
        assert (a["emp4"], a["emp5"]) == (m["emp4"], m["emp5"])


def test_tuner_lists_the_tile(cc):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    with cc.Job(arrays, 8, 40, kchunk) as job:
        best, table = job.autotune(trip, mode="ozaki_int8", slices=4, reps=1)
        assert {r["variant"] for r in table} == {0, ROWS}
        _, table5 = job.autotune(trip, mode="ozaki_int8", slices=5, reps=1)
        assert {r["variant"] for r in table5} == {0}


def test_rejected_forms(cc):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    with cc.Job(arrays, 8, 40, kchunk) as job:
        for kw in (dict(mode="ozaki_int8", slices=5), dict(mode="bf16x3"), dict(mode="fp32"),
                   dict(mode="fp64"), dict(mode="ozaki_int8", slices=4, variant=3)):
            kw.setdefault("variant", ROWS)
            with pytest.raises(cc.CcsdtError) as e:
                job.run(trip[:3], **kw)
            assert e.value.code == -1
    with pytest.raises(cc.CcsdtError):
        cc.Server(mode="ozaki_int8", slices=5, variant=ROWS)
    # K above 768 (INT32 pairing of digit levels): rejected, and not offered by the tuner
    nocc, nvir = 2, 769
    arrays = cc.make_inputs(nocc, nvir, 1, 5)
    trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, 1))[:1]
    with cc.Job(arrays, nocc, nvir, 1) as job:
        with pytest.raises(cc.CcsdtError) as e:
            job.run(trip, mode="ozaki_int8", slices=4, variant=ROWS)
        assert e.value.code == -1
        _, table = job.autotune(trip, mode="ozaki_int8", slices=4, reps=1)
        assert {r["variant"] for r in table} == {0}
