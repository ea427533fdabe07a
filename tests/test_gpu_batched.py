"""Persistent grids and captured graphs give the same results as plain launches
(needs a GPU).  Graph replays are checked one by one against eager runs, with
eager work and synchronizations between the replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device(cc):
    cc.init(0)


@pytest.mark.parametrize("mode,slices", [("fp64", 0), ("ozaki_int8", 4)])
def test_persistent_equals_plain(cc, mode, slices):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    with cc.Job(arrays, 8, 40, kchunk) as job:
        a = job.run(trip, mode=mode, slices=slices, terms=True)
        b = job.run(trip, mode=mode, slices=slices, terms=True, persistent=True)
    assert np.array_equal(a["terms"], b["terms"]) and a["emp4"] == b["emp4"]


@pytest.mark.parametrize("mode,slices,persistent", [("fp64", 0, True), ("ozaki_int8", 3, True),
                                                    ("fp64", 0, False)])
def test_graph_replays_match_eager(cc, mode, slices, persistent):
    nocc, nvir, kchunk = 6, 45, 3
    arrays = cc.make_inputs(nocc, nvir, kchunk, 77)
    full = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
    rng = np.random.default_rng(0)
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        for rep in range(4):
            sub = full[np.sort(rng.choice(len(full), 20, replace=False))]
            eager = job.run(sub, mode=mode, slices=slices, terms=True)
            g = job.run(sub, mode=mode, slices=slices, terms=True, persistent=persistent, graph=True)
            assert np.array_equal(eager["terms"], g["terms"]), rep
            assert eager["emp4"] == g["emp4"] and eager["emp5"] == g["emp5"]
            # other eager work between replays, including a different list size
            job.run(full[:7], mode="fp64")


def test_graph_recaptures_after_growth(cc):
    arrays, blocks, trip, kchunk = cc.config_case(0)
    with cc.Job(arrays, 8, 40, kchunk) as job:
        a = job.run(trip[:30], terms=True, graph=True, persistent=True)
        job.run(trip, terms=True)               # grows the task and partial buffers
        b = job.run(trip[:30], terms=True, graph=True, persistent=True)
        c = job.run(trip[:30], terms=True)
    assert np.array_equal(a["terms"], c["terms"]) and np.array_equal(b["terms"], c["terms"])


def test_uneven_blocks_batched(cc, orc):
    """Uneven k blocks (lengths 8..48 over nocc) at a reduced size: one launch over
    all triples, FP64 and split, against the oracle."""
    nocc, nvir = 12, 48
    blocks = cc.blocks_uneven(nocc, 5, lo=2, hi=6)
    kchunk = max(b - a + 1 for a, b in blocks)
    arrays = cc.make_inputs(nocc, nvir, kchunk, 55)
    trip = cc.triple_list(nocc, blocks)
    ref = orc.run_list(arrays, 0, nocc, nvir, trip, with_abs=True)
    with cc.Job(arrays, nocc, nvir, kchunk) as job:
        f = job.run(trip, terms=True, persistent=True)
        s = job.run(trip, mode="ozaki_int8", slices=5, terms=True, persistent=True, graph=True)
    scale = np.maximum(np.abs(ref["terms"]), ref["abs4"])
    assert np.max(np.abs(f["terms"] - ref["terms"]) / scale) < 1e-11
    E = ref["emp4"] + ref["emp5"]
    assert abs(s["emp4"] + s["emp5"] - E) / abs(E) < 1e-8
