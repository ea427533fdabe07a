"""The committed golden fixtures (terms computed by the reference's own tengy
and body glue, see tests/golden/make_golden.py) against the regenerated inputs
and the oracle."""
import hashlib
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def checksums(arrays, names):
    return np.array([[float(np.sum(arrays[n])),
                      float(np.sum(arrays[n] * (np.arange(arrays[n].size) % 7)))] for n in names])


def digests(arrays, names):
    return [hashlib.sha256(np.ascontiguousarray(arrays[n], dtype="<f8").tobytes()).hexdigest()
            for n in names]


def test_inputs_regenerate(cc):
    for cfg, fname in [(0, "config0.npz"), (1, "config1_head.npz")]:
        g = load(fname)
        arrays = cc.config_case(cfg)[0]
        names = list(g["names"])
        assert np.array_equal(checksums(arrays, names), g["sums"])
        # every bit of every regenerated input
        assert digests(arrays, names) == list(g["digests"])


def test_checksum_sees_a_one_ulp_change(cc):
    """The fixture's pins catch a change of one input element by one ulp."""
    g = load("config0.npz")
    arrays = dict(cc.config_case(0)[0])
    names = list(g["names"])
    a = arrays["tkj"].copy()
    a[77] = np.nextafter(a[77], np.inf)
    arrays["tkj"] = a
    assert digests(arrays, names) != list(g["digests"])


def test_oracle_matches_config0(cc, orc):
    g = load("config0.npz")
    arrays, blocks, trip, kchunk = cc.config_case(0)
    assert np.array_equal(trip, g["triples"])
    out = orc.run_list(arrays, 0, 8, 40, trip, with_abs=True)
    scale = np.maximum(np.abs(g["terms"]), out["abs4"])
    assert np.max(np.abs(out["terms"] - g["terms"]) / scale) < 1e-13
    assert np.allclose([out["emp4_plain"], out["emp5_plain"]], g["emp"], rtol=1e-13, atol=0)
    f7 = orc.fmats(arrays, 8, 40, int(trip[7, 3]), int(trip[7, 4]))
    assert np.allclose(f7, g["f7"], rtol=0, atol=1e-16)


def test_oracle_matches_config1_head(cc, orc):
    g = load("config1_head.npz")
    arrays = cc.config_case(1)[0]
    out = orc.run_list(arrays, 0, 30, 300, g["triples"], with_abs=True)
    scale = np.maximum(np.abs(g["terms"]), out["abs4"])
    assert np.max(np.abs(out["terms"] - g["terms"]) / scale) < 1e-13
