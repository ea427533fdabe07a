"""C-ABI surface and host logic, without a GPU.

The library must load, export every function include/ccsdt.h declares, and
fail loudly (an error code, never a silent CPU result) when no gfx950 device is
present.  Also covers the input generator and the triple lists of the
workload configurations.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ccsdt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \t\*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


def test_header_symbols_exported(cc):
    L = cc.lib()
    names = header_functions()
    assert "ccsd_trpdrv_omp_cbody_" in names and "ccsd_tengy_omp" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert len(names) >= 20


def test_version(cc):
    assert b"gfx950" in cc.lib().ccsdt_version()


def _has_gpu(cc):
    return cc.lib().ccsdt_init(0) == 0


def test_no_device_is_an_error(cc):
    if _has_gpu(cc):
        pytest.skip("a GPU is present")
    L = cc.lib()
    assert L.ccsdt_init(0) == -2
    assert "device" in L.ccsdt_last_error_string().decode().lower()
    assert L.ccsdt_debug_load_rate(16, 0) == -1.0


def test_dropin_without_device_reports_nan(cc):
    if _has_gpu(cc):
        pytest.skip("a GPU is present")
    L = cc.lib()
    L.ccsdt_set_error_policy(0)
    nocc, nvir = 3, 8
    arrays = cc.make_inputs(nocc, nvir, 1, 1)
    f = [np.zeros(nvir * nvir) for _ in range(8)]
    e4, e5 = cc.cbody(f, arrays["eorb"], 0, nocc, nvir, 0.0, 0.0, nocc + 1, 1, 1, 1, 1, arrays)
    assert math.isnan(e4) and math.isnan(e5)
    assert L.ccsdt_last_error() == -2
    vec = {n: arrays[n] for n in ("dintc1", "dintx1", "t1v1", "dintc2", "dintx2", "t1v2")}
    out = cc.tengy(f, vec, arrays["eorb"], 1.0, 0, nocc, nvir)
    assert all(math.isnan(x) for x in out)
    L.ccsdt_set_error_policy(1)


def test_job_argument_errors(cc):
    arrays = cc.make_inputs(3, 8, 1, 1)
    with pytest.raises(cc.CcsdtError) as e:
        cc.Job(arrays, nocc=3, nvir=8, kchunk=4)  # kchunk > nocc
    assert e.value.code == -1
    with pytest.raises(cc.CcsdtError) as e:
        cc.Job(arrays, nocc=0, nvir=8, kchunk=1)
    assert e.value.code == -1
    L = cc.lib()
    assert L.ccsdt_job_run(None, None, 0, None, None, None, None) == -5
    assert L.ccsdt_job_destroy(None) == -5


def test_footprint(cc):
    small = cc.footprint(8, 40, 8, 120)
    big = cc.footprint(30, 300, 16, 4960)
    assert 0 < small < big
    assert cc.footprint(0, 40, 1, 1) == -1
    # a full-tensor job holds nocc^2 T2 slices and 2(nocc + nocc^2) panels: at
    # nocc=40, nvir=400 that is several GB, so a 200 GB budget admits tens of jobs
    t = cc.tensor_footprint(40, 400, 11480)
    assert 5e9 < t < 12e9 and t > cc.footprint(40, 400, 40, 11480)
    assert cc.tensor_footprint(0, 400, 1) == -1
    # driver tensors: nocc + 2 nocc^2 pool rows of nvp doubles replace the six
    # fixed vectors (six rows, and their 256-byte aligned raw copies)
    nocc, nvir, nvp = 40, 400, 400
    d = cc.tensor_footprint(nocc, nvir, 11480, driver=True)
    assert d - t == ((nocc + 2 * nocc * nocc - 6) * nvp - 6 * 416) * 8


def test_gather_triple_reads_the_tensor_slices(cc):
    """The host gather used to check tensor jobs picks slice (j*nocc + i) of t2
    for Tij, (j*nocc + k) for Tkj, slice i / k of jv for Jia / Jka, and so on."""
    nocc, nvir = 3, 5
    vv, ov = nvir * nvir, nocc * nvir
    t = {"eorb": np.zeros(nocc + nvir)}
    for name, n, per in (("t2", nocc * nocc, vv), ("jv", nocc, vv), ("kv", nocc, vv),
                         ("t1o", nocc, ov), ("xo", nocc, ov), ("jo", nocc * nocc, ov),
                         ("ko", nocc * nocc, ov)):
        t[name] = np.repeat(np.arange(n, dtype=float), per)   # slice index as value
    for name in cc.VECTOR_NAMES:
        t[name] = np.ones(nvir)
    a = cc.gather_triple(t, nocc, nvir, 3, 2, 1)   # i=3, j=2, k=1 (1-based)
    i0, j0, k0 = 2, 1, 0
    assert a["tij"][0] == j0 * nocc + i0 and a["tkj"][0] == j0 * nocc + k0
    assert a["jia"][0] == i0 and a["jka"][0] == k0 and a["kka"][0] == k0
    assert a["tia"][0] == i0 and a["xka"][0] == k0
    assert a["jij"][0] == j0 * nocc + i0 and a["kkj"][0] == j0 * nocc + k0
    assert a["tia"].size == ov * nocc and a["t1v1"].size == vv


def test_triple_vectors_read_the_driver_rows(cc):
    """With driver tensors, triple (i,j,k) reads t1 rows i, k and the (j, k),
    (j, i) rows of dja and djx; without them, the job's fixed vectors."""
    nocc, nvir = 3, 4
    t = {"t1": np.repeat(np.arange(nocc, dtype=float), nvir),
         "dja": np.repeat(np.arange(nocc * nocc, dtype=float), nvir),
         "djx": np.repeat(100 + np.arange(nocc * nocc, dtype=float), nvir)}
    v = cc.triple_vectors(t, nocc, nvir, 3, 2, 1)   # i=3, j=2, k=1 (1-based)
    i0, j0, k0 = 2, 1, 0
    assert v["t1v1"][0] == i0 and v["t1v2"][0] == k0
    assert v["dintc1"][0] == j0 * nocc + k0 and v["dintc2"][0] == j0 * nocc + i0
    assert v["dintx1"][0] == 100 + j0 * nocc + k0 and v["dintx2"][0] == 100 + j0 * nocc + i0
    assert all(len(x) == nvir for x in v.values())
    fixed = {name: np.full(nvir, float(q)) for q, name in enumerate(cc.VECTOR_NAMES)}
    w = cc.triple_vectors(fixed, nocc, nvir, 3, 2, 1)
    assert all(np.array_equal(w[n], fixed[n]) for n in cc.VECTOR_NAMES)
    a = cc.gather_triple(dict(cc.make_tensor_inputs(nocc, nvir, 3, driver=True)), nocc, nvir, 3, 2, 1)
    assert len(a["t1v1"]) == nvir * nvir and not np.any(a["t1v1"][nvir:])


def test_tensor_job_without_device(cc):
    if _has_gpu(cc):
        pytest.skip("a GPU is present")
    t = cc.make_tensor_inputs(3, 8, 1)
    with pytest.raises(cc.CcsdtError):
        cc.Job.from_tensors(t, 3, 8)


def test_triple_lists(cc):
    assert len(cc.triple_list(8, cc.blocks_uniform(8, 8), "ijk")) == 120
    assert len(cc.triple_list(8, cc.blocks_uniform(8, 8), "ref")) == 288
    t1 = cc.triple_list(30, cc.blocks_uniform(30, 16), "ijk")
    assert len(t1) == 4960
    assert cc.blocks_uniform(30, 16) == [(1, 16), (17, 30)]
    i, j, k, klo = t1[:, 1], t1[:, 2], t1[:, 3], t1[:, 4]
    assert np.all(i >= j) and np.all(j >= k) and np.all(k >= klo) and np.all(k - klo < 16)
    assert np.all(t1[:, 0] == 31)
    # every (i >= j >= k) triple exactly once
    assert len({tuple(r) for r in t1[:, 1:4]}) == 4960
    bl = cc.blocks_uneven(40, cc.config_seed(3))
    assert bl[0][0] == 1 and bl[-1][1] == 40
    assert all(b[1] + 1 == c[0] for b, c in zip(bl, bl[1:]))
    assert len(cc.triple_list(40, bl, "ijk")) == 11480


def test_inputs_shapes_and_ranges(cc):
    nocc, nvir, kc = 8, 40, 8
    a = cc.make_inputs(nocc, nvir, kc, cc.config_seed(0))
    vv, ov = nvir * nvir, nocc * nvir
    assert a["tkj"].size == kc * vv and a["tia"].size == ov * nocc and a["kkj"].size == kc * ov
    e = a["eorb"]
    assert np.all(np.diff(e[:nocc]) >= 0) and np.all(np.diff(e[nocc:]) >= 0)
    assert e[:nocc].min() >= -1.5 and e[:nocc].max() < -0.3
    assert e[nocc:].min() >= 0.1 and e[nocc:].max() < 4.0
    for n in ("tij", "jia", "kkj", "t1v2"):
        assert np.abs(a[n]).max() < 0.05
    ill = cc.make_inputs(nocc, nvir, kc, cc.config_seed(2), illcond=True)
    assert np.abs(ill["jia"]).min() < 1e-7 and np.abs(ill["tij"]).max() < 0.05


def test_struct_layouts_match_header(cc, tmp_path):
    """Every ctypes mirror of an include/ccsdt.h struct has the C size and field
    offsets (a field added on one side only would shift the rest)."""
    pairs = [("ccsdt_job_desc", cc.JobDesc), ("ccsdt_tensor_desc", cc.TensorDesc),
             ("ccsdt_run_opts", cc.RunOpts), ("ccsdt_tune_entry", cc.TuneEntry),
             ("ccsdt_run_result", cc.RunResult), ("ccsdt_server_opts", cc.ServerOpts),
             ("ccsdt_job_result", cc.JobResult)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ccsdt.h"', "int main(void) {"]
    for cname, py in pairs:
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                         text=True).stdout.splitlines())
    for cname, py in pairs:
        assert int(got["%s size" % cname]) == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert int(got["%s %s" % (cname, f[0])]) == getattr(py, f[0]).offset, (cname, f[0])


def _host_copy_check(cc, rng, sizes):
    L = cc.lib()
    src = [rng.integers(0, 256, n, dtype=np.uint8) for n in sizes]
    dst = [np.zeros(n, np.uint8) for n in sizes]
    vp = ctypes.c_void_p
    D = (vp * len(sizes))(*[d.ctypes.data for d in dst])
    S = (vp * len(sizes))(*[s.ctypes.data for s in src])
    B = (ctypes.c_int64 * len(sizes))(*sizes)
    nthreads = L.ccsdt_debug_host_copy(D, S, B, len(sizes))
    assert nthreads >= 0
    for d, s in zip(dst, src):
        assert np.array_equal(d, s)
    return nthreads


def test_host_copy_pool(cc):
    """The drop-ins' host copy pool (no device): every byte lands, for lists of
    many odd-sized buffers above and below the parallel threshold, over many
    calls in a row (a chunk list reused while a worker still read it would show
    as a wrong byte here)."""
    rng = np.random.default_rng(7)
    for it in range(40):
        n = int(rng.integers(1, 22))
        sizes = [int(x) for x in rng.integers(0, 1 << int(rng.integers(4, 22)), n)]
        _host_copy_check(cc, rng, sizes)
    # the drop-in's shape at nocc=30, nvir=300: 21 arrays, 4.9 MB in all
    sizes = [8 * x for x in (330, 90000, 90000, 9000, 9000, 9000, 9000, 90000, 90000, 90000, 90000,
                             9000, 9000, 9000, 9000, 300, 300, 300, 300, 300, 300)]
    assert _host_copy_check(cc, rng, sizes) == 4  # default worker count
    L = cc.lib()
    assert L.ccsdt_debug_host_copy(None, None, None, 1) == -1
    neg = (ctypes.c_int64 * 1)(-1)
    one = (ctypes.c_void_p * 1)(1)
    assert L.ccsdt_debug_host_copy(one, one, neg, 1) == -1


def test_host_copy_pool_thread_count(cc):
    """CCSDT_COPY_THREADS sets the worker count (0: the caller copies alone)."""
    import subprocess
    import sys
    code = ("import ctypes, numpy as np, sys; sys.path.insert(0, %r); import __graft_entry__ as g; "
            "cc = g.package(); L = cc.lib(); a = np.arange(1 << 19, dtype=np.float64); b = np.zeros_like(a); "
            "vp = ctypes.c_void_p; n = L.ccsdt_debug_host_copy((vp*1)(b.ctypes.data), (vp*1)(a.ctypes.data), "
            "(ctypes.c_int64*1)(a.nbytes), 1); assert np.array_equal(a, b); print(n)" % ROOT)
    for env, want in (("0", 0), ("3", 3)):
        out = subprocess.run([sys.executable, "-c", code], check=True, capture_output=True, text=True,
                             env=dict(os.environ, CCSDT_COPY_THREADS=env)).stdout.split()
        assert int(out[-1]) == want


def test_server_options_validated_before_device(cc):
    """Server options are checked on the host before any device work: occupied
    digit planes need the INT8 form at 3 slices (CCSDT_E_ARG otherwise); valid
    options get past that check (then, without a GPU, fail as a device error)."""
    L = cc.lib()
    h = ctypes.c_void_p()
    for mode, slices, occ in (("ozaki_int8", 4, 1), ("ozaki_int8", 3, 2), ("ozaki_int8", 9, 0)):
        o = cc.ServerOpts(0, 0, cc.MODES[mode], slices, 0, occ)
        assert L.ccsdt_server_create(ctypes.byref(o), ctypes.byref(h)) == -1, (mode, slices, occ)
    if not _has_gpu(cc):
        for mode, slices, occ in (("ozaki_int8", 3, 1), ("fp64", 0, 1)):
            o = cc.ServerOpts(0, 0, cc.MODES[mode], slices, 0, occ)
            assert L.ccsdt_server_create(ctypes.byref(o), ctypes.byref(h)) == -2, (mode, slices, occ)


def test_tune_cache_file_round_trip(cc, tmp_path):
    """The tuning cache's file: well-formed lines are read (later lines win),
    malformed ones skipped; clearing forgets the in-memory picks only.  Needs no
    device."""
    path = tmp_path / "tune.txt"
    good = ("v1 mode=2 slices=0 occ=0 graph=1 kind=0 ncor=0 nocc=30 nvir=300 kchunk=16 n=4960 "
            "hash=0123456789abcdef dev=gfx950/256 -> mode=2 slices=3 flags=2 variant=0 "
            "wg_per_cu=0 order=1 occ=1 ms=10.900000\n")
    other = good.replace("nvir=300", "nvir=400")
    path.write_text(good + "this is not a pick\n" + other + good.replace("ms=10.9", "ms=10.7"))
    try:
        cc.tune_cache_clear()
        assert cc.tune_cache_file(str(path)) == 3
        assert cc.tune_cache_size() == 2  # the repeated key counts once
        cc.tune_cache_clear()
        assert cc.tune_cache_size() == 0
        assert path.read_text().count("\n") == 4  # the file is kept
        assert cc.tune_cache_file(str(tmp_path / "absent.txt")) == 0
    finally:
        cc.tune_cache_file(None)
        cc.tune_cache_clear()


def test_tune_cache_server_picks(cc, tmp_path):
    """Server picks (tile and tick size per mode and slice policy) share the
    cache file: a "v1 srv" line needs its items field and a job line may not
    carry one.  Storing a pick checks the options on the host first: the tile
    and tick size must be set.  Needs no device."""
    path = tmp_path / "tune.txt"
    srv = "v1 srv mode=2 slices=-1 occ=0 dev=gfx950/256 -> mode=2 slices=-1 flags=0 variant=0 " \
          "wg_per_cu=0 order=0 occ=0 ms=47.100000 items=131072\n"
    job = ("v1 mode=0 slices=0 occ=0 graph=0 kind=0 ncor=0 nocc=8 nvir=40 kchunk=8 n=120 "
           "hash=0123456789abcdef dev=gfx950/256 -> mode=0 slices=0 flags=0 variant=1 "
           "wg_per_cu=0 order=1 occ=0 ms=1.000000\n")
    bad = [srv.replace(" items=131072", ""), job.replace("ms=1.000000", "ms=1.000000 items=4096"),
           srv.replace("items=131072", "items=0")]
    path.write_text(srv + job + "".join(bad))
    try:
        cc.tune_cache_clear()
        assert cc.tune_cache_file(str(path)) == 2
        assert cc.tune_cache_size() == 2
    finally:
        cc.tune_cache_file(None)
        cc.tune_cache_clear()
    L = cc.lib()
    for variant, items in ((cc.TILE_AUTO, 65536), (0, 0), (7, 65536)):
        o = cc.ServerOpts(0, items, cc.MODES["fp64"], 0, variant, 0)
        assert L.ccsdt_server_tune_store(ctypes.byref(o), 1.0) == -1, (variant, items)
    assert cc.tune_cache_size() == 0


def test_f_index_tiled_layout(cc):
    """The tiled f layout of the separate energy pass holds every element of a
    triple's eight nv x nv matrices exactly once in 8 nv^2 places (no padding);
    each 32x32 tile pair (the tile at block row bi, block column ci <= bi and its
    mirror at (ci, bi), of all eight matrices; the last block row and column
    narrower) is one contiguous run, its tiles column-major; and the
    column-major layout is the reference's m nv^2 + b + c nv."""
    L = cc.lib()
    for nv in (1, 5, 31, 32, 33, 64, 65, 300):
        idx = np.array([[[L.ccsdt_debug_f_index(m, b, c, nv, 1) for c in range(nv)] for b in range(nv)]
                        for m in range(8)])
        assert np.array_equal(np.sort(idx.ravel()), np.arange(8 * nv * nv)), nv
        cm = np.array([[[L.ccsdt_debug_f_index(m, b, c, nv, 0) for c in range(nv)] for b in range(nv)]
                       for m in range(8)])
        want = (np.arange(8)[:, None, None] * nv * nv + np.arange(nv)[None, :, None]
                + nv * np.arange(nv)[None, None, :])
        assert np.array_equal(cm, want), nv
        start = 0
        for bi in range(0, nv, 32):
            for ci in range(0, bi + 1, 32):
                D = idx[:, bi:bi + 32, ci:ci + 32]
                M = idx[:, ci:ci + 32, bi:bi + 32]
                cells = np.concatenate([D.ravel(), M.ravel()]) if bi != ci else D.ravel()
                # the pair's run follows the previous pair's
                assert cells.min() == start and cells.max() == start + cells.size - 1, (nv, bi, ci)
                start += cells.size
                for T in ([D] + ([M] if bi != ci else [])):
                    for m in range(8):
                        t = T[m]
                        rows, cols = t.shape
                        assert np.array_equal(t, t[0, 0] + np.arange(rows)[:, None]
                                              + rows * np.arange(cols)[None, :])
        assert start == 8 * nv * nv
    assert L.ccsdt_debug_f_index(0, 300, 0, 300, 1) == -1 and L.ccsdt_debug_f_index(8, 0, 0, 300, 0) == -1
