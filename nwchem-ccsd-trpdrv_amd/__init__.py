"""Host-side interface to libccsdt.so, the MI355X (gfx950) CCSD(T) triples body.

The library exports the reference's C entry points (``ccsd_trpdrv_omp_cbody_``,
``ccsd_tengy_omp`` and their aliases, see include/ccsdt.h) plus a job API with
explicit device residency.  This module binds both with ctypes and adds the
synthetic inputs and triple lists of the benchmark harness (omp/test.c).

There is no CPU fallback: every compute call goes to the HIP library, and a
missing library or device raises ``CcsdtError``.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CCSDT_LIB: another build of the library (the sanitizer builds of sanitize.py)
LIB_PATH = os.environ.get("CCSDT_LIB") or os.path.join(HERE, "libccsdt.so")

# Job array order = the body's argument order (omp/ccsd_trpdrv_omp_cbody.c:31-44).
ARRAY_NAMES = ["eorb", "tij", "tkj", "tia", "tka", "xia", "xka", "jia", "jka", "kia",
               "kka", "jij", "jkj", "kij", "kkj", "dintc1", "dintx1", "t1v1", "dintc2",
               "dintx2", "t1v2"]
# PCG stream of each synthetic array: its 1-based position in the body's
# argument list; Dja/Djka/Djia (allocated by omp/test.c:137-139, never passed)
# follow at 40-42.
STREAM = {"eorb": 9, "tij": 20, "tkj": 21, "tia": 22, "tka": 23, "xia": 24, "xka": 25,
          "jia": 26, "jka": 27, "kia": 28, "kka": 29, "jij": 30, "jkj": 31, "kij": 32,
          "kkj": 33, "dintc1": 34, "dintx1": 35, "t1v1": 36, "dintc2": 37, "dintx2": 38,
          "t1v2": 39, "dja": 40, "djka": 41, "djia": 42}
ILLCOND_STREAM_BASE = 100   # exponent draws of the ill-conditioned variant
JK_NAMES = ["jia", "jka", "kia", "kka", "jij", "jkj", "kij", "kkj"]

# Full-tensor jobs (ccsdt_job_create_tensors): every occupied-indexed slice of
# the body's operands.  Each family draws from the stream of the body argument
# it generalises.
TENSOR_NAMES = ["t2", "jv", "kv", "t1o", "xo", "jo", "ko"]
TENSOR_STREAM = {"t2": STREAM["tij"], "jv": STREAM["jia"], "kv": STREAM["kia"],
                 "t1o": STREAM["tia"], "xo": STREAM["xia"], "jo": STREAM["jij"],
                 "ko": STREAM["kij"]}
VECTOR_NAMES = ["dintc1", "dintx1", "t1v1", "dintc2", "dintx2", "t1v2"]
# Driver tensors of a full-tensor job (optional): t1 [m][c], dja and djx
# [j][m][c]; the six vectors of each triple are rows of them (ccsdt_tensor_desc).
DRIVER_NAMES = ["t1", "dja", "djx"]
DRIVER_STREAM = {"t1": STREAM["t1v1"], "dja": STREAM["dja"], "djx": STREAM["djia"]}

MODES = {"fp64": 0, "fp64_split": 1, "ozaki_int8": 2, "bf16x3": 3, "fp32": 4}
# Largest Ozaki INT8 slice count whose K = nocc half may also run as INT8 digit
# planes with row scales of its own (run option occ_digits; kTail8MaxSlices in the
# library); otherwise that half is an FP64 tail
TAIL8_MAX_SLICES = 3
MODE_NAMES = {v: k for k, v in MODES.items()}
TILE_ONE_WAVE, TILE_TWO_WAVE, TILE_ROW_PAIR = 0, 1, 2
TILE_AUTO = -1  # servers: the tuning cache's pick, else 0
ORDER_TASK, ORDER_PAIR = 0, 1
RUN_PERSISTENT, RUN_GRAPH = 1, 2
SLICES_PER_JOB = -1  # server option: each job's split form chosen from its size
E_CODES = {0: "OK", -1: "E_ARG", -2: "E_DEVICE", -3: "E_NOMEM", -4: "E_BUDGET",
           -5: "E_STATE", -6: "E_INJECTED"}


class CcsdtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (E_CODES.get(code, "?"), code, msg))
        self.code = code


class JobDesc(ctypes.Structure):
    _fields_ = [("ncor", ctypes.c_int), ("nocc", ctypes.c_int), ("nvir", ctypes.c_int),
                ("kchunk", ctypes.c_int), ("arrays", ctypes.c_void_p * len(ARRAY_NAMES))]


class TensorDesc(ctypes.Structure):
    _fields_ = ([("ncor", ctypes.c_int), ("nocc", ctypes.c_int), ("nvir", ctypes.c_int),
                 ("pad", ctypes.c_int), ("eorb", ctypes.c_void_p)]
                + [(n, ctypes.c_void_p) for n in TENSOR_NAMES]
                + [(n, ctypes.c_void_p) for n in VECTOR_NAMES]
                + [(n, ctypes.c_void_p) for n in DRIVER_NAMES])


def has_driver_tensors(t):
    return any(t.get(name) is not None for name in DRIVER_NAMES)


def _tensor_desc(t, nocc, nvir, ncor):
    """TensorDesc over contiguous float64 copies of t (returned to keep alive).
    With driver tensors the six vectors are left null (they are not read)."""
    d = TensorDesc(ncor, nocc, nvir, 0)
    keep = []
    names = ["eorb"] + TENSOR_NAMES
    names += DRIVER_NAMES if has_driver_tensors(t) else VECTOR_NAMES
    for name in names:
        if t.get(name) is None:
            continue
        a = np.ascontiguousarray(t[name], dtype=np.float64)
        keep.append(a)
        setattr(d, name, a.ctypes.data)
    return d, keep


class RunOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("slices", ctypes.c_int), ("flags", ctypes.c_int),
                ("variant", ctypes.c_int), ("wg_per_cu", ctypes.c_int), ("order", ctypes.c_int),
                ("occ_digits", ctypes.c_int), ("reserved", ctypes.c_int)]

    def as_dict(self):
        return {"mode": MODE_NAMES[self.mode], "slices": self.slices,
                "persistent": bool(self.flags & RUN_PERSISTENT),
                "graph": bool(self.flags & RUN_GRAPH), "variant": self.variant,
                "wg_per_cu": self.wg_per_cu, "order": self.order, "occ_digits": self.occ_digits}


class TuneEntry(ctypes.Structure):
    _fields_ = [("opts", RunOpts), ("ms", ctypes.c_double), ("status", ctypes.c_int),
                ("pad", ctypes.c_int)]


class RunResult(ctypes.Structure):
    _fields_ = [("emp4", ctypes.c_double), ("emp5", ctypes.c_double),
                ("ms_total", ctypes.c_double), ("ms_pack", ctypes.c_double),
                ("ms_main", ctypes.c_double), ("ms_tengy", ctypes.c_double),
                ("ms_reduce", ctypes.c_double), ("ntriples", ctypes.c_int64),
                ("launches", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerOpts(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_int64), ("items_per_tick", ctypes.c_int),
                ("mode", ctypes.c_int), ("slices", ctypes.c_int), ("variant", ctypes.c_int),
                ("occ_digits", ctypes.c_int)]


class JobResult(ctypes.Structure):
    _fields_ = [("emp4", ctypes.c_double), ("emp5", ctypes.c_double),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_finish", ctypes.c_double), ("ntriples", ctypes.c_int64),
                ("footprint", ctypes.c_int64), ("status", ctypes.c_int), ("slices", ctypes.c_int),
                ("occ_digits", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    """Loads libccsdt.so (built in-tree by build.py); raises if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise CcsdtError(-2, "libccsdt.so not built: run nwchem-ccsd-trpdrv_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        sig = {
            "ccsdt_last_error": (ctypes.c_int, []),
            "ccsdt_last_error_string": (ctypes.c_char_p, []),
            "ccsdt_clear_error": (None, []),
            "ccsdt_set_error_policy": (None, [ctypes.c_int]),
            "ccsdt_debug_inject_failure": (None, [ctypes.c_int]),
            "ccsdt_version": (ctypes.c_char_p, []),
            "ccsdt_init": (ctypes.c_int, [ctypes.c_int]),
            "ccsdt_set_f_writeback": (None, [ctypes.c_int]),
            "ccsdt_pcg_raw": (None, [ctypes.c_uint64, ctypes.c_uint64, vp, i64]),
            "ccsdt_pcg_uniform": (None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                         ctypes.c_double, vp, i64]),
            "ccsdt_job_create": (ctypes.c_int, [ctypes.POINTER(JobDesc), ctypes.POINTER(vp)]),
            "ccsdt_job_run": (ctypes.c_int, [vp, vp, i64, ctypes.POINTER(RunOpts),
                                             ctypes.POINTER(RunResult), vp, vp]),
            "ccsdt_job_destroy": (ctypes.c_int, [vp]),
            "ccsdt_job_autotune": (ctypes.c_int, [vp, vp, i64, ctypes.POINTER(RunOpts), ctypes.c_int,
                                                  ctypes.POINTER(RunOpts), ctypes.POINTER(TuneEntry),
                                                  ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
            "ccsdt_job_footprint": (i64, [ctypes.c_int] * 4 + [i64]),
            "ccsdt_tune_cache_file": (ctypes.c_int, [ctypes.c_char_p]),
            "ccsdt_tune_cache_clear": (None, []),
            "ccsdt_tune_cache_size": (ctypes.c_int, []),
            "ccsdt_job_create_tensors": (ctypes.c_int, [ctypes.POINTER(TensorDesc), ctypes.POINTER(vp)]),
            "ccsdt_tensor_job_footprint": (i64, [ctypes.c_int] * 3 + [i64, ctypes.c_int]),
            "ccsdt_server_submit_tensors": (ctypes.c_int, [vp, ctypes.POINTER(TensorDesc), vp, i64,
                                                           ctypes.POINTER(i64)]),
            "ccsdt_selftest_mfma": (ctypes.c_int, [vp, vp, vp]),
            "ccsdt_selftest_mfma_i8": (ctypes.c_int, [vp, vp, vp, ctypes.c_int]),
            "ccsdt_debug_copy": (ctypes.c_double, [i64, ctypes.c_int]),
            "ccsdt_debug_read": (ctypes.c_double, [i64, ctypes.c_int]),
            "ccsdt_debug_mfma_rate": (ctypes.c_double, [ctypes.c_int, ctypes.c_int]),
            "ccsdt_debug_load_rate": (ctypes.c_double, [ctypes.c_int, ctypes.c_int]),
            "ccsdt_debug_host_copy": (ctypes.c_int, [vp, vp, vp, ctypes.c_int]),
            "ccsdt_debug_occ_fold": (None, [ctypes.c_int]),
            "ccsdt_debug_f_index": (i64, [ctypes.c_int] * 5),
            "ccsdt_server_create": (ctypes.c_int, [ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)]),
            "ccsdt_server_submit": (ctypes.c_int, [vp, ctypes.POINTER(JobDesc), vp, i64,
                                                   ctypes.POINTER(i64)]),
            "ccsdt_server_progress": (ctypes.c_int, [vp, ctypes.c_int]),
            "ccsdt_server_result": (ctypes.c_int, [vp, i64, ctypes.POINTER(JobResult)]),
            "ccsdt_server_counts": (ctypes.c_int, [vp] + [ctypes.POINTER(i64)] * 4),
            "ccsdt_server_clock": (ctypes.c_double, [vp]),
            "ccsdt_server_options": (ctypes.c_int, [vp, ctypes.POINTER(ServerOpts)]),
            "ccsdt_server_tune_store": (ctypes.c_int, [ctypes.POINTER(ServerOpts), ctypes.c_double]),
            "ccsdt_server_destroy": (ctypes.c_int, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        body_args = [_dp] * 9 + [_ip] * 3 + [_dp] * 2 + [_ip] * 5 + [_dp] * 20
        for name in ("ccsd_trpdrv_omp_cbody_", "ccsd_trpdrv_cbody_", "ccsd_trpdrv_omp_fbody_"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = body_args
        tengy_args = [_dp] * 15 + [ctypes.c_double] + [_dp] * 4 + [ctypes.c_int] * 3
        for name in ("ccsd_tengy_omp", "ccsd_tengy"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = tengy_args
        L.ccsd_tengy_omp_.restype = None
        L.ccsd_tengy_omp_.argtypes = [_dp] * 20 + [_ip] * 3
        _LIB = L
    return _LIB


def last_error():
    L = lib()
    return L.ccsdt_last_error(), L.ccsdt_last_error_string().decode()


def _check(rc):
    if rc != 0:
        raise CcsdtError(rc, lib().ccsdt_last_error_string().decode())


def init(device=0):
    _check(lib().ccsdt_init(device))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dptr(a):
    return a.ctypes.data_as(_dp)


# ------------------------------------------------------------------ inputs
def pcg_raw(seed, stream, n):
    out = np.empty(n, dtype=np.uint32)
    lib().ccsdt_pcg_raw(seed, stream, _ptr(out), n)
    return out


def pcg_uniform(seed, stream, lo, hi, n):
    out = np.empty(n, dtype=np.float64)
    lib().ccsdt_pcg_uniform(seed, stream, lo, hi, _ptr(out), n)
    return out


def config_seed(cfg):
    """Seed of the synthetic inputs of a workload configuration."""
    return 0xC0DE0000 + int(cfg)


def make_inputs(nocc, nvir, kchunk, seed, ncor=0, illcond=False):
    """Seeded arrays at the harness's shapes (omp/test.c:112-146).

    eorb: ncor core energies U[-20,-10), nocc occupied U[-1.5,-0.3) and nvir
    virtual U[0.1,4.0), each segment sorted ascending.  Every T/X/J/K/D/dint/t1v
    array is U(-0.05,0.05).  ``illcond`` scales the J and K arrays elementwise by
    10**u, u ~ U(-6,0), drawn from stream 100 + the array's stream.
    """
    vv, ov = nvir * nvir, nocc * nvir
    nbf = ncor + nocc + nvir
    u = pcg_uniform(seed, STREAM["eorb"], 0.0, 1.0, nbf)
    eorb = np.empty(nbf)
    eorb[:ncor] = np.sort(-20.0 + 10.0 * u[:ncor])
    eorb[ncor:ncor + nocc] = np.sort(-1.5 + 1.2 * u[ncor:ncor + nocc])
    eorb[ncor + nocc:] = np.sort(0.1 + 3.9 * u[ncor + nocc:])
    sizes = {"tij": vv, "tkj": kchunk * vv, "tia": ov * nocc, "tka": kchunk * ov,
             "xia": ov * nocc, "xka": kchunk * ov, "jia": vv, "jka": kchunk * vv, "kia": vv,
             "kka": kchunk * vv, "jij": ov * nocc, "jkj": kchunk * ov, "kij": ov * nocc,
             "kkj": kchunk * ov, "dja": ov, "djka": nvir * kchunk, "djia": nvir * nocc}
    for name in ("dintc1", "dintx1", "t1v1", "dintc2", "dintx2", "t1v2"):
        sizes[name] = vv
    arrays = {"eorb": eorb}
    for name, n in sizes.items():
        arrays[name] = pcg_uniform(seed, STREAM[name], -0.05, 0.05, n)
    if illcond:
        for name in JK_NAMES:
            e = pcg_uniform(seed, ILLCOND_STREAM_BASE + STREAM[name], -6.0, 0.0, arrays[name].size)
            arrays[name] = arrays[name] * np.power(10.0, e)
    return arrays


def make_tensor_inputs(nocc, nvir, seed, ncor=0, illcond=False, driver=False):
    """Seeded full-tensor job inputs (see ccsdt_tensor_desc): eorb as in
    make_inputs; t2 (nocc*nocc*nvir^2), jv, kv (nocc*nvir^2), t1o, xo
    (nocc*nocc*nvir), jo, ko (nocc*nocc*nocc*nvir) and the six nvir vectors, all
    U(-0.05,0.05).  ``illcond`` scales jv, kv, jo, ko by 10**u, u ~ U(-6,0).
    ``driver`` replaces the six vectors by the driver tensors t1 (nocc*nvir),
    dja and djx (nocc*nocc*nvir), U(-0.05,0.05)."""
    vv, ov = nvir * nvir, nocc * nvir
    base = make_inputs(nocc, nvir, 1, seed, ncor=ncor)
    sizes = {"t2": nocc * nocc * vv, "jv": nocc * vv, "kv": nocc * vv, "t1o": nocc * ov,
             "xo": nocc * ov, "jo": nocc * nocc * ov, "ko": nocc * nocc * ov}
    t = {"eorb": base["eorb"]}
    for name, n in sizes.items():
        t[name] = pcg_uniform(seed, TENSOR_STREAM[name], -0.05, 0.05, n)
        if illcond and name in ("jv", "kv", "jo", "ko"):
            e = pcg_uniform(seed, ILLCOND_STREAM_BASE + TENSOR_STREAM[name], -6.0, 0.0, n)
            t[name] = t[name] * np.power(10.0, e)
    if driver:
        sizes = {"t1": nocc * nvir, "dja": nocc * nocc * nvir, "djx": nocc * nocc * nvir}
        for name, n in sizes.items():
            t[name] = pcg_uniform(seed, DRIVER_STREAM[name], -0.05, 0.05, n)
        return t
    for name in VECTOR_NAMES:
        t[name] = base[name][:nvir].copy()
    return t


def triple_vectors(t, nocc, nvir, i, j, k):
    """The six nvir vectors triple (i,j,k) (1-based) reads: the job's fixed
    ones, or rows of the driver tensors (ccsdt_tensor_desc):
    t1v1 = t1[i], t1v2 = t1[k], dintc1 = dja[j][k], dintc2 = dja[j][i],
    dintx1 = djx[j][k], dintx2 = djx[j][i] (0-based rows)."""
    if not has_driver_tensors(t):
        return {name: np.asarray(t[name])[:nvir] for name in VECTOR_NAMES}
    i0, j0, k0 = i - 1, j - 1, k - 1

    def row(name, r):
        return np.asarray(t[name])[r * nvir:(r + 1) * nvir]
    return {"t1v1": row("t1", i0), "t1v2": row("t1", k0),
            "dintc1": row("dja", j0 * nocc + k0), "dintc2": row("dja", j0 * nocc + i0),
            "dintx1": row("djx", j0 * nocc + k0), "dintx2": row("djx", j0 * nocc + i0)}


def gather_triple(t, nocc, nvir, i, j, k):
    """Harness-shaped arrays (kchunk 1, klo = k) holding the slices triple
    (i,j,k) (1-based) reads from a full-tensor job -- the host gather the
    tensor job makes unnecessary, used to check it against the oracle."""
    vv, ov = nvir * nvir, nocc * nvir
    i0, j0, k0 = i - 1, j - 1, k - 1

    def sl(name, idx, n):
        return t[name][idx * n:(idx + 1) * n]
    a = {"eorb": t["eorb"],
         "tij": sl("t2", j0 * nocc + i0, vv), "tkj": sl("t2", j0 * nocc + k0, vv),
         "jia": sl("jv", i0, vv), "jka": sl("jv", k0, vv),
         "kia": sl("kv", i0, vv), "kka": sl("kv", k0, vv),
         "tia": sl("t1o", i0, ov), "tka": sl("t1o", k0, ov),
         "xia": sl("xo", i0, ov), "xka": sl("xo", k0, ov),
         "jij": sl("jo", j0 * nocc + i0, ov), "jkj": sl("jo", j0 * nocc + k0, ov),
         "kij": sl("ko", j0 * nocc + i0, ov), "kkj": sl("ko", j0 * nocc + k0, ov)}
    # at the harness allocation sizes (omp/test.c:125-146), of which the body reads
    # the first nocc*nvir (or nvir) elements
    for name in ("tia", "xia", "jij", "kij"):
        a[name] = np.concatenate([a[name], np.zeros(ov * (nocc - 1))])
    for name, v in triple_vectors(t, nocc, nvir, i, j, k).items():
        a[name] = np.concatenate([v, np.zeros(vv - nvir)])
    return a


def tensor_footprint(nocc, nvir, ntriples, ncor=0, driver=False):
    """Device bytes of a full-tensor job (ccsdt_tensor_job_footprint)."""
    return lib().ccsdt_tensor_job_footprint(ncor, nocc, nvir, ntriples, 1 if driver else 0)


def make_harness_inputs(nocc, nvir, nkpass=1, ncor=0):
    """The reference harness's own inputs (omp/test.c:55-62, 94-146): every array
    filled by make_array, a[i] = 1/(100+i), at the harness's allocation sizes,
    with kchunk = (nocc-1)/nkpass + 1 (omp/test.c:97).  Returns (arrays, kchunk)."""
    kchunk = (nocc - 1) // nkpass + 1
    vv, ov = nvir * nvir, nocc * nvir
    sizes = {"eorb": ncor + nocc + nvir, "tij": vv, "tkj": kchunk * vv, "tia": ov * nocc,
             "tka": kchunk * ov, "xia": ov * nocc, "xka": kchunk * ov, "jia": vv,
             "jka": kchunk * vv, "kia": vv, "kka": kchunk * vv, "jij": ov * nocc,
             "jkj": kchunk * ov, "kij": ov * nocc, "kkj": kchunk * ov}
    for name in VECTOR_NAMES:
        sizes[name] = vv
    return {name: 1.0 / (100.0 + np.arange(n)) for name, n in sizes.items()}, kchunk


def harness_triples(nocc, kchunk, maxiter=100):
    """The calls of omp/test.c:160-182 in order: k blocks of kchunk, j and i over
    all occupied, k from klo to min(khi, i), a = 1, stopping after maxiter calls
    (maxiter < 0: all)."""
    trip = triple_list(nocc, blocks_uniform(nocc, kchunk), "ref", a=1)
    return trip if maxiter < 0 else trip[:maxiter]


def blocks_uniform(nocc, length):
    """k blocks klo..khi of a fixed length (kchunk, omp/test.c:97,160-161)."""
    length = max(1, min(int(length), nocc))
    return [(klo, min(nocc, klo + length - 1)) for klo in range(1, nocc + 1, length)]


def blocks_uneven(nocc, seed, lo=8, hi=48):
    """k blocks of lengths drawn uniformly from {lo..hi} until nocc is covered."""
    draws = pcg_raw(seed, 77, 4 * nocc + 8)
    blocks, klo, q = [], 1, 0
    while klo <= nocc:
        length = lo + int(draws[q] % (hi - lo + 1))
        q += 1
        blocks.append((klo, min(nocc, klo + length - 1)))
        klo += length
    return blocks


def triple_list(nocc, blocks, order="ijk", a=None, ncor=0):
    """(a,i,j,k,klo) quintuples, 1-based, as int32 array of shape (n,5).

    order "ijk": i >= j >= k within each k block; "ref": the loop of
    omp/test.c:160-165 (j and i over all occupied, k <= i).  ``a`` defaults to
    the first virtual orbital ncor+nocc+1 (a virtual a keeps the denominators
    away from zero, unlike omp/test.c:162's a=1).
    """
    if a is None:
        a = ncor + nocc + 1
    out = []
    for klo, khi in blocks:
        if order == "ijk":
            for i in range(1, nocc + 1):
                for j in range(1, i + 1):
                    for k in range(klo, min(khi, j) + 1):
                        out.append((a, i, j, k, klo))
        elif order == "ref":
            for j in range(1, nocc + 1):
                for i in range(1, nocc + 1):
                    for k in range(klo, min(khi, i) + 1):
                        out.append((a, i, j, k, klo))
        else:
            raise ValueError("order must be 'ijk' or 'ref'")
    return np.ascontiguousarray(np.array(out, dtype=np.int32).reshape(-1, 5))


# Workload configurations (sizes and block scheme of each measured case).
CONFIGS = {
    0: dict(name="fp64-cpu-case", nocc=8, nvir=40, block=8),
    1: dict(name="fp64-production", nocc=30, nvir=300, block=16),
    2: dict(name="split-production", nocc=30, nvir=300, block=16),
    3: dict(name="batched-uneven", nocc=40, nvir=400, block="uneven"),
}


def config_case(cfg, illcond=False, order="ijk"):
    """(inputs, blocks, triples, kchunk) of a configuration."""
    c = CONFIGS[cfg]
    seed = config_seed(cfg)
    if c["block"] == "uneven":
        blocks = blocks_uneven(c["nocc"], seed)
    else:
        blocks = blocks_uniform(c["nocc"], c["block"])
    kchunk = max(khi - klo + 1 for klo, khi in blocks)
    arrays = make_inputs(c["nocc"], c["nvir"], kchunk, seed, illcond=illcond)
    return arrays, blocks, triple_list(c["nocc"], blocks, order), kchunk


def gemm_flops_per_triple(nocc, nvir):
    """16*nvir^2*(nvir+nocc): the 16 GEMMs at 2 flops per multiply-add."""
    return 16.0 * nvir * nvir * (nvir + nocc)


def tengy_bytes_per_triple(nvir):
    """64*nvir^2: the eight f matrices read once by the energy pass."""
    return 64.0 * nvir * nvir


# ------------------------------------------------------------------ job API
class Job:
    """Device-resident arrays of one harness-shaped job (see ccsdt_job_create)."""

    def __init__(self, arrays, nocc, nvir, kchunk, ncor=0):
        self.ncor, self.nocc, self.nvir, self.kchunk = ncor, nocc, nvir, kchunk
        d = JobDesc(ncor, nocc, nvir, kchunk)
        self._keep = []
        for q, name in enumerate(ARRAY_NAMES):
            a = np.ascontiguousarray(arrays[name], dtype=np.float64)
            self._keep.append(a)
            d.arrays[q] = a.ctypes.data
        h = ctypes.c_void_p()
        _check(lib().ccsdt_job_create(ctypes.byref(d), ctypes.byref(h)))
        self._keep = None
        self.handle = h

    @classmethod
    def from_tensors(cls, tensors, nocc, nvir, ncor=0):
        """A full-tensor job (ccsdt_job_create_tensors); tensors as made by
        make_tensor_inputs.  Any triple (a,i,j,k,klo) with 1 <= klo <= k runs."""
        self = cls.__new__(cls)
        self.ncor, self.nocc, self.nvir, self.kchunk = ncor, nocc, nvir, nocc
        d, keep = _tensor_desc(tensors, nocc, nvir, ncor)
        h = ctypes.c_void_p()
        _check(lib().ccsdt_job_create_tensors(ctypes.byref(d), ctypes.byref(h)))
        del keep
        self._keep = None
        self.handle = h
        return self

    def run(self, triples, mode="fp64", slices=0, terms=False, fout=False, persistent=False,
            graph=False, variant=0, wg_per_cu=0, order=0, occ_digits=0):
        """Runs the triples; persistent / graph select CCSDT_RUN_PERSISTENT and
        CCSDT_RUN_GRAPH, variant the FP64 tile (CCSDT_TILE_*), wg_per_cu the
        persistent grid (0 = as many workgroups per CU as fit), order the work
        item order (ORDER_TASK or ORDER_PAIR), occ_digits (ozaki_int8, slices 3)
        the K = nocc half as INT8 digit planes instead of an FP64 tail."""
        trip = np.ascontiguousarray(triples, dtype=np.int32).reshape(-1, 5)
        n = trip.shape[0]
        flags = (RUN_PERSISTENT if persistent else 0) | (RUN_GRAPH if graph else 0)
        opts = RunOpts(MODES[mode], slices, flags, variant, wg_per_cu, order, int(occ_digits))
        res = RunResult()
        t = np.empty((n, 4)) if terms else None
        f = np.empty((8, self.nvir, self.nvir)) if fout else None
        _check(lib().ccsdt_job_run(self.handle, _ptr(trip), n, ctypes.byref(opts),
                                   ctypes.byref(res), _ptr(t) if terms else None,
                                   _ptr(f) if fout else None))
        out = res.as_dict()
        if terms:
            out["terms"] = t
        if fout:
            # f[m] is column major (nvir x nvir): f[m][c, b] holds element (b, c)
            out["f"] = np.ascontiguousarray(np.swapaxes(f, 1, 2))
        return out

    def autotune(self, triples, mode="fp64", slices=0, graph=True, reps=3, occ_digits=0):
        """Times every launch configuration (tile variant x schedule) on these
        triples (ccsdt_job_autotune); with ozaki_int8 and slices 0 it first picks
        the precision (slices, occ_digits).  Returns (best, table): best is a dict
        of run() keyword arguments, table one dict per candidate with its median
        ms (-1 when it failed) and status -- empty when the pick came from the
        tuning cache (tune_cache_file)."""
        trip = np.ascontiguousarray(triples, dtype=np.int32).reshape(-1, 5)
        base = RunOpts(MODES[mode], slices, RUN_GRAPH if graph else 0, 0, 0, 0, int(occ_digits))
        best = RunOpts()
        cap = 16
        table = (TuneEntry * cap)()
        count = ctypes.c_int()
        _check(lib().ccsdt_job_autotune(self.handle, _ptr(trip), trip.shape[0], ctypes.byref(base),
                                        int(reps), ctypes.byref(best), table, cap,
                                        ctypes.byref(count)))
        rows = []
        for q in range(min(count.value, cap)):
            d = table[q].opts.as_dict()
            d.update(ms=table[q].ms, status=table[q].status)
            rows.append(d)
        b = best.as_dict()
        del b["mode"]
        b["mode"] = mode
        return b, rows

    def close(self):
        if self.handle:
            lib().ccsdt_job_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server over the device (see ccsdt_server_* in ccsdt.h).
    variant=TILE_AUTO and items_per_tick=0 take the tuning cache's server pick
    for this mode and slice policy (tune_server), else tile 0 and 131072."""

    def __init__(self, budget_bytes=0, items_per_tick=0, mode="fp64", slices=0, variant=TILE_AUTO,
                 occ_digits=0):
        o = ServerOpts(int(budget_bytes), int(items_per_tick), MODES[mode], int(slices),
                       int(variant), int(occ_digits))
        h = ctypes.c_void_p()
        _check(lib().ccsdt_server_create(ctypes.byref(o), ctypes.byref(h)))
        self.handle = h

    def submit(self, arrays, nocc, nvir, kchunk, triples, ncor=0):
        d = JobDesc(ncor, nocc, nvir, kchunk)
        keep = []
        for q, name in enumerate(ARRAY_NAMES):
            a = np.ascontiguousarray(arrays[name], dtype=np.float64)
            keep.append(a)
            d.arrays[q] = a.ctypes.data
        trip = np.ascontiguousarray(triples, dtype=np.int32).reshape(-1, 5)
        jid = ctypes.c_int64()
        _check(lib().ccsdt_server_submit(self.handle, ctypes.byref(d), _ptr(trip), trip.shape[0],
                                         ctypes.byref(jid)))
        return jid.value

    def submit_tensors(self, tensors, nocc, nvir, triples, ncor=0):
        d, keep = _tensor_desc(tensors, nocc, nvir, ncor)
        trip = np.ascontiguousarray(triples, dtype=np.int32).reshape(-1, 5)
        jid = ctypes.c_int64()
        _check(lib().ccsdt_server_submit_tensors(self.handle, ctypes.byref(d), _ptr(trip),
                                                 trip.shape[0], ctypes.byref(jid)))
        del keep
        return jid.value

    def progress(self, block=False):
        rc = lib().ccsdt_server_progress(self.handle, 1 if block else 0)
        if rc < 0:
            _check(rc)
        return rc

    def result(self, jid):
        r = JobResult()
        rc = lib().ccsdt_server_result(self.handle, jid, ctypes.byref(r))
        if rc < 0:
            _check(rc)
        return r.as_dict() if rc == 1 else None

    def counts(self):
        v = [ctypes.c_int64() for _ in range(4)]
        _check(lib().ccsdt_server_counts(self.handle, *[ctypes.byref(x) for x in v]))
        return {"waiting": v[0].value, "running": v[1].value, "finished": v[2].value,
                "bytes_in_use": v[3].value}

    def clock(self):
        return lib().ccsdt_server_clock(self.handle)

    def options(self):
        """The options the server runs with (defaults and cached picks filled in)."""
        o = ServerOpts()
        _check(lib().ccsdt_server_options(self.handle, ctypes.byref(o)))
        return {k: getattr(o, k) for k, _ in ServerOpts._fields_}

    def drain(self):
        """Progresses until every submitted job has finished."""
        while True:
            c = self.counts()
            if c["waiting"] == 0 and c["running"] == 0:
                return
            self.progress(block=True)

    def close(self):
        if self.handle:
            lib().ccsdt_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def serving_jobs(njobs, seed, nocc_range=(5, 40), nvir_range=(50, 400), tensors=False):
    """Independent jobs of the serving workload: nocc and nvir log-uniform in
    their ranges, inputs seeded per job as in the 8/40 case (one k block, all
    i >= j >= k triples).  tensors=True makes full-tensor jobs (key "tensors")."""
    u = pcg_uniform(seed, 90, 0.0, 1.0, 2 * njobs)
    jobs = []
    for q in range(njobs):
        nocc = int(round(np.exp(np.log(nocc_range[0]) + u[2 * q] *
                                (np.log(nocc_range[1]) - np.log(nocc_range[0])))))
        nvir = int(round(np.exp(np.log(nvir_range[0]) + u[2 * q + 1] *
                                (np.log(nvir_range[1]) - np.log(nvir_range[0])))))
        trip = triple_list(nocc, blocks_uniform(nocc, nocc))
        jb = {"nocc": nocc, "nvir": nvir, "kchunk": nocc, "triples": trip}
        if tensors:
            jb["tensors"] = make_tensor_inputs(nocc, nvir, seed + 1000 + q)
        else:
            jb["arrays"] = make_inputs(nocc, nvir, nocc, seed + 1000 + q)
        jobs.append(jb)
    return jobs


def poisson_arrivals(njobs, rate, seed):
    """Arrival times (seconds) of a Poisson stream of the given rate."""
    u = pcg_uniform(seed, 91, 0.0, 1.0, njobs)
    return np.cumsum(-np.log1p(-u) / rate)


def submit_job(srv, jb):
    """Submits one job dict of serving_jobs() (harness arrays or tensors)."""
    if "tensors" in jb:
        return srv.submit_tensors(jb["tensors"], jb["nocc"], jb["nvir"], jb["triples"])
    return srv.submit(jb["arrays"], jb["nocc"], jb["nvir"], jb["kchunk"], jb["triples"])


def serve(jobs, arrivals, **server_opts):
    """Feeds the jobs to a new Server (server_opts as for Server) at their
    arrival times (seconds after the start) and returns (results, latencies in
    ms, span in s from the first submit to the last finish, the server's
    options)."""
    import time
    with Server(**server_opts) as srv:
        opts = srv.options()
        ids = [None] * len(jobs)
        nxt = 0
        start = srv.clock()
        while True:
            now = srv.clock() - start
            while nxt < len(jobs) and arrivals[nxt] <= now:
                ids[nxt] = submit_job(srv, jobs[nxt])
                nxt += 1
            srv.progress()
            c = srv.counts()
            if nxt == len(jobs) and c["waiting"] == 0 and c["running"] == 0:
                break
            if c["running"] == 0 and nxt < len(jobs):
                time.sleep(max(0.0, min(arrivals[nxt] - now, 0.002)))
        res = [srv.result(i) for i in ids]
    for r in res:
        if r["status"] != 0:
            raise CcsdtError(r["status"], "a served job failed")
    lat = np.array([(r["t_finish"] - r["t_submit"]) * 1e3 for r in res])
    span = max(r["t_finish"] for r in res) - min(r["t_submit"] for r in res)
    return res, lat, span, opts


def tune_server(jobs, arrivals, mode="fp64", slices=0, occ_digits=0, budget_bytes=0,
                candidates=None, warm=True):
    """Serves the same jobs and arrivals once per candidate (tile variant, items
    per tick) and keeps the one with the lowest p99 job latency as the server
    pick for this mode and slice policy on this device
    (ccsdt_server_tune_store: memory and the cache file).  Servers created
    afterwards with variant=TILE_AUTO or items_per_tick=0 use it.  Returns
    (pick, table).  Default candidates: tile 0 at 32k, 64k, 128k and 256k items
    per tick, and the other tiles of the mode at 128k."""
    if candidates is None:
        rows = slices in (0, 3, 4, SLICES_PER_JOB)  # the row-pair tile runs at most 4 slices
        other = {"fp64": [TILE_TWO_WAVE], "ozaki_int8": [TILE_ROW_PAIR] if rows else []}
        candidates = [(0, n) for n in (32768, 65536, 131072, 262144)]
        candidates += [(v, 131072) for v in other.get(mode, [])]
    base = dict(mode=mode, slices=slices, occ_digits=occ_digits, budget_bytes=budget_bytes)
    if warm:  # kernels loaded, pinned and device pools populated before the first timing
        serve(jobs, arrivals, variant=candidates[0][0], items_per_tick=candidates[0][1], **base)
    table = []
    for v, n in candidates:
        _, lat, span, _ = serve(jobs, arrivals, variant=v, items_per_tick=n, **base)
        table.append({"variant": v, "items_per_tick": n, "p50_ms": float(np.percentile(lat, 50)),
                      "p99_ms": float(np.percentile(lat, 99)), "jobs_per_s": len(jobs) / span})
    best = min(table, key=lambda r: r["p99_ms"])
    o = ServerOpts(int(budget_bytes), best["items_per_tick"], MODES[mode], int(slices),
                   best["variant"], int(occ_digits))
    _check(lib().ccsdt_server_tune_store(ctypes.byref(o), best["p99_ms"]))
    return best, table


def tune_cache_file(path):
    """Sets the autotuner's cache file (None: memory only) and returns the number
    of picks read from it (ccsdt_tune_cache_file)."""
    return lib().ccsdt_tune_cache_file(None if path is None else os.fsencode(path))


def tune_cache_clear():
    lib().ccsdt_tune_cache_clear()


def tune_cache_size():
    return lib().ccsdt_tune_cache_size()


def footprint(nocc, nvir, kchunk, ntriples, ncor=0):
    return lib().ccsdt_job_footprint(ncor, nocc, nvir, kchunk, ntriples)


# ------------------------------------------------------------------ drop-ins
def cbody(f, eorb, ncor, nocc, nvir, emp4, emp5, a, i, j, k, klo, arrays,
          symbol="ccsd_trpdrv_omp_cbody_"):
    """Calls the drop-in body with the reference's argument list.

    ``f`` is a list of eight float64 arrays (f1n,f1t,f2n,f2t,f3n,f3t,f4n,f4t),
    overwritten in place (column major, nvir*nvir each); returns (emp4, emp5)
    after the accumulation.
    """
    ints = [ctypes.c_int(v) for v in (ncor, nocc, nvir, a, i, j, k, klo)]
    e4, e5 = ctypes.c_double(emp4), ctypes.c_double(emp5)
    arr = [np.ascontiguousarray(arrays[n], dtype=np.float64) for n in ARRAY_NAMES[1:]]
    fn = getattr(lib(), symbol)
    fn(*[_dptr(x) for x in f], _dptr(np.ascontiguousarray(eorb)),
       ctypes.byref(ints[0]), ctypes.byref(ints[1]), ctypes.byref(ints[2]),
       ctypes.byref(e4), ctypes.byref(e5),
       *[ctypes.byref(v) for v in ints[3:]], *[_dptr(x) for x in arr])
    return e4.value, e5.value


class BodyCall:
    """The drop-in body with its pointer arguments built once, for callers that
    make many calls on the same buffers (the timing loop of a C or Fortran caller
    pays no per-call argument conversion either).  ``call(a, i, j, k, klo)``
    accumulates into ``emp4``/``emp5`` as the body does."""

    def __init__(self, f, eorb, ncor, nocc, nvir, arrays, symbol="ccsd_trpdrv_omp_cbody_"):
        self._keep = [np.ascontiguousarray(arrays[n], dtype=np.float64) for n in ARRAY_NAMES[1:]]
        self._keep.append(np.ascontiguousarray(eorb, dtype=np.float64))
        self._f = f
        self._ints = [ctypes.c_int(v) for v in (ncor, nocc, nvir, 0, 0, 0, 0, 0)]
        self._e4, self._e5 = ctypes.c_double(0.0), ctypes.c_double(0.0)
        self._fn = getattr(lib(), symbol)
        self._args = ([_dptr(x) for x in f] + [_dptr(self._keep[-1])]
                      + [ctypes.byref(v) for v in self._ints[:3]]
                      + [ctypes.byref(self._e4), ctypes.byref(self._e5)]
                      + [ctypes.byref(v) for v in self._ints[3:]]
                      + [_dptr(x) for x in self._keep[:-1]])

    @property
    def emp4(self):
        return self._e4.value

    @property
    def emp5(self):
        return self._e5.value

    def call(self, a, i, j, k, klo):
        for v, x in zip(self._ints[3:], (a, i, j, k, klo)):
            v.value = int(x)
        self._fn(*self._args)


def tengy(f, vec, eorb, eaijk, ncor, nocc, nvir, symbol="ccsd_tengy_omp"):
    """Drop-in energy reduction; f as in cbody, vec = dict of the six vectors.

    Returns (emp4i, emp5i, emp4k, emp5k)."""
    out = [ctypes.c_double() for _ in range(4)]
    L = lib()
    fs = [_dptr(np.ascontiguousarray(x)) for x in f]
    v = {n: _dptr(np.ascontiguousarray(vec[n])) for n in
         ("dintc1", "dintx1", "t1v1", "dintc2", "dintx2", "t1v2")}
    e = _dptr(np.ascontiguousarray(eorb))
    if symbol == "ccsd_tengy_omp_":
        ea = ctypes.c_double(eaijk)
        sz = [ctypes.c_int(x) for x in (ncor, nocc, nvir)]
        L.ccsd_tengy_omp_(*fs, v["dintc1"], v["dintc2"], v["dintx1"], v["dintx2"], v["t1v1"],
                          v["t1v2"], e, ctypes.byref(ea), *[ctypes.byref(o) for o in out],
                          *[ctypes.byref(s) for s in sz])
    else:
        getattr(L, symbol)(*fs, v["dintc1"], v["dintx1"], v["t1v1"], v["dintc2"], v["dintx2"],
                           v["t1v2"], e, eaijk, *[ctypes.byref(o) for o in out], ncor, nocc, nvir)
    return tuple(o.value for o in out)


def selftest_mfma_i8(a_bytes, b_bytes, shape):
    """a_bytes, b_bytes: (64, 16) int8 per-lane operand bytes."""
    a = np.ascontiguousarray(a_bytes, dtype=np.int8).view(np.int32).ravel()
    b = np.ascontiguousarray(b_bytes, dtype=np.int8).view(np.int32).ravel()
    d = np.empty(64 * (16 if shape == 1 else 4), dtype=np.int32)
    _check(lib().ccsdt_selftest_mfma_i8(_ptr(a), _ptr(b), _ptr(d), shape))
    return d.reshape(64, -1)


def selftest_mfma(a, b):
    d = np.empty(256)
    _check(lib().ccsdt_selftest_mfma(_ptr(np.ascontiguousarray(a, dtype=np.float64)),
                                     _ptr(np.ascontiguousarray(b, dtype=np.float64)), _ptr(d)))
    return d
