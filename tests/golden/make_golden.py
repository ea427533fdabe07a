"""Writes the golden fixtures under tests/golden/ (run in a container where
oracle/_ref is built from the reference sources):

    make -C oracle && python tests/golden/make_golden.py

config0.npz -- the 8-occupied / 40-virtual case, 120 triples (i >= j >= k, one
  k block).  f matrices come from the oracle's DGEMMs (pinned against exact and
  numpy GEMMs in tests/test_oracle.py); every per-triple term is then computed
  by the reference's own ccsd_tengy_omp (omp/ccsd_tengy_c99.c, single thread)
  and the running totals by the reference body compiled with its SKIP_DGEMM
  switch (eaijk + tengy + the i != k rule, omp/ccsd_trpdrv_omp_cbody.c:121-142).
config1_head.npz -- the 30-occupied / 300-virtual case: reference terms of the
  first 6 triples of each k block.
fp32_ref.npz -- the FP32 reference body (gpu-omp/ccsd_trpdrv_omp_cbody.c built
  with its SKIP_DGEMM switch) on FP32-rounded inputs and FP32-rounded oracle f
  matrices: per triple its emp4 and emp5 for the two cases of
  tests/test_gpu_fp32.py, which compares with them when oracle/_ref is absent.
  `python tests/golden/make_golden.py fp32` writes this file only.
Inputs are not stored: they are regenerated from their seeds, and the fixture
keeps checksums of them.
"""
import ctypes
import ctypes.util
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
from conftest import load_package  # noqa: E402


def omp_threads(n):
    """Thread count of every OpenMP runtime in the process (MKL brings its own libgomp)."""
    paths = {ctypes.util.find_library("gomp") or "libgomp.so.1"}
    with open("/proc/self/maps") as fh:
        for line in fh:
            p = line.split()[-1]
            if "/libgomp" in p or "/libomp" in p:
                paths.add(p)
    for p in paths:
        ctypes.CDLL(p).omp_set_num_threads(n)


def checksums(arrays):
    """Plain and position-weighted sums of each input, and a SHA-256 of its
    bytes (which pins every bit of the regenerated array)."""
    names = sorted(arrays)
    sums = np.array([[float(np.sum(arrays[n])),
                      float(np.sum(arrays[n] * (np.arange(arrays[n].size) % 7)))]
                     for n in names])
    digests = np.array([hashlib.sha256(np.ascontiguousarray(arrays[n], dtype="<f8").tobytes())
                        .hexdigest() for n in names])
    return names, sums, digests


def reference_terms(arrays, nocc, nvir, triples, Lt, Lb):
    terms = np.zeros((len(triples), 4))
    e4 = e5 = 0.0
    keep_f = {}
    for t, (a, i, j, k, klo) in enumerate(triples.tolist()):
        f = oracle.fmats(arrays, nocc, nvir, k, klo)
        fcm = [np.ascontiguousarray(f[m].T) for m in range(8)]
        ea = oracle.eaijk(arrays["eorb"], 0, a, i, j, k)
        terms[t] = oracle.ref_tengy(Lt, fcm, arrays, ea, 0, nocc, nvir)
        e4, e5 = oracle.ref_cbody_glue(Lb, fcm, arrays, 0, nocc, nvir, e4, e5, a, i, j, k, klo)
        if t == 7:
            keep_f[t] = f
    return terms, np.array([e4, e5]), keep_f


FP32_CASES = ((8, 40, 8, 808), (5, 37, 2, 808))  # nocc, nvir, kchunk, input seed


def fp32_reference(cc):
    """Per-triple (emp4, emp5) of the FP32 reference body for FP32_CASES."""
    L = oracle.ref_cbody_f32_lib()
    if L is None:
        raise SystemExit("oracle/_ref not built: run `make -C oracle` with the reference mounted")
    out = {}
    for c, (nocc, nvir, kchunk, seed) in enumerate(FP32_CASES):
        arrays = cc.make_inputs(nocc, nvir, kchunk, seed)
        a32 = {n: np.asarray(v, dtype=np.float32).astype(np.float64) for n, v in arrays.items()}
        trip = cc.triple_list(nocc, cc.blocks_uniform(nocc, kchunk))
        e45 = np.zeros((len(trip), 2))
        for t, (a, i, j, k, klo) in enumerate(trip.tolist()):
            f = oracle.fmats(a32, nocc, nvir, k, klo)
            fcm = [np.ascontiguousarray(f[m].T).astype(np.float32) for m in range(8)]
            e45[t] = oracle.ref_cbody_f32_glue(L, fcm, a32, 0, nocc, nvir, a, i, j, k, klo)
        names, sums, digests = checksums(arrays)
        out["case%d_shape" % c] = np.array([nocc, nvir, kchunk, seed])
        out["case%d_triples" % c] = trip
        out["case%d_e45" % c] = e45
        out["case%d_names" % c] = np.array(names)
        out["case%d_digests" % c] = digests
    np.savez_compressed(os.path.join(HERE, "fp32_ref.npz"), **out)
    print("wrote fp32_ref.npz")


def main():
    cc = load_package()
    if len(sys.argv) > 1 and sys.argv[1] == "fp32":
        oracle.blas_select(0)
        omp_threads(1)
        fp32_reference(cc)
        return
    Lt, Lb = oracle.ref_tengy_lib(), oracle.ref_cbody_skipdgemm_lib()
    if Lt is None or Lb is None:
        raise SystemExit("oracle/_ref not built: run `make -C oracle` with the reference mounted")
    oracle.blas_select(0)
    omp_threads(1)

    arrays, blocks, trip, kchunk = cc.config_case(0)
    names, sums, digests = checksums(arrays)
    terms, emp, keep = reference_terms(arrays, 8, 40, trip, Lt, Lb)
    np.savez_compressed(os.path.join(HERE, "config0.npz"), triples=trip, terms=terms, emp=emp,
                        f7=keep[7], names=np.array(names), sums=sums, digests=digests,
                        provider=np.array(oracle.blas_name()))

    arrays, blocks, trip, kchunk = cc.config_case(1)
    pick = []
    for klo, _ in blocks:
        pick += [q for q in range(len(trip)) if trip[q, 4] == klo][:6]
    sub = trip[pick]
    names, sums, digests = checksums(arrays)
    omp_threads(8)
    terms, emp, _ = reference_terms(arrays, 30, 300, sub, Lt, Lb)
    np.savez_compressed(os.path.join(HERE, "config1_head.npz"), triples=sub, index=np.array(pick),
                        terms=terms, names=np.array(names), sums=sums, digests=digests,
                        provider=np.array(oracle.blas_name()))
    print("wrote config0.npz, config1_head.npz with", oracle.blas_name())
    fp32_reference(cc)


if __name__ == "__main__":
    main()
