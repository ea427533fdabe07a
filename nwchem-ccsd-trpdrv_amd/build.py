"""Builds libccsdt.so in-tree for gfx950 with hipcc.

    python nwchem-ccsd-trpdrv_amd/build.py

The library holds the HIP kernels and the C-ABI layer declared in
include/ccsdt.h.  It is compiled every time this runs (seconds).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["csrc/ccsdt_api.hip", "csrc/ccsdt_kernels.hip"]
LIB = os.path.join(HERE, "libccsdt.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 ops stay off (the host pass warns it does not know the feature)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs: the fused energy epilogue reads them with VALU
    # instructions, and the AGPR form made the compiler copy all 128 accumulator
    # registers between the two files on every K step.
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
    "-I" + os.path.join(ROOT, "include"),
    # roctx ranges around job runs, drop-in calls and serving ticks
    # (rocprofv3 --marker-trace)
    "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx",
]


def build(verbose=False):
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in p.stderr.splitlines()
                    if "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + err)
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
