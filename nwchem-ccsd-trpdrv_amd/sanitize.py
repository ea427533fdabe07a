"""Sanitizer builds of the host side, run on the CPU test suites.

    python nwchem-ccsd-trpdrv_amd/sanitize.py [asan|tsan|all]

asan: libccsdt_asan.so -- the library with its host code (ccsdt_api.hip's C ABI,
  job and serving layers, copy pool; the kernels' host stubs) instrumented by
  AddressSanitizer and UndefinedBehaviorSanitizer -- and oracle/liboracle_asan.so,
  the oracle likewise (clang, so one sanitizer runtime serves both); then the
  CPU tests test_abi.py, test_oracle.py and test_golden.py against them, with the
  runtime preloaded into the interpreter.  Leak checking is off (the interpreter
  itself is not instrumented and keeps its allocations); any other report fails
  the run.
tsan: libccsdt_tsan.so with ThreadSanitizer, and the tests of the host-copy pool
  (the worker threads the synchronous drop-ins stage their arrays with), which
  run without a device.  The serving loop's admission thread needs a device and
  is not reached on the CPU.
Device code is never instrumented: every -fsanitize= is host-only (-Xarch_host).
Outputs stay in-tree beside the regular builds (git-ignored).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
import build  # noqa: E402

LLVM = "/opt/rocm/lib/llvm"
RT = os.path.join(LLVM, "lib/clang/22/lib/linux")
SAN = {
    "asan": (["address", "undefined"], "libclang_rt.asan-x86_64.so",
             ["tests/test_abi.py", "tests/test_oracle.py", "tests/test_golden.py"]),
    "tsan": (["thread"], "libclang_rt.tsan-x86_64.so", ["tests/test_abi.py", "-k", "copy"]),
}


def build_lib(kind):
    """The library with host-only sanitizers."""
    flags = []
    for f in SAN[kind][0]:
        flags += ["-Xarch_host", "-fsanitize=" + f]
    flags += ["-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-g"]
    out = os.path.join(HERE, "libccsdt_%s.so" % kind)
    srcs = [os.path.join(HERE, s) for s in build.SOURCES]
    cmd = [build.HIPCC] + build.FLAGS + flags + srcs + ["-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + p.stderr)
    return out


def build_oracle_asan():
    """The oracle with AddressSanitizer and UBSan, by clang (the runtime of the
    library's build), OpenMP from LLVM's runtime."""
    src = os.path.join(ROOT, "oracle", "ccsdt_oracle.c")
    out = os.path.join(ROOT, "oracle", "liboracle_asan.so")
    cmd = [os.path.join(LLVM, "bin/clang"), "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-shared-libsan", "-ffp-contract=off", "-fopenmp",
           "-fPIC", "-shared", src, "-o", out, "-ldl", "-lm",
           "-L" + os.path.join(LLVM, "lib"), "-Wl,-rpath," + os.path.join(LLVM, "lib")]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("clang failed:\n" + " ".join(cmd) + "\n" + p.stderr)
    return out


CONTROL = """
import ctypes, numpy as np, oracle
L = oracle.olib()
out = np.zeros(16)
L.orc_pcg_uniform(1, 2, 0.0, 1.0, out.ctypes.data_as(ctypes.c_void_p), 4096)  # writes past out
print("not caught")
"""


def control(env):
    """Positive control: the sanitizer must stop a heap overflow that the oracle is
    made to commit (on the CPU).  Returns True when it did."""
    p = subprocess.run([sys.executable, "-c", CONTROL], cwd=ROOT, env=env, capture_output=True, text=True)
    caught = p.returncode != 0 and "heap-buffer-overflow" in p.stderr
    print("sanitize asan control: %s" % ("overflow caught" if caught else "NOT caught\n" + p.stdout + p.stderr[-2000:]),
          flush=True)
    return caught


def run(kind):
    lib = build_lib(kind)
    env = dict(os.environ)
    env["CCSDT_LIB"] = lib
    env["LD_PRELOAD"] = os.path.join(RT, SAN[kind][1])
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    if kind == "asan":
        env["CCSDT_ORACLE_LIB"] = build_oracle_asan()
        env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:detect_odr_violation=0"
        if not control(env):
            return 1
    else:
        env["TSAN_OPTIONS"] = "halt_on_error=1:second_deadlock_stack=1"
    cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + SAN[kind][2]
    print("sanitize %s: %s" % (kind, " ".join(cmd)), flush=True)
    return subprocess.run(cmd, cwd=ROOT, env=env).returncode


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    kinds = ["asan", "tsan"] if which == "all" else [which]
    rc = 0
    for k in kinds:
        r = run(k)
        print("sanitize %s: exit %d" % (k, r), flush=True)
        rc = rc or r
    sys.exit(rc)
