"""Build libfmma_hip.so (gfx950) in-tree with hipcc.

    python -m fmma_amd.build          # or fmma_amd.build.build()

Every HIP source under fmma_amd/csrc is compiled for gfx950 into
fmma_amd/_build/*.o (in parallel) and linked into fmma_amd/libfmma_hip.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfmma_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# (source, object name, extra flags).  builtins.hip is compiled once per
# (TYPE, DIM) so its kernel instantiations build in parallel, and once more
# as the dispatcher over them.
SOURCES = [("tree.hip", "tree", []), ("boxops.hip", "boxops", []), ("m2l.hip", "m2l", []), ("batch.hip", "batch", []),
           ("plan.hip", "plan", []), ("builtins.hip", "builtins", [])] + [
    ("builtins.hip", "builtins_%s%d" % (t[0], d), ["-DFMMA_BI_TYPE=" + t, "-DFMMA_BI_DIM=%d" % d])
    for t in ("double", "float") for d in (1, 2, 3)]

# Packed FP32 instructions are switched off for all device code.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _newest_header():
    newest = 0.0
    for d in (CSRC, os.path.join(ROOT, "include")):
        for dp, _, files in os.walk(d):
            for f in files:
                if f.endswith((".hpp", ".h")):
                    newest = max(newest, os.path.getmtime(os.path.join(dp, f)))
    return newest


def _compile(entry, incremental=False, newest_header=0.0):
    src, name, extra = entry
    obj = os.path.join(OBJ, name + ".o")
    if incremental and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if t > os.path.getmtime(os.path.join(CSRC, src)) and t > newest_header:
            return obj
    cmd = [HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, r.stderr[-6000:]))
    return obj


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s" % (what, r.stderr[-6000:]))


def build_pyfmma():
    """The pybind11 drop-in module fmma_amd/pyfmma*.so (host C++ over the C-ABI)."""
    import sysconfig
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "pyfmma" + suffix)
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
           os.path.join(CSRC, "pyfmma.cpp"), "-L" + HERE, "-lfmma_hip", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    _run(cmd, "pyfmma build")
    os.replace(out + ".tmp", out)
    return out


def build_cpp_tests():
    """tests/cpp/dropin_test: fmma::FMMA used as the reference's C++ users use it."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "dropin_test")
    cmd = [HIPCC] + HIP_FLAGS + [src, "-L" + HERE, "-lfmma_hip", "-Wl,-rpath,$ORIGIN/../../fmma_amd",
                                 "-o", out + ".tmp"]
    _run(cmd, "dropin_test build")
    os.replace(out + ".tmp", out)
    return out


def build(verbose=False, incremental=False):
    """Compile every source (incremental=True: only objects older than their
    source or any header) and link the library, pyfmma and the C++ test."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(16, os.cpu_count() or 4)))
    newest = _newest_header()
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda e: _compile(e, incremental, newest), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    extra = [build_pyfmma(), build_cpp_tests()]
    if verbose:
        print("built", LIB, *extra)
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
