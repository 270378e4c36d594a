"""Build libaoa.so (gfx950 HIP kernels + C-ABI + host shim) in-tree with hipcc.

python -m aoa_amd.build   (or __graft_entry__.build())
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaoa.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["params.cpp", "ntt.hip", "rns.hip", "encoder.hip", "sampler.hip", "capi.cpp", "batch.cpp", "tune.cpp", "server.cpp", "noalloc.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
HIP_FLAGS = ["--offload-arch=gfx950", "-x", "hip",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [HIPCC] + COMMON + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_newckks(verbose)
    return LIB


NEWCKKS_LIB = os.path.join(HERE, "libnewckks.so")
CPP_TEST = os.path.join(ROOT, "tests", "cpp", "_build", "test_newckks")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))


def build_newckks(verbose=False):
    """libnewckks.so: the newckks-compatible C++ API, a host library over the
    C-ABI of libaoa.so; plus the C++ test program written against it."""
    inc = ["-I" + os.path.join(ROOT, "include")]
    objs = []
    for src in ("newckks.cpp", "newckks_ops.cpp", "newckks_io.cpp"):
        obj = os.path.join(OBJ, src.replace(".cpp", ".o"))
        _run([HIPCC, "-O2", "-fPIC", "-std=c++17", "-Wall"] + inc +
             ["-c", os.path.join(CSRC, "newckks", src), "-o", obj])
        objs.append(obj)
    tmp = NEWCKKS_LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "-o", tmp] + objs + ["-L" + HERE, "-laoa", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, NEWCKKS_LIB)
    os.makedirs(os.path.dirname(CPP_TEST), exist_ok=True)
    _run([HIPCC, "-O2", "-std=c++17"] + inc + [os.path.join(ROOT, "tests", "cpp", "test_newckks.cpp"), "-o",
                                               CPP_TEST + ".tmp", "-L" + HERE, "-lnewckks", "-laoa",
                                               "-Wl,-rpath," + os.path.relpath(HERE, os.path.dirname(CPP_TEST))
                                               .join(["$ORIGIN/", ""])])
    os.replace(CPP_TEST + ".tmp", CPP_TEST)
    if verbose:
        print("built", NEWCKKS_LIB, CPP_TEST)
    build_issue_probe(verbose)


ISSUE_PROBE = os.path.join(ROOT, "tests", "native", "_build", "issue_rate")


def build_issue_probe(verbose=False):
    """tests/native/_build/issue_rate: the issue-rate probe behind the VALU
    reference rates quoted with the roofline (run by a GPU test)."""
    os.makedirs(os.path.dirname(ISSUE_PROBE), exist_ok=True)
    _run([HIPCC, "-O3", "-std=c++17"] + HIP_FLAGS[:1] + HIP_FLAGS[3:] +
         [os.path.join(ROOT, "tests", "native", "issue_rate.hip"), "-o", ISSUE_PROBE + ".tmp"])
    os.replace(ISSUE_PROBE + ".tmp", ISSUE_PROBE)
    if verbose:
        print("built", ISSUE_PROBE)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
