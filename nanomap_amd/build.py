"""Build libnanomap_amd.so for gfx950 with hipcc (no CMake).

Every .hip file under csrc/ is compiled to an object in _build/ and the
objects are linked into nanomap_amd/libnanomap_amd.so, next to this file.
Run as `python -m nanomap_amd.build` or through __graft_entry__.build().
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build", "nanomap_amd")
LIB = os.path.join(HERE, "libnanomap_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the cast reproduces IEEE operation sequences exactly,
#   so no a*b+c may be fused.
# -fhip-fp32-correctly-rounded-divide-sqrt: IEEE FP32 division and sqrt.
# -packed-fp32-ops off: no packed FP32 instructions; hipcc forms them from
#   ordinary float code and some of their forms have given intermittently
#   wrong results on this hardware.
COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f.endswith(".cpp"))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    extra = []
    if src.endswith(".cpp"):
        # host-only code; its FP set-up runs under fesetround()
        extra = ["-x", "c++", "-frounding-math"]
    cmd = [HIPCC] + COMMON + extra + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for o, w in results:
            if w:
                print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    return LIB


HOST_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_host.cpp")
HOST_TEST = os.path.join(ROOT, "tests", "cpp", "test_host")


def build_host_test():
    """The C++ driver of the host-API tests, linked against the library."""
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall",
           "-I" + os.path.join(HERE, "host"), "-I" + os.path.join(ROOT, "include"),
           HOST_TEST_SRC, "-L" + HERE, "-lnanomap_amd",
           "-Wl,-rpath,$ORIGIN/../../nanomap_amd", "-o", HOST_TEST]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host test build failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    return HOST_TEST


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_host_test())
