"""Builds libpotree_gpu.so (gfx950 kernels + host drop-ins) in-tree.

hipcc compiles every HIP and host C++ source for --offload-arch=gfx950 with
IEEE binary64 semantics (-ffp-contract=off, no fast-math) so that the FP64
geometry matches the reference's x86-64 release build, and with packed FP32
instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpotree_gpu.so")

SOURCES = [
    "scan.hip",
    "chunker.hip",
    "octree.hip",
    "sampler.hip",
    "chunkgen.hip",
    "chunking.cpp",
    "indexing.cpp",
    "writers.cpp",
    "encoding.cpp",
    "attributes.cpp",
    "lasgen_host.cpp",
    "chunkgen_host.cpp",
    "capi.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{msgs}")
    return obj, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]))
    objs = [o for o, _ in results]
    if verbose:
        for _, m in results:
            if m.strip():
                print(m)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
