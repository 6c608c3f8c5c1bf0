"""Builds libgcs_amd.so (gfx950 HIP kernels, host engine, C ABI and the
cuda_engine.h entry points) in-tree with hipcc.

    python -m gcs_amd._build        (from galaxy-collision-simulator_amd/)
"""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(PKG_DIR, "libgcs_amd.so")
BAKE = os.path.join(PKG_DIR, "gcs_bake")

SOURCES = ["gcs_engine.hip", "gcs_generator.cpp", "gcs_cuda_engine.cpp", "gcs_server.cpp"]

# -packed-fp32-ops: no packed FP32 instructions are generated.
# Correctly rounded float divide and sqrt: the Morton scale and the
# aggregation COM divide must round as IEEE does.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.isabs(c) and os.path.exists(c):
            return c
    return "hipcc"


def build(verbose=False):
    """Compiles every source (always) and links libgcs_amd.so."""
    objs = []
    obj_dir = os.path.join(ROOT, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + HIP_FLAGS + ["-I", INCLUDE, "-I", CSRC, "-c",
                                       os.path.join(CSRC, src), "-o", obj]
        _run(cmd, verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lpthread"], verbose)
    os.replace(tmp, LIB)
    # Bake driver linked against the library (cuda_engine.h entry points).
    bake_obj = os.path.join(obj_dir, "gcs_bake_main.o")
    _run([hipcc(), "-O2", "-std=c++17", "-fopenmp", "-I", INCLUDE, "-c",
          os.path.join(CSRC, "gcs_bake_main.cpp"), "-o", bake_obj], verbose)
    _run([hipcc(), "-fopenmp", "-o", BAKE + ".tmp", bake_obj, LIB, "-Wl,-rpath,$ORIGIN",
          "-lpthread"], verbose)
    os.replace(BAKE + ".tmp", BAKE)
    return LIB


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
