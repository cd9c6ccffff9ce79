"""Build libdmrjhip.so in-tree for gfx950 with hipcc (no cmake/ninja).

Usage: python dmrjtensor.jl_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, "csrc", f) for f in ("api.hip", "kernels.hip", "svd_small.hip", "kernels_lp.hip", "autotune.hip", "complex.hip", "svd_jacobi.hip", "svd.hip")]
OUT = os.path.join(HERE, "lib", "libdmrjhip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROCM, "include"),
]
LIBS = ["-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
        "-Wl,-rpath," + os.path.join(ROCM, "lib")]


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [s for s in SRC if os.path.exists(s)]
    tmp = OUT + ".tmp"
    cmd = [os.path.join(ROCM, "bin", "hipcc")] + FLAGS + srcs + LIBS + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libdmrjhip.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
