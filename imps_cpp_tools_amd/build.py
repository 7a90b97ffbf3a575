"""Build recipe for libimps_hip.so (HIP kernels + C ABI) for gfx950.

Run as ``python -m imps_cpp_tools_amd.build``; the shared library is written in-tree next to this
file so that it travels with the source tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libimps_hip.so")

SOURCES = ["gemm_kernels.hip", "panel_kernels.hip", "blas_kernels.hip", "qr_kernels.hip", "cplx_kernels.hip", "static_scratch.hip", "plan.cpp",
           "krylov.cpp",
           "heff.cpp", "adjoint.cpp", "capi.cpp", "models.cpp"]
HEADERS = ["engine.hpp", "panel.hpp", "layout.hpp", "blas.hpp", "common.hpp", "plan.hpp", "krylov.hpp", "qr.hpp", "cplx.hpp", "heff.hpp", "adjoint.hpp", "static_scratch.hpp"]

# -packed-fp32-ops: keep the compiler from forming packed FP32 (v_pk_*_f32) instructions.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; ROCm is required to build libimps_hip.so")
    return exe


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (res.stderr or "").splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + err)
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
