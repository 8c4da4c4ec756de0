"""Build recipe of libvortexhip.so for gfx950 (hipcc, in-tree: vortexpanel.jl_amd/lib/)."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "lib", "libvortexhip.so")
SOURCES = ["vp_assemble.hip", "vp_lu.hip", "vp_field.hip", "vp_capi.hip", "vp_freewake.hip", "vp_mixed.hip", "vp_tree.hip"]

# -munsafe-fp-atomics: hardware FP64 atomic adds (global and LDS) instead of compare-and-swap loops.
# -packed-fp32-ops off: no v_pk_*_f32 instructions anywhere.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# rocBLAS: the LU autotuner's library candidate for the trailing update (a plain GEMM)
LIBS = ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocblas"]


def build_library(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = ["hipcc", *HIPFLAGS, "-o", tmp, *srcs, *LIBS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
