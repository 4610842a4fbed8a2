"""Build recipe for the gfx950 libraries (in-tree, no JIT cache).

Produces, next to this file:
  lib/libpf_kernels_hip.so   the drop-in C-ABI library (pfc_*, pfgl_*, extensions)
  lib/libpf_kernels.so       symlink, the name the reference's build.rs links (build.rs:15)
  lib/libpf_host.so          C++ host layer (Backend/ParticleFilter/GlobalLoc, pfh_*)

Every HIP translation unit is compiled with -ffp-contract=off (the deterministic-draw
code relies on explicit fma) and with packed FP32 disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "lib", "obj")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PF_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["pf_kernels.hip", "pf_batched.hip", "gl_kernels.hip", "mlp_kernels.hip"]
HOST_SOURCES = ["pf_host.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    noisy = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", out])
    return out


def build(verbose: bool = False) -> dict:
    """Compile every HIP/C++ source and link the shared libraries. Returns their paths."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    kern = os.path.join(LIB, "libpf_kernels_hip.so")
    tmp = kern + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, kern)
    alias = os.path.join(LIB, "libpf_kernels.so")
    if os.path.lexists(alias):
        os.remove(alias)
    os.symlink("libpf_kernels_hip.so", alias)
    host = os.path.join(LIB, "libpf_host.so")
    host_srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if host_srcs:
        tmp = host + ".tmp"
        _run([HIPCC, "-O2", "-fPIC", "-shared", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}", *host_srcs,
              f"-L{LIB}", "-lpf_kernels_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp])
        os.replace(tmp, host)
    if verbose:
        print("built", kern, host)
    return {"kernels": kern, "host": host}


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
