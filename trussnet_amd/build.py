"""Build the gfx950 library (trussnet_amd/lib/libtrussnet_amd.so) with hipcc.

Floating-point flags: the relaxation is compared bit for bit with the CPU path,
which is IEEE single precision without contraction (-ffp-contract=off), so the
device code keeps correctly rounded fp32 division and square root, keeps fp32
denormals, and never contracts a*b+c into an FMA. Packed fp32 instructions are
disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libtrussnet_amd.so")
ARCH = os.environ.get("TN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form=1", "-Wall", "-Wno-unused-function",
]
SOURCES = ["tn_relax.hip"]
DEPS = ["tn_relax.hip", "tn_kernels.hpp", "tn_device.hpp", "tn_seed.hpp", "tn_thin.hpp", "tn_adjoint.hpp", "tn_grid.hpp", os.path.join("..", "..", "include", "trussnet_amd.h")]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, out: str | None = None, defines: tuple = ()) -> str:
    """Compile the library (to `out`, default lib/libtrussnet_amd.so) with extra -D defines."""
    out = out or LIB_PATH
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, *[f"-D{d}" for d in defines], "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    if verbose and r.stderr.strip():
        # the host pass reports the device-only target feature; everything else is shown
        lines = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if lines:
            print("\n".join(lines))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
