"""Build libpfx_chem.so (HIP, gfx950) in-tree with hipcc.

    python -m pflotran_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpfx_chem.so")
SOURCES = ["pfx_react.hip", "pfx_api.hip", "pfx_mlp.hip", "pfx_gi.hip", "pfx_krylov.hip"]
ARCH = os.environ.get("PFX_OFFLOAD_ARCH", "gfx950")


def hipcc_cmd(out=LIB, defines=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    return [hipcc, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared"] + [f"-D{d}" for d in defines] + [
            # no packed-FP32 instructions in device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
            "-o", out] + [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose=True):
    cmd = hipcc_cmd(LIB + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
