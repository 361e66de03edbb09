"""Builds the HIP library (gfx950) in-tree.

    python -m discrete_adjoint_amd.build

produces discrete_adjoint_amd/libtg.so from csrc/.  The FP64 kernels are
compiled with -ffp-contract=off so that no multiply-add is fused behind the
source's back (the FP64 path is bit-compared against a CPU restatement), and
packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["tg_api.hip", "tg_kernels.hip", "tg_kernels32.hip", "tg_leaves.hip", "tg_noalloc.hip",
           "tg_error.hip", "tg_host.cpp",
           "tg_flatten.cpp", "tg_solrec.cpp", "tg_adapt.cpp",
           "tg_gis.cpp", "tg_io.cpp", "tg_error_host.cpp", "tg_funky.cpp"]
HEADERS = ["tg_kernels.h", "tg_leafops.h", "tg_math.h", "tg_math32.h"]

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), *srcs, "-lz", "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libtg.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
