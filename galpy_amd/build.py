"""Build the HIP engine (gfx950) and the CPU oracle in-tree.

    python -m galpy_amd.build            # both
    python -m galpy_amd.build --hip-only

The HIP library is compiled with hipcc straight to a shared object next to
this file (galpy_amd/libgalpy_hip.so), so it travels with the source tree.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgalpy_hip.so")

HIP_SOURCES = ["galpy_hip.hip"]
HIP_HEADERS = ["potential_dev.h", "integrators_dev.h", "dop853_dev.h", "dop853_rkn.h", "batch_dev.h",
               "adjoint_dev.h", "gh_math.h"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    # Machine-level LICM hoists every FP64 literal of the integrators' loops
    # (tableau and polynomial coefficients, 2 SGPRs or VGPRs each) out of
    # the loop, where at two waves per SIMD they no longer fit the register
    # budget and spill to scratch; materialised at each use they cost only
    # scalar moves (DOP853 kernel: 381 spilled VGPRs -> none).
    "-mllvm", "-disable-machine-licm",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_hip(verbose=True, extra_flags=()):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *extra_flags, "-I", os.path.join(ROOT, "include"), *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=True):
    cmd = ["make", "-C", os.path.join(ROOT, "oracle"), "-s"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return os.path.join(ROOT, "oracle", "liboracle.so")


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_hip()
    if "--hip-only" not in argv:
        build_oracle()


if __name__ == "__main__":
    main()
