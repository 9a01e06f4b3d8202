"""Build the HIP engine (gfx950) and the CPU oracle in-tree.

    python -m galpy_amd.build                  # both, every unit compiled
    python -m galpy_amd.build --hip-only
    python -m galpy_amd.build --incremental    # recompile only stale units

The HIP library is compiled as several translation units in parallel
(hipcc -c for gfx950 into galpy_amd/_build/), then linked into a shared
object next to this file (galpy_amd/libgalpy_hip.so), so it travels with the
source tree.  gh_instances.hip is compiled once per force-composition / orbit
kind pair (-DGH_INSTANCE=k).
"""

import concurrent.futures
import glob
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgalpy_hip.so")

# Number of (composition, orbit kind) pairs in gh_instances.hip; must match
# GH_LAUNCH_SPEC_ALL in gh_internal.h.
NUM_INSTANCES = 16
HEAVY_INSTANCES = (5, 4, 11, 10, 13, 15)

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
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


def units():
    """(object name, source, extra defines) of every translation unit."""
    inst = [(f"gh_instances_{k}.o", "gh_instances.hip", [f"-DGH_INSTANCE={k}"])
            for k in range(NUM_INSTANCES)]
    other = [("gh_batch.o", "gh_batch.hip", []), ("gh_adjoint.o", "gh_adjoint.hip", []),
             ("gh_host.o", "gh_host.hip", []), ("gh_noalloc.o", "gh_noalloc.hip", [])]
    # longest compiles first (the run-time dispatch and wrapper-aware
    # instances, the batch kernel), so the parallel build ends sooner
    heavy = [inst[k] for k in HEAVY_INSTANCES] + other[:1]
    return heavy + [u for u in inst + other if u not in heavy]


def _deps_mtime():
    files = glob.glob(os.path.join(CSRC, "*.h")) + [os.path.join(ROOT, "include", "galpy_hip.h")]
    return max(os.path.getmtime(f) for f in files)


def _compile(obj, src, defs, extra_flags, verbose):
    out = os.path.join(OBJDIR, obj)
    cmd = [hipcc(), *HIPCC_FLAGS, *extra_flags, *defs, "-I", os.path.join(ROOT, "include"), "-c",
           os.path.join(CSRC, src), "-o", out + ".tmp"]
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(line for line in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops' is not a recognized feature" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {obj}:\n{' '.join(cmd)}\n{msgs}")
    os.replace(out + ".tmp", out)
    if verbose:
        print(f"  {obj}: {time.time() - t0:.0f} s" + (f"\n{msgs}" if msgs.strip() else ""),
              flush=True)
    return out


def build_hip(verbose=True, extra_flags=(), incremental=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    deps = _deps_mtime()
    todo, objs = [], []
    for obj, src, defs in units():
        path = os.path.join(OBJDIR, obj)
        objs.append(path)
        if incremental and os.path.exists(path):
            m = os.path.getmtime(path)
            if m > deps and m > os.path.getmtime(os.path.join(CSRC, src)):
                continue
        todo.append((obj, src, defs))
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 4, 16)
    if verbose:
        print(f"hipcc: {len(todo)} of {len(objs)} units, {jobs} in parallel", flush=True)
    t0 = time.time()
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, o, s, d, extra_flags, verbose) for o, s, d in todo]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"linked {LIB} ({time.time() - t0:.0f} s)", flush=True)
    return LIB


def build_oracle(verbose=True):
    cmd = ["make", "-C", os.path.join(ROOT, "oracle"), "-s"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return os.path.join(ROOT, "oracle", "liboracle.so")


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    build_hip(incremental="--incremental" in argv)
    if "--hip-only" not in argv:
        build_oracle()


if __name__ == "__main__":
    main()
