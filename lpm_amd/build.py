"""Builds liblpm_mi355x.so (hipcc, gfx950) and the C++ example driver.

Usage: python -m lpm_amd.build [--jobs N]

Every HIP translation unit is compiled for gfx950 only, with packed-FP32
instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
Objects go to lpm_amd/_build/, the shared library to lpm_amd/.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shlex
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblpm_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-x", "hip",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-flush-denormals-to-zero",
    # matrix-core operands and results in VGPRs: the VALU consumes every
    # MFMA result, so accumulation registers would only add copies
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]
# host-only sources: strict IEEE operation order for the mesh generator
HOST_SOURCES = {"mesh.cpp": ["-ffp-contract=off"], "abi.cpp": [], "order.cpp": []}
HIP_SOURCES = ["bve_sums.hip", "ensemble.hip", "bve_ensemble.hip", "bve_rk4.hip", "swe_sums.hip",
               "swe_rk4.hip", "ic2d.hip", "diagnostics.hip", "noalloc.hip"]
EXAMPLES = ["bve_rotation", "sphere_swe_tc2", "sphere_rh54",
            "sphere_gaussian_vortex"]


def _sources():
    srcs = []
    for name, extra in HOST_SOURCES.items():
        srcs.append((name, [HIPCC, *COMMON, *extra]))
    for name in HIP_SOURCES:
        srcs.append((name, [HIPCC, *COMMON, *HIP_FLAGS]))
    return srcs


def _compile(name, cmd_prefix, verbose):
    src = os.path.join(CSRC, name)
    obj = os.path.join(BUILD, name + ".o")
    cmd = [*cmd_prefix, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    t0 = time.time()
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile of {name} failed:\n{r.stdout}\n{r.stderr}")
    return obj, time.time() - t0, r.stderr


def build(jobs: int = 4, verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, n, c, verbose) for n, c in _sources()]
        for f in futs:
            obj, dt, err = f.result()
            objs.append(obj)
            if verbose:
                print(f"  {os.path.basename(obj)}: {dt:.1f}s", flush=True)
            warn = [l for l in err.splitlines() if "warning" in l
                    and "packed-fp32-ops" not in l and "not a recognized feature" not in l]
            if warn and verbose:
                print("\n".join(warn))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for ex in EXAMPLES:
        _build_example(ex, verbose)
    return LIB


def _build_example(name, verbose=False):
    """C++ drivers over the C++ API (host code only), linked to the library."""
    src = os.path.join(HERE, "examples", name + ".cpp")
    out = os.path.join(BUILD, name)
    cmd = [HIPCC, "-std=c++17", "-O2", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(HERE, 'cpp', 'include')}", src, "-o", out,
           f"-L{HERE}", "-llpm_mi355x", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build of example {name} failed:\n{r.stderr}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    t0 = time.time()
    path = build(a.jobs, a.verbose)
    print(f"built {path} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    sys.exit(main())
