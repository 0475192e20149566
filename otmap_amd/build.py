"""Build recipe of libotmap_amd.so (gfx950) -- run by __graft_entry__.build().

Every translation unit under csrc/ is compiled by hipcc for gfx950 into
csrc/_build/*.o and linked into otmap_amd/libotmap_amd.so (in-tree, so the
library travels with the repository).  Flags:
  -ffp-contract=off   keep the reference's rounding (no implicit FMA); kernels
                      that want FMA call fma() explicitly
  -packed-fp32-ops    disabled: no v_pk_* FP32 instructions are generated
  -amdgpu-kernarg-preload-count=16
                      the first 16 dwords of kernel arguments arrive in SGPRs
                      (no kernarg load before the first dependent load); the
                      solve iterations are chains of short dependent launches,
                      1024^2 solves measured 1-2% faster
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libotmap_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "stencil.hip",
    "dct.hip",
    "solver.hip",
    "inverse.hip",
    "raster.hip",
    "dct_mixed.hip",
    "driver.hip",
    "batch.hip",
    "pipeline.hip",
    "bvh_build.hip",
    "fft.hip",
]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


# Kernels whose hand-placed waits (s_waitcnt vmcnt(N) after a known number of
# issued transfers, dct_mixed.hip) are only exact when the compiler adds no
# vector memory operation of its own -- a register spill would add scratch
# loads and stores.  The build fails if any of them reports scratch use.
NO_SCRATCH = {"dct_mixed.hip": ("k_mgemm",)}


def _resource_usage(stderr: str) -> dict:
    """kernel (mangled) name -> scratch bytes per lane, from hipcc's
    -Rpass-analysis=kernel-resource-usage remarks."""
    usage, name = {}, None
    for line in stderr.splitlines():
        if "Function Name:" in line:
            name = line.split("Function Name:")[1].split()[0]
        elif "ScratchSize [bytes/lane]:" in line and name:
            usage[name] = int(line.split("ScratchSize [bytes/lane]:")[1].split()[0])
    return usage


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src in NO_SCRATCH:
        cmd.append("-Rpass-analysis=kernel-resource-usage")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    if src in NO_SCRATCH:
        usage = _resource_usage(r.stderr)
        checked = [k for k in usage if any(p in k for p in NO_SCRATCH[src])]
        if not checked:
            raise RuntimeError(f"{src}: no resource-usage remarks for {NO_SCRATCH[src]}")
        spilled = {k: v for k, v in usage.items() if k in checked and v}
        if spilled:
            raise RuntimeError(f"{src}: kernels with hand-placed vmcnt waits use scratch: {spilled}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
