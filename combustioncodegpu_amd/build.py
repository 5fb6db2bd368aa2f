"""Build of the step library libcfx.so for gfx950 (in-tree).

Compiles every source under csrc/ with hipcc into objects under
combustioncodegpu_amd/_build/ (in parallel) and links
combustioncodegpu_amd/libcfx.so.  Called by __graft_entry__.build(); can be
run directly: ``python -m combustioncodegpu_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcfx.so")
ARCH = os.environ.get("CFX_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for every HIP translation unit.
# MFMA results are kept in ordinary VGPRs (the surrogate applies SiLU to
# every accumulator; in AGPRs each one would first need a v_accvgpr_read).
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form"]

# Species counts with compiled kernels (the species loops are unrolled at
# compile time); kernels_ns.hip is built once per count and precision.
SPECIES_COUNTS = (9, 5, 3, 2, 1, 0)
SOURCES = ["cfx_api.cpp", "surrogate.cpp", "mesh.cpp", "deck_io.cpp", "kernels_common.hip", "mlp_kernels.hip",
           "noalloc.hip"]


def _units():
    units = [(f"kernels_ns{n}_f{prec}", "kernels_ns.hip", [f"-DCFX_NS={n}", f"-DCFX_PREC={prec}"])
             for n in SPECIES_COUNTS for prec in (64, 32)]
    units += [(os.path.splitext(s)[0], s, []) for s in SOURCES]
    return units


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.exists(cand) or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(unit, extra):
    name, src, defs = unit
    out = os.path.join(OBJ, name + ".o")
    cmd = [_hipcc(), *COMMON, *defs, *extra, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out, r.stderr


def build(verbose=False, extra=(), jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    units = _units()
    jobs = jobs or min(16, os.cpu_count() or 4, len(units))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, list(extra)), units))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
