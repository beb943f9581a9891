"""Build ``libdiffbmp_amd.so`` in-tree with hipcc for gfx950."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", f) for f in ("capi.hip", "batch.hip", "psd.hip")]
OUT = os.path.join(HERE, "libdiffbmp_amd.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # exact fp32 decision geometry: no implicit FMA contraction
    "-ffp-contract=off",
    # hardware fp32 atomics on coarse-grained (hipMalloc / torch) memory
    "-munsafe-fp-atomics",
    # no packed-fp32 instruction selection
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no SLP vectorisation: it packs the per-lane push predicates into
    # small integer vectors (VALU shifts and compares) instead of lane masks
    "-fno-slp-vectorize",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


OUT_LANE_STATS = os.path.join(HERE, "libdiffbmp_amd_lanestats.so")


def build(verbose: bool = False, lane_stats: bool = False) -> str:
    """Compile the HIP sources into ``diffbmp_amd/libdiffbmp_amd.so``, or with
    ``lane_stats`` the profiling build ``libdiffbmp_amd_lanestats.so`` whose
    raster kernels count pushed lanes per (wave, list entry) (load it with
    DBMP_LIBRARY)."""
    out = OUT_LANE_STATS if lane_stats else OUT
    tmp = out + ".tmp"
    extra = ["-DDBMP_LANE_STATS"] if lane_stats else []
    cmd = [hipcc(), *HIPCC_FLAGS, *extra, "-o", tmp, *SRCS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    import sys
    print(build(verbose=True, lane_stats="--lane-stats" in sys.argv))
