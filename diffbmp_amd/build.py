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


def build(verbose: bool = False) -> str:
    """Compile the HIP sources into ``diffbmp_amd/libdiffbmp_amd.so``."""
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *SRCS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
