"""Writes the golden fixtures in this folder from the FP64 oracle.

The reference ships no golden vectors for this path, so these fixtures are the
oracle's own outputs on small seeded scenes, frozen so that any later change
to the oracle or to the HIP path is caught against fixed data.  Each .npz
holds the inputs (parameters, templates, template selection, optional colour
maps and blend), the tile lists, the image (premultiplied colour and alpha),
the upstream gradients and the parameter gradients.

    python tests/golden/make_golden.py
"""
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from diffbmp_amd import scenes as S  # noqa: E402


def scene_cases():
    cases = {}
    cases["plain_small"] = S.make_scene(0, H=64, W=64, N=24, tsz=32, with_target=False)
    sc = S.make_scene(1, H=96, W=80, N=40, tsz=32, with_target=False)
    cases["multi_template_ragged"] = sc
    sc = S.make_scene(3, H=64, W=64, N=30, tsz=32, with_target=False)
    sc.radii = (sc.radii * 6).astype(np.float32)
    sc.templates = sc.templates[:4].copy()               # four RGBA bitmaps
    sc.colors_orig_tmpl = sc.colors_orig_tmpl[:4].copy()
    sc.bmp_sel = S.cyclic_sel(sc.N, 4)
    cases["colour_map_blend"] = sc
    sc = S.make_scene(0, H=48, W=48, N=12, tsz=16, with_target=False)
    sc.means2D = sc.means2D.copy()
    sc.means2D[:3] = [[-40, 10], [10, -30], [60, 70]]
    sc.bmp_sel = sc.bmp_sel.copy()
    sc.bmp_sel[4] = 7
    sc.radii = sc.radii.copy()
    sc.radii[5] = 1e-7
    sc.opacities = sc.opacities.copy()
    sc.opacities[6] = 30.0          # sigmoid rounds to 1 in fp32: an opaque layer
    cases["edge_cases"] = sc
    return cases


def make(name, sc):
    offs, idx = O.bin_scene(sc)
    osc = O.scene_from(sc, offs, idx)
    col, alp, cnt = O.forward(osc)
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    # upstream gradients are stored as fp32: derive the expected values from those
    gC = rng.standard_normal((sc.H, sc.W, 3)).astype(np.float32)
    gA = rng.standard_normal((sc.H, sc.W)).astype(np.float32)
    g = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    co = sc.colors_orig_tmpl if sc.colors_orig_tmpl is not None else np.zeros((0,), np.float32)
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        H=sc.H, W=sc.W, tile_size=sc.tile_size, alpha_ub=sc.alpha_ub, c_blend=sc.c_blend,
        means2D=sc.means2D, radii=sc.radii, rotations=sc.rotations, opacities=sc.opacities,
        colors=sc.colors, templates=sc.templates, bmp_sel=sc.bmp_sel, colors_orig_tmpl=co,
        tile_offsets=offs, tile_indices=idx, color=col.astype(np.float32),
        alpha=alp.astype(np.float32), push_count=cnt, gC=gC.astype(np.float32),
        gA=gA.astype(np.float32),
        **{"g_" + k: v for k, v in g.items()})


if __name__ == "__main__":
    for n, s in scene_cases().items():
        make(n, s)
        print("wrote", n)
