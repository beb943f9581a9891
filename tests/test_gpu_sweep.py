"""Seeded random sweep against the oracle.

Each case draws one configuration and checks the images and all five
gradient groups against the oracle on the same (storage-rounded) inputs. A
case draws:
- a canvas side in [20, 150) px, mostly not a multiple of the tile;
- 0, 1 or 5 primitives (one case in eight each), else 20-400, with radii
  scaled by U[0.3, 3], rotations in [-7, 13) rad, and 15 % of the centres
  anywhere in a 40 px margin around the canvas;
- 8, 16, 33 or 64 texel bitmaps, a third of them with a hard-edged square as
  template 0;
- tile size 16 or 32, and early termination on or off;
- the workload family: plain disk, four shapes, or RGBA colour maps with
  c_blend 0.5 under either gradient mode.

fp32 cases go through the drop-in ``TileRasterizer`` under torch autograd. That
path bins lazily on the GPU: super-tile lists and push masks for tile size 16,
per-tile lists with binning ranges for 32. fp16 and mixed-storage cases go
through the C ABI with per-tile lists, colour maps per primitive or per
template.

Bars, as in the fixed-scene tests:
- images within 1e-5 of the oracle's full walk;
- gradients within 1e-4 normwise.
"""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import scenes as S
from diffbmp_amd import tile_rasterizer as TR
from helpers import DeviceRun, normwise, rand_grads, rounded

pytestmark = pytest.mark.gpu

CASES = 32
NAMES = ("means2D", "radii", "rotations", "opacities", "colors")


def random_case(seed):
    """(scene, options) of sweep case ``seed`` (numpy draws)."""
    rng = np.random.default_rng(10_000 + seed)
    cfg = int(rng.choice([0, 1, 3]))
    H, W = int(rng.integers(20, 150)), int(rng.integers(20, 150))
    tsz = int(rng.choice([8, 16, 33, 64]))
    n = int(rng.choice([0, 1, 5] + [int(rng.integers(20, 400))] * 5))
    sc = S.make_scene(cfg, H=H, W=W, N=max(n, 1), tsz=tsz, with_target=False)
    cut = lambda a: np.ascontiguousarray(a[:n])
    sc.means2D, sc.radii, sc.rotations = cut(sc.means2D), cut(sc.radii), cut(sc.rotations)
    sc.opacities, sc.colors, sc.bmp_sel = cut(sc.opacities), cut(sc.colors), cut(sc.bmp_sel)
    if n:
        sc.radii = (sc.radii * rng.uniform(0.3, 3.0, n)).astype(np.float32)
        sc.rotations = rng.uniform(-7.0, 13.0, n).astype(np.float32)
        off = rng.random(n) < 0.15
        sc.means2D[off] = rng.uniform(-40.0, max(H, W) + 40.0, (int(off.sum()), 2)).astype(np.float32)
    if rng.random() < 1 / 3:
        sc.templates = sc.templates.copy()
        sc.templates[0] = S.hard_square(tsz)
    sc.tile_size = int(rng.choice([16, 32]))
    opts = dict(cfg=cfg, eps=float(rng.choice([0.0, 1e-6])),
                storage=str(rng.choice(["f32", "f32", "f16", "mixed"])),
                exact=bool(rng.random() < 0.5), co_mode=int(rng.choice([1, 2])) if cfg == 3 else 0)
    if opts["storage"] == "f32":
        opts["co_mode"] = 1 if cfg == 3 else 0   # the drop-in takes the per-primitive gather
    return sc, opts


def _oracle(sc, opts, gC, gA):
    offs, idx = O.bin_scene(sc)
    co = sc.colors_orig_per_prim() if opts["co_mode"] == 1 else None
    osc = O.scene_from(sc, offs, idx, colors_orig=co, blend_grad_exact=opts["exact"])
    rc, ra, _ = O.forward(osc)
    return rc, ra, O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))


def _dropin(sc, opts, dev, gC, gA):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TR.TileRasterizer(sc.H, sc.W, sc.tile_size, 0.0, sc.alpha_ub, 151, sc.N)
    rast.eps_T = opts["eps"]
    rast.exact_color_map_grad = opts["exact"]
    P = [t(a).requires_grad_() for a in (sc.means2D, sc.radii, sc.rotations, sc.opacities, sc.colors)]
    co = t(sc.colors_orig_per_prim()) if opts["co_mode"] == 1 else torch.empty(0, device=dev)
    color, alpha = rast(*P, co, t(sc.templates), t(sc.bmp_sel), sc.c_blend, torch.zeros(7, device=dev))
    ((color * t(gC)).sum() + (alpha * t(gA)).sum()).backward()
    return (color.detach().cpu().numpy(), alpha.detach().cpu().numpy(),
            {k: (p.grad if p.grad is not None else torch.zeros_like(p)).cpu().numpy()
             for k, p in zip(NAMES, P)})


def _capi(sc, opts, dev, gC, gA):
    run = DeviceRun(sc, dev, dtype=torch.float16, eps_T=opts["eps"], ts=sc.tile_size,
                 co_mode=opts["co_mode"], exact_blend=opts["exact"],
                 fp32_geometry=opts["storage"] == "mixed")
    color, alpha = run.forward()
    return color, alpha, run.backward(gC, gA)


@pytest.mark.parametrize("seed", range(CASES))
def test_random_case_matches_oracle(dev, seed):
    sc, opts = random_case(seed)
    print(f"case {seed}: {sc.H}x{sc.W} N={sc.N} tmpl={sc.templates.shape} ts={sc.tile_size} {opts}",
          flush=True)
    gC, gA = rand_grads(sc.H, sc.W, seed=seed)
    if opts["storage"] == "f32":
        got = _dropin(sc, opts, dev, gC, gA)
        ref_sc = sc
    else:
        got = _capi(sc, opts, dev, gC, gA)
        ref_sc = rounded(sc, torch.float16, fp32_geometry=opts["storage"] == "mixed")
    rc, ra, g = _oracle(ref_sc, opts, gC, gA)
    assert np.abs(got[0] - rc).max() <= 1e-5
    assert np.abs(got[1] - ra).max() <= 1e-5
    if sc.N:
        err = normwise(got[2], g)
        assert max(err.values()) <= 1e-4, err
