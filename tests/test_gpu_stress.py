"""Stress workload on the GPU: 2048^2 canvas, 32 768 primitives, 16 RGBA
256^2 bitmaps stored as fp16 (alpha channel = template, RGB = colour map,
per template), colour-map blend c_blend = 0.5, parameters fp32.

* binning is bit-exact against the oracle (lists from the fp16 bitmaps);
* forward and backward match the fp64 oracle on sampled tiles, with the
  reference's colour gradient and with the exact one;
* the exact gradient matches central differences of the fp64 oracle
  (h = 1e-3 and a small step) on 64 random scalars, on the same scene with
  its bitmaps windowed to an exactly-zero margin.
"""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import scenes as S
from diffbmp_amd._native import Handle
from helpers import DeviceRun, fd_scalars, fd_window, half_rounded, normwise, rand_grads

pytestmark = pytest.mark.gpu

N_SAMPLED = 24


def _stress_scene(window=False):
    sc = S.make_scene(3, with_target=False)
    tm = fd_window(sc.templates) if window else sc.templates
    sc.templates = half_rounded(tm)               # the GPU stores fp16 bitmaps
    sc.colors_orig_tmpl = half_rounded(sc.colors_orig_tmpl)
    return sc


def _gpu(sc, dev, eps, exact):
    run = DeviceRun(sc, dev, eps_T=eps, co_mode=2, bm_dtype=torch.float16)
    run.bm = Handle.bitmaps(run.d["templates"], run.co, 2, sc.c_blend, -1, exact)
    return run


@pytest.fixture(scope="module")
def stress(dev):
    sc = _stress_scene()
    run = _gpu(sc, dev, 1e-6, False)
    col, alp = run.forward()
    offs, idx = run.offs.cpu().numpy(), run.idx.cpu().numpy()
    rng = np.random.default_rng(31)
    tX = sc.W // 16
    tiles = np.sort(rng.choice(tX * tX, N_SAMPLED, replace=False)).astype(np.int32)
    mask = np.zeros((tX, tX), bool)
    mask.flat[tiles] = True
    pm = np.kron(mask, np.ones((16, 16), bool))
    return sc, run, col, alp, offs, idx, tiles, pm


def test_stress_binning_bit_exact(stress):
    sc, run, _, _, offs, idx, _, _ = stress
    ext = O.template_extents_np(sc.templates)
    assert np.array_equal(run.ext.cpu().numpy(), ext)
    r_offs, r_idx = O.bin_lists_c(sc.means2D, sc.radii, sc.rotations, sc.bmp_sel, ext,
                                  sc.H, sc.W, sc.tile_size)
    print(f"stress: P = {len(r_idx)} pairs, mean list {len(r_idx) / (len(r_offs) - 1):.1f}")
    assert np.array_equal(offs, r_offs)
    assert np.array_equal(idx, r_idx)


def test_stress_forward_sampled_tiles(stress):
    sc, run, col, alp, offs, idx, tiles, pm = stress
    rc, ra, _ = O.forward(O.scene_from(sc, offs, idx, tile_subset=tiles))
    e = max(np.abs(col - rc)[pm].max(), np.abs(alp - ra)[pm].max())
    print(f"stress sampled-tile image error {e:.3e}")
    assert e <= 1e-5


@pytest.mark.parametrize("exact", [False, True])
def test_stress_backward_sampled_tiles(stress, exact):
    sc, run, _, _, offs, idx, tiles, pm = stress
    run.bm = Handle.bitmaps(run.d["templates"], run.co, 2, sc.c_blend, -1, exact)
    gC, gA = rand_grads(sc.H, sc.W, seed=33, scale=1e-3)
    gC[~pm] = 0
    gA[~pm] = 0
    g = run.backward(gC, gA)
    osc = O.scene_from(sc, offs, idx, tile_subset=tiles, blend_grad_exact=exact)
    ref = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    err = normwise(g, ref)
    print(f"stress exact={exact}: sampled-tile normwise gradient error", err)
    assert max(err.values()) <= 1e-4


@pytest.fixture(scope="module")
def stress_fd(dev):
    sc = _stress_scene(window=True)
    run = _gpu(sc, dev, 0.0, True)
    run.forward()
    gC, gA = rand_grads(sc.H, sc.W, seed=35)
    g = run.backward(gC, gA)
    offs, idx = run.offs.cpu().numpy(), run.idx.cpu().numpy()
    rng = np.random.default_rng(36)
    keys = [("means2D", 2), ("radii", 1), ("rotations", 1), ("opacities", 1), ("colors", 3)]
    valid = np.nonzero(np.isin(np.arange(sc.N), np.unique(idx)))[0]
    picks = []
    for i in range(64):
        key, w = keys[i % 5]
        picks.append((key, int(rng.choice(valid)), int(rng.integers(w))))
    return sc, offs, idx, g, gC.astype(np.float64), gA.astype(np.float64), picks


def test_stress_fd_h1e3(stress_fd):
    """h = 1e-3: normwise over the 64 scalars.  Bilinear slope jumps across
    texel lines cost O(h) per crossing, and a 5 px primitive maps 1e-3 px to
    0.025 texel of a 256^2 bitmap, so this step carries a few per cent of
    noise (2.3e-2 for these scalars with the fp64 oracle's own gradient).
    Leaving out the colour map's geometry term moves these gradients by 69%."""
    sc, offs, idx, g, gC, gA, picks = stress_fd
    fd, an = fd_scalars(sc, offs, idx, g, gC, gA, picks, 1e-3, blend_grad_exact=True)
    err = np.linalg.norm(fd - an) / np.linalg.norm(an)
    print(f"stress FD h=1e-3 normwise error {err:.3e}")
    assert err <= 5e-2


def test_stress_fd_small_step(stress_fd):
    """h = 1e-5 (at least 2 ulp of the fp32 parameter): the difference to the
    analytic gradient shrinks with the step, as it must for the derivative
    (1.0e-3 for these scalars with the fp64 oracle's own gradient)."""
    sc, offs, idx, g, gC, gA, picks = stress_fd
    fd, an = fd_scalars(sc, offs, idx, g, gC, gA, picks, 1e-5, blend_grad_exact=True)
    err = np.linalg.norm(fd - an) / np.linalg.norm(an)
    print(f"stress FD small-step normwise error {err:.3e}")
    assert err <= 3e-3
