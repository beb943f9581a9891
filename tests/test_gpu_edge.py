"""Edge cases of the raster path against the oracle: no primitives, every
primitive off the canvas, canvases smaller than one tile, and tile sizes the
kernels cannot take."""
import copy

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle as O
from diffbmp_amd import _native
from diffbmp_amd import scenes as S
from diffbmp_amd.fit import FitJob
from diffbmp_amd.tile_rasterizer import TileRasterizer
from helpers import DeviceRun, normwise, rand_grads

pytestmark = pytest.mark.gpu

PARAMS = ("means2D", "radii", "rotations", "opacities", "colors")


def subset(sc, keep):
    """The scene restricted to the primitives in ``keep`` (index array)."""
    s = copy.copy(sc)
    for k in PARAMS + ("bmp_sel",):
        setattr(s, k, np.ascontiguousarray(getattr(sc, k)[keep]))
    return s


def test_no_primitives_autograd_and_fit(dev):
    """N = 0: the image is empty (colour 0, alpha 0), the five gradients are
    empty tensors, and a fit step composites the white background only."""
    sc = subset(S.make_scene(0, H=48, W=80, with_target=False), np.arange(0))
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    leaves = [t(getattr(sc, k)).requires_grad_() for k in PARAMS]
    rast = TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, 0)
    empty = torch.empty(0, device=dev)
    color, alpha = rast(*leaves, empty, t(sc.templates), t(sc.bmp_sel), 0.0, torch.zeros(7, device=dev))
    assert color.shape == (sc.H, sc.W, 3) and alpha.shape == (sc.H, sc.W)
    assert color.abs().max().item() == 0.0 and alpha.abs().max().item() == 0.0
    (color.sum() + alpha.sum()).backward()
    for k, leaf in zip(PARAMS, leaves):
        assert leaf.grad is not None and leaf.grad.shape == leaf.shape, k
    target = torch.rand((sc.H, sc.W, 3), device=dev, generator=torch.Generator(dev).manual_seed(2))
    job = FitJob(*[t(getattr(sc, k)) for k in PARAMS], t(sc.templates), t(sc.bmp_sel), target,
                 lr=1e-2, max_steps=4)
    job.step()
    job.step()
    losses = job.loss_history().cpu().numpy()
    ref = F.mse_loss(torch.ones_like(target), target).item()
    assert np.allclose(losses, ref, rtol=1e-5), (losses, ref)


def test_primitives_off_canvas(dev):
    """Half the primitives moved far off the canvas: they are in no tile list
    (the lists equal the oracle's), and image and gradients equal the oracle's
    for the rest; with all of them off, the lists are empty and the image is
    empty."""
    sc = S.make_scene(0, with_target=False)
    off = copy.copy(sc)
    off.means2D = sc.means2D.copy()
    far = np.arange(0, sc.N, 2)
    off.means2D[far] += np.array([4.0 * sc.W, -3.0 * sc.H], np.float32)
    run = DeviceRun(off, dev, eps_T=1e-6)
    ext = O.template_extents_np(off.templates)
    offs, idx = O.bin_lists_c(off.means2D, off.radii, off.rotations, off.bmp_sel, ext, off.H, off.W,
                              off.tile_size)
    assert np.array_equal(run.offs.cpu().numpy(), offs) and np.array_equal(run.idx.cpu().numpy(), idx)
    assert not np.isin(idx, far).any()
    col, alp = run.forward()
    osc = O.scene_from(off, offs, idx)
    rc, ra, _ = O.forward(osc)
    assert max(np.abs(col - rc).max(), np.abs(alp - ra).max()) <= 1e-5
    gC, gA = rand_grads(off.H, off.W, seed=4)
    g = run.backward(gC, gA)
    ref = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    assert max(normwise(g, ref).values()) <= 1e-4
    for k in PARAMS:
        assert np.abs(g[k][far]).max() == 0.0, k
    gone = copy.copy(sc)
    gone.means2D = sc.means2D + np.array([4.0 * sc.W, 0.0], np.float32)
    run = DeviceRun(gone, dev, eps_T=1e-6)
    assert run.idx.numel() == 0 and int(run.offs[-1].item()) == 0
    col, alp = run.forward()
    assert np.abs(col).max() == 0.0 and np.abs(alp).max() == 0.0


@pytest.mark.parametrize("H,W", [(5, 11), (16, 1), (1, 40), (17, 9)])
def test_canvas_smaller_than_a_block(dev, H, W):
    """Canvases narrower or shorter than one 16x16 block (and a one-pixel
    row and column): lists, image and gradients against the oracle."""
    sc = S.make_scene(0, H=H, W=W, N=40, with_target=False)
    sc.means2D = (np.random.default_rng(8).uniform(-0.2, 1.2, (sc.N, 2)) *
                  np.array([W, H])).astype(np.float32)
    sc.radii = np.full(sc.N, max(H, W) / 3.0, np.float32)
    run = DeviceRun(sc, dev, eps_T=0.0)
    ext = O.template_extents_np(sc.templates)
    offs, idx = O.bin_lists_c(sc.means2D, sc.radii, sc.rotations, sc.bmp_sel, ext, H, W, sc.tile_size)
    assert np.array_equal(run.offs.cpu().numpy(), offs) and np.array_equal(run.idx.cpu().numpy(), idx)
    col, alp = run.forward()
    osc = O.scene_from(sc, offs, idx)
    rc, ra, _ = O.forward(osc)
    assert max(np.abs(col - rc).max(), np.abs(alp - ra).max()) <= 1e-5
    gC, gA = rand_grads(H, W, seed=5)
    g = run.backward(gC, gA)
    ref = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    assert max(normwise(g, ref).values()) <= 1e-4


@pytest.mark.parametrize("ts", [8, 24, 0, -16])
def test_tile_sizes_the_kernels_cannot_take(dev, ts):
    """Tile sizes that are not a positive multiple of 16 raise; nothing is
    launched."""
    with pytest.raises((ValueError, _native.NativeError)):
        TileRasterizer(64, 64, ts, 0.0, 1.0, 151, 10)


@pytest.mark.parametrize("dtype,exact", [(torch.float16, False), (torch.float16, True),
                                         (torch.bfloat16, False), (torch.float32, True)])
def test_colour_map_samples_on_the_last_texel(dev, dtype, exact):
    """Samples that land exactly on a bitmap's last column and row (u = 1 or
    v = 1: r = 4, theta = 0, integer centres) read the right neighbour past the
    bitmap (alpha 0) and the colour map's clamp-to-edge neighbours: the
    padding of the RGBA records.  Images and gradients (both colour-gradient
    modes) against the oracle on the same storage-rounded inputs."""
    from helpers import rounded
    rng = np.random.default_rng(21)
    Ht = Wt = 5
    N = 6
    means = np.array([[10, 10], [30, 12], [19, 27], [41, 31], [10, 36], [33, 40]], np.float32)
    T = 2
    # hw = Wt/2 * r/(Wt/2) = r: the push box reaches exactly the u = +-1 pixels
    templates = np.stack([np.ones((Ht, Wt)), 0.25 + 0.75 * rng.random((Ht, Wt))]).astype(np.float32)
    sc = S.Scene(H=48, W=52, means2D=means, radii=np.full(N, 4.0, np.float32),
                 rotations=np.zeros(N, np.float32),
                 opacities=rng.normal(size=N).astype(np.float32),
                 colors=rng.normal(size=(N, 3)).astype(np.float32), templates=templates,
                 bmp_sel=(np.arange(N) % T).astype(np.int32), c_blend=0.5,
                 colors_orig_tmpl=rng.random((T, Ht, Wt, 3)).astype(np.float32))
    sc = rounded(sc, dtype) if dtype != torch.float32 else sc
    run = DeviceRun(sc, dev, dtype=dtype, eps_T=0.0, co_mode=2, exact_blend=exact)
    color, alpha = run.forward()
    offs, idx = O.bin_scene(sc)
    osc = O.scene_from(sc, offs, idx, blend_grad_exact=exact)
    rc, ra, _ = O.forward(osc)
    # the edge pixels are pushed: the last column / row of each primitive is drawn
    assert ra[10, 14] > 0.1 and ra[14, 10] > 0.1
    assert np.abs(color - rc).max() <= 1e-5 and np.abs(alpha - ra).max() <= 1e-5
    gC, gA = rand_grads(sc.H, sc.W, seed=22)
    g = run.backward(gC, gA)
    err = normwise(g, O.backward(osc, gC.astype(np.float64), gA.astype(np.float64)))
    assert max(err.values()) <= 1e-4, err
