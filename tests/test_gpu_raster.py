"""GPU parity of binning, forward and backward against the CPU oracle."""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import scenes as S
from helpers import GpuRun, normwise, rand_grads, rounded

pytestmark = pytest.mark.gpu


def _oracle_lists(scene):
    ext = O.template_extents_np(scene.templates)
    return O.bin_lists_c(scene.means2D, scene.radii, scene.rotations, scene.bmp_sel, ext,
                         scene.H, scene.W, scene.tile_size)


@pytest.mark.parametrize("cfg", [0, 1, 3])
def test_binning_bit_exact(dev, cfg):
    sc = S.make_scene(cfg, with_target=False)
    run = GpuRun(sc, dev)
    ext_ref = O.template_extents_np(sc.templates)
    assert np.array_equal(run.ext.cpu().numpy(), ext_ref)
    offs, idx = _oracle_lists(sc)
    g_offs, g_idx = run.offs.cpu().numpy(), run.idx.cpu().numpy()
    print(f"cfg {cfg}: P = {len(idx)} pairs")
    assert np.array_equal(g_offs, offs)
    assert np.array_equal(g_idx, idx)


@pytest.mark.parametrize("eps", [0.0, 1e-6])
def test_forward_plain(dev, eps):
    sc = S.make_scene(0, with_target=False)
    run = GpuRun(sc, dev, eps_T=eps)
    col, alp = run.forward()
    osc = O.scene_from(sc, run.offs.cpu().numpy(), run.idx.cpu().numpy())
    rc, ra, _ = O.forward(osc)
    e = max(np.abs(col - rc).max(), np.abs(alp - ra).max())
    print(f"eps {eps}: max abs image error {e:.3e}")
    assert e <= 1e-5


@pytest.mark.parametrize("eps", [0.0, 1e-6])
def test_backward_plain(dev, eps):
    sc = S.make_scene(0, with_target=False)
    run = GpuRun(sc, dev, eps_T=eps)
    run.forward()
    gC, gA = rand_grads(sc.H, sc.W, seed=3)
    g = run.backward(gC, gA)
    osc = O.scene_from(sc, run.offs.cpu().numpy(), run.idx.cpu().numpy())
    ref = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    err = normwise(g, ref)
    print("normwise gradient error", err)
    assert max(err.values()) <= 1e-4


def test_scale_sampled_tiles(dev):
    """Full 1024^2 / 8192-primitive workload; oracle on a sample of tiles."""
    sc = S.make_scene(1, with_target=False)
    run = GpuRun(sc, dev, eps_T=1e-6)
    col, alp = run.forward()
    rng = np.random.default_rng(7)
    tiles = np.sort(rng.choice(64 * 64, 24, replace=False)).astype(np.int32)
    osc = O.scene_from(sc, run.offs.cpu().numpy(), run.idx.cpu().numpy(), tile_subset=tiles)
    rc, ra, _ = O.forward(osc)
    mask = np.zeros((64, 64), bool)
    mask.flat[tiles] = True
    pm = np.kron(mask, np.ones((16, 16), bool))
    e = max(np.abs(col - rc)[pm].max(), np.abs(alp - ra)[pm].max())
    print(f"sampled-tile image error {e:.3e}")
    assert e <= 1e-5
    gC, gA = rand_grads(sc.H, sc.W, seed=5, scale=1e-3)
    gC[~pm] = 0
    gA[~pm] = 0
    g = run.backward(gC, gA)
    ref = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    err = normwise(g, ref)
    print("sampled-tile normwise gradient error", err)
    assert max(err.values()) <= 1e-4
