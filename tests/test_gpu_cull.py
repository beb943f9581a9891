"""The raster kernels skip (wave, entry) pairs whose rotated template content
box misses the wave's pixels.  These scenes make that cull do most of the
work -- small off-centre content, texels just above and just below the 1e-6
mask threshold at the content edge, every rotation -- and check the image and
the gradients against the oracle, which walks every list entry."""
import numpy as np
import pytest

import oracle as O
from diffbmp_amd import scenes as S
from helpers import DeviceRun, normwise, rand_grads

pytestmark = pytest.mark.gpu


def _templates(tsz, rng):
    T = 4
    tm = np.zeros((T, tsz, tsz), np.float32)
    # 0: a dot in one corner (content box much smaller than the square)
    tm[0] = S.shape_bitmap("disk", tsz, rng, scale=0.2, offset=(3.5, -3.5))
    # 1: a bar along one edge, with a ring of texels straddling the threshold
    tm[1, 2:6, 4:tsz - 4] = 1.0
    tm[1, 1, 4:tsz - 4] = 1.01e-6      # above 1e-6: part of the content box
    tm[1, 6, 4:tsz - 4] = 1.0e-6       # at 1e-6: outside it, never pushed alone
    # 2: an off-centre ring
    tm[2] = S.shape_bitmap("ring", tsz, rng, scale=0.4, offset=(-1.2, 1.0))
    # 3: empty (no texel above the threshold)
    tm[3, :, :] = 5e-7
    return tm


@pytest.mark.parametrize("seed", [0, 1])
def test_content_cull_exact(dev, seed):
    rng = np.random.default_rng(seed)
    H, W, N, tsz = 96, 128, 300, 32
    sc = S.make_scene(0, H=H, W=W, N=N, tsz=tsz, with_target=False)
    sc.templates = _templates(tsz, rng)
    sc.bmp_sel = (np.arange(N) % 4).astype(np.int32)
    sc.rotations = rng.uniform(-np.pi, np.pi, N).astype(np.float32)
    sc.radii = np.exp(rng.uniform(np.log(3.0), np.log(40.0), N)).astype(np.float32)
    sc.means2D = np.stack([rng.uniform(-10, W + 10, N), rng.uniform(-10, H + 10, N)], 1).astype(np.float32)
    run = DeviceRun(sc, dev, eps_T=0.0)
    col, alp = run.forward()
    osc = O.scene_from(sc, run.offs.cpu().numpy(), run.idx.cpu().numpy())
    rc, ra, _ = O.forward(osc)
    e = max(np.abs(col - rc).max(), np.abs(alp - ra).max())
    print(f"seed {seed}: {len(run.idx)} pairs, max abs image error {e:.3e}")
    assert e <= 1e-5
    gC, gA = rand_grads(H, W, seed=seed + 7)
    g = run.backward(gC, gA)
    ref = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    err = normwise(g, ref)
    print("normwise gradient error", err)
    assert max(err.values()) <= 1e-4
