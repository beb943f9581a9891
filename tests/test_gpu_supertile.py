"""Super-tile binning (dbmp_bin_tiles): lists per f x f block of 16-px tiles,
filtered per 16-px tile at staging by the binning's tile ranges.  The raster
results must be those of the per-tile lists."""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import _native
from diffbmp_amd import scenes as S
from helpers import DeviceRun, rand_grads

pytestmark = pytest.mark.gpu


def _raster(run, lists, gC, gA):
    sc, dev = run.scene, run.dev
    H, W, N = sc.H, sc.W, sc.N
    col = torch.empty((H, W, 3), device=dev)
    alp = torch.empty((H, W), device=dev)
    T = torch.empty((H, W), device=dev)
    last = torch.empty((H, W), dtype=torch.int32, device=dev)
    run.h.forward(run.prims, run.bm, lists, run.eps, col, alp, T, last)
    gg = [torch.empty(s, device=dev) for s in ((N, 2), (N,), (N,), (N,), (N, 3))]
    run.h.backward(run.prims, run.bm, lists, run.eps, torch.from_numpy(gC).to(dev),
                   torch.from_numpy(gA).to(dev), col, T, last, *gg)
    torch.cuda.synchronize()
    names = ["means2D", "radii", "rotations", "opacities", "colors"]
    return col.cpu().numpy(), alp.cpu().numpy(), T.cpu().numpy(), \
        {k: t.cpu().numpy() for k, t in zip(names, gg)}


def _check_same(a, b):
    ca, aa, Ta, ga = a
    cb, ab, Tb, gb = b
    assert np.array_equal(ca, cb)
    assert np.array_equal(aa, ab)
    assert np.array_equal(Ta, Tb)
    for k in ga:
        den = max(np.linalg.norm(ga[k]), 1e-30)
        # same per-pixel terms and the same per-workgroup sums; only the order
        # of the cross-workgroup float atomics can differ
        assert np.linalg.norm(ga[k] - gb[k]) <= 1e-5 * den, k


@pytest.mark.parametrize("cfg,H,W,factor", [
    (0, None, None, 4), (0, None, None, 2), (0, None, None, 8),
    (0, 100, 70, 4), (0, 17, 33, 8), (0, 200, 330, 4), (1, None, None, 4)])
def test_supertile_lists_match_per_tile(dev, cfg, H, W, factor):
    kw = {} if H is None else dict(H=H, W=W, N=120, tsz=32)
    sc = S.make_scene(cfg, with_target=False, **kw)
    run = DeviceRun(sc, dev, eps_T=1e-6)
    gC, gA = rand_grads(sc.H, sc.W, seed=5)
    ref = _raster(run, _native.Lists(run.offs, run.idx), gC, gA)
    offs, idx, rg = run.h.bin(run.prims, run.ext, tuple(run.d["templates"].shape), factor)
    # the super-tile lists are the per-tile binning at tile size 16 * factor
    o_offs, o_idx = O.bin_lists_c(sc.means2D, sc.radii, sc.rotations, sc.bmp_sel,
                                  O.template_extents_np(sc.templates), sc.H, sc.W, 16 * factor)
    assert np.array_equal(offs.cpu().numpy(), o_offs)
    assert np.array_equal(idx.cpu().numpy(), o_idx)
    print(f"cfg {cfg} {sc.H}x{sc.W} f={factor}: {run.idx.numel()} -> {idx.numel()} pairs")
    got = _raster(run, _native.Lists(offs, idx, factor=factor, ranges=rg), gC, gA)
    _check_same(ref, got)


def test_supertile_blend_rgba(dev):
    """Colour-map blending with RGBA texel records (fp16 bitmaps)."""
    sc = S.make_scene(3, H=192, W=160, N=300, tsz=64, with_target=False)
    run = DeviceRun(sc, dev, eps_T=1e-6, co_mode=2, bm_dtype=torch.float16)
    gC, gA = rand_grads(sc.H, sc.W, seed=6)
    ref = _raster(run, _native.Lists(run.offs, run.idx), gC, gA)
    offs, idx, rg = run.h.bin(run.prims, run.ext, tuple(run.d["templates"].shape), 4)
    _check_same(ref, _raster(run, _native.Lists(offs, idx, factor=4, ranges=rg), gC, gA))


def test_supertile_overflow_fallback(dev):
    """Super-tile binning into a too-small buffer: the raster kernels walk
    every primitive, filtered by the tile ranges, with the same results."""
    sc = S.make_scene(0, with_target=False)
    run = DeviceRun(sc, dev, eps_T=1e-6)
    gC, gA = rand_grads(sc.H, sc.W, seed=7)
    ref = _raster(run, _native.Lists(run.offs, run.idx), gC, gA)
    _, full, _ = run.h.bin(run.prims, run.ext, tuple(run.d["templates"].shape), 4)
    P = full.numel()
    offs = torch.empty(run.h.list_count(4) + 1, dtype=torch.int32, device=dev)
    idx = torch.empty(P // 3, dtype=torch.int32, device=dev)
    cnt = torch.empty(1, dtype=torch.int64, device=dev)
    rg = torch.empty((sc.N, 4), dtype=torch.int32, device=dev)
    run.h.bin_async(run.prims, run.ext, tuple(run.d["templates"].shape), P // 3, offs, idx, cnt,
                    4, rg)
    assert int(cnt.item()) == P
    _check_same(ref, _raster(run, _native.Lists(offs, idx, cnt, P // 3, 4, rg), gC, gA))


def test_supertile_argument_checks(dev):
    sc = S.make_scene(0, H=64, W=64, N=10, tsz=32, with_target=False)
    run = DeviceRun(sc, dev)
    shape = tuple(run.d["templates"].shape)
    offs = torch.empty(64, dtype=torch.int32, device=dev)
    idx = torch.empty(1024, dtype=torch.int32, device=dev)
    cnt = torch.empty(1, dtype=torch.int64, device=dev)
    rg = torch.empty((sc.N, 4), dtype=torch.int32, device=dev)
    with pytest.raises(RuntimeError, match="bin_factor"):
        run.h.bin_async(run.prims, run.ext, shape, 1024, offs, idx, cnt, 3, rg)
    with pytest.raises(RuntimeError, match="ranges"):
        run.h.bin_async(run.prims, run.ext, shape, 1024, offs, idx, cnt, 4, None)
    out = [torch.empty((64, 64, 3), device=dev)] + [torch.empty((64, 64), device=dev)] * 2 + \
        [torch.empty((64, 64), dtype=torch.int32, device=dev)]
    with pytest.raises(RuntimeError, match="ranges"):
        run.h.forward(run.prims, run.bm, _native.Lists(offs, idx, cnt, 1024, 4, None), 0.0, *out)
