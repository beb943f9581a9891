"""Push masks (dbmp_tile_lists.push_masks): the forward records, per 64-entry
batch of a list, 16x16 block and wave, which entries pushed a pixel of the
wave; the backward of that forward walks only those entries instead of
culling.  Images and gradients must be those of the backward that culls for
itself (and of the oracle)."""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import _native
from diffbmp_amd import scenes as S
from helpers import DeviceRun, normwise, rand_grads

pytestmark = pytest.mark.gpu

NAMES = ["means2D", "radii", "rotations", "opacities", "colors"]


def _bin(run, factor, cap):
    """Lists binned into a buffer of `cap` entries, with a push-mask buffer."""
    h, dev, sc = run.h, run.dev, run.scene
    offs = torch.empty(h.list_count(factor) + 1, dtype=torch.int32, device=dev)
    idx = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
    cnt = torch.empty(1, dtype=torch.int64, device=dev)
    rg = torch.empty((max(sc.N, 1), 4), dtype=torch.int32, device=dev)
    h.bin_async(run.prims, run.ext, tuple(run.d["templates"].shape), cap, offs, idx, cnt, factor, rg)
    nw = int(h._L.dbmp_push_mask_words(h._h, cap, factor))
    # filled with a pattern the forward must overwrite wherever it is read
    pm = torch.full((nw,), -1, dtype=torch.int64, device=dev)
    return _native.Lists(offs, idx, cnt, cap, factor, rg, pm)


def _without_masks(L):
    return _native.Lists(L.offsets, L.indices, L.count, L.capacity, L.factor, L.ranges, None)


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
    return col.cpu().numpy(), alp.cpu().numpy(), {k: t.cpu().numpy() for k, t in zip(NAMES, gg)}


def _check_same(a, b):
    ca, aa, ga = a
    cb, ab, gb = b
    assert np.array_equal(ca, cb)
    assert np.array_equal(aa, ab)
    for k in ga:
        den = max(np.linalg.norm(ga[k]), 1e-30)
        # the same per-entry sums; only the order of cross-workgroup atomics differs
        assert np.linalg.norm(ga[k] - gb[k]) <= 1e-5 * den, k


def _hard_ring_scene(**kw):
    """Templates whose content leaves holes and corners inside the content box
    (a ring and a hard centre square): many culled entries push no pixel."""
    sc = S.make_scene(0, with_target=False, **kw)
    t = sc.templates.shape[1]
    yy, xx = np.mgrid[0:t, 0:t].astype(np.float64)
    rho = np.hypot(xx - (t - 1) / 2, yy - (t - 1) / 2)
    ring = ((rho > 0.25 * t) & (rho < 0.4 * t)).astype(np.float32)
    sc.templates = np.stack([ring, S.hard_square(t)]).astype(np.float32)
    sc.bmp_sel = (np.arange(sc.N)[::-1] % 2).astype(np.int32)
    return sc


@pytest.mark.parametrize("kind,factor,eps", [("plain", 1, 1e-6), ("plain", 4, 1e-6), ("plain", 4, 0.0),
                                             ("hard", 1, 1e-6), ("hard", 4, 1e-6), ("scale", 4, 1e-6)])
def test_push_masks_equal_culling(dev, kind, factor, eps):
    if kind == "scale":
        sc = S.make_scene(1, H=512, W=512, N=2048, with_target=False)
    elif kind == "hard":
        sc = _hard_ring_scene()
    else:
        sc = S.make_scene(0, with_target=False)
    run = DeviceRun(sc, dev, eps_T=eps)
    gC, gA = rand_grads(sc.H, sc.W, seed=11)
    P = run.idx.numel() if factor == 1 else \
        run.h.bin(run.prims, run.ext, tuple(run.d["templates"].shape), factor)[1].numel()
    L = _bin(run, factor, P + 100)
    got = _raster(run, L, gC, gA)
    words = L.masks.cpu().numpy()
    written = words[words != -1]
    print(f"{kind} f={factor} eps={eps}: {P} pairs, {written.size} mask words written, "
          f"{int(sum(bin(int(w) & (2**64 - 1)).count('1') for w in written))} entries marked")
    assert written.size > 0 and np.any(written != 0)
    ref = _raster(run, _without_masks(L), gC, gA)
    _check_same(ref, got)
    if kind != "scale":
        offs, idx = O.bin_scene(sc)
        osc = O.scene_from(sc, offs, idx)
        rc, ra, _ = O.forward(osc)   # the reference's full walk: early termination is within 1e-5
        assert np.abs(got[0] - rc).max() <= 1e-5 and np.abs(got[1] - ra).max() <= 1e-5
        g = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
        err = normwise(got[2], g)
        assert max(err.values()) <= 1e-4, err


def test_push_masks_blend_rgba(dev):
    """Colour-map blending (RGBA texel records, fp16 bitmaps), super-tile lists."""
    sc = S.make_scene(3, H=192, W=160, N=300, tsz=64, with_target=False)
    run = DeviceRun(sc, dev, eps_T=1e-6, co_mode=2, bm_dtype=torch.float16)
    gC, gA = rand_grads(sc.H, sc.W, seed=12)
    P = run.h.bin(run.prims, run.ext, tuple(run.d["templates"].shape), 4)[1].numel()
    L = _bin(run, 4, P + 64)
    got = _raster(run, L, gC, gA)
    _check_same(_raster(run, _without_masks(L), gC, gA), got)


def test_push_masks_unused_after_overflow(dev):
    """A pair overflow walks every primitive: the forward writes no masks and
    the backward culls, with the same results as complete lists."""
    sc = _hard_ring_scene()
    run = DeviceRun(sc, dev, eps_T=1e-6)
    gC, gA = rand_grads(sc.H, sc.W, seed=13)
    P = run.h.bin(run.prims, run.ext, tuple(run.d["templates"].shape), 4)[1].numel()
    full = _bin(run, 4, P + 64)
    ref = _raster(run, full, gC, gA)
    small = _bin(run, 4, P // 3)
    got = _raster(run, small, gC, gA)
    assert int(small.count.item()) == P
    assert np.all(small.masks.cpu().numpy() == -1)   # untouched
    _check_same(ref, got)


def test_push_masks_need_a_capacity(dev):
    """Lists without a pair count (a caller's complete mapping) carry no
    capacity to size the masks by: the mask buffer is ignored."""
    sc = S.make_scene(0, with_target=False)
    run = DeviceRun(sc, dev, eps_T=1e-6)
    gC, gA = rand_grads(sc.H, sc.W, seed=14)
    pm = torch.full((1 << 16,), -1, dtype=torch.int64, device=dev)
    L = _native.Lists(run.offs, run.idx, masks=pm)
    got = _raster(run, L, gC, gA)
    assert np.all(pm.cpu().numpy() == -1)
    _check_same(_raster(run, _native.Lists(run.offs, run.idx), gC, gA), got)
