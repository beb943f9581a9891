"""GPU parity of binning, forward and backward against the CPU oracle."""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import scenes as S
from helpers import DeviceRun, normwise, rand_grads, rounded

pytestmark = pytest.mark.gpu


def _oracle_lists(scene):
    ext = O.template_extents_np(scene.templates)
    return O.bin_lists_c(scene.means2D, scene.radii, scene.rotations, scene.bmp_sel, ext,
                         scene.H, scene.W, scene.tile_size)


@pytest.mark.parametrize("cfg", [0, 1, 3])
def test_binning_bit_exact(dev, cfg):
    sc = S.make_scene(cfg, with_target=False)
    run = DeviceRun(sc, dev)
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
    run = DeviceRun(sc, dev, eps_T=eps)
    col, alp = run.forward()
    osc = O.scene_from(sc, run.offs.cpu().numpy(), run.idx.cpu().numpy())
    rc, ra, _ = O.forward(osc)
    e = max(np.abs(col - rc).max(), np.abs(alp - ra).max())
    print(f"eps {eps}: max abs image error {e:.3e}")
    assert e <= 1e-5


@pytest.mark.parametrize("eps", [0.0, 1e-6])
def test_backward_plain(dev, eps):
    sc = S.make_scene(0, with_target=False)
    run = DeviceRun(sc, dev, eps_T=eps)
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
    run = DeviceRun(sc, dev, eps_T=1e-6)
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


def _hard_scene(**kw):
    """The plain workload with a hard-edged template (binary centre square):
    its bilinear fringe reaches past the binned box, so only the binning's
    tile ranges keep a walk over every primitive equal to the lists."""
    sc = S.make_scene(0, with_target=False, **kw)
    sc.templates = S.hard_square(sc.templates.shape[1])[None]
    return sc


@pytest.mark.parametrize("kind,ts,explicit", [("soft", 16, False), ("hard", 16, False),
                                              ("hard", 16, True), ("hard", 32, False),
                                              ("hard", 32, True)])
def test_overflow_fallback_is_exact(dev, kind, ts, explicit):
    """Binning into a too-small buffer: the raster kernels walk every
    primitive instead, each within its binning tile range, and produce the
    same image and gradients (and the oracle's, for the binned lists).
    ``explicit``: the lists carry their own ranges buffer; otherwise the
    handle's, from its latest binning."""
    import torch
    from diffbmp_amd import _native
    sc = _hard_scene() if kind == "hard" else S.make_scene(0, with_target=False)
    sc.tile_size = ts
    run = DeviceRun(sc, dev, eps_T=1e-6, ts=ts)
    col, alp = run.forward()
    gC, gA = rand_grads(sc.H, sc.W, seed=9)
    g = run.backward(gC, gA)
    osc = O.scene_from(sc, *_oracle_lists(sc))
    rc, ra, _ = O.forward(osc)
    assert max(np.abs(col - rc).max(), np.abs(alp - ra).max()) <= 1e-5
    assert max(normwise(g, O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))).values()) <= 1e-4
    P = run.idx.numel()
    offs = torch.empty_like(run.offs)
    idx = torch.empty(P // 3, dtype=torch.int32, device=dev)
    cnt = torch.empty(1, dtype=torch.int64, device=dev)
    rg = torch.empty((sc.N, 4), dtype=torch.int32, device=dev) if explicit else None
    run.h.bin_async(run.prims, run.ext, tuple(run.d["templates"].shape), P // 3, offs, idx, cnt,
                    1, rg)
    assert int(cnt.item()) == P
    assert int(offs.abs().sum().item()) == 0
    run.offs, run.idx = offs, idx
    lists = _native.Lists(offs, idx, cnt, P // 3, 1, rg)
    H, W = sc.H, sc.W
    c2 = torch.empty((H, W, 3), device=dev); a2 = torch.empty((H, W), device=dev)
    T2 = torch.empty((H, W), device=dev); l2 = torch.empty((H, W), dtype=torch.int32, device=dev)
    run.h.forward(run.prims, run.bm, lists, 1e-6, c2, a2, T2, l2)
    torch.cuda.synchronize()
    assert np.abs(c2.cpu().numpy() - col).max() <= 1e-6
    assert np.abs(a2.cpu().numpy() - alp).max() <= 1e-6
    N = sc.N
    gg = [torch.empty(s, device=dev) for s in ((N, 2), (N,), (N,), (N,), (N, 3))]
    run.h.backward(run.prims, run.bm, lists, 1e-6, torch.from_numpy(gC).to(dev), torch.from_numpy(gA).to(dev),
                   c2, T2, l2, *gg)
    torch.cuda.synchronize()
    for k, t in zip(["means2D", "radii", "rotations", "opacities", "colors"], gg):
        ref = g[k]
        assert np.linalg.norm(t.cpu().numpy() - ref) <= 1e-5 * max(np.linalg.norm(ref), 1e-30), k


@pytest.mark.parametrize("ts", [16, 32])
def test_lazy_binning_overflow_then_grows(dev, ts):
    """Lazy binning with too small a pair buffer (hard-edged template; super-
    tile lists at tile size 16, per-tile lists at 32): every call still gives
    the exact image (the raster kernels walk all primitives within their
    binning ranges), and the capacity grows once an asynchronous read-back
    (one per READBACK_EVERY calls) has seen the count."""
    from diffbmp_amd import _native
    sc = _hard_scene()
    sc.tile_size = ts
    run = DeviceRun(sc, dev, eps_T=1e-6, ts=ts)
    col, alp = run.forward()
    h, shape = run.h, tuple(run.d["templates"].shape)
    P = h.bin(run.prims, run.ext, shape, h.bin_factor)[1].numel()   # pairs of the lists bin_lazy builds
    H, W = sc.H, sc.W
    h.pair_capacity = P // 4
    grown_at = None
    for call in range(1, 3 * _native.Handle.READBACK_EVERY + 1):
        lists = h.bin_lazy(run.prims, run.ext, shape, h.bin_factor)
        c2 = torch.empty((H, W, 3), device=dev); a2 = torch.empty((H, W), device=dev)
        T2 = torch.empty((H, W), device=dev); l2 = torch.empty((H, W), dtype=torch.int32, device=dev)
        h.forward(run.prims, run.bm, lists, 1e-6, c2, a2, T2, l2)
        torch.cuda.synchronize()
        assert np.abs(c2.cpu().numpy() - col).max() <= 1e-6, call
        assert np.abs(a2.cpu().numpy() - alp).max() <= 1e-6, call
        if grown_at is None and h.pair_capacity >= P:
            grown_at = call
    print("capacity grown at call", grown_at, "of", 3 * _native.Handle.READBACK_EVERY)
    assert grown_at is not None and grown_at <= 2 * _native.Handle.READBACK_EVERY + 1
    assert int(lists.count.item()) <= lists.capacity


@pytest.mark.parametrize("H,W,ts", [(100, 70, 16), (96, 160, 32), (17, 33, 16)])
def test_ragged_canvas_and_tile_sizes(dev, H, W, ts):
    sc = S.make_scene(0, H=H, W=W, N=40, tsz=32, with_target=False)
    sc.tile_size = ts
    run = DeviceRun(sc, dev, eps_T=0.0, ts=ts)
    offs, idx = _oracle_lists(sc)
    assert np.array_equal(run.offs.cpu().numpy(), offs) and np.array_equal(run.idx.cpu().numpy(), idx)
    col, alp = run.forward()
    osc = O.scene_from(sc, offs, idx)
    rc, ra, _ = O.forward(osc)
    assert max(np.abs(col - rc).max(), np.abs(alp - ra).max()) <= 1e-5
    gC, gA = rand_grads(H, W, seed=H)
    g = run.backward(gC, gA)
    ref = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    assert max(normwise(g, ref).values()) <= 1e-4


def test_empty_and_degenerate_scenes(dev):
    """No primitives; primitives entirely off canvas; invalid template index;
    tiny radius below the 1e-6 reciprocal guard."""
    import copy
    sc = S.make_scene(0, H=64, W=64, N=12, tsz=16, with_target=False)
    sc.means2D[:4] = [[-500, 10], [10, -400], [900, 20], [20, 900]]
    sc.bmp_sel = sc.bmp_sel.copy()
    sc.bmp_sel[5] = 3             # only one template: invalid, never drawn
    sc.radii[6] = 1e-7
    run = DeviceRun(sc, dev, eps_T=0.0)
    col, alp = run.forward()
    offs, idx = _oracle_lists(sc)
    assert np.array_equal(run.offs.cpu().numpy(), offs) and np.array_equal(run.idx.cpu().numpy(), idx)
    osc = O.scene_from(sc, offs, idx)
    rc, ra, _ = O.forward(osc)
    assert max(np.abs(col - rc).max(), np.abs(alp - ra).max()) <= 1e-5
    gC, gA = rand_grads(64, 64, seed=1)
    g = run.backward(gC, gA)
    ref = O.backward(osc, gC.astype(np.float64), gA.astype(np.float64))
    assert max(normwise(g, ref).values()) <= 1e-4
    assert np.all(g["opacities"][:4] == 0) and np.all(g["colors"][5] == 0)
    empty = copy.copy(sc)
    for k in ("means2D", "radii", "rotations", "opacities", "colors", "bmp_sel"):
        setattr(empty, k, getattr(sc, k)[:0])
    run0 = DeviceRun(empty, dev)
    col0, alp0 = run0.forward()
    assert run0.idx.numel() == 0 and col0.max() == 0 and alp0.max() == 0


def test_repeat_forward_is_bit_identical(dev):
    """The second forward of a handle runs its workgroups heaviest first (the
    order measured by the first); every pixel's result is unchanged."""
    sc = S.make_scene(0, with_target=False)
    run = DeviceRun(sc, dev, eps_T=1e-6)
    c1, a1 = run.forward()
    T1, l1 = run.T.cpu().numpy(), run.last.cpu().numpy()
    gC, gA = rand_grads(sc.H, sc.W, seed=3)
    g1 = run.backward(gC, gA)
    c2, a2 = run.forward()
    assert np.array_equal(c1, c2) and np.array_equal(a1, a2)
    assert np.array_equal(T1, run.T.cpu().numpy()) and np.array_equal(l1, run.last.cpu().numpy())
    g2 = run.backward(gC, gA)
    for k in g1:
        assert np.linalg.norm(g2[k] - g1[k]) <= 1e-6 * max(np.linalg.norm(g1[k]), 1e-30), k
