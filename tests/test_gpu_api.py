"""Reference-compatible API on the GPU: TileRasterizer / autograd Functions."""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import scenes as S
from diffbmp_amd import tile_rasterizer as TR
from helpers import normwise

pytestmark = pytest.mark.gpu


def _leaves(sc, dev):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return [t(a).requires_grad_() for a in (sc.means2D, sc.radii, sc.rotations, sc.opacities, sc.colors)]


def test_autograd_matches_oracle_and_mapping_roundtrip(dev):
    sc = S.make_scene(0, with_target=False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TR.TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N)
    rast.eps_T = 0.0
    lr = torch.zeros(7, device=dev)
    P = _leaves(sc, dev)
    color, alpha = rast(*P, torch.empty(0, device=dev), t(sc.templates), t(sc.bmp_sel).long(), 0.0, lr)
    assert color.shape == (sc.H, sc.W, 3) and alpha.shape == (sc.H, sc.W)
    # the reference's composite over a white background and an MSE loss
    target = torch.rand((sc.H, sc.W, 3), device=dev, generator=torch.Generator(dev).manual_seed(0))
    out = color + (1 - alpha.unsqueeze(-1)) * torch.ones_like(target)
    loss = torch.nn.functional.mse_loss(out, target)
    loss.backward()
    offs, idx = O.bin_scene(sc)
    osc = O.scene_from(sc, offs, idx)
    rc, ra, _ = O.forward(osc)
    assert np.abs(color.detach().cpu().numpy() - rc).max() <= 1e-5
    # closed-form dL/dC and dL/dA of the composite + MSE
    res = rc + (1 - ra)[..., None] - target.cpu().numpy()
    gC = 2 * res / res.size
    gA = -(gC).sum(-1)
    ref = O.backward(osc, gC, gA)
    got = dict(means2D=P[0].grad.cpu().numpy(), radii=P[1].grad.cpu().numpy(),
               rotations=P[2].grad.cpu().numpy(), opacities=P[3].grad.cpu().numpy(),
               colors=P[4].grad.cpu().numpy())
    assert max(normwise(got, ref).values()) <= 1e-4
    # a packed mapping (reference wire format) gives the same result
    mapping = TR.pack_mapping(torch.from_numpy(offs).to(dev), torch.from_numpy(idx).to(dev))
    P2 = _leaves(sc, dev)
    c2, a2 = TR.TileRasterizerFunction.apply(*P2, torch.empty(0, device=dev), t(sc.templates),
                                             t(sc.bmp_sel), 0.0, lr, mapping, sc.H, sc.W, 16, 0.0, True)
    assert torch.equal(c2, color.detach()) and torch.equal(a2, alpha.detach())


def test_fp16_function_returns_fp16_and_accepts_fp32_params(dev):
    sc = S.make_scene(0, with_target=False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TR.TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N, use_fp16=True)
    P = _leaves(sc, dev)
    color, alpha = rast(*P, torch.empty(0, device=dev), t(sc.templates), t(sc.bmp_sel), 0.0,
                        torch.zeros(7, device=dev))
    assert color.dtype == torch.float16 and alpha.dtype == torch.float16
    (color.float().sum() + alpha.float().sum()).backward()
    assert all(p.grad is not None and p.grad.dtype == torch.float32 for p in P)


def _batch_case(sc, dev, B, seed, with_maps):
    """B perturbed copies of a scene's parameters (and per-candidate colour
    maps) through TileRasterizerBatchFunction; -> params, outputs, grads."""
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rng = np.random.default_rng(seed)
    P = [np.stack([a + rng.normal(0, s, a.shape).astype(np.float32) for _ in range(B)])
         for a, s in ((sc.means2D, 3.0), (sc.radii, 0.5), (sc.rotations, 0.3), (sc.opacities, 0.5),
                      (sc.colors, 0.5))]
    P[1] = np.abs(P[1]) + 1.0
    leaves = [t(a).requires_grad_() for a in P]
    T, Ht, Wt = sc.templates.shape
    co = rng.uniform(0, 1, (B, sc.N, Ht, Wt, 3)).astype(np.float32) if with_maps else None
    c_blend = 0.4 if with_maps else 0.0
    rast = TR.TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N)
    rast.eps_T = 0.0
    lr = torch.zeros(7, device=dev)
    cb, ab = rast.forward_batch(*leaves, t(co) if with_maps else torch.empty(0, device=dev),
                                t(sc.templates), t(sc.bmp_sel), c_blend, lr)
    gC = rng.standard_normal((B, sc.H, sc.W, 3)).astype(np.float32)
    gA = rng.standard_normal((B, sc.H, sc.W)).astype(np.float32)
    ((cb * t(gC)).sum() + (ab * t(gA)).sum()).backward()
    return P, co, c_blend, cb.detach().cpu().numpy(), ab.detach().cpu().numpy(), gC, gA, leaves


@pytest.mark.parametrize("with_maps", [False, True])
def test_batch_function_matches_oracle_per_candidate(dev, with_maps):
    """Every candidate of one batched launch equals the oracle on that
    candidate's own parameters (image and gradients); the reference's batch
    backward reuses the last candidate's cache for all (D3)."""
    sc = S.make_scene(0, H=96, W=112, N=40, tsz=32, with_target=False)
    B = 3
    P, co, cbl, cb, ab, gC, gA, leaves = _batch_case(sc, dev, B, 4, with_maps)
    names = ("means2D", "radii", "rotations", "opacities", "colors")
    for b in range(B):
        import copy
        s2 = copy.copy(sc)
        for k, a in zip(names, P):
            setattr(s2, k, a[b])
        s2.c_blend = cbl
        offs, idx = O.bin_scene(s2)
        osc = O.scene_from(s2, offs, idx, colors_orig=co[b] if with_maps else None)
        rc, ra, _ = O.forward(osc)
        e = max(np.abs(cb[b] - rc).max(), np.abs(ab[b] - ra).max())
        ref = O.backward(osc, gC[b].astype(np.float64), gA[b].astype(np.float64))
        got = {k: leaf.grad[b].cpu().numpy() for k, leaf in zip(names, leaves)}
        err = normwise(got, ref)
        print(f"candidate {b}: image {e:.2e}, gradients {max(err.values()):.2e}")
        assert e <= 1e-5 and max(err.values()) <= 1e-4, (b, e, err)


def test_batch_function_beyond_one_launch(dev):
    """35 candidates: two launches of at most 32; each candidate equals its
    single-candidate render and gradients."""
    sc = S.make_scene(0, H=48, W=64, N=10, tsz=16, with_target=False)
    B = 35
    P, _, _, cb, ab, gC, gA, leaves = _batch_case(sc, dev, B, 9, False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TR.TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N)
    rast.eps_T = 0.0
    lr = torch.zeros(7, device=dev)
    for b in (0, 17, 31, 32, 34):
        single = [t(a[b]).requires_grad_() for a in P]
        cs, as_ = rast(*single, torch.empty(0, device=dev), t(sc.templates), t(sc.bmp_sel), 0.0, lr)
        ((cs * t(gC[b])).sum() + (as_ * t(gA[b])).sum()).backward()
        assert np.array_equal(cs.detach().cpu().numpy(), cb[b]), b
        assert np.array_equal(as_.detach().cpu().numpy(), ab[b]), b
        for pb, ps in zip(leaves, single):
            assert torch.allclose(pb.grad[b], ps.grad, rtol=1e-5, atol=1e-6), b


def test_errors_raise(dev):
    sc = S.make_scene(0, H=32, W=32, N=4, tsz=16, with_target=False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TR.TileRasterizer(32, 32, 16, 0.0, 1.0, 151, 4)
    P = _leaves(sc, dev)
    with pytest.raises(ValueError):
        TR.TileRasterizerFunction.apply(*P, torch.empty(0, device=dev), t(sc.templates), t(sc.bmp_sel), 0.0,
                                        torch.zeros(7, device=dev), None, 64, 64, 16, 0.0, True)
    with pytest.raises(RuntimeError):
        rast(*[p.detach().cpu() for p in P], torch.empty(0), torch.from_numpy(sc.templates),
             torch.from_numpy(sc.bmp_sel), 0.0, torch.zeros(7))
    with pytest.raises(Exception):
        TR.TileRasterizer(32, 32, 24)   # tile size must be a multiple of 16


def test_interleaved_calls_recompute_replaced_records(dev):
    """Two forwards on one rasterizer before either backward: the second
    replaces the handle's primitive records, so the first call's backward must
    prepare its own again.  Its gradients equal those of a lone forward and
    backward of the same parameters."""
    sc = S.make_scene(0, with_target=False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TR.TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N)
    lr = torch.zeros(7, device=dev)
    args = (torch.empty(0, device=dev), t(sc.templates), t(sc.bmp_sel).long(), 0.0, lr)
    target = torch.rand((sc.H, sc.W, 3), device=dev, generator=torch.Generator(dev).manual_seed(1))

    def loss(color, alpha):
        return torch.nn.functional.mse_loss(color + (1 - alpha.unsqueeze(-1)), target)

    A = _leaves(sc, dev)
    loss(*rast(*A, *args)).backward()
    lone = [p.grad.clone() for p in A]
    A2 = _leaves(sc, dev)
    B = _leaves(sc, dev)
    with torch.no_grad():
        B[0].add_(3.0)          # different parameters: different records
        B[2].add_(0.5)
    la = loss(*rast(*A2, *args))
    lb = loss(*rast(*B, *args))
    la.backward()
    lb.backward()
    for g, p in zip(lone, A2):
        err = ((p.grad - g).norm() / g.norm().clamp_min(1e-30)).item()
        assert err <= 1e-5, err


@pytest.mark.parametrize("ts", [32, 16])
def test_hard_template_overflow_through_autograd(dev, ts):
    """TileRasterizer at the reference's default tile size 32 (per-tile
    lists) and at 16 (super-tile lists), hard-edged template, lazy binning
    forced to overflow its pair buffer, and a second forward of other
    parameters before the first backward: every primitive is walked within
    its own call's binning ranges, so image and gradients equal the oracle's
    on the binned lists."""
    sc = S.make_scene(0, with_target=False)
    sc.templates = S.hard_square(sc.templates.shape[1])[None]
    sc.tile_size = ts
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TR.TileRasterizer(sc.H, sc.W, ts, 0.0, 1.0, 151, sc.N)
    rast.eps_T = 0.0
    lr = torch.zeros(7, device=dev)
    args = (torch.empty(0, device=dev), t(sc.templates), t(sc.bmp_sel), 0.0, lr)
    rast(*_leaves(sc, dev), *args)          # sizes the pair buffer
    torch.cuda.synchronize()
    rast.handle.pair_capacity = 64          # the next calls overflow
    A = _leaves(sc, dev)
    cA, aA = rast(*A, *args)
    B = _leaves(sc, dev)
    with torch.no_grad():
        B[0].add_(5.0)
        B[2].add_(0.7)
    cB, aB = rast(*B, *args)
    assert int(rast.handle.pair_capacity) == 64
    gC, gA = (torch.from_numpy(a).to(dev) for a in
              (np.random.default_rng(2).standard_normal((sc.H, sc.W, 3)).astype(np.float32),
               np.random.default_rng(3).standard_normal((sc.H, sc.W)).astype(np.float32)))
    ((cA * gC).sum() + (aA * gA).sum()).backward()
    ((cB * gC).sum() + (aB * gA).sum()).backward()
    ext = O.template_extents_np(sc.templates)
    offs, idx = O.bin_lists_c(sc.means2D, sc.radii, sc.rotations, sc.bmp_sel, ext, sc.H, sc.W, ts)
    osc = O.scene_from(sc, offs, idx)
    rc, ra, _ = O.forward(osc)
    e = max(np.abs(cA.detach().cpu().numpy() - rc).max(), np.abs(aA.detach().cpu().numpy() - ra).max())
    print(f"ts {ts}: image error {e:.2e}")
    assert e <= 1e-5
    ref = O.backward(osc, gC.cpu().numpy().astype(np.float64), gA.cpu().numpy().astype(np.float64))
    got = dict(means2D=A[0].grad.cpu().numpy(), radii=A[1].grad.cpu().numpy(),
               rotations=A[2].grad.cpu().numpy(), opacities=A[3].grad.cpu().numpy(),
               colors=A[4].grad.cpu().numpy())
    err = normwise(got, ref)
    print("gradient error", err)
    assert max(err.values()) <= 1e-4
