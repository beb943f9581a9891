"""Fused device step (loss kernel, Adam kernel, FitJob) against torch."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from diffbmp_amd import scenes as S
from diffbmp_amd._native import Handle
from diffbmp_amd.fit import FitJob, GROUPS
from diffbmp_amd.tile_rasterizer import TileRasterizer

pytestmark = pytest.mark.gpu


def test_loss_kernel_matches_autograd(dev):
    g = torch.Generator(dev).manual_seed(1)
    H, W = 64, 48
    color = torch.rand((H, W, 3), device=dev, generator=g).requires_grad_()
    alpha = torch.rand((H, W), device=dev, generator=g).requires_grad_()
    target = torch.rand((H, W, 3), device=dev, generator=g)
    for bg in (None, torch.rand((H, W, 3), device=dev, generator=g)):
        b = torch.ones_like(target) if bg is None else bg
        loss = F.mse_loss(color + (1 - alpha).unsqueeze(-1) * b, target)
        gc, ga = torch.autograd.grad(loss, (color, alpha))
        h = Handle(H, W, 16)
        gC = torch.empty_like(target)
        gA = torch.empty((H, W), device=dev)
        acc = torch.zeros(1, device=dev)
        h.loss_mse(color.detach(), alpha.detach(), target, gC, gA, acc, bg=bg)
        torch.cuda.synchronize()
        assert abs(acc.item() - loss.item()) <= 1e-6 * loss.item()
        assert torch.allclose(gC, gc, rtol=1e-5, atol=1e-10)
        assert torch.allclose(gA, ga, rtol=1e-5, atol=1e-10)


@pytest.mark.parametrize("with_bg", [False, True])
def test_fused_backward_mse_matches_two_kernels(dev, with_bg):
    """dbmp_backward_mse equals dbmp_loss_mse followed by dbmp_backward."""
    from helpers import DeviceRun
    sc = S.make_scene(0, with_target=False)
    run = DeviceRun(sc, dev, eps_T=1e-6)
    run.forward()
    g = torch.Generator(dev).manual_seed(7)
    target = torch.rand((sc.H, sc.W, 3), device=dev, generator=g)
    bg = torch.rand((sc.H, sc.W, 3), device=dev, generator=g) if with_bg else None
    gC = torch.empty_like(target)
    gA = torch.empty((sc.H, sc.W), device=dev)
    l1 = torch.zeros(1, device=dev)
    run.h.loss_mse(run.color, run.alpha, target, gC, gA, l1, bg=bg)
    ref = run.backward(gC.cpu().numpy(), gA.cpu().numpy())
    N = sc.N
    out = [torch.empty(sh, device=dev) for sh in ((N, 2), (N,), (N,), (N,), (N, 3))]
    l2 = torch.zeros(1, device=dev)
    from diffbmp_amd import _native
    run.h.backward_mse(run.prims, run.bm, _native.Lists(run.offs, run.idx), run.eps, target, run.color,
                       run.T, run.last, *out, loss_accum=l2, bg=bg)
    torch.cuda.synchronize()
    assert abs(l1.item() - l2.item()) <= 1e-5 * l1.item()   # fp32 sums in another order
    for k, o in zip(("means2D", "radii", "rotations", "opacities", "colors"), out):
        o = o.cpu().numpy()
        err = np.linalg.norm(o - ref[k]) / np.linalg.norm(ref[k])
        assert err <= 1e-6, (k, err)


def test_backward_rows_accumulate_and_adam_clears_them(dev):
    """dbmp_backward_mse_rows adds the (N,8) gradient rows that
    dbmp_backward_mse unpacks; an Adam step with zero_grad reads the rows as
    the unpacked arrays and leaves them zero."""
    from helpers import DeviceRun
    from diffbmp_amd import _native
    sc = S.make_scene(0, with_target=False)
    run = DeviceRun(sc, dev, eps_T=1e-6)
    run.forward()
    target = torch.rand((sc.H, sc.W, 3), device=dev, generator=torch.Generator(dev).manual_seed(8))
    N = sc.N
    lists = _native.Lists(run.offs, run.idx)
    out = [torch.empty(sh, device=dev) for sh in ((N, 2), (N,), (N,), (N,), (N, 3))]
    run.h.backward_mse(run.prims, run.bm, lists, run.eps, target, run.color, run.T, run.last, *out)
    ref = torch.cat([out[0], out[1][:, None], out[2][:, None], out[3][:, None], out[4]], 1)
    rows = torch.zeros((N, 8), device=dev)
    run.h.backward_mse_rows(run.prims, run.bm, lists, run.eps, target, run.color, run.T, run.last, rows)
    torch.cuda.synchronize()
    assert torch.linalg.norm(rows - ref) <= 1e-6 * torch.linalg.norm(ref)
    run.h.backward_mse_rows(run.prims, run.bm, lists, run.eps, target, run.color, run.T, run.last, rows)
    torch.cuda.synchronize()
    assert torch.linalg.norm(rows - 2 * ref) <= 1e-6 * torch.linalg.norm(ref)
    # Adam over the row columns == Adam over the unpacked arrays; rows cleared
    cols = {"means2D": (0, 2), "radii": (2, 1), "rotations": (3, 1), "opacities": (4, 1), "colors": (5, 3)}
    params = [torch.rand(o.shape, device=dev) + 1.0 for o in out]
    results = []
    for use_rows in (False, True):
        adam = _native.Adam()
        adam.num_groups, adam.beta1, adam.beta2, adam.eps, adam.step = 5, 0.9, 0.999, 1e-8, 1
        adam.store_dtype = _native.DBMP_F32
        keep, ps = [], []
        for i, (k, o) in enumerate(zip(cols, out)):
            p = params[i].clone()
            m, v = torch.zeros_like(p), torch.zeros_like(p)
            keep += [m, v]
            ps.append(p)
            g = adam.groups[i]
            g.param, g.exp_avg, g.exp_avg_sq = p.data_ptr(), m.data_ptr(), v.data_ptr()
            g.numel, g.lr, g.stride, g.clamp = p.numel(), 0.01, 2 if k == "means2D" else 1, 0
            if use_rows:
                g.grad = rows.data_ptr() + 4 * cols[k][0]
                g.grad_row, g.grad_comps, g.zero_grad = 8, cols[k][1], 1
            else:
                keep.append((2 * o).contiguous())
                g.grad = keep[-1].data_ptr()
        run.h.adam_step(adam, dev)
        torch.cuda.synchronize()
        results.append([q.clone() for q in ps])
    for a, b in zip(*results):
        assert torch.allclose(a, b, rtol=1e-6, atol=1e-7)
    assert rows.abs().max().item() == 0.0


def test_fitjob_tracks_autograd_adam(dev):
    sc = S.make_scene(0, with_target=False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    target = torch.rand((sc.H, sc.W, 3), device=dev, generator=torch.Generator(dev).manual_seed(3))
    init = [t(a) for a in (sc.means2D, sc.radii, sc.rotations, sc.opacities, sc.colors)]
    job = FitJob(*init, t(sc.templates), t(sc.bmp_sel), target, lr=1e-2, eps_T=1e-6)
    rast = TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N)
    leaves = [a.clone().requires_grad_() for a in init]
    opt = torch.optim.Adam(leaves, lr=1e-2)
    empty = torch.empty(0, device=dev)
    lr_conf = torch.zeros(7, device=dev)
    ref_losses = []
    steps = 5
    for _ in range(steps):
        job.step()
        opt.zero_grad()
        c, a = rast(*leaves, empty, t(sc.templates), t(sc.bmp_sel), 0.0, lr_conf)
        loss = F.mse_loss(c + (1 - a).unsqueeze(-1), target)
        loss.backward()
        opt.step()
        with torch.no_grad():
            leaves[0][:, 0].clamp_(0, sc.W - 1)
            leaves[0][:, 1].clamp_(0, sc.H - 1)
            leaves[1].clamp_(2, sc.W // 4)
        ref_losses.append(loss.item())
    got = job.loss_history().cpu().numpy()
    assert np.allclose(got, ref_losses, rtol=1e-5), (got, ref_losses)
    for k, leaf in zip(GROUPS, leaves):
        d = (job.params[k] - leaf.detach()).abs()
        # Adam's first steps move each scalar by ~lr regardless of |g|; a scalar
        # whose gradient is at rounding level may step the other way
        frac_far = (d > 1e-4).float().mean().item()
        assert frac_far <= 0.01, (k, frac_far)


@pytest.mark.parametrize("store", [torch.float16, torch.bfloat16])
def test_fitjob_reduced_precision_runs(dev, store):
    sc = S.make_scene(0, with_target=False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    target = torch.full((sc.H, sc.W, 3), 0.5, device=dev)
    init = [t(a) for a in (sc.means2D, sc.radii, sc.rotations, sc.opacities, sc.colors)]
    job = FitJob(*init, t(sc.templates), t(sc.bmp_sel), target, lr=1e-2, store_dtype=store)
    for _ in range(10):
        job.step()
    l = job.loss_history().cpu().numpy()
    assert np.all(np.isfinite(l)) and l[-1] < l[0]
    for k in GROUPS:
        assert torch.equal(job.store[k], job.params[k].to(store))


@pytest.mark.parametrize("store", [torch.float16, torch.bfloat16])
def test_fitjob_mixed_storage_tracks_fp32(dev, store):
    """Mixed storage (fp32 positions, radii, rotations; half-precision
    opacities, colours, bitmaps): the half copies follow the fp32 masters, the
    geometry is the master itself, and the first image is within the storage
    rounding of the fp32 job's."""
    sc = S.make_scene(0, with_target=False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    target = torch.full((sc.H, sc.W, 3), 0.5, device=dev)
    init = [t(a) for a in (sc.means2D, sc.radii, sc.rotations, sc.opacities, sc.colors)]
    mixed = FitJob(*init, t(sc.templates), t(sc.bmp_sel), target, lr=1e-2, store_dtype=store,
                   fp32_geometry=True)
    ref = FitJob(*init, t(sc.templates), t(sc.bmp_sel), target, lr=1e-2)
    mixed.step()
    ref.step()
    torch.cuda.synchronize()
    e = (mixed.color - ref.color).abs().max().item()
    print(f"{store}: first image vs fp32 {e:.2e}")
    assert e <= (2e-3 if store == torch.float16 else 1.6e-2)
    for _ in range(9):
        mixed.step()
    for k in ("means2D", "radii", "rotations"):
        assert mixed.store[k] is mixed.params[k]
    for k in ("opacities", "colors"):
        assert torch.equal(mixed.store[k], mixed.params[k].to(store))
    l = mixed.loss_history().cpu().numpy()
    assert np.all(np.isfinite(l)) and l[-1] < l[0]


def c0_target(dev):
    """The plain workload's target: the render over white of a second random
    256-primitive parameter set (role-1 seed) through the drop-in rasterizer."""
    sc = S.make_scene(0, with_target=False)
    tp = S.make_target_params(0, sc.H, sc.W, sc.N)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rast = TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N)
    with torch.no_grad():
        c, a = rast(*[t(x) for x in tp], torch.empty(0, device=dev), t(sc.templates), t(sc.bmp_sel),
                    0.0, torch.zeros(7, device=dev))
    return sc, (c + (1 - a).unsqueeze(-1)).contiguous()


def test_c0_end_to_end_fit(dev):
    """The plain workload end to end: 100 Adam steps (lr 1e-2) of MSE to the
    second random render, as the device-resident FitJob step and as the
    reference loop (drop-in TileRasterizer under torch autograd,
    torch.optim.Adam, the reference's clamp), step by step."""
    sc, target = c0_target(dev)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    init = [t(a) for a in (sc.means2D, sc.radii, sc.rotations, sc.opacities, sc.colors)]
    steps = 100
    job = FitJob(*init, t(sc.templates), t(sc.bmp_sel), target, lr=1e-2, eps_T=1e-6, max_steps=steps)
    rast = TileRasterizer(sc.H, sc.W, 16, 0.0, 1.0, 151, sc.N)
    leaves = [a.clone().requires_grad_() for a in init]
    opt = torch.optim.Adam(leaves, lr=1e-2)
    empty, lr_conf = torch.empty(0, device=dev), torch.zeros(7, device=dev)
    ref = []
    for _ in range(steps):
        job.step()
        opt.zero_grad()
        c, a = rast(*leaves, empty, t(sc.templates), t(sc.bmp_sel), 0.0, lr_conf)
        loss = F.mse_loss(c + (1 - a).unsqueeze(-1), target)
        loss.backward()
        opt.step()
        with torch.no_grad():
            leaves[0][:, 0].clamp_(0, sc.W - 1)
            leaves[0][:, 1].clamp_(0, sc.H - 1)
            leaves[1].clamp_(2, sc.W // 4)
        ref.append(loss.item())
    got = job.loss_history().cpu().numpy()
    ref = np.array(ref)
    dev_rel = np.abs(got - ref) / ref
    print(f"loss {ref[0]:.5f} -> {ref[-1]:.5f}; max relative deviation {dev_rel.max():.2e}")
    assert ref[-1] < 0.6 * ref[0]
    assert dev_rel.max() <= 1e-3, dev_rel.max()
    # Adam steps a scalar by about lr whatever the size of its gradient, so a
    # scalar whose gradient sits at rounding level may walk either way: the
    # parameters agree in the bulk, a few drift by up to steps * lr.  (The
    # plain workload's template is a radially symmetric disk: every rotation
    # gradient is rounding noise, and the rotations random-walk in both loops.)
    for k, leaf in zip(GROUPS, leaves):
        d = (job.params[k] - leaf.detach()).abs().flatten()
        med, far = d.median().item(), (d > 0.25).float().mean().item()
        print(f"{k}: median |diff| {med:.2e}, fraction > 0.25: {far:.4f}, max {d.max().item():.3f}")
        assert med <= 5e-3 and far <= 0.02, k


@pytest.mark.parametrize("with_bg", [False, True])
def test_composite_mse_loss_matches_torch(dev, with_bg):
    """The drop-in fused loss op equals torch's composite + F.mse_loss, value
    and gradients, also through an fp16 rasterizer output and a scaled loss."""
    from diffbmp_amd.tile_rasterizer import composite_mse_loss
    g = torch.Generator(dev).manual_seed(11)
    # 48x80: the float4 path; 7x9 (63 pixels): float4 groups plus a 3-pixel
    # tail; the offset view: pointers off 16-byte alignment, the scalar path
    for H, W, dt, offset in ((48, 80, torch.float32, 0), (48, 80, torch.float16, 0),
                             (7, 9, torch.float32, 0), (48, 80, torch.float32, 1)):
        h = Handle(H, W, 16)
        target = torch.rand((H, W, 3), device=dev, generator=g)
        bg = torch.rand((H, W, 3), device=dev, generator=g) if with_bg else None
        if offset:
            buf = torch.rand(H * W * 3 + offset, device=dev, generator=g)
            c0 = buf[offset:].view(H, W, 3)
            assert c0.data_ptr() % 16 != 0
        else:
            c0 = torch.rand((H, W, 3), device=dev, generator=g).to(dt)
        a0 = torch.rand((H, W), device=dev, generator=g).to(dt)
        c1, a1 = c0.clone().requires_grad_(), a0.clone().requires_grad_()
        if offset:
            c2 = buf.clone()[offset:].view(H, W, 3).requires_grad_()
        else:
            c2 = c0.clone().requires_grad_()
        a2 = a0.clone().requires_grad_()
        b = torch.ones_like(target) if bg is None else bg
        ref = F.mse_loss(c1.float() + (1 - a1.float()).unsqueeze(-1) * b, target)
        (3.0 * ref).backward()
        got = composite_mse_loss(h, c2, a2, target, bg)
        assert got.dim() == 0
        (3.0 * got).backward()
        assert abs(got.item() - ref.item()) <= 1e-5 * ref.item()
        # the kernel forms the composite with one fused multiply-add, torch
        # rounds the product and the sum: the residual may differ by ~1 ulp of
        # the composite (<= 2, ulp 2.4e-7), the gradient by 2*3/n of that
        ulp_g = 6.0 / (3 * H * W) * 2.4e-7
        tol = dict(rtol=1e-5, atol=2 * ulp_g) if dt == torch.float32 else dict(rtol=2e-3, atol=1e-7)
        assert c2.grad.dtype == dt and a2.grad.dtype == dt
        assert torch.allclose(c2.grad.float(), c1.grad.float(), **tol)
        assert torch.allclose(a2.grad.float(), a1.grad.float(), **tol)
    with pytest.raises(ValueError):
        composite_mse_loss(h, c2, a2, target[:-1])
    with pytest.raises(ValueError):
        composite_mse_loss(h, c2, a2, target.clone().requires_grad_())


def test_composite_mse_loss_repeated_backward(dev):
    """The loss op's backward scales its saved gradient images in place from the
    factor they carry to the incoming gradient: an incoming 1 leaves them as
    they are, and backwards through a retained graph with different incoming
    gradients (1, then 2.5, then 1 again) accumulate what torch's composite +
    F.mse_loss accumulates."""
    from diffbmp_amd.tile_rasterizer import composite_mse_loss
    g = torch.Generator(dev).manual_seed(12)
    H, W = 32, 48
    h = Handle(H, W, 16)
    target = torch.rand((H, W, 3), device=dev, generator=g)
    c0 = torch.rand((H, W, 3), device=dev, generator=g)
    a0 = torch.rand((H, W), device=dev, generator=g)
    c1, a1 = c0.clone().requires_grad_(), a0.clone().requires_grad_()
    c2, a2 = c0.clone().requires_grad_(), a0.clone().requires_grad_()
    ref = F.mse_loss(c1 + (1 - a1).unsqueeze(-1), target)
    got = composite_mse_loss(h, c2, a2, target)
    for k, s in enumerate((1.0, 2.5, 1.0)):
        keep = k < 2
        (s * ref).backward(retain_graph=keep)
        (s * got).backward(retain_graph=keep) if s != 1.0 else got.backward(retain_graph=keep)
        torch.cuda.synchronize()
        for x, y in ((c2, c1), (a2, a1)):
            assert torch.allclose(x.grad, y.grad, rtol=1e-5, atol=1e-9), (k, s)


def test_fused_adam_matches_torch_adam(dev):
    """FusedAdam against torch.optim.Adam followed by the reference's clamp:
    six tensors (two launches per step) in parameter groups with different
    lr, per-column and scalar clamps, fresh random gradients each step, 6
    steps; then a tensor added later, whose step count differs."""
    from diffbmp_amd.optim import FusedAdam
    g = torch.Generator(dev).manual_seed(3)
    shapes = [(300, 2), (300,), (300,), (300,), (300, 3), (77,)]
    init = [torch.rand(s, device=dev, generator=g) * 20 for s in shapes]
    a = [x.clone().requires_grad_() for x in init]
    b = [x.clone().requires_grad_() for x in init]
    lo, hi = (1.0, 2.0), (15.0, 12.0)
    ref = torch.optim.Adam([{"params": a[:2], "lr": 0.5}, {"params": a[2:], "lr": 0.1}],
                           betas=(0.9, 0.99), eps=1e-8)
    fused = FusedAdam([{"params": [b[0]], "lr": 0.5, "clamp": (lo, hi)},
                       {"params": [b[1]], "lr": 0.5, "clamp": (2.0, 10.0)},
                       {"params": b[2:], "lr": 0.1}], betas=(0.9, 0.99), eps=1e-8)
    for _ in range(6):
        grads = [torch.randn(s, device=dev, generator=g) for s in shapes]
        for x, y, gr in zip(a, b, grads):
            x.grad, y.grad = gr.clone(), gr.clone()
        ref.step()
        with torch.no_grad():
            a[0][:, 0].clamp_(lo[0], hi[0])
            a[0][:, 1].clamp_(lo[1], hi[1])
            a[1].clamp_(2.0, 10.0)
        fused.step()
        for x, y in zip(a, b):
            assert torch.allclose(x, y, rtol=1e-5, atol=1e-5), (x - y).abs().max()
    # a tensor added later: its own step count, its own launch
    c = torch.rand(100, device=dev, generator=g).requires_grad_()
    c2 = c.detach().clone().requires_grad_()
    fused.add_param_group({"params": [c], "lr": 0.2})
    ref.add_param_group({"params": [c2], "lr": 0.2})
    for x, y in zip(a + [c2], b + [c]):
        gr = torch.randn(x.shape, device=dev, generator=g)
        x.grad, y.grad = gr.clone(), gr.clone()
    ref.step()
    with torch.no_grad():
        a[0][:, 0].clamp_(lo[0], hi[0])
        a[0][:, 1].clamp_(lo[1], hi[1])
        a[1].clamp_(2.0, 10.0)
    fused.step()
    for x, y in zip(a + [c2], b + [c]):
        assert torch.allclose(x, y, rtol=1e-5, atol=1e-5), (x - y).abs().max()
