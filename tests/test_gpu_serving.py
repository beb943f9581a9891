"""Batched multi-job steps on the GPU.

* one job in the batch follows the single-canvas FitJob step by step;
* a job in a mix of other jobs gets the same first forward image (bit for
  bit: binning and forward are deterministic) and the same trajectory as
  alone (within fp32 atomic-order noise);
* a graph-captured step replays the eager one, also across a job-set change;
* a pair-capacity overflow still gives the exact image;
* the server admits by budget and completes every job.
"""
import numpy as np
import pytest
import torch

from diffbmp_amd import serve as SV
from diffbmp_amd.fit import FitJob

pytestmark = pytest.mark.gpu


def spec(name, H, W, N, T, steps=8, seed=0, **kw):
    return SV.JobSpec(H, W, N, tuple(range(T)), steps=steps, seed=seed, name=name,
                      template_size=64, **kw)


A = spec("A", 128, 128, 500, 2, seed=11)
B = spec("B", 160, 208, 900, 4, seed=12)
C = spec("C", 64, 64, 200, 1, seed=13)


def engine(dtype=torch.float16, graph=False, **kw):
    return SV.BatchEngine(store_dtype=dtype, use_graph=graph, **kw)


def states(eng, specs):
    out = [SV.FitState(s, eng) for s in specs]
    for s in out:
        eng.add(s)
    return out


def test_single_job_matches_fitjob(dev):
    s = spec("F", 96, 80, 300, 3, seed=5, steps=6)
    eng = engine(torch.float32)
    (js,) = states(eng, [s])
    p0 = js.params.clone()
    fit = FitJob(p0[:, 0:2], p0[:, 2], p0[:, 3], p0[:, 4], p0[:, 5:8], js.templates, js.sel,
                 js.target, lr=s.lr, store_dtype=torch.float32, eps_T=s.eps_T, max_steps=s.steps,
                 colors_orig=js.colors_orig, c_blend=s.c_blend, colors_orig_per_template=True)
    eng.step()
    fit.step()
    torch.cuda.synchronize()
    assert torch.equal(js.color, fit.color), (js.color - fit.color).abs().max()
    assert torch.equal(js.alpha, fit.alpha)
    for _ in range(s.steps - 1):
        eng.step()
        fit.step()
    torch.cuda.synchronize()
    lb, lf = js.loss.cpu().numpy(), fit.losses[:s.steps].cpu().numpy()
    print("batch loss", lb, "\nfit loss  ", lf)
    assert np.allclose(lb, lf, rtol=1e-5, atol=0)
    got = torch.cat([fit.params[k].reshape(s.num_primitives, -1) for k in
                     ("means2D", "radii", "rotations", "opacities", "colors")], 1)
    assert torch.allclose(js.params, got, rtol=1e-4, atol=1e-5), (js.params - got).abs().max()


def test_job_alone_equals_in_mix(dev):
    e1, e2 = engine(), engine()
    (a1,) = states(e1, [A])
    _, a2, _ = states(e2, [B, A, C])
    e1.step()
    e2.step()
    torch.cuda.synchronize()
    assert torch.equal(a1.color, a2.color) and torch.equal(a1.alpha, a2.alpha)
    for _ in range(A.steps - 1):
        e1.step()
        e2.step()
    torch.cuda.synchronize()
    l1, l2 = a1.loss.cpu().numpy(), a2.loss.cpu().numpy()
    print("alone", l1, "\nmix  ", l2)
    assert np.allclose(l1, l2, rtol=1e-4, atol=0)
    d = (a1.params - a2.params).abs().max().item()
    print("max |param difference|", d)
    assert torch.allclose(a1.params, a2.params, rtol=1e-3, atol=1e-3)
    assert l2[-1] < l2[0]


def test_shared_bitmaps_equal_per_job_copies(dev):
    """Jobs indexing the engine's one copy of the library's texel records
    render what they render from their own copies: the same first images bit
    for bit and the same trajectories within atomic-order noise (bitmap ids
    out of order and not starting at 0)."""
    specs = [SV.JobSpec(112, 96, 400, (5, 2, 11), steps=6, seed=21, name="S1", template_size=64),
             SV.JobSpec(80, 144, 700, (15, 3), steps=6, seed=22, name="S2", template_size=64),
             SV.JobSpec(64, 64, 150, (7,), steps=6, seed=23, name="S3", template_size=64)]
    shared, own = engine(share_bitmaps=True), engine(share_bitmaps=False)
    js, jo = states(shared, specs), states(own, specs)
    assert js[0].templates.shape[0] == 16 and jo[0].templates.shape[0] == 3
    shared.step()
    own.step()
    torch.cuda.synchronize()
    for a, b in zip(js, jo):
        assert torch.equal(a.color, b.color) and torch.equal(a.alpha, b.alpha), a.spec.name
    for _ in range(5):
        shared.step()
        own.step()
    torch.cuda.synchronize()
    for a, b in zip(js, jo):
        la, lb = a.loss.cpu().numpy(), b.loss.cpu().numpy()
        assert np.allclose(la, lb, rtol=1e-4, atol=0), (a.spec.name, la, lb)


def test_graph_replay_matches_eager(dev):
    eg, ee = engine(graph=True), engine(graph=False)
    sg, se = states(eg, [A, B]), states(ee, [A, B])
    junk = torch.randn(1 << 20, device=dev)
    eg.step()
    ee.step()
    torch.cuda.synchronize()
    for x, y in zip(sg, se):
        assert torch.equal(x.color, y.color)
    for k in range(4):
        eg.step()
        torch.cuda.synchronize()
        junk.mul_(1.0001).add_(float(junk.sum().item()) * 1e-9)   # eager work between replays
        ee.step()
        torch.cuda.synchronize()
        for x, y in zip(sg, se):
            t = x.steps_done - 1
            assert np.isclose(x.loss[t].item(), y.loss[t].item(), rtol=1e-4), (k, t)
    # a set change re-captures
    cg, ce = SV.FitState(C, eg), SV.FitState(C, ee)
    eg.add(cg)
    ee.add(ce)
    eg.step()
    ee.step()
    torch.cuda.synchronize()
    assert torch.equal(cg.color, ce.color)
    for x, y in zip(sg + [cg], se + [ce]):
        assert np.allclose(x.loss.cpu().numpy(), y.loss.cpu().numpy(), rtol=1e-4, atol=0)
    assert eg.set_count == 2


@pytest.mark.parametrize("library", ["stress", "hard"])
def test_pair_overflow_is_exact(dev, library):
    """A pair-capacity overflow walks every primitive of the job, each within
    its binning tile range, so even a hard-edged template (whose bilinear
    fringe reaches past its binned box) renders exactly as with the lists."""
    jb = spec("B", 160, 208, 900, 4, seed=12, library=library)
    small, big = engine(pairs_per_primitive=0.05), engine()
    (x,), (y,) = states(small, [jb]), states(big, [jb])
    small.step()
    big.step()
    torch.cuda.synchronize()
    assert torch.equal(x.color, y.color) and torch.equal(x.alpha, y.alpha)
    assert small.pair_count() > small.pair_capacity
    assert small.grow_if_needed()
    assert not big.grow_if_needed()
    small.step()
    torch.cuda.synchronize()
    assert small.pair_count() <= small.pair_capacity


def test_server_completes_every_job(dev):
    rng = np.random.default_rng(3)
    t = SV.poisson_arrivals(200.0, 5, rng)
    specs = [spec(f"s{i}", int(rng.choice([64, 128, 256])), int(rng.choice([64, 128, 256])),
                  int(rng.integers(100, 2000)), int(rng.integers(1, 5)), steps=20, seed=100 + i,
                  arrival=float(t[i])) for i in range(5)]
    eng = engine(graph=True)
    need = max(SV.job_bytes(s, torch.float16) for s in specs)
    srv = SV.Server(eng, SV.HbmBudget(3 * need), max_jobs=4)
    stats = srv.run(specs)
    summ = stats.summary()
    print(summ)
    assert summ["jobs"] == 5 and summ["max_active"] <= 4
    assert srv.budget.used == 0
    for r in stats.results:
        assert r.latency > 0 and r.final_loss < r.first_loss
