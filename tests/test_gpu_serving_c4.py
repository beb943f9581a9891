"""The batched engine at the serving workload's shapes, against the oracle.

Jobs of 1024^2 with 8192 primitives and 2048^2 with 32768 primitives over
256^2 RGBA bitmaps (alpha templates with per-template colour maps, c_blend
0.5), fp16 storage / fp32 arithmetic, batched with a smaller job:

* each job's first image equals the oracle's (on the job's fp16-rounded
  parameters and bitmaps) on a sample of its tiles;
* each job's gradient rows equal the oracle's backward: a second step (at
  learning rate 0, no clamp, so the parameters and the image stay the same)
  runs against a target equal to the job's own render except on the sampled
  tiles, so its loss gradient lives on those tiles only;
* a job's first image is the same alone and in the mix, bit for bit.
"""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import scenes as S
from diffbmp_amd import serve as SV
from helpers import normwise

pytestmark = pytest.mark.gpu

NAMES = ("means2D", "radii", "rotations", "opacities", "colors")


def c4_spec(name, side, N, T, seed):
    return SV.JobSpec(side, side, N, tuple(range(T)), steps=4, seed=seed, name=name, lr=0.0,
                      clamp=False, template_size=256)


J1 = c4_spec("c4-1024", 1024, 8192, 4, 31)
J2 = c4_spec("c4-2048", 2048, 32768, 16, 32)
J3 = c4_spec("c4-512", 512, 2000, 2, 33)


def oracle_scene(js):
    s = js.spec
    p = js.store.float().cpu().numpy()
    return S.Scene(H=s.height, W=s.width, means2D=p[:, 0:2].copy(), radii=p[:, 2].copy(),
                   rotations=p[:, 3].copy(), opacities=p[:, 4].copy(), colors=p[:, 5:8].copy(),
                   templates=js.templates.float().cpu().numpy(), bmp_sel=js.sel.cpu().numpy(),
                   c_blend=s.c_blend, colors_orig_tmpl=js.colors_orig.float().cpu().numpy())


def sampled(sc, k, seed):
    tX, tY = -(-sc.W // 16), -(-sc.H // 16)
    rng = np.random.default_rng(seed)
    tiles = np.sort(rng.choice(tX * tY, k, replace=False)).astype(np.int32)
    m = np.zeros(tX * tY, bool)
    m[tiles] = True
    pm = np.kron(m.reshape(tY, tX), np.ones((16, 16), bool))[:sc.H, :sc.W]
    return tiles, pm


def test_c4_jobs_match_oracle(dev):
    eng = SV.BatchEngine(store_dtype=torch.float16, use_graph=True)
    jobs = [SV.FitState(s, eng) for s in (J3, J1, J2)]
    for j in jobs:
        eng.add(j)
    eng.step()
    torch.cuda.synchronize()
    targets = []
    checks = []
    for js in jobs[1:]:
        sc = oracle_scene(js)
        ext = O.template_extents_np(sc.templates)
        offs, idx = O.bin_lists_c(sc.means2D, sc.radii, sc.rotations, sc.bmp_sel, ext, sc.H, sc.W, 16)
        tiles, pm = sampled(sc, 12, js.spec.seed)
        osc = O.scene_from(sc, offs, idx, tile_subset=tiles)
        rc, ra, _ = O.forward(osc)
        col, alp = js.color.cpu().numpy(), js.alpha.cpu().numpy()
        e = max(np.abs(col - rc)[pm].max(), np.abs(alp - ra)[pm].max())
        print(f"{js.spec.name}: P = {len(idx)}, sampled-tile image error {e:.2e}", flush=True)
        assert e <= 1e-5
        # the job's own render over white, except on the sampled tiles
        own = js.color + (1.0 - js.alpha).unsqueeze(-1)
        tgt = torch.where(torch.from_numpy(pm).to(dev).unsqueeze(-1), js.target, own)
        js.target.copy_(tgt)
        checks.append((js, sc, osc, pm))
    eng.step()
    torch.cuda.synchronize()
    for js, sc, osc, pm in checks:
        res = (js.color + (1.0 - js.alpha).unsqueeze(-1) - js.target).cpu().numpy().astype(np.float64)
        assert np.abs(res[~pm]).max() == 0.0
        gC = 2.0 * res / res.size
        gA = -gC.sum(-1)
        ref = O.backward(osc, gC, gA)
        rows = eng.gradients(js).cpu().numpy()
        got = dict(means2D=rows[:, 0:2], radii=rows[:, 2], rotations=rows[:, 3], opacities=rows[:, 4],
                   colors=rows[:, 5:8])
        err = normwise(got, ref)
        print(f"{js.spec.name}: gradient error {err}", flush=True)
        assert max(err.values()) <= 1e-4


def test_c4_job_alone_equals_in_mix(dev):
    alone = SV.BatchEngine(store_dtype=torch.float16, use_graph=True)
    mix = SV.BatchEngine(store_dtype=torch.float16, use_graph=True)
    a = SV.FitState(J1, alone)
    alone.add(a)
    m = [SV.FitState(s, mix) for s in (J2, J1, J3)]
    for j in m:
        mix.add(j)
    alone.step()
    mix.step()
    torch.cuda.synchronize()
    assert torch.equal(a.color, m[1].color) and torch.equal(a.alpha, m[1].alpha)
    ga, gm = alone.gradients(a), mix.gradients(m[1])
    rel = ((ga - gm).norm() / ga.norm()).item()
    print(f"gradient rows alone vs in mix: {rel:.2e}")
    assert rel <= 1e-5
