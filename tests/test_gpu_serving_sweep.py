"""Seeded random job mixes through the batched engine, against the oracle.

Each case builds one ``BatchEngine`` and two to four jobs, then runs one step
at learning rate 0 (so the parameters, and the image, stay put). Every job's
image, and its gradient rows of that step, are compared with the oracle on
the job's stored (storage-rounded) parameters and bitmaps. A case draws:
- fp16 or fp32 storage, the exact or the reference colour-map gradient, and
  shared or per-job texel records;
- per job: a canvas of 24-160 px a side, 1-300 primitives with radius
  scales from one of two ranges, one to four templates picked from the
  16-template library (not in order), 32 or 64 texel bitmaps from the
  stress or the hard-edged library, c_blend 0 or 0.5, eps_T 0 or 1e-6, and
  alpha_upper_bound 1 or 0.8.

Bars, as in the fixed-scene tests: images within 1e-5 of the oracle's full
walk, gradients within 1e-4 normwise.
"""
import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import scenes as S
from diffbmp_amd import serve as SV
from helpers import normwise

pytestmark = pytest.mark.gpu

CASES = 16


def random_mix(seed):
    """(engine options, job specs) of case ``seed``."""
    rng = np.random.default_rng(20_000 + seed)
    eng = dict(store_dtype=torch.float16 if rng.random() < 0.5 else torch.float32,
               exact_blend=bool(rng.random() < 0.5), share_bitmaps=bool(rng.random() < 0.5))
    specs = []
    for k in range(int(rng.integers(2, 5))):
        T = int(rng.integers(1, 5))
        ids = tuple(int(i) for i in rng.choice(16, T, replace=False))
        specs.append(SV.JobSpec(
            int(rng.integers(24, 161)), int(rng.integers(24, 161)), int(rng.integers(1, 301)), ids,
            steps=2, seed=int(rng.integers(1 << 20)), lr=0.0, clamp=False, name=f"job{k}",
            c_blend=float(rng.choice([0.0, 0.5])), eps_T=float(rng.choice([0.0, 1e-6])),
            alpha_upper_bound=float(rng.choice([1.0, 0.8])),
            ls_range=[(0.005, 0.1), (0.05, 0.4)][int(rng.integers(2))],
            template_size=int(rng.choice([32, 64])), library=str(rng.choice(["stress", "hard"]))))
    return eng, specs


def oracle_scene(js):
    s = js.spec
    p = js.store.float().cpu().numpy()
    return S.Scene(H=s.height, W=s.width, means2D=p[:, 0:2].copy(), radii=p[:, 2].copy(),
                   rotations=p[:, 3].copy(), opacities=p[:, 4].copy(), colors=p[:, 5:8].copy(),
                   templates=js.templates.float().cpu().numpy(), bmp_sel=js.sel.cpu().numpy(),
                   alpha_ub=s.alpha_upper_bound, c_blend=s.c_blend,
                   colors_orig_tmpl=js.colors_orig.float().cpu().numpy())


@pytest.mark.parametrize("seed", range(CASES))
def test_random_mix_matches_oracle(dev, seed):
    opts, specs = random_mix(seed)
    print(f"case {seed}: {opts}", flush=True)
    for s in specs:
        print(f"  {s.name}: {s.height}x{s.width} N={s.num_primitives} ids={s.template_ids} "
              f"tsz={s.template_size} {s.library} c_blend={s.c_blend} eps={s.eps_T} "
              f"aub={s.alpha_upper_bound} ls={s.ls_range}", flush=True)
    eng = SV.BatchEngine(use_graph=False, **opts)
    jobs = [SV.FitState(s, eng) for s in specs]
    for j in jobs:
        eng.add(j)
    eng.step()
    torch.cuda.synchronize()
    for js in jobs:
        sc = oracle_scene(js)
        offs, idx = O.bin_scene(sc)
        osc = O.scene_from(sc, offs, idx, blend_grad_exact=opts["exact_blend"])
        rc, ra, _ = O.forward(osc)
        col, alp = js.color.cpu().numpy(), js.alpha.cpu().numpy()
        e = max(np.abs(col - rc).max(), np.abs(alp - ra).max())
        res = (col + (1.0 - alp)[..., None] - js.target.cpu().numpy()).astype(np.float64)
        gC = 2.0 * res / res.size
        ref = O.backward(osc, gC, -gC.sum(-1))
        rows = eng.gradients(js).cpu().numpy()
        got = dict(means2D=rows[:, 0:2], radii=rows[:, 2], rotations=rows[:, 3], opacities=rows[:, 4],
                   colors=rows[:, 5:8])
        err = normwise(got, ref)
        print(f"  {js.spec.name}: P = {len(idx)}, image error {e:.2e}, gradient error "
              f"{max(err.values()):.2e}", flush=True)
        assert e <= 1e-5
        assert max(err.values()) <= 1e-4, err
