"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every fixture (bit-exact tile lists, image
and gradients to 1e-12).  GPU: the HIP path reproduces them to the parity
bar (bit-exact tile lists, image <= 1e-5 abs, gradients <= 1e-4 normwise).
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from diffbmp_amd import scenes as S

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "*.npz")))
GRADS = ("means2D", "radii", "rotations", "opacities", "colors")


def load(path):
    z = np.load(path, allow_pickle=False)
    co = z["colors_orig_tmpl"]
    sc = S.Scene(H=int(z["H"]), W=int(z["W"]), means2D=z["means2D"], radii=z["radii"],
                 rotations=z["rotations"], opacities=z["opacities"], colors=z["colors"],
                 templates=z["templates"], bmp_sel=z["bmp_sel"], tile_size=int(z["tile_size"]),
                 alpha_ub=float(z["alpha_ub"]), c_blend=float(z["c_blend"]),
                 colors_orig_tmpl=co if co.size else None)
    return sc, z


def test_fixtures_present():
    assert len(FILES) >= 4


@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_fixture(path):
    sc, z = load(path)
    offs, idx = O.bin_scene(sc)
    assert np.array_equal(offs, z["tile_offsets"]) and np.array_equal(idx, z["tile_indices"])
    osc = O.scene_from(sc, offs, idx)
    col, alp, cnt = O.forward(osc)
    assert np.abs(col - z["color"]).max() < 1e-6 and np.abs(alp - z["alpha"]).max() < 1e-6
    assert np.array_equal(cnt, z["push_count"])
    g = O.backward(osc, z["gC"].astype(np.float64), z["gA"].astype(np.float64))
    for k in GRADS:
        assert np.allclose(g[k], z["g_" + k], rtol=1e-9, atol=1e-12), k


@pytest.mark.gpu
@pytest.mark.parametrize("eps", [0.0, 1e-6])
@pytest.mark.parametrize("path", FILES, ids=lambda p: os.path.basename(p)[:-4])
def test_hip_matches_fixture(dev, path, eps):
    from helpers import DeviceRun, normwise
    sc, z = load(path)
    run = DeviceRun(sc, dev, eps_T=eps, co_mode=2 if sc.c_blend > 0 else 0)
    assert np.array_equal(run.offs.cpu().numpy(), z["tile_offsets"])
    assert np.array_equal(run.idx.cpu().numpy(), z["tile_indices"])
    col, alp = run.forward()
    err = max(np.abs(col - z["color"]).max(), np.abs(alp - z["alpha"]).max())
    assert err <= 1e-5, err
    g = run.backward(z["gC"], z["gA"])
    ref = {k: z["g_" + k].astype(np.float64) for k in GRADS}
    e = normwise(g, ref)
    assert max(e.values()) <= 1e-4, e
