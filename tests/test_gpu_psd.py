"""GPU layer export (dbmp_psd_bounds / dbmp_psd_layers) against the oracle:
boxes, offsets and every RGBA byte bit-exact."""
import time

import numpy as np
import pytest
import torch

from diffbmp_amd import scenes as S
from diffbmp_amd.psd_export_uint8 import export_psd_layers_cuda_uint8, layer_bounds
from oracle import psd as P

pytestmark = pytest.mark.gpu

f32 = np.float32


def _scene(N=160, H=96, W=128, T=3, tsz=24, seed=0):
    rng = np.random.default_rng(seed)
    m = np.stack([rng.uniform(-20, W + 20, N), rng.uniform(-20, H + 20, N)], 1).astype(f32)
    r = np.exp(rng.uniform(np.log(0.6), np.log(24), N)).astype(f32)
    th = rng.uniform(-np.pi, 2 * np.pi, N).astype(f32)
    c = rng.uniform(0, 1, (N, 3)).astype(f32)
    v = rng.uniform(0, 1, N).astype(f32)
    sel = rng.integers(0, T, N).astype(np.int32)
    sel[:3] = [T, -1, T + 5]            # no template: transparent layers
    tm = np.stack([S.shape_bitmap(S.SHAPES[i % 4], tsz, rng) for i in range(T)]).astype(f32)
    cm = rng.uniform(0, 1, (T, tsz, tsz, 3)).astype(f32)
    return m, r, th, c, v, sel, tm, cm


def _gpu_layers(dev, m, r, th, c, v, sel, tm, H, W, scale, ub, cb, cmaps, mode):
    from diffbmp_amd import _native
    import ctypes
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    md, rd, thd, cd, vd, seld, tmd = t(m), t(r), t(th), t(c), t(v), t(sel), t(tm)
    boxes, sizes, offs, total = layer_bounds(md, rd, H, W, scale)
    buf = torch.zeros(max(total, 16), dtype=torch.uint8, device=dev)
    amax = torch.zeros(len(r), dtype=torch.int32, device=dev)
    p = _native.LayerPrims(md.data_ptr(), rd.data_ptr(), thd.data_ptr(), cd.data_ptr(),
                           vd.data_ptr(), seld.data_ptr(), len(r))
    cmd = t(cmaps) if cmaps is not None else None
    _native.check(_native.lib().dbmp_psd_layers(
        ctypes.byref(p), tmd.data_ptr(), tm.shape[0], tm.shape[1], tm.shape[2],
        cmd.data_ptr() if cmd is not None else None, mode, float(cb), H, W, float(scale), float(ub),
        boxes.data_ptr(), offs.data_ptr(), total, buf.data_ptr(), amax.data_ptr(),
        _native._stream(dev)))
    torch.cuda.synchronize()
    return (boxes.cpu().numpy(), sizes.cpu().numpy(), offs.cpu().numpy(), buf[:total].cpu().numpy(),
            amax.cpu().numpy())


@pytest.mark.parametrize("scale,ub,cb,mode", [(1.0, 1.0, 0.0, 0), (1.7, 0.8, 0.0, 0),
                                             (1.0, 1.0, 0.35, 2), (0.6, 0.9, 0.5, 1)])
def test_layers_bit_exact(dev, scale, ub, cb, mode):
    m, r, th, c, v, sel, tm, cm = _scene()
    H, W = 96, 128
    if mode == 1:   # per-primitive maps, the reference's gathered layout
        cm = cm[np.clip(sel, 0, tm.shape[0] - 1)]
    cmaps = cm if mode else None
    boxes, sizes, offs, buf, amax = _gpu_layers(dev, m, r, th, c, v, sel, tm, H, W, scale, ub, cb,
                                                cmaps, mode)
    rb, rs, ro = P.psd_bounds(m, r, H, W, scale)
    assert np.array_equal(boxes, rb) and np.array_equal(sizes, rs) and np.array_equal(offs, ro)
    ref, ramax = P.psd_layers(m, r, th, c, v, sel, tm, H, W, scale, ub, cb, cmaps, mode)
    diff = np.flatnonzero(buf != ref)
    print(f"scale {scale} mode {mode}: {len(buf)} bytes, {np.count_nonzero(ramax)} layers with alpha,"
          f" {len(diff)} differing bytes")
    assert len(diff) == 0
    assert np.array_equal(amax, ramax)


def test_export_api_matches_reference_filtering(dev):
    """export_psd_layers_cuda_uint8: the layers with a crop larger than 1x1 and
    some alpha, in index order, as (h, w, 4) arrays with their boxes."""
    m, r, th, c, v, sel, tm, cm = _scene(seed=4)
    H, W = 96, 128
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    arrays, bounds, idx = export_psd_layers_cuda_uint8(
        t(tm), t(m[:, 0]), t(m[:, 1]), t(r), t(th), t(v), t(c), t(sel), H, W, 1.3, 0.9, 0.4, t(cm))
    rb, rs, ro = P.psd_bounds(m, r, H, W, 1.3)
    ref, ramax = P.psd_layers(m, r, th, c, v, sel, tm, H, W, 1.3, 0.9, 0.4, cm, 2)
    want = [i for i in range(len(r)) if rs[i, 0] > 1 and rs[i, 1] > 1 and ramax[i] > 0]
    assert idx == want
    for a, b, i in zip(arrays, bounds, idx):
        assert b == tuple(int(q) for q in rb[i])
        assert a.shape == (rs[i, 1], rs[i, 0], 4)
        assert np.array_equal(a.reshape(-1), ref[ro[i]:ro[i + 1]])
    # (P,3) colours become constant maps, as the reference expands them
    avg = cm.mean(axis=(1, 2)).astype(f32)
    arrays2, _, idx2 = export_psd_layers_cuda_uint8(
        t(tm), t(m[:, 0]), t(m[:, 1]), t(r), t(th), t(v), t(c), t(sel), H, W, 1.3, 0.9, 0.4, t(avg))
    full = np.broadcast_to(avg[:, None, None, :], cm.shape).astype(f32)
    ref2, _ = P.psd_layers(m, r, th, c, v, sel, tm, H, W, 1.3, 0.9, 0.4, full, 2)
    for a, i in zip(arrays2, idx2):
        assert np.array_equal(a.reshape(-1), ref2[ro[i]:ro[i + 1]])


def test_scale_workload_sampled_layers(dev):
    """The 1024^2, 8192-primitive scene: boxes and offsets of all layers, and
    the bytes of 40 sampled layers (the whole buffer is ~4 GB); reports the
    export's write bandwidth."""
    sc = S.make_scene(1, with_target=False)
    N = sc.N
    rng = np.random.default_rng(5)
    v = rng.uniform(0, 1, N).astype(f32)
    c = rng.uniform(0, 1, (N, 3)).astype(f32)
    boxes, sizes, offs, total = layer_bounds(torch.from_numpy(sc.means2D).to(dev),
                                             torch.from_numpy(sc.radii).to(dev), sc.H, sc.W, 1.0)
    rb, rs, ro = P.psd_bounds(sc.means2D, sc.radii, sc.H, sc.W, 1.0)
    assert np.array_equal(boxes.cpu().numpy(), rb) and np.array_equal(offs.cpu().numpy(), ro)
    assert total == ro[-1]
    from diffbmp_amd import _native
    import ctypes
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    md, rd, thd, cd, vd, seld, tmd = (t(sc.means2D), t(sc.radii), t(sc.rotations), t(c), t(v),
                                      t(sc.bmp_sel), t(sc.templates))
    buf = torch.empty(total, dtype=torch.uint8, device=dev)
    p = _native.LayerPrims(md.data_ptr(), rd.data_ptr(), thd.data_ptr(), cd.data_ptr(),
                           vd.data_ptr(), seld.data_ptr(), N)
    T, Ht, Wt = sc.templates.shape
    args = (ctypes.byref(p), tmd.data_ptr(), T, Ht, Wt, None, 0, 0.0, sc.H, sc.W, 1.0, 1.0,
            boxes.data_ptr(), offs.data_ptr(), total, buf.data_ptr(), None, _native._stream(dev))
    _native.check(_native.lib().dbmp_psd_layers(*args))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        _native.check(_native.lib().dbmp_psd_layers(*args))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    print(f"layer export: {total / 1e9:.2f} GB in {dt * 1e3:.2f} ms = {total / dt / 1e12:.2f} TB/s")
    pick = rng.choice(N, 40, replace=False)
    ref, _ = P.psd_layers(sc.means2D, sc.radii, sc.rotations, c, v, sc.bmp_sel, sc.templates, sc.H,
                          sc.W, boxes=rb, offs=ro, only=pick)
    for i in pick:
        got = buf[int(ro[i]):int(ro[i + 1])].cpu().numpy()
        assert np.array_equal(got, ref[ro[i]:ro[i + 1]]), i
