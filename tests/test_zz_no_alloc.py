"""dbmp_forward_noalloc (no device allocation, static scratch) against
dbmp_forward through the C ABI: bit-identical outputs on consecutive calls,
the second with a smaller input, plus the size-limit error."""
import ctypes

import numpy as np
import pytest
import torch

from diffbmp_amd import _native
from diffbmp_amd import scenes as S
from diffbmp_amd._native import Handle
from helpers import to_dev

pytestmark = pytest.mark.gpu


def _outputs(H, W, dev):
    return (torch.full((H, W, 3), -7.0, device=dev), torch.full((H, W), -7.0, device=dev),
            torch.full((H, W), -7.0, device=dev), torch.full((H, W), -7, dtype=torch.int32, device=dev))


def _case(sc, dev, dtype=torch.float32, ts=16, factor=1, co_mode=0):
    """Outputs of dbmp_forward (a handle) and dbmp_forward_noalloc for one scene."""
    d = to_dev(sc, dev, dtype)
    h = Handle(sc.H, sc.W, ts, 0.0, sc.alpha_ub, 151, sc.N)
    prims = Handle.prims(d["means2D"], d["radii"], d["rotations"], d["opacities"], d["colors"],
                         d["bmp_sel"])
    co = None
    if co_mode:
        co = torch.from_numpy(np.ascontiguousarray(sc.colors_orig_tmpl)).to(dev).to(dtype)
    bm = Handle.bitmaps(d["templates"], co, co_mode, sc.c_blend)
    ext = h.template_extents(d["templates"])
    out = h.bin(prims, ext, tuple(d["templates"].shape), factor)
    lists = _native.Lists(out[0], out[1], factor=factor, ranges=out[2] if factor > 1 else None)
    ref = _outputs(sc.H, sc.W, dev)
    h.forward(prims, bm, lists, 1e-6, *ref)
    got = _outputs(sc.H, sc.W, dev)
    L = lists.c()
    st = _native._stream(dev)
    _native.check(_native.lib().dbmp_forward_noalloc(
        sc.H, sc.W, ts, float(sc.alpha_ub), ctypes.byref(prims), ctypes.byref(bm), ctypes.byref(L),
        1e-6, *(ctypes.c_void_p(t.data_ptr()) for t in got), st))
    torch.cuda.synchronize()
    return ref, got


def _same(ref, got, name):
    for a, b, what in zip(ref, got, ("color", "alpha", "T_final", "last")):
        assert torch.equal(a, b), (name, what, (a.float() - b.float()).abs().max().item())


def test_noalloc_matches_forward_consecutive_calls(dev):
    print("case: plain 256^2, 256 primitives, fp32, per-tile lists", flush=True)
    big = S.make_scene(0, with_target=False)
    _same(*_case(big, dev), "plain")
    print("case: smaller 96x128, 40 primitives, tile 32", flush=True)
    small = S.make_scene(0, H=96, W=128, N=40, tsz=32, with_target=False)
    small.tile_size = 32
    _same(*_case(small, dev, ts=32), "small")
    print("case: stress-like 128^2, fp16 RGBA colour maps, super-tile lists", flush=True)
    rgba = S.make_scene(3, H=128, W=128, N=300, tsz=64, with_target=False)
    _same(*_case(rgba, dev, dtype=torch.float16, factor=4, co_mode=2), "rgba")
    print("case: smaller again, 64^2, 12 primitives", flush=True)
    tiny = S.make_scene(0, H=64, W=64, N=12, tsz=16, with_target=False)
    _same(*_case(tiny, dev), "tiny")


def test_noalloc_rejects_sizes_beyond_its_maximum(dev):
    N = 65537
    z = torch.zeros(N, device=dev)
    m = torch.zeros((N, 2), device=dev)
    c = torch.zeros((N, 3), device=dev)
    sel = torch.zeros(N, dtype=torch.int32, device=dev)
    prims = Handle.prims(m, z, z, z, c, sel)
    tm = torch.ones((1, 8, 8), device=dev)
    bm = Handle.bitmaps(tm)
    offs = torch.zeros(17, dtype=torch.int32, device=dev)
    L = _native.Lists(offs, offs).c()
    outs = _outputs(64, 64, dev)
    rc = _native.lib().dbmp_forward_noalloc(64, 64, 16, 1.0, ctypes.byref(prims), ctypes.byref(bm),
                                            ctypes.byref(L), 0.0,
                                            *(ctypes.c_void_p(t.data_ptr()) for t in outs),
                                            _native._stream(dev))
    assert rc == 1 and b"65536" in _native.lib().dbmp_last_error()
    big_tm = torch.ones((2, 1024, 1024), device=dev)   # 2M texels x 16 B > 16 MiB
    prims = Handle.prims(m[:4], z[:4], z[:4], z[:4], c[:4], sel[:4])
    bm = Handle.bitmaps(big_tm)
    rc = _native.lib().dbmp_forward_noalloc(64, 64, 16, 1.0, ctypes.byref(prims), ctypes.byref(bm),
                                            ctypes.byref(L), 0.0,
                                            *(ctypes.c_void_p(t.data_ptr()) for t in outs),
                                            _native._stream(dev))
    assert rc == 1 and b"texel records" in _native.lib().dbmp_last_error()
