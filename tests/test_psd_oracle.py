"""Pins the layer-export oracle (oracle/psd.py) with checks that do not
retype its formulas: hand-computed boxes, closed forms for constant and
texel-aligned templates, a quarter turn that must rotate an asymmetric
template, the truncating byte conversion and the colour blend."""
import numpy as np

from oracle import psd as P

f32 = np.float32


def _one(x, y, r, th=0.0, v=1.0, c=(0.2, 0.4, 0.6), t=0):
    return (np.array([[x, y]], f32), np.array([r], f32), np.array([th], f32),
            np.array([c], f32), np.array([v], f32), np.array([t], np.int32))


def test_bounds_by_hand():
    m = np.array([[10.0, 20.0], [0.0, 0.0], [-50.0, 5.0], [99.5, 49.2]], f32)
    r = np.array([2.0, 3.0, 4.0, 1.0], f32)
    boxes, sizes, offs = P.psd_bounds(m, r, 50, 100, 1.0)
    # 1.5 r padding, floor / ceil, clipped to the canvas
    assert boxes[0].tolist() == [7, 17, 13, 23]
    assert boxes[1].tolist() == [0, 0, 5, 5]            # ceil(4.5) = 5
    assert boxes[2].tolist() == [0, 0, 1, 1]            # off the canvas: 1x1 placeholder
    assert boxes[3].tolist() == [98, 47, 100, 50]
    assert sizes.tolist() == [[6, 6], [5, 5], [1, 1], [2, 3]]
    assert offs.tolist() == [0, 144, 244, 248, 272]
    # the scale factor scales centre and radius
    b2, _, _ = P.psd_bounds(m[:1], r[:1], 100, 200, 2.0)
    assert b2[0].tolist() == [14, 34, 26, 46]


def test_constant_template_closed_form():
    """All-ones template, no rotation: inside the template square the alpha
    byte is trunc(255 v ub) and RGB trunc(255 c); outside it is 0."""
    Ht = Wt = 9
    tm = np.ones((1, Ht, Wt), f32)
    x, y, r, v, ub = 20.0, 15.0, 4.0, 0.7, 0.9
    args = _one(x, y, r, v=v)
    out, amax = P.psd_layers(*args, tm, 40, 40, 1.0, ub)
    boxes, sizes, _ = P.psd_bounds(args[0], args[1], 40, 40, 1.0)
    x0, y0, x1, y1 = boxes[0]
    img = out.reshape(y1 - y0, x1 - x0, 4)
    a_in = int(np.float32(np.float32(np.float32(1.0) * f32(v)) * f32(ub)) * f32(255))
    for j in range(y1 - y0):
        for i in range(x1 - x0):
            px, py = x0 + i, y0 + j
            un, vn = (px - x) / r, (py - y) / r
            if max(abs(un - (-1.0)), abs(un - 1.0), abs(vn + 1.0), abs(vn - 1.0)) < 1e-6:
                continue   # on the square's edge: decided by rounding
            inside = -1.0 <= un < 1.0 and -1.0 <= vn < 1.0
            if inside:
                assert img[j, i].tolist() == [51, 102, 153, a_in], (i, j)
            else:
                assert img[j, i].tolist() == [0, 0, 0, 0], (i, j)
    assert amax[0] == a_in


def test_texel_aligned_sampling_returns_texels():
    """Centre on a pixel and r = (Wt-1)/2: pixel (x+k, y+l) lands exactly on
    texel (k + (Wt-1)/2, l + (Ht-1)/2), so the mask is that texel's value."""
    rng = np.random.default_rng(3)
    Ht = Wt = 11
    tm = rng.uniform(0.2, 1.0, (1, Ht, Wt)).astype(f32)
    x, y = 30.0, 25.0
    r = (Wt - 1) / 2
    args = _one(x, y, r, v=1.0)
    out, _ = P.psd_layers(*args, tm, 60, 60, 1.0, 1.0)
    boxes, _, _ = P.psd_bounds(args[0], args[1], 60, 60, 1.0)
    x0, y0, x1, y1 = boxes[0]
    img = out.reshape(y1 - y0, x1 - x0, 4)
    h = (Wt - 1) // 2
    for l in range(-h, h):          # the last texel row/column is never sampled
        for k in range(-h, h):
            j, i = int(y) + l - y0, int(x) + k - x0
            want = int(f32(tm[0, l + h, k + h]) * f32(255))
            assert int(img[j, i, 3]) == want, (k, l)


def test_quarter_turn_rotates_the_template():
    """theta = pi/2 samples the template rotated by a quarter turn: a bar
    along the template's u axis shows up along the canvas's y axis (a sign
    error in the rotation would leave it along x)."""
    Ht = Wt = 21
    tm = np.zeros((1, Ht, Wt), f32)
    tm[0, 9:12, :] = 1.0          # horizontal bar through the middle rows
    x = y = 40.0
    r = 10.0
    for th, along_y in ((0.0, False), (np.pi / 2, True)):
        args = _one(x, y, r, th=th, v=1.0)
        out, _ = P.psd_layers(*args, tm, 80, 80, 1.0, 1.0)
        boxes, _, _ = P.psd_bounds(args[0], args[1], 80, 80, 1.0)
        x0, y0, x1, y1 = boxes[0]
        a = out.reshape(y1 - y0, x1 - x0, 4)[..., 3].astype(int)
        cx, cy = int(x) - x0, int(y) - y0
        row_hits = (a[cy, cx - 8:cx + 8] > 0).sum()
        col_hits = (a[cy - 8:cy + 8, cx] > 0).sum()
        if along_y:
            assert col_hits == 16 and row_hits <= 4
        else:
            assert row_hits == 16 and col_hits <= 4


def test_byte_conversion_truncates():
    assert P._u8(np.array([0.5, 1.0, 2.0, -1.0, 0.999], f32)).tolist() == [127, 255, 255, 0, 254]
    # NaN: fminf(255, NaN) = 255
    assert P._u8(np.array([np.nan], f32)).tolist() == [255]


def test_alpha_floor_and_blend():
    """Alpha below 1e-3 gives a transparent pixel; a constant colour map
    blends as (1 - k) c + k c_o."""
    Ht = Wt = 9
    tm = np.ones((1, Ht, Wt), f32)
    args = _one(20.0, 20.0, 4.0, v=0.0009)
    out, amax = P.psd_layers(*args, tm, 40, 40)
    assert out.max() == 0 and amax[0] == 0
    cm = np.full((1, Ht, Wt, 3), 0.5, f32)
    k = 0.25
    args = _one(20.0, 20.0, 4.0, v=1.0, c=(0.2, 0.4, 0.6))
    out, _ = P.psd_layers(*args, tm, 40, 40, c_blend=k, cmaps=cm, cmap_mode=2)
    boxes, _, _ = P.psd_bounds(args[0], args[1], 40, 40, 1.0)
    x0, y0, x1, y1 = boxes[0]
    img = out.reshape(y1 - y0, x1 - x0, 4)
    centre = img[20 - y0, 20 - x0]
    want = [int((1 - k) * c * 255 + k * 0.5 * 255) for c in (0.2, 0.4, 0.6)]
    assert all(abs(int(a) - b) <= 1 for a, b in zip(centre[:3], want)), (centre, want)
    # per-primitive maps (mode 1) index by primitive, per-template (mode 2) by template
    cm2 = np.stack([np.full((Ht, Wt, 3), 0.0, f32), np.full((Ht, Wt, 3), 1.0, f32)])
    tm2 = np.ones((2, Ht, Wt), f32)
    a1 = _one(20.0, 20.0, 4.0, v=1.0, t=1)
    o1, _ = P.psd_layers(*a1, tm2, 40, 40, c_blend=1.0, cmaps=cm2, cmap_mode=2)
    o0, _ = P.psd_layers(*a1, tm2, 40, 40, c_blend=1.0, cmaps=cm2, cmap_mode=1)
    i1 = o1.reshape(y1 - y0, x1 - x0, 4)[20 - y0, 20 - x0]
    i0 = o0.reshape(y1 - y0, x1 - x0, 4)[20 - y0, 20 - x0]
    assert i1[:3].tolist() == [255, 255, 255] and i0[:3].tolist() == [0, 0, 0]


def test_bad_template_index_is_transparent():
    tm = np.ones((1, 9, 9), f32)
    out, amax = P.psd_layers(*_one(20.0, 20.0, 4.0, t=5), tm, 40, 40)
    assert out.max() == 0 and amax[0] == 0
