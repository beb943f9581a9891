"""Pins the CPU oracle (no GPU).

The reference ships no golden vectors, known-answer tests or fixtures for
this path, so each oracle function is pinned by checks a plausible mistake
would fail:

* template extents   -- closed form (single texel, empty, full), numpy == C;
* binning            -- per-primitive tile ranges == the reference's dense
                        (tiles x N) test written verbatim == the C version;
                        theta = 0 closed form; closed-interval edge cases;
* forward            -- closed forms: one layer, two layers (OVER), bilinear
                        at integer texel coordinates, push-box edge, mask gate,
                        invalid template, list order;
* colour-map sample  -- closed forms: constant map, integer texel, linear
                        ramp (value and slopes) at fractional coordinates,
                        clamp to the edge; through the forward with a rotated
                        primitive;
* list membership    -- walking every primitive within its binning range
                        equals the binned lists; without the ranges a
                        hard-edged template's fringe changes the image;
* backward           -- central finite differences (h = 1e-3, fp64) on
                        random scalars, with binning frozen; closed form for
                        the colour-logit gradient;
* fp32 port          -- agrees with the fp64 oracle (it is the CPU baseline).
"""
import math

import numpy as np
import pytest

import oracle as O
from diffbmp_amd import scenes as S

f32 = np.float32


def sig(x):
    """Sigmoid of the fp32-stored parameter (the oracle reads fp32 inputs)."""
    x = float(np.float32(x))
    return 1.0 / (1.0 + math.exp(-x))


# ---------------------------------------------------------------- extents
def test_extents_closed_form():
    T = np.zeros((3, 9, 7), f32)
    T[0, 2, 5] = 0.5
    T[2] = 1.0
    e = O.template_extents_np(T)
    u = (f32(5) / f32(6)) * f32(2) - f32(1)
    v = (f32(2) / f32(8)) * f32(2) - f32(1)
    assert e[0].tolist() == [u, u, v, v]
    assert e[1].tolist() == [-1, 1, -1, 1]          # empty -> full fallback
    assert e[2].tolist() == [-1, 1, -1, 1]
    T[1, 0, 0] = 1e-6                                 # not > 1e-6
    assert O.template_extents_np(T)[1].tolist() == [-1, 1, -1, 1]


def test_extents_numpy_matches_c():
    rng = np.random.default_rng(0)
    T = (rng.uniform(0, 1, (6, 33, 17)) > 0.97).astype(f32) * rng.uniform(0, 1, (6, 33, 17)).astype(f32)
    ext_c = np.zeros((6, 4), f32)
    O.lib().oracle_template_extents(O._p(T), 6, 33, 17, O._p(ext_c))
    assert np.array_equal(O.template_extents_np(T), ext_c)


# ---------------------------------------------------------------- binning
def _random_prims(rng, N, H, W, spread=1.5):
    m = np.stack([rng.uniform(-0.25 * W, spread * W, N), rng.uniform(-0.25 * H, spread * H, N)], 1).astype(f32)
    r = np.exp(rng.uniform(math.log(0.5), math.log(0.6 * W), N)).astype(f32)
    th = rng.uniform(-7, 7, N).astype(f32)
    return m, r, th


@pytest.mark.parametrize("H,W,ts", [(64, 64, 16), (100, 70, 16), (96, 160, 32), (33, 17, 16)])
def test_binning_three_restatements_agree(H, W, ts):
    rng = np.random.default_rng(H * W + ts)
    N, T = 300, 3
    m, r, th = _random_prims(rng, N, H, W)
    ext = np.array([[-1, 1, -1, 1], [-0.5, 0.25, -1, 0.75], [0, 0, 0, 0]], f32)
    sel = S.cyclic_sel(N, T)
    bb = O.bin_bboxes_np(m, r, th, sel, ext)
    dense = O.bin_dense_np(bb, H, W, ts)
    o1, i1 = O.lists_from_dense(dense)
    o2, i2 = O.bin_lists_np(m, r, th, sel, ext, H, W, ts)
    o3, i3 = O.bin_lists_c(m, r, th, sel, ext, H, W, ts)
    assert np.array_equal(o1, o2) and np.array_equal(i1, i2)
    assert np.array_equal(o1, o3) and np.array_equal(i1, i3)
    # within each tile ids ascend (depth = primitive index)
    for t in range(len(o1) - 1):
        seg = i1[o1[t]:o1[t + 1]]
        assert np.all(np.diff(seg) > 0)


def test_binning_theta_zero_closed_form():
    m = np.array([[20.0, 20.0], [16.0, 40.0], [7.25, 31.0]], f32)
    r = np.array([5.0, 2.0, 8.75], f32)
    th = np.zeros(3, f32)
    ext = np.array([[-1, 1, -1, 1]], f32)
    sel = np.zeros(3, np.int32)
    bb = O.bin_bboxes_np(m, r, th, sel, ext)
    for n in range(3):
        assert bb[n, 0] == f32(m[n, 0] + f32(r[n] * f32(-1)))
        assert bb[n, 1] == f32(m[n, 0] + r[n])
    offs, idx = O.bin_lists_np(m, r, th, sel, ext, 64, 64, 16)
    tiles_of = lambda n: sorted(t for t in range(16) if n in idx[offs[t]:offs[t + 1]])
    # [15,25]^2 touches tiles (0,0),(1,0),(0,1),(1,1)
    assert tiles_of(0) == [0, 1, 4, 5]
    # x in [14,18] (x-tiles 0,1), y in [38,42] (y-tile 2)
    assert tiles_of(1) == [8, 9]
    # closed intervals: xmax = 16 exactly still touches tile 1
    m2 = np.array([[8.0, 8.0]], f32)
    r2 = np.array([8.0], f32)
    o2, i2 = O.bin_lists_np(m2, r2, np.zeros(1, f32), np.zeros(1, np.int32), ext, 64, 64, 16)
    assert [t for t in range(16) if o2[t + 1] > o2[t]] == [0, 1, 4, 5]


def test_binning_offcanvas_and_huge():
    ext = np.array([[-1, 1, -1, 1]], f32)
    m = np.array([[-50.0, 10.0], [200.0, 200.0], [32.0, 32.0]], f32)
    r = np.array([10.0, 5.0, 1000.0], f32)
    offs, idx = O.bin_lists_np(m, r, np.zeros(3, f32), np.zeros(3, np.int32), ext, 64, 64, 16)
    assert 0 not in idx and 1 not in idx
    assert (idx == 2).sum() == 16


# ---------------------------------------------------------------- forward closed forms
def _scene(H, W, m, r, th, v, c, tmpl, sel=None, geom64=False, **kw):
    N = len(r)
    sel = np.zeros(N, np.int32) if sel is None else sel
    ts = 16
    tiles = (-(-H // ts)) * (-(-W // ts))
    offs = (np.arange(tiles + 1) * N).astype(np.int32)        # every primitive in every tile
    idx = np.tile(np.arange(N, dtype=np.int32), tiles)
    return O.OracleScene(H, W, np.asarray(m, f32), np.asarray(r, f32), np.asarray(th, f32),
                         np.asarray(v, f32), np.asarray(c, f32), np.asarray(tmpl, f32), sel, offs, idx,
                         geom64=geom64, **kw)


def test_forward_single_layer_closed_form():
    tm = np.ones((1, 8, 8), f32)
    sc = _scene(32, 32, [[10, 12]], [4], [0.3], [0.7], [[0.2, -1.0, 2.0]], tm)
    col, a, cnt = O.forward(sc)
    al = sig(0.7)
    inside = np.zeros((32, 32), bool)
    inside[8:17, 6:15] = True                              # |x-10|<=4, |y-12|<=4
    # rotated unit square covers the axis box only partly; check the box centre
    assert np.abs(col[12, 10] - [al * sig(0.2), al * sig(-1.0), al * sig(2.0)]).max() < 1e-14
    assert abs(a[12, 10] - al) < 1e-14
    assert np.all(a[~inside] == 0)


def test_forward_two_layers_over():
    tm = np.ones((1, 8, 8), f32)
    sc = _scene(32, 32, [[10, 10], [11, 10]], [6, 6], [0, 0], [0.4, -0.3],
                [[1.0, 0, 0], [-2.0, 0.5, 0]], tm)
    col, a, _ = O.forward(sc)
    a1, a2 = sig(0.4), sig(-0.3)
    c1, c2 = sig(1.0), sig(-2.0)
    assert abs(col[10, 10, 0] - (a1 * c1 + (1 - a1) * a2 * c2)) < 1e-14
    assert abs(a[10, 10] - (1 - (1 - a1) * (1 - a2))) < 1e-14
    # depth = index: swapping order changes the colour as OVER predicts
    sc2 = _scene(32, 32, [[11, 10], [10, 10]], [6, 6], [0, 0], [-0.3, 0.4],
                 [[-2.0, 0.5, 0], [1.0, 0, 0]], tm)
    col2, _, _ = O.forward(sc2)
    assert abs(col2[10, 10, 0] - (a2 * c2 + (1 - a2) * a1 * c1)) < 1e-14


def test_forward_bilinear_integer_texel_and_push_edge():
    # Wt = 5, r = 2, theta = 0: pixel offset dx -> tex = (dx/2 + 1) * 2 = dx + 2
    tm = np.arange(25, dtype=f32).reshape(1, 5, 5) / 24.0
    sc = _scene(16, 16, [[8, 8]], [2], [0], [50.0], [[50.0, 50, 50]], tm)
    col, a, cnt = O.forward(sc)
    for dy in (-2, -1, 0, 1, 2):
        for dx in (-2, -1, 0, 1, 2):
            want = tm[0, dy + 2, dx + 2]
            if want <= 1e-6:
                assert a[8 + dy, 8 + dx] == 0
            else:
                assert abs(a[8 + dy, 8 + dx] - want) < 1e-12    # sigmoid(50) == 1 in fp64 to 2e-22
    assert a[8, 11] == 0 and a[11, 8] == 0                      # |dx| = 3 > r: outside the push box


def test_forward_mask_gate_and_invalid_template():
    tm = np.full((1, 4, 4), 1e-6, f32)
    sc = _scene(16, 16, [[8, 8]], [3], [0], [0], [[0, 0, 0]], tm)
    _, a, cnt = O.forward(sc)
    assert a.max() == 0 and cnt.max() == 0
    tm = np.ones((1, 4, 4), f32)
    sc = _scene(16, 16, [[8, 8]], [3], [0], [0], [[0, 0, 0]], tm, sel=np.array([5], np.int32))
    _, a, cnt = O.forward(sc)
    assert a.max() == 0 and cnt.max() == 0


# ---------------------------------------------------------------- backward
def _fd_scene(seed, N=12, H=48, W=48):
    rng = np.random.default_rng(seed)
    tm = S.smooth_bump_bitmaps(2, 32, rng)
    m = np.stack([rng.uniform(8, W - 8, N), rng.uniform(8, H - 8, N)], 1)
    r = rng.uniform(6, 14, N)
    th = rng.uniform(0, 2 * math.pi, N)
    v = rng.standard_normal(N)
    c = rng.standard_normal((N, 3))
    return dict(H=H, W=W, m=m, r=r, th=th, v=v, c=c, tm=tm, sel=S.cyclic_sel(N, 2))


def _loss(p, gC, gA):
    sc = _scene(p["H"], p["W"], p["m"], p["r"], p["th"], p["v"], p["c"], p["tm"], sel=p["sel"], geom64=True)
    col, a, _ = O.forward(sc)
    return float((col * gC).sum() + (a * gA).sum())


def _fd_picks(p, g, gC, gA, h, picks):
    fds, ans = [], []
    for pk, gk, n, j in picks:
        arr = p[pk]
        ix = (n,) if arr.ndim == 1 else (n, j % arr.shape[1])
        base = arr.copy()
        plus, minus = base.copy(), base.copy()
        plus[ix] = base[ix] + h
        minus[ix] = base[ix] - h
        p[pk] = plus
        lp = _loss(p, gC, gA)
        p[pk] = minus
        lm = _loss(p, gC, gA)
        p[pk] = base
        # the oracle reads fp32 parameters: use the step actually taken
        h_eff = (float(np.float32(plus[ix])) - float(np.float32(minus[ix]))) / 2
        fds.append((lp - lm) / (2 * h_eff))
        ans.append(float(g[gk][ix]))
    return np.array(fds), np.array(ans)


@pytest.fixture(scope="module")
def fd_case():
    p = _fd_scene(11)
    rng = np.random.default_rng(12)
    gC = rng.standard_normal((p["H"], p["W"], 3))
    gA = rng.standard_normal((p["H"], p["W"]))
    sc = _scene(p["H"], p["W"], p["m"], p["r"], p["th"], p["v"], p["c"], p["tm"], sel=p["sel"], geom64=True)
    g = O.backward(sc, gC, gA)
    N = len(p["r"])
    keys = [("m", "means2D"), ("r", "radii"), ("th", "rotations"), ("v", "opacities"), ("c", "colors")]
    picks = []
    for _ in range(64):
        pk, gk = keys[int(rng.integers(5))]
        picks.append((pk, gk, int(rng.integers(N)), int(rng.integers(3))))
    return p, g, gC, gA, picks


def test_backward_fd_h1e3_normwise(fd_case):
    """64 random scalars, central differences h = 1e-3 in fp64.  Bilinear
    masks are only C0 across texel lines, which costs O(h f'') per sample, so
    the criterion is normwise over the 64 scalars."""
    p, g, gC, gA, picks = fd_case
    fd, an = _fd_picks(p, g, gC, gA, 1e-3, picks)
    err = np.linalg.norm(fd - an) / np.linalg.norm(an)
    assert err <= 1.5e-2, err


def test_backward_fd_h1e5_elementwise(fd_case):
    """The same scalars with h = 1e-5: kink crossings become 100x rarer, and
    each derivative must match to 1e-3 relative (floor 1e-4 of the largest)."""
    p, g, gC, gA, picks = fd_case
    fd, an = _fd_picks(p, g, gC, gA, 1e-5, picks)
    tol = 1e-3 * np.abs(an) + 1e-4 * np.abs(an).max()
    bad = np.abs(fd - an) > tol
    assert not bad.any(), [(picks[i][1], fd[i], an[i]) for i in np.nonzero(bad)[0]]


def test_backward_colour_logit_closed_form():
    tm = np.ones((1, 8, 8), f32)
    sc = _scene(16, 16, [[8, 8]], [3], [0], [0.5], [[0.3, -0.2, 1.1]], tm)
    gC = np.zeros((16, 16, 3))
    gC[8, 8] = [1.0, 2.0, -1.0]
    g = O.backward(sc, gC, np.zeros((16, 16)))
    al = sig(0.5)
    for ch, cl in enumerate([0.3, -0.2, 1.1]):
        s = sig(cl)
        assert abs(g["colors"][0, ch] - gC[8, 8, ch] * al * s * (1 - s)) < 1e-14


@pytest.fixture(scope="module")
def blend_case():
    """A colour-mapped scene (c_blend = 0.6, low-frequency per-primitive maps)."""
    p = _fd_scene(21)
    rng = np.random.default_rng(22)
    N = len(p["r"])
    p["co"] = np.stack([S.smooth_field(32, 32, rng, terms=4) for _ in range(N)])
    gC = rng.standard_normal((p["H"], p["W"], 3))
    gA = rng.standard_normal((p["H"], p["W"]))
    return p, gC, gA


def _blend_scene(p, exact):
    return _scene(p["H"], p["W"], p["m"], p["r"], p["th"], p["v"], p["c"], p["tm"], sel=p["sel"],
                  geom64=True, colors_orig=p["co"], c_blend=0.6, blend_grad_exact=exact)


def test_backward_fd_exact_colour_map(blend_case):
    """With the exact colour-map gradient, every parameter's derivative matches
    central differences (h = 1e-5, fp64, 1e-3 relative, floor 1e-4 of the
    largest) through the blended colour (1 - k) sig(c) + k c_o(tx, ty)."""
    p, gC, gA = blend_case
    g = O.backward(_blend_scene(p, True), gC, gA)
    rng = np.random.default_rng(23)
    N = len(p["r"])
    keys = [("m", "means2D"), ("r", "radii"), ("th", "rotations"), ("v", "opacities"), ("c", "colors")]
    picks = [keys[i % 5] + (int(rng.integers(N)), int(rng.integers(3))) for i in range(40)]

    def loss(q):
        col, a, _ = O.forward(_blend_scene(q, True))
        return float((col * gC).sum() + (a * gA).sum())

    fd, an = [], []
    h = 1e-5
    for pk, gk, n, j in picks:
        arr = p[pk]
        ix = (n,) if arr.ndim == 1 else (n, j % arr.shape[1])
        base = arr.copy()
        plus, minus = base.copy(), base.copy()
        plus[ix] += h
        minus[ix] -= h
        p[pk] = plus
        lp = loss(p)
        p[pk] = minus
        lm = loss(p)
        p[pk] = base
        h_eff = (float(np.float32(plus[ix])) - float(np.float32(minus[ix]))) / 2
        fd.append((lp - lm) / (2 * h_eff))
        an.append(float(g[gk][ix]))
    fd, an = np.array(fd), np.array(an)
    tol = 1e-3 * np.abs(an) + 1e-4 * np.abs(an).max()
    bad = np.abs(fd - an) > tol
    assert not bad.any(), [(picks[i][1], fd[i], an[i]) for i in np.nonzero(bad)[0]]


def test_backward_reference_colour_gradient_under_blend(blend_case):
    """The reference's gradient under c_blend > 0 drops the (1 - k) factor on
    the colour logits and the colour map's dependence on geometry: its colour
    gradient is exactly the exact one / (1 - k), and its geometry gradient
    differs from the exact one."""
    p, gC, gA = blend_case
    ge = O.backward(_blend_scene(p, True), gC, gA)
    gr = O.backward(_blend_scene(p, False), gC, gA)
    k1 = 1.0 - float(np.float32(0.6))
    assert np.allclose(gr["colors"] * k1, ge["colors"], rtol=1e-12, atol=1e-15)
    assert np.allclose(gr["opacities"], ge["opacities"], rtol=1e-12, atol=1e-15)
    d = np.linalg.norm(gr["means2D"] - ge["means2D"]) / np.linalg.norm(ge["means2D"])
    assert d > 1e-3, d


# ---------------------------------------------------------------- fp32 port
def test_fp32_port_agrees_with_fp64_oracle():
    sc = S.make_scene(0, H=96, W=96, N=64, with_target=False)
    offs, idx = O.bin_scene(sc)
    osc = O.scene_from(sc, offs, idx)
    c64, a64, _ = O.forward(osc)
    c32, a32, _ = O.forward(osc, "f32")
    assert np.abs(c64 - c32).max() < 1e-5 and np.abs(a64 - a32).max() < 1e-5
    gC, gA = np.ones((96, 96, 3)), np.ones((96, 96))
    g64 = O.backward(osc, gC, gA)
    g32 = O.backward(osc, gC.astype(f32), gA.astype(f32), "f32")
    for k in g64:
        assert np.linalg.norm(g64[k] - g32[k]) <= 1e-4 * np.linalg.norm(g64[k])


# ---------------------------------------------------------------- colour-map sample
def _cmap_sample(cm, cx, cy):
    import ctypes
    L = O.lib("f64")
    Ht, Wt, _ = cm.shape
    cm = np.ascontiguousarray(cm, f32)
    co, dx, dy = np.zeros(3), np.zeros(3), np.zeros(3)
    L.oracle_color_map_sample(O._p(cm), Ht, Wt, ctypes.c_double(cx), ctypes.c_double(cy),
                              O._p(co), O._p(dx), O._p(dy))
    return co, dx, dy


def _ramp_map(Ht, Wt):
    """c(x, y, ch) = a_ch x + b_ch y + d_ch with different slopes along x and y
    (and per channel), so a swapped x/y corner or a transposed map changes
    the value off the diagonal."""
    a = np.array([0.05, -0.03, 0.011])
    b = np.array([0.02, 0.07, -0.013])
    d = np.array([0.1, 0.5, 0.9])
    yy, xx = np.mgrid[0:Ht, 0:Wt]
    return (a * xx[..., None] + b * yy[..., None] + d).astype(f32), a, b, d


def test_colour_map_sample_closed_forms():
    """sample_color_map (tile_forward.cu:17-89): clamp-to-edge bilinear RGB at
    texel coordinates.  Pinned by (i) a constant map returns the constant,
    (ii) integer coordinates return that texel, (iii) bilinear interpolation
    reproduces a linear ramp exactly at fractional coordinates, with the
    ramp's slopes as derivative (pins the x/y orientation of the four
    corners), (iv) coordinates outside [0, W-1] x [0, H-1] take the edge
    texel's value (clamp, not zero padding)."""
    Ht, Wt = 5, 7
    const = np.zeros((Ht, Wt, 3), f32) + np.array([0.25, 0.5, 0.75], f32)
    for cx, cy in ((0.0, 0.0), (3.3, 1.7), (6.0, 4.0), (-2.0, 9.0)):
        co, _, _ = _cmap_sample(const, cx, cy)
        assert np.array_equal(co, [0.25, 0.5, 0.75])
    cm, a, b, d = _ramp_map(Ht, Wt)
    for y in range(Ht):
        for x in range(Wt):
            co, _, _ = _cmap_sample(cm, float(x), float(y))
            assert np.array_equal(co, cm[y, x].astype(np.float64)), (x, y)
    rng = np.random.default_rng(4)
    for _ in range(50):
        cx, cy = rng.uniform(0, Wt - 1), rng.uniform(0, Ht - 1)
        co, dcx, dcy = _cmap_sample(cm, cx, cy)
        assert np.abs(co - (a * cx + b * cy + d)).max() < 1e-6
        assert np.abs(dcx - a).max() < 1e-6 and np.abs(dcy - b).max() < 1e-6
    for cx, cy, ex, ey in ((-3.5, 2.25, 0, 2.25), (Wt + 4.0, 1.5, Wt - 1, 1.5),
                           (2.5, -1.0, 2.5, 0), (4.25, Ht + 0.5, 4.25, Ht - 1),
                           (-1.0, -1.0, 0, 0), (Wt + 1.0, Ht + 1.0, Wt - 1, Ht - 1)):
        co, _, _ = _cmap_sample(cm, cx, cy)
        assert np.abs(co - (a * ex + b * ey + d)).max() < 1e-6, (cx, cy)


def test_forward_colour_map_closed_form_rotated():
    """One primitive with an all-ones template (mask 1 inside the texture
    bounds), theta = pi/2 and a linear ramp colour map blended with k: the
    pixel colour is alpha ((1-k) sig(c) + k ramp(tx, ty)) with, for theta =
    pi/2, tx = (dy/r + 1)(Wt-1)/2 and ty = (-dx/r + 1)(Ht-1)/2 -- the map
    turns with the primitive.  A mirrored rotation, a transposed map or an
    unblended colour fails."""
    Ht, Wt = 9, 9
    cm, a, b, d = _ramp_map(Ht, Wt)
    tm = np.ones((1, Ht, Wt), f32)
    k = 0.6
    r = 8.0
    cl = [0.3, -0.4, 1.1]
    sc = _scene(48, 48, [[20, 22]], [r], [math.pi / 2], [1.5], [cl], tm,
                colors_orig=cm[None], c_blend=k)
    col, al, _ = O.forward(sc)
    alpha = sig(1.5)
    checked = 0
    for py in range(16, 29):
        for px in range(14, 27):
            dx, dy = px - 20.0, py - 22.0
            if abs(dx) > r - 0.5 or abs(dy) > r - 0.5:
                continue
            tx = (dy / r + 1.0) * (Wt - 1) / 2.0
            ty = (-dx / r + 1.0) * (Ht - 1) / 2.0
            ramp = a * tx + b * ty + d
            want = alpha * ((1 - k) * np.array([sig(c) for c in cl]) + k * ramp)
            assert np.abs(col[py, px] - want).max() < 2e-6, (px, py)
            checked += 1
    assert checked >= 100


# ---------------------------------------------------------------- walk-all fallback
def test_walk_all_needs_binning_ranges_for_hard_edges():
    """The tile lists decide membership (vector_renderer.py:401-444): a
    hard-edged template's bilinear fringe reaches up to one texel past the
    box its extents give, and the reference does not draw that fringe in a
    tile whose list lacks the primitive.  Walking every primitive in every
    tile therefore changes the image of a hard-edged template (so the GPU
    overflow tests, which use one, fail a fallback that ignores the binning
    ranges), while the soft disk is unchanged; walking every primitive
    within its binning tile range gives exactly the binned lists."""
    for tmpl, differs in ((S.hard_square(64)[None], True), (None, False)):
        sc = S.make_scene(0, with_target=False)
        if tmpl is not None:
            sc.templates = tmpl
        offs, idx = O.bin_scene(sc)
        tiles = len(offs) - 1
        col, al, _ = O.forward(O.scene_from(sc, offs, idx))
        all_offs = (np.arange(tiles + 1) * sc.N).astype(np.int32)
        all_idx = np.tile(np.arange(sc.N, dtype=np.int32), tiles)
        col2, al2, _ = O.forward(O.scene_from(sc, all_offs, all_idx))
        e = max(np.abs(col - col2).max(), np.abs(al - al2).max())
        print("hard" if differs else "soft", "walk-all vs lists:", e)
        assert (e > 0.05) if differs else (e < 1e-9)
        # every primitive, kept to the tiles of its binning range
        bb = O.bin_bboxes_np(sc.means2D, sc.radii, sc.rotations, sc.bmp_sel,
                             O.template_extents_np(sc.templates))
        dense = O.bin_dense_np(bb, sc.H, sc.W, sc.tile_size)
        r_offs, r_idx = O.lists_from_dense(dense)
        assert np.array_equal(r_offs, offs) and np.array_equal(r_idx, idx)
