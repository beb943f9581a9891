"""CPU oracle for the DiffBMP tile rasterizer -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package; the product path (``diffbmp_amd``) never
does.  It has two halves:

* a numpy restatement of the binning (template extents, rotated-corner AABB,
  closed tile test, CSR tile lists), written op-for-op in fp32 like the
  reference's torch code -- ``vector_renderer.py:151-252,401-444`` and
  ``simple_tile_renderer.py:721-766`` (paths relative to the reference tree);
* ctypes bindings to ``dbmp_oracle.c`` (built by ``oracle/Makefile``): an FP64
  restatement of the forward/backward kernels
  (``cuda_kernels/tile_forward.cu:96-398``, ``tile_backward.cu:101-452``) and
  the same code in fp32 + OpenMP as the CPU baseline.

How each function is pinned is recorded in ``tests/test_oracle.py`` (closed
forms, brute force, finite differences, cross-restatement agreement).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "build")
_LIBS: dict[str, ctypes.CDLL] = {}

f32 = np.float32


def build() -> None:
    """Compile both oracle libraries (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _Scene(ctypes.Structure):
    _fields_ = [
        ("H", ctypes.c_int), ("W", ctypes.c_int), ("ts", ctypes.c_int),
        ("alpha_ub", ctypes.c_float),
        ("N", ctypes.c_int), ("T", ctypes.c_int), ("Ht", ctypes.c_int), ("Wt", ctypes.c_int),
        ("means2D", ctypes.c_void_p), ("radii", ctypes.c_void_p), ("rotations", ctypes.c_void_p),
        ("opacities", ctypes.c_void_p), ("colors", ctypes.c_void_p),
        ("colors_orig", ctypes.c_void_p), ("c_blend", ctypes.c_float),
        ("templates", ctypes.c_void_p), ("bmp_sel", ctypes.c_void_p),
        ("tile_offsets", ctypes.c_void_p), ("tile_indices", ctypes.c_void_p),
        ("geom64", ctypes.c_int),
        ("tile_subset", ctypes.c_void_p), ("n_subset", ctypes.c_int),
        ("blend_grad_exact", ctypes.c_int), ("colors_orig_per_template", ctypes.c_int),
    ]


def lib(real: str = "f64") -> ctypes.CDLL:
    if real not in _LIBS:
        path = os.path.join(_BUILD, f"liboracle_{real}.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_bin_lists.restype = ctypes.c_longlong
        L.oracle_count_samples.restype = ctypes.c_longlong
        _LIBS[real] = L
    return _LIBS[real]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# --------------------------------------------------------------------------
# binning (numpy, fp32 op-for-op)
# --------------------------------------------------------------------------

def template_extents_np(templates: np.ndarray) -> np.ndarray:
    """(T,4) fp32 ``(min_u, max_u, min_v, max_v)`` of texels > 1e-6 with
    ``u = (col/(W-1))*2 - 1``; empty template -> (-1,1,-1,1)
    (``vector_renderer.py:151-192``)."""
    T, Ht, Wt = templates.shape
    out = np.empty((T, 4), f32)
    for t in range(T):
        nz = templates[t] > f32(1e-6)
        if not nz.any():
            out[t] = (-1, 1, -1, 1)
            continue
        vv, uu = np.nonzero(nz)
        un = (uu.astype(f32) / f32(Wt - 1)) * f32(2) - f32(1)
        vn = (vv.astype(f32) / f32(Ht - 1)) * f32(2) - f32(1)
        out[t] = (un.min(), un.max(), vn.min(), vn.max())
    return out


def cos_sin(theta: np.ndarray):
    """fp32 cos/sin of ``theta`` by ``dbmp_sincosf`` (``diffbmp_amd/csrc/dbmp_trig.h``,
    compiled into the C oracle): the same values the GPU kernels use, so the
    binning decisions below are the kernels'."""
    t = np.ascontiguousarray(theta, f32).ravel()
    s = np.empty_like(t)
    c = np.empty_like(t)
    lib("f64").oracle_sincos_array(len(t), _p(t), _p(s), _p(c))
    return c.reshape(np.shape(theta)), s.reshape(np.shape(theta))


def bin_bboxes_np(means2D, radii, rotations, sel, ext) -> np.ndarray:
    """(N,4) xmin,xmax,ymin,ymax of the rotated extent corners, every torch op
    rounded separately (``vector_renderer.py:217-250``)."""
    x = means2D[:, 0].astype(f32)[:, None]
    y = means2D[:, 1].astype(f32)[:, None]
    r = radii.astype(f32)[:, None]
    c, s = cos_sin(rotations.astype(f32))
    c, s = c[:, None], s[:, None]
    e = ext[np.clip(sel, 0, ext.shape[0] - 1)]
    cu = np.stack([e[:, 0], e[:, 1], e[:, 0], e[:, 1]], 1).astype(f32)
    cv = np.stack([e[:, 2], e[:, 2], e[:, 3], e[:, 3]], 1).astype(f32)
    wx = x + r * (cu * c - cv * s)
    wy = y + r * (cu * s + cv * c)
    with np.errstate(invalid="ignore"):
        return np.stack([wx.min(1), wx.max(1), wy.min(1), wy.max(1)], 1).astype(f32)


def tile_bounds(H: int, W: int, ts: int):
    """Row-major tile starts/ends; ends clamp to the canvas (closed intervals,
    ``simple_tile_renderer.py:485-490``)."""
    tX, tY = -(-W // ts), -(-H // ts)
    ty = np.repeat(np.arange(tY), tX)
    tx = np.tile(np.arange(tX), tY)
    ys = ty * ts
    xs = tx * ts
    return xs, np.minimum(xs + ts, W), ys, np.minimum(ys + ts, H)


def bin_dense_np(bb: np.ndarray, H: int, W: int, ts: int) -> np.ndarray:
    """(tiles, N) bool: the reference's dense intersection test verbatim
    (``vector_renderer.py:436-444``)."""
    xs, xe, ys, ye = (a.astype(f32)[None, :] for a in tile_bounds(H, W, ts))
    bxmin, bxmax, bymin, bymax = (bb[:, i][:, None] for i in range(4))
    with np.errstate(invalid="ignore"):
        xi = ~((bxmax < xs) | (bxmin > xe))
        yi = ~((bymax < ys) | (bymin > ye))
    return (xi & yi).T


def lists_from_dense(mask: np.ndarray):
    """CSR (offsets (tiles+1), indices (P)) with ascending ids per tile
    (``simple_tile_renderer.py:736-751``)."""
    counts = mask.sum(1)
    offs = np.zeros(mask.shape[0] + 1, np.int32)
    offs[1:] = np.cumsum(counts)
    idx = np.nonzero(mask)[1].astype(np.int32)   # row-major => tile, then ascending id
    return offs, idx


def bin_lists_np(means2D, radii, rotations, sel, ext, H, W, ts):
    """Tile lists via per-primitive tile ranges (no dense mask).  A primitive
    whose template index is out of range is never drawn and is in no list."""
    bb = bin_bboxes_np(means2D, radii, rotations, sel, ext)
    tX, tY = -(-W // ts), -(-H // ts)
    xs, xe, ys, ye = tile_bounds(H, W, ts)
    cx = np.arange(tX)
    cy = np.arange(tY)
    xs1, xe1 = (cx * ts).astype(f32), np.minimum(cx * ts + ts, W).astype(f32)
    ys1, ye1 = (cy * ts).astype(f32), np.minimum(cy * ts + ts, H).astype(f32)
    tiles_of = []
    T = ext.shape[0]
    with np.errstate(invalid="ignore"):
        for n in range(bb.shape[0]):
            if not (0 <= sel[n] < T):        # never drawn: in no list
                tiles_of.append(np.zeros(0, np.int64))
                continue
            hx = np.nonzero(~((bb[n, 1] < xs1) | (bb[n, 0] > xe1)))[0]
            hy = np.nonzero(~((bb[n, 3] < ys1) | (bb[n, 2] > ye1)))[0]
            tiles_of.append((hy[:, None] * tX + hx[None, :]).ravel())
    ntiles = tX * tY
    counts = np.zeros(ntiles, np.int64)
    for t in tiles_of:
        counts[t] += 1
    offs = np.zeros(ntiles + 1, np.int64)
    offs[1:] = np.cumsum(counts)
    idx = np.empty(int(offs[-1]), np.int32)
    fill = offs[:-1].copy()
    for n, t in enumerate(tiles_of):
        idx[fill[t]] = n
        fill[t] += 1
    return offs.astype(np.int32), idx


def tie_band(bb: np.ndarray, H: int, W: int, ts: int, band: float = 1e-3) -> np.ndarray:
    """(N,) bool: primitives with an AABB edge within ``band`` px of a tile
    boundary, where a 1-ulp trig difference could flip membership."""
    def near(v, lim):
        q = np.clip(np.round(v / ts) * ts, 0, lim)
        return np.abs(v - q) <= band
    with np.errstate(invalid="ignore"):
        return near(bb[:, 0], W) | near(bb[:, 1], W) | near(bb[:, 2], H) | near(bb[:, 3], H)


def pack_mapping(offsets: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Reference wire format ``[tiles+1, P, offsets..., indices...]`` int32
    (``simple_tile_renderer.py:757-766``)."""
    return np.concatenate([np.array([len(offsets), len(indices)], np.int32),
                           offsets.astype(np.int32), indices.astype(np.int32)])


def bin_lists_c(means2D, radii, rotations, sel, ext, H, W, ts):
    """Tile lists from the C restatement (fp32 decisions)."""
    L = lib("f64")
    N = radii.shape[0]
    tX, tY = -(-W // ts), -(-H // ts)
    offs = np.zeros(tX * tY + 1, np.int32)
    m = np.ascontiguousarray(means2D, f32)
    r = np.ascontiguousarray(radii, f32)
    th = np.ascontiguousarray(rotations, f32)
    sl = np.ascontiguousarray(sel, np.int32)
    e = np.ascontiguousarray(ext, f32)
    P = L.oracle_bin_lists(N, _p(m), _p(r), _p(th), _p(sl), _p(e), e.shape[0], H, W, ts, _p(offs), None)
    idx = np.empty(P, np.int32)
    L.oracle_bin_lists(N, _p(m), _p(r), _p(th), _p(sl), _p(e), e.shape[0], H, W, ts, _p(offs), _p(idx))
    return offs, idx


# --------------------------------------------------------------------------
# forward / backward (C)
# --------------------------------------------------------------------------

class OracleScene:
    """Keeps the contiguous arrays alive while the C struct points at them."""

    def __init__(self, H, W, means2D, radii, rotations, opacities, colors, templates, bmp_sel,
                 tile_offsets, tile_indices, *, ts=16, alpha_ub=1.0, colors_orig=None,
                 c_blend=0.0, geom64=False, tile_subset=None, blend_grad_exact=False,
                 colors_orig_per_template=False):
        c = lambda a, dt=f32: None if a is None else np.ascontiguousarray(a, dt)
        self.arrs = dict(
            means2D=c(means2D), radii=c(radii), rotations=c(rotations), opacities=c(opacities),
            colors=c(colors), colors_orig=c(colors_orig), templates=c(templates),
            bmp_sel=c(bmp_sel, np.int32), tile_offsets=c(tile_offsets, np.int32),
            tile_indices=c(tile_indices, np.int32), tile_subset=c(tile_subset, np.int32))
        a = self.arrs
        T, Ht, Wt = a["templates"].shape
        self.H, self.W, self.N = H, W, a["radii"].shape[0]
        self.s = _Scene(H=H, W=W, ts=ts, alpha_ub=alpha_ub, N=self.N, T=T, Ht=Ht, Wt=Wt,
                        means2D=_p(a["means2D"]), radii=_p(a["radii"]), rotations=_p(a["rotations"]),
                        opacities=_p(a["opacities"]), colors=_p(a["colors"]),
                        colors_orig=_p(a["colors_orig"]), c_blend=c_blend,
                        templates=_p(a["templates"]), bmp_sel=_p(a["bmp_sel"]),
                        tile_offsets=_p(a["tile_offsets"]), tile_indices=_p(a["tile_indices"]),
                        geom64=int(bool(geom64)), tile_subset=_p(a["tile_subset"]),
                        n_subset=0 if tile_subset is None else len(a["tile_subset"]),
                        blend_grad_exact=int(bool(blend_grad_exact)),
                        colors_orig_per_template=int(bool(colors_orig_per_template)))


def forward(sc: OracleScene, real: str = "f64"):
    """-> color (H,W,3) premultiplied over black, alpha (H,W), push count (H,W)."""
    dt = np.float64 if real == "f64" else np.float32
    col = np.zeros((sc.H, sc.W, 3), dt)
    alp = np.zeros((sc.H, sc.W), dt)
    cnt = np.zeros((sc.H, sc.W), np.int32)
    lib(real).oracle_forward(ctypes.byref(sc.s), _p(col), _p(alp), _p(cnt))
    return col, alp, cnt


def backward(sc: OracleScene, gC: np.ndarray, gA: np.ndarray | None, real: str = "f64"):
    """-> dict of gradients means2D (N,2), radii, rotations, opacities, colors (N,3)."""
    dt = np.float64 if real == "f64" else np.float32
    N = sc.N
    g = dict(means2D=np.zeros((N, 2), dt), radii=np.zeros(N, dt), rotations=np.zeros(N, dt),
             opacities=np.zeros(N, dt), colors=np.zeros((N, 3), dt))
    gC = np.ascontiguousarray(gC, dt)
    gA = None if gA is None else np.ascontiguousarray(gA, dt)
    lib(real).oracle_backward(ctypes.byref(sc.s), _p(gC), _p(gA), _p(g["means2D"]), _p(g["radii"]),
                              _p(g["rotations"]), _p(g["opacities"]), _p(g["colors"]))
    return g


def count_samples(sc: OracleScene) -> int:
    return int(lib("f64").oracle_count_samples(ctypes.byref(sc.s)))


def scene_from(scene, offs, idx, *, geom64=False, tile_subset=None, colors_orig=None,
               blend_grad_exact=False, **over):
    """OracleScene from a ``diffbmp_amd.scenes.Scene`` and tile lists.  The
    scene's colour maps are passed per template (indexed by bmp_sel) unless
    ``colors_orig`` gives them per primitive."""
    kw = dict(means2D=scene.means2D, radii=scene.radii, rotations=scene.rotations,
              opacities=scene.opacities, colors=scene.colors, templates=scene.templates)
    kw.update(over)
    co, per_tmpl = colors_orig, False
    if co is None and scene.c_blend > 0 and scene.colors_orig_tmpl is not None:
        co, per_tmpl = scene.colors_orig_tmpl, True
    return OracleScene(scene.H, scene.W, kw["means2D"], kw["radii"], kw["rotations"], kw["opacities"],
                       kw["colors"], kw["templates"], scene.bmp_sel, offs, idx, ts=scene.tile_size,
                       alpha_ub=scene.alpha_ub, colors_orig=co, c_blend=scene.c_blend,
                       geom64=geom64, tile_subset=tile_subset, blend_grad_exact=blend_grad_exact,
                       colors_orig_per_template=per_tmpl)


def bin_scene(scene):
    ext = template_extents_np(scene.templates)
    return bin_lists_np(scene.means2D, scene.radii, scene.rotations, scene.bmp_sel, ext,
                        scene.H, scene.W, scene.tile_size)
