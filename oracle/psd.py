"""CPU oracle for the layer export -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module; the product path never does.  A
numpy restatement, in fp32 with every operation rounded separately (no fused
multiply-add), of the reference's two export kernels and the host prefix sum
between them (paths relative to the reference tree):

* ``psd_bounds``  -- ``compute_bounding_boxes_kernel``,
  ``cuda_tile_rasterizer/cuda_kernels/psd_export_kernels.cu:12-55``, and the
  layer offsets of ``cuda_tile_rasterizer/psd_export_uint8.py:64-75``;
* ``psd_layers``  -- ``generate_cropped_layers_kernel`` (``:61-198``) with
  ``psd_grid_sample_bilinear`` (``psd_export_common.h:58-108``),
  ``psd_sample_color_map`` (``:129-196``) and ``float_to_uint8`` (``:123-125``).

cos/sin are the library's fp32 ``dbmp_sincosf`` (the source the GPU kernels
compile, through the C oracle) and the divisions are IEEE; the reference
builds with fast math, whose approximate
trig and division this restatement does not reproduce.  How each function is
pinned: ``tests/test_psd_oracle.py`` (hand-computed boxes, closed forms for
constant and texel-aligned templates, a rotation that must mirror an
asymmetric template, truncating byte conversion, blend weights).
"""
from __future__ import annotations

import numpy as np

from . import cos_sin

f32 = np.float32


def _to_int(v: np.ndarray) -> np.ndarray:
    return np.clip(v, f32(-1073741824.0), f32(1073741824.0)).astype(np.int64)


def psd_bounds(means2D, radii, H: int, W: int, scale: float):
    """Crop boxes (N,4) int32 (left, top, right, bottom), sizes (N,2) int32
    and byte offsets (N+1) int64 of the packed RGBA buffer
    (``psd_export_kernels.cu:22-54``, ``psd_export_uint8.py:66-72``)."""
    m = np.asarray(means2D, f32)
    sc = f32(scale)
    x = m[:, 0] * sc
    y = m[:, 1] * sc
    r = np.asarray(radii, f32) * sc
    tr = r * f32(1.5)
    x0 = np.maximum(0, _to_int(np.floor(x - tr)))
    y0 = np.maximum(0, _to_int(np.floor(y - tr)))
    x1 = np.minimum(W, _to_int(np.ceil(x + tr)))
    y1 = np.minimum(H, _to_int(np.ceil(y + tr)))
    empty = (x1 <= x0) | (y1 <= y0)
    boxes = np.stack([np.where(empty, 0, x0), np.where(empty, 0, y0),
                      np.where(empty, 1, x1), np.where(empty, 1, y1)], 1).astype(np.int32)
    sizes = np.stack([boxes[:, 2] - boxes[:, 0], boxes[:, 3] - boxes[:, 1]], 1).astype(np.int32)
    offs = np.zeros(len(boxes) + 1, np.int64)
    offs[1:] = np.cumsum(4 * sizes[:, 0].astype(np.int64) * sizes[:, 1])
    return boxes, sizes, offs


def _u8(v: np.ndarray) -> np.ndarray:
    """``(uint8_t)fmaxf(0, fminf(255, v * 255))``: truncation; fmin/fmax
    return the other operand for NaN, as C's do."""
    return np.fmax(f32(0), np.fmin(f32(255), v * f32(255))).astype(np.uint8)


def _template_value(tm, Ht, Wt, px, py, x, y, r, c, s):
    un = (px.astype(f32) - x) / r
    vn = (py.astype(f32) - y) / r
    ur = c * un + s * vn
    vr = -s * un + c * vn
    ut = (ur + f32(1.0)) * f32(0.5) * f32(Wt - 1)
    vt = (vr + f32(1.0)) * f32(0.5) * f32(Ht - 1)
    inside = (ut >= 0) & (ut < f32(Wt - 1)) & (vt >= 0) & (vt < f32(Ht - 1))
    uts = np.where(inside, ut, f32(0))
    vts = np.where(inside, vt, f32(0))
    u0 = np.clip(np.floor(uts).astype(np.int64), 0, Wt - 1)
    v0 = np.clip(np.floor(vts).astype(np.int64), 0, Ht - 1)
    u1 = np.minimum(u0 + 1, Wt - 1)
    v1 = np.minimum(v0 + 1, Ht - 1)
    wu = uts - u0.astype(f32)
    wv = vts - v0.astype(f32)
    wui = f32(1.0) - wu
    wvi = f32(1.0) - wv
    a00, a01 = tm[v0, u0], tm[v0, u1]
    a10, a11 = tm[v1, u0], tm[v1, u1]
    val = wui * wvi * a00 + wu * wvi * a01 + wui * wv * a10 + wu * wv * a11
    return np.where(inside, val, f32(0))


def _color_map(cm, Ht, Wt, dx, dy, radius, c, s):
    rinv = (f32(1.0) / radius) if radius > f32(1e-6) else f32(1e6)
    nx = dx * rinv
    ny = dy * rinv
    rx = c * nx + s * ny
    ry = -s * nx + c * ny
    mw1, mh1 = f32(Wt - 1), f32(Ht - 1)
    cx = np.fmax(f32(0), np.fmin(mw1, (rx + f32(1.0)) * f32(0.5) * mw1))
    cy = np.fmax(f32(0), np.fmin(mh1, (ry + f32(1.0)) * f32(0.5) * mh1))
    x0 = cx.astype(np.int64)
    y0 = cy.astype(np.int64)
    x1 = np.minimum(x0 + 1, Wt - 1)
    y1 = np.minimum(y0 + 1, Ht - 1)
    fx = (cx - x0.astype(f32))[..., None]
    fy = (cy - y0.astype(f32))[..., None]
    c00, c01 = cm[y0, x0], cm[y1, x0]
    c10, c11 = cm[y0, x1], cm[y1, x1]
    a = c00 * (f32(1.0) - fx) + c10 * fx
    b = c01 * (f32(1.0) - fx) + c11 * fx
    return a * (f32(1.0) - fy) + b * fy


def psd_layers(means2D, radii, rotations, colors, visibility, bmp_sel, templates, H, W,
               scale=1.0, alpha_ub=1.0, c_blend=0.0, cmaps=None, cmap_mode=0, boxes=None,
               offs=None, only=None):
    """Packed uint8 RGBA crops of every primitive (the buffer the reference's
    stage 2 fills), and each layer's largest alpha byte.  ``only``: compute
    just these layers (the others stay zero)."""
    m = np.asarray(means2D, f32)
    rad = np.asarray(radii, f32)
    rot = np.asarray(rotations, f32)
    col = np.asarray(colors, f32)
    vis = np.asarray(visibility, f32)
    sel = np.asarray(bmp_sel, np.int64)
    tms = np.asarray(templates, f32)
    T, Ht, Wt = tms.shape
    if boxes is None:
        boxes, _, offs = psd_bounds(m, rad, H, W, scale)
    sc = f32(scale)
    cb = f32(c_blend)
    out = np.zeros(int(offs[-1]), np.uint8)
    amax = np.zeros(len(rad), np.int32)
    for n in (range(len(rad)) if only is None else only):
        x0, y0, x1, y1 = (int(v) for v in boxes[n])
        t = int(sel[n])
        if not (0 <= t < T):
            continue   # no template: transparent
        py, px = np.mgrid[y0:y1, x0:x1]
        x, y, r = m[n, 0] * sc, m[n, 1] * sc, rad[n] * sc
        c, s = cos_sin(rot[n:n + 1])
        c, s = c[0], s[0]
        tv = _template_value(tms[t], Ht, Wt, px, py, x, y, r, c, s)
        alpha = tv * vis[n] * f32(alpha_ub)
        rgb = np.broadcast_to(col[n], tv.shape + (3,)).astype(f32)
        if cmaps is not None and cmap_mode != 0 and c_blend > 0.0:
            cm = np.asarray(cmaps, f32)[n if cmap_mode == 1 else t]
            co = _color_map(cm, Ht, Wt, px.astype(f32) / sc - x / sc, py.astype(f32) / sc - y / sc,
                            r / sc, c, s)
            rgb = (f32(1.0) - cb) * rgb + cb * co
        mask = np.where(tv > 0, f32(1.0), f32(0.0))[..., None]
        px_rgba = np.concatenate([_u8(mask * rgb), _u8(alpha)[..., None]], -1)
        px_rgba[~(alpha >= f32(0.001)) & ~np.isnan(alpha)] = 0   # alpha < 0.001: transparent
        out[offs[n]:offs[n + 1]] = px_rgba.reshape(-1)
        amax[n] = int(px_rgba[..., 3].max())
    return out, amax
