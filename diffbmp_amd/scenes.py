"""Seeded synthetic workloads for the tile rasterizer.

Every array is generated with ``numpy.random.default_rng(1000 * cfg + role)``
(role 0 = primitive parameters, 1 = target parameters / target field,
2 = bitmaps, 3 = extra), so a workload is fully determined by its config id
and shapes.  Parameter conventions follow the reference renderer:

* ``means2D`` (N,2) pixel coordinates [x, y] (pixel centres are integers,
  ``pydiffbmp/core/renderer/vector_renderer.py:135-149``);
* ``radii`` (N,) half-extent in pixels, ``r = exp(ls) * W / 2`` with
  ``ls ~ U(log a, log b)``;
* ``rotations`` (N,) radians; ``opacities`` (N,) logits ~ N(0,1);
* ``colors`` (N,3) logits, ``logit(u)`` with ``u ~ U(0.001, 0.999)``;
* ``bmp_sel`` (N,) the reference's cyclic assignment
  ``flip(arange(N) % T)`` (``simple_tile_renderer.py:204-207``), so the binning
  template choice (``vector_renderer.py:221``) and the kernel's agree.

Workloads (canvas, primitive count, bitmaps):

* ``plain``      256x256, 256 primitives, one 64x64 soft disk (radius 28 px,
  sigmoid edge of 10-90 % width 2 px), log-scale in [log .05, log .5];
* ``scale``      1024x1024, 8192 primitives, four 128x128 bitmaps (soft disk,
  five-point star, ring, blurred thresholded noise), same distributions;
* ``precision``  same as ``scale``, parameters and bitmaps stored as fp16/bf16;
* ``stress``     2048x2048, 32768 primitives, sixteen 256x256 RGBA bitmaps
  (RGB smoothed noise, alpha from the four shape families at random size and
  offset), log-scale in [log .005, log .1], colour-map blend on.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Scene", "make_scene", "shape_bitmap", "smooth_field", "cyclic_sel",
    "CONFIGS", "plain_disk", "margin_bitmaps", "hard_square",
]


@dataclass
class Scene:
    """One rasterizer workload held in host memory (numpy, fp32)."""
    H: int
    W: int
    means2D: np.ndarray
    radii: np.ndarray
    rotations: np.ndarray
    opacities: np.ndarray
    colors: np.ndarray
    templates: np.ndarray
    bmp_sel: np.ndarray
    tile_size: int = 16
    alpha_ub: float = 1.0
    c_blend: float = 0.0
    colors_orig_tmpl: np.ndarray | None = None   # (T,Ht,Wt,3) per-template colour maps
    target: np.ndarray | None = None             # (H,W,3)
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.radii.shape[0])

    def colors_orig_per_prim(self) -> np.ndarray:
        """(N,Ht,Wt,3) gather ``c_o[global_bmp_sel]`` as the reference passes it
        (``simple_tile_renderer.py:593``)."""
        T, Ht, Wt = self.templates.shape
        if self.colors_orig_tmpl is None:
            return np.zeros((self.N, Ht, Wt, 3), np.float32)
        return self.colors_orig_tmpl[self.bmp_sel]


def cyclic_sel(N: int, T: int) -> np.ndarray:
    """``torch.arange(N) % T`` flipped (``simple_tile_renderer.py:204-207``)."""
    return ((np.arange(N) % T)[::-1]).astype(np.int32).copy()


def _sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def _gauss_blur(img: np.ndarray, sigma: float, wrap: bool = True) -> np.ndarray:
    """Separable Gaussian blur (truncated at 4 sigma)."""
    rad = int(math.ceil(4 * sigma))
    xs = np.arange(-rad, rad + 1, dtype=np.float64)
    k = np.exp(-0.5 * (xs / sigma) ** 2)
    k /= k.sum()
    out = img.astype(np.float64)
    for ax in (0, 1):
        if wrap:
            acc = np.zeros_like(out)
            for i, w in enumerate(k):
                acc += w * np.roll(out, i - rad, axis=ax)
            out = acc
        else:
            pad = [(0, 0)] * out.ndim
            pad[ax] = (rad, rad)
            p = np.pad(out, pad, mode="constant")
            acc = np.zeros_like(out)
            n = out.shape[ax]
            for i, w in enumerate(k):
                sl = [slice(None)] * out.ndim
                sl[ax] = slice(i, i + n)
                acc += w * p[tuple(sl)]
            out = acc
    return out


def plain_disk(size: int = 64, radius: float = 28.0, edge: float = 2.0) -> np.ndarray:
    """Soft disk: ``1/(1+exp((rho-R)/s))`` with ``s = edge / (2 ln 9)`` so the
    10-90 % transition is ``edge`` px wide; centre ((size-1)/2, (size-1)/2)."""
    c = (size - 1) / 2.0
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float64)
    rho = np.hypot(xx - c, yy - c)
    s = edge / (2.0 * math.log(9.0))
    return _sigmoid((radius - rho) / s).astype(np.float32)


def shape_bitmap(kind: str, size: int, rng: np.random.Generator | None = None,
                 scale: float = 1.0, offset=(0.0, 0.0)) -> np.ndarray:
    """Procedural alpha bitmap in [0,1] of one of four families.

    ``scale`` shrinks the shape about the centre, ``offset`` shifts it
    (fractions of the half-size).  The edge is a sigmoid ``size/64`` px wide.
    """
    h = (size - 1) / 2.0
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float64)
    x = (xx - h) / h / scale - offset[0]
    y = (yy - h) / h / scale - offset[1]
    rho = np.hypot(x, y)
    s = (size / 64.0) / h / scale * 0.5
    if kind == "disk":
        d = 0.8 - rho
    elif kind == "star":
        phi = np.arctan2(y, x)
        rb = 0.45 + 0.4 * np.cos(5.0 * phi - math.pi / 2) ** 2
        d = rb - rho
    elif kind == "ring":
        d = 0.14 - np.abs(rho - 0.6)
    elif kind == "noise":
        if rng is None:
            rng = np.random.default_rng(0)
        white = rng.standard_normal((size, size))
        thr = (white > 0).astype(np.float64)
        blob = _gauss_blur(thr, 4.0 * size / 128.0)
        z = (blob - blob.mean()) / (blob.std() + 1e-12)
        win = np.clip((0.85 - rho) / (2 * s + 1e-12), -30, 30)
        return (_sigmoid(2.5 * z) * _sigmoid(win)).astype(np.float32)
    else:
        raise ValueError(f"unknown shape family {kind!r}")
    return _sigmoid(np.clip(d / s, -60, 60)).astype(np.float32)


SHAPES = ("disk", "star", "ring", "noise")


def hard_square(size: int = 64, inset: int | None = None) -> np.ndarray:
    """Binary centre square: 1 on texels ``[inset, size-1-inset]`` in both
    axes, 0 elsewhere (``inset`` defaults to ``size // 4``).  Its content does
    not reach the bitmap's edge and its bilinear fringe is hard: a sample up to
    one texel outside the content rows and columns still has a mask > 1e-6,
    so a primitive draws a little beyond the box its template extents give
    (the binning box).  The reference's own bitmaps are hard-edged like this
    (e.g. ``assets/primitives/circle.png``)."""
    k = size // 4 if inset is None else int(inset)
    out = np.zeros((size, size), np.float32)
    out[k:size - k, k:size - k] = 1.0
    return out


def smooth_field(H: int, W: int, rng: np.random.Generator, terms: int = 32) -> np.ndarray:
    """``clamp(0.5 + sum_i (a_i/8) sin(2 pi (kx_i x/W + ky_i y/H) + phi_i), 0, 1)``
    per channel with a ~ U[0,1], k ~ U[0.5, 8], phi ~ U[0, 2 pi)."""
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    out = np.empty((H, W, 3), np.float32)
    for ch in range(3):
        a = rng.uniform(0, 1, terms)
        kx = rng.uniform(0.5, 8, terms)
        ky = rng.uniform(0.5, 8, terms)
        ph = rng.uniform(0, 2 * math.pi, terms)
        f = np.full((H, W), 0.5)
        for i in range(terms):
            f += (a[i] / 8.0) * np.sin(2 * math.pi * (kx[i] * xx / W + ky[i] * yy / H) + ph[i])
        out[..., ch] = np.clip(f, 0, 1)
    return out


def _params(rng: np.random.Generator, N: int, H: int, W: int, ls_lo: float, ls_hi: float):
    x = rng.uniform(0, W, N)
    y = rng.uniform(0, H, N)
    ls = rng.uniform(math.log(ls_lo), math.log(ls_hi), N)
    r = np.exp(ls) * W / 2.0
    th = rng.uniform(0, 2 * math.pi, N)
    u = rng.uniform(0.001, 0.999, (N, 3))
    c = np.log(u / (1 - u))
    v = rng.standard_normal(N)
    f = np.float32
    return (np.stack([x, y], 1).astype(f), r.astype(f), th.astype(f), v.astype(f), c.astype(f))


def margin_bitmaps(T: int, size: int, rng: np.random.Generator, margin: int = 3) -> np.ndarray:
    """Smooth bitmaps whose outer ``margin`` texels are exactly zero and whose
    support is either exactly zero or above 1e-3, so the push box and the
    ``mask <= 1e-6`` gate never cut a nonzero mask (finite-difference scenes)."""
    out = np.zeros((T, size, size), np.float32)
    inner = size - 2 * margin
    for t in range(T):
        b = shape_bitmap(SHAPES[t % 4], inner, rng)
        b = np.where(b < 1e-3, 0.0, b).astype(np.float32)
        out[t, margin:margin + inner, margin:margin + inner] = b
    return out


def smooth_bump_bitmaps(T: int, size: int, rng: np.random.Generator, support: float = 0.62) -> np.ndarray:
    """C2-smooth bitmaps with compact support (exactly zero beyond radius
    ``support`` of the half size): ``(1 - (rho/R)^2)^3`` times a smooth
    harmonic modulation.  Bilinear interpolation of such a bitmap has small
    slope jumps across texel lines, so central differences with h = 1e-3
    resolve its gradients; the support lies inside every rotation of the push
    box, so neither the box nor the texture bounds ever cut a nonzero mask."""
    h = (size - 1) / 2.0
    yy, xx = np.mgrid[0:size, 0:size].astype(np.float64)
    x, y = (xx - h) / h / support, (yy - h) / h / support
    rho2 = x * x + y * y
    base = np.where(rho2 < 1.0, (1.0 - rho2) ** 3, 0.0)
    out = np.zeros((T, size, size), np.float32)
    for t in range(T):
        a, b = rng.uniform(-0.35, 0.35, 2)
        mod = 0.6 + a * (x ** 3 - 3 * x * y * y) + b * (3 * x * x * y - y ** 3)
        out[t] = np.clip(base * mod, 0.0, 1.0)
    out[out < 1e-5] = 0.0
    return out


CONFIGS = {
    0: dict(name="plain", H=256, W=256, N=256, T=1, tsz=64, ls=(0.05, 0.5)),
    1: dict(name="scale", H=1024, W=1024, N=8192, T=4, tsz=128, ls=(0.05, 0.5)),
    2: dict(name="precision", H=1024, W=1024, N=8192, T=4, tsz=128, ls=(0.05, 0.5)),
    3: dict(name="stress", H=2048, W=2048, N=32768, T=16, tsz=256, ls=(0.005, 0.1)),
}


def make_scene(cfg: int, H: int | None = None, W: int | None = None, N: int | None = None,
               tsz: int | None = None, with_target: bool = True) -> Scene:
    """Build workload ``cfg`` (0..3); sizes may be overridden to shrink it for
    tests while keeping the same generators."""
    spec = dict(CONFIGS[cfg])
    H = H or spec["H"]
    W = W or spec["W"]
    N = N or spec["N"]
    tsz = tsz or spec["tsz"]
    T = spec["T"]
    rp = np.random.default_rng(1000 * cfg + 0)
    rb = np.random.default_rng(1000 * cfg + 2)
    means, r, th, v, c = _params(rp, N, H, W, *spec["ls"])
    colors_orig = None
    c_blend = 0.0
    if cfg == 0:
        tm = plain_disk(tsz, radius=28.0 * tsz / 64.0, edge=2.0 * tsz / 64.0)[None]
    elif cfg in (1, 2):
        tm = np.stack([shape_bitmap(SHAPES[i], tsz, rb) for i in range(T)])
    else:
        tm = np.zeros((T, tsz, tsz), np.float32)
        colors_orig = np.zeros((T, tsz, tsz, 3), np.float32)
        for i in range(T):
            sc = rb.uniform(0.5, 1.0)
            off = rb.uniform(-0.3, 0.3, 2) * (1 - sc)
            tm[i] = shape_bitmap(SHAPES[i % 4], tsz, rb, scale=sc, offset=tuple(off))
            for ch in range(3):
                f = _gauss_blur(rb.uniform(0, 1, (tsz, tsz)), 8.0 * tsz / 256.0)
                f = (f - f.min()) / (f.max() - f.min() + 1e-12)
                colors_orig[i, :, :, ch] = f
        c_blend = 0.5
    sel = cyclic_sel(N, T)
    target = None
    if with_target:
        rt = np.random.default_rng(1000 * cfg + 1)
        if cfg == 0:
            target = None  # a second random render; built by the caller with make_target_params
        else:
            target = smooth_field(H, W, rt)
    return Scene(H=H, W=W, means2D=means, radii=r, rotations=th, opacities=v, colors=c,
                 templates=tm.astype(np.float32), bmp_sel=sel, c_blend=c_blend,
                 colors_orig_tmpl=colors_orig, target=target, name=spec["name"],
                 meta=dict(cfg=cfg))


def make_target_params(cfg: int, H: int, W: int, N: int):
    """Second random parameter set (role 1) for the plain workload's target
    render: a different seed, the same distributions."""
    spec = CONFIGS[cfg]
    rt = np.random.default_rng(1000 * cfg + 1)
    return _params(rt, N, H, W, *spec["ls"])
