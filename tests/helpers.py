"""Shared test utilities: scene -> device tensors, oracle comparisons."""
from __future__ import annotations

import numpy as np
import torch

import oracle as O
from diffbmp_amd import _native
from diffbmp_amd._native import Handle


def to_dev(scene, dev, dtype=torch.float32, fp32_geometry=False):
    t = lambda a, dt=dtype: torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(dt)
    g = torch.float32 if fp32_geometry else dtype
    return dict(means2D=t(scene.means2D, g), radii=t(scene.radii, g), rotations=t(scene.rotations, g),
                opacities=t(scene.opacities), colors=t(scene.colors),
                templates=t(scene.templates), bmp_sel=t(scene.bmp_sel, torch.int32))


def rounded(scene, dtype, fp32_geometry=False):
    """Scene whose float arrays are rounded to `dtype` storage and back to fp32
    (positions, radii and rotations left in fp32 with ``fp32_geometry``)."""
    import copy
    s = copy.copy(scene)
    r = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dtype).float().numpy()
    if not fp32_geometry:
        s.means2D, s.radii, s.rotations = r(scene.means2D), r(scene.radii), r(scene.rotations)
    s.opacities, s.colors, s.templates = r(scene.opacities), r(scene.colors), r(scene.templates)
    if scene.colors_orig_tmpl is not None:
        s.colors_orig_tmpl = r(scene.colors_orig_tmpl)
    return s


class DeviceRun:
    """Bins, renders and back-propagates one scene through the C ABI."""

    def __init__(self, scene, dev, dtype=torch.float32, eps_T=0.0, ts=16, co_mode=0,
                 bm_dtype=None, exact_blend=False, fp32_geometry=False):
        """``dtype``: parameter storage (positions, radii and rotations fp32
        with ``fp32_geometry``); ``bm_dtype``: bitmap and colour-map storage
        (default ``dtype``)."""
        self.scene, self.dev, self.dtype, self.eps = scene, dev, dtype, eps_T
        bm_dtype = dtype if bm_dtype is None else bm_dtype
        self.h = Handle(scene.H, scene.W, ts, 0.0, scene.alpha_ub, 151, scene.N)
        self.d = to_dev(scene, dev, dtype, fp32_geometry)
        d = self.d
        d["templates"] = d["templates"].to(bm_dtype)
        self.prims = Handle.prims(d["means2D"], d["radii"], d["rotations"], d["opacities"],
                                  d["colors"], d["bmp_sel"])
        self.co = None
        if co_mode:
            src = scene.colors_orig_tmpl if co_mode == 2 else scene.colors_orig_per_prim()
            self.co = torch.from_numpy(np.ascontiguousarray(src)).to(dev).to(bm_dtype)
        self.bm = Handle.bitmaps(d["templates"], self.co, co_mode, scene.c_blend, -1, exact_blend)
        self.ext = self.h.template_extents(d["templates"])
        self.offs, self.idx = self.h.bin(self.prims, self.ext, tuple(d["templates"].shape))

    def forward(self):
        H, W, dev = self.scene.H, self.scene.W, self.dev
        self.color = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
        self.alpha = torch.empty((H, W), dtype=torch.float32, device=dev)
        self.T = torch.empty((H, W), dtype=torch.float32, device=dev)
        self.last = torch.empty((H, W), dtype=torch.int32, device=dev)
        self.h.forward(self.prims, self.bm, _native.Lists(self.offs, self.idx), self.eps, self.color,
                       self.alpha, self.T, self.last)
        torch.cuda.synchronize()
        return self.color.cpu().numpy(), self.alpha.cpu().numpy()

    def backward(self, gC, gA):
        N, dev = self.scene.N, self.dev
        gCt = torch.from_numpy(np.ascontiguousarray(gC, np.float32)).to(dev)
        gAt = None if gA is None else torch.from_numpy(np.ascontiguousarray(gA, np.float32)).to(dev)
        g = dict(means2D=torch.empty((N, 2), device=dev), radii=torch.empty(N, device=dev),
                 rotations=torch.empty(N, device=dev), opacities=torch.empty(N, device=dev),
                 colors=torch.empty((N, 3), device=dev))
        self.h.backward(self.prims, self.bm, _native.Lists(self.offs, self.idx), self.eps, gCt, gAt, self.color,
                        self.T, self.last, g["means2D"], g["radii"], g["rotations"],
                        g["opacities"], g["colors"])
        torch.cuda.synchronize()
        return {k: v.cpu().numpy() for k, v in g.items()}


def normwise(g, ref):
    """Per-group ||g - ref|| / ||ref||."""
    out = {}
    for k in ref:
        den = np.linalg.norm(ref[k].ravel())
        out[k] = float(np.linalg.norm((g[k] - ref[k]).ravel()) / (den if den > 0 else 1.0))
    return out


def rand_grads(H, W, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((H, W, 3)) * scale).astype(np.float32), \
           (rng.standard_normal((H, W)) * scale).astype(np.float32)


def half_rounded(a):
    """fp32 array rounded to fp16 storage and back."""
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).half().float().numpy()


def fd_window(templates, support=0.95):
    """Templates times the C2 window ``(1 - (rho/R)^2)^3`` (zero beyond R =
    ``support`` of the half size), so that every bitmap is exactly zero over
    its outer texels and the push box never cuts a nonzero mask."""
    T, Ht, Wt = templates.shape
    yy, xx = np.mgrid[0:Ht, 0:Wt].astype(np.float64)
    hy, hx = (Ht - 1) / 2, (Wt - 1) / 2
    rho2 = ((xx - hx) / (hx * support)) ** 2 + ((yy - hy) / (hy * support)) ** 2
    win = np.where(rho2 < 1.0, (1.0 - rho2) ** 3, 0.0)
    return (templates * win[None]).astype(np.float32)


def fd_scalars(scene, offs, idx, g, gC, gA, picks, h, **oracle_kw):
    """Central differences of L = sum(gC * C) + sum(gA * A) (fp64 oracle,
    geometry in fp64, tile lists frozen) for each (key, n, j) in ``picks``.
    L is summed over the tiles whose lists hold primitive n, the only pixels
    its parameters reach.  The step is max(h, 2 ulp) of the fp32 parameter
    and the step actually stored is used.  -> (fd, analytic) arrays."""
    tiles = len(offs) - 1
    tX = -(-scene.W // scene.tile_size)
    tile_of = np.repeat(np.arange(tiles), np.diff(offs))
    ts = scene.tile_size
    fd, an = [], []
    for key, n, j in picks:
        sub = np.ascontiguousarray(tile_of[idx == n], np.int32)
        m = np.zeros(tiles, bool)
        m[sub] = True
        pm = np.kron(m.reshape(-1, tX), np.ones((ts, ts), bool))[:scene.H, :scene.W]
        arr = getattr(scene, key)
        base = arr.copy()
        ix = (n,) if arr.ndim == 1 else (n, j)
        step = max(h, 2.0 * float(np.spacing(np.float32(base[ix]))))
        vals = []
        for sgn in (1.0, -1.0):
            a2 = base.copy()
            a2[ix] = base[ix] + sgn * step
            setattr(scene, key, a2)
            osc = O.scene_from(scene, offs, idx, geom64=True, tile_subset=sub, **oracle_kw)
            col, al, _ = O.forward(osc)
            vals.append(((col * gC)[pm].sum() + (al * gA)[pm].sum(), float(np.float32(a2[ix]))))
        setattr(scene, key, base)
        fd.append((vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1]))
        an.append(float(g[key][ix]))
    return np.array(fd), np.array(an)
