"""Device-resident fitting of primitive parameters to a target image.

A :class:`FitJob` keeps everything of one fit on the GPU -- fp32 master
parameters (and their fp16 / bf16 storage copies when the job runs at reduced
precision), Adam moments, the target, tile-list and image buffers -- and one
:meth:`FitJob.step` is the reference's optimisation step
(``pydiffbmp/core/renderer/simple_tile_renderer.py:1002-1131``): binning,
forward composite, background composite + MSE + gradient, backward adjoint,
Adam, clamp.  Every stage is a call into ``libdiffbmp_amd`` on the current
stream; there are no torch elementwise kernels in the step and no host
synchronisation once the tile-list capacity has settled.
"""
from __future__ import annotations

import math

import torch

from . import _native
from ._native import Adam, AdamGroup, Handle

__all__ = ["FitJob", "GROUPS"]

GROUPS = ("means2D", "radii", "rotations", "opacities", "colors")
# first column and width of each group in the (N,8) gradient rows
_ROW_COLUMN = {"means2D": (0, 2), "radii": (2, 1), "rotations": (3, 1), "opacities": (4, 1),
               "colors": (5, 3)}


class FitJob:
    """One fit: ``H x W`` canvas, N primitives over ``templates``.

    ``lr`` is a float (every group) or a dict per group name.  ``store_dtype``
    is the storage of parameters and bitmaps seen by the rasterizer
    (torch.float32 / float16 / bfloat16); arithmetic, accumulation and the
    Adam master weights stay fp32.  ``bg`` is a (H,W,3) tensor or None for white.
    ``max_steps`` sizes the on-device loss history.  ``exact_color_map_grad``
    selects the exact gradient of the colour-map blend (``c_blend > 0``)
    instead of the reference's.  ``fp32_geometry`` keeps positions, radii and
    rotations in fp32 when ``store_dtype`` is fp16 / bf16 (mixed storage).
    """

    def __init__(self, means2D, radii, rotations, opacities, colors, templates, bmp_sel, target,
                 *, lr=1e-2, betas=(0.9, 0.999), eps=1e-8, store_dtype=torch.float32,
                 eps_T: float = 1e-6, tile_size: int = 16, alpha_upper_bound: float = 1.0,
                 clamp: bool = True, bg=None, max_steps: int = 1000, handle: Handle | None = None,
                 colors_orig=None, c_blend: float = 0.0, colors_orig_per_template: bool = True,
                 exact_color_map_grad: bool = False, fp32_geometry: bool = False):
        dev = means2D.device
        _native.require_gpu(means2D, "means2D")
        H, W = int(target.shape[0]), int(target.shape[1])
        self.H, self.W, self.N = H, W, int(radii.shape[0])
        self.device = dev
        self.store_dtype = store_dtype
        self.eps_T = float(eps_T)
        self.handle = handle or Handle(H, W, tile_size, 0.0, alpha_upper_bound, 151, self.N, device=dev)
        if (self.handle.H, self.handle.W) != (H, W):
            raise ValueError("handle canvas does not match the target")
        f32 = torch.float32
        self.params = {k: t.detach().to(dev, f32).contiguous().clone()
                       for k, t in zip(GROUPS, (means2D, radii, rotations, opacities, colors))}
        self.exp_avg = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.exp_avg_sq = {k: torch.zeros_like(v) for k, v in self.params.items()}
        # gradient rows (N,8): x, y, r, theta, opacity logit, 3 colour logits.
        # The backward adds into them and the Adam step zeroes what it reads.
        self.grad_rows = torch.zeros((self.N, 8), dtype=f32, device=dev)
        # fp32_geometry: positions, radii and rotations keep their fp32 master
        # values (the rasterizer reads them directly); only opacities, colours
        # and the bitmaps are stored in store_dtype
        geom = ("means2D", "radii", "rotations")
        self.fp32_geometry = bool(fp32_geometry) and store_dtype != f32
        if store_dtype == f32:
            self.store = self.params
        else:
            self.store = {k: (v if self.fp32_geometry and k in geom else v.to(store_dtype))
                          for k, v in self.params.items()}
        self.sel = bmp_sel.to(dev, torch.int32).contiguous()
        self.templates, self.ext, self.tmpl_version = self.handle.prepared_templates(
            templates.to(dev), store_dtype)
        self.colors_orig, self.co_mode, self.c_blend = None, 0, float(c_blend)
        if colors_orig is not None and c_blend > 0:
            # owned copy: the colour-map quads built from it are cached for the job
            self.colors_orig = colors_orig.to(dev, store_dtype).contiguous().clone()
            self.co_mode = 2 if colors_orig_per_template else 1
        self.target = target.to(dev, f32).contiguous()
        self.bg = None if bg is None else bg.to(dev, f32).contiguous()
        self.color = torch.empty((H, W, 3), dtype=f32, device=dev)
        self.alpha = torch.empty((H, W), dtype=f32, device=dev)
        self.T = torch.empty((H, W), dtype=f32, device=dev)
        self.last = torch.empty((H, W), dtype=torch.int32, device=dev)
        self.losses = torch.zeros(max_steps, dtype=f32, device=dev)
        self.t = 0
        lrs = lr if isinstance(lr, dict) else {k: float(lr) for k in GROUPS}
        self.adam = Adam()
        self.adam.num_groups = 5
        self.adam.beta1, self.adam.beta2, self.adam.eps = float(betas[0]), float(betas[1]), float(eps)
        self.adam.store_dtype = _native._DT[store_dtype]
        rmax = float(min(H, W) // 4)
        for i, k in enumerate(GROUPS):
            g = self.adam.groups[i]
            p = self.params[k]
            g.param, g.exp_avg, g.exp_avg_sq = p.data_ptr(), self.exp_avg[k].data_ptr(), self.exp_avg_sq[k].data_ptr()
            col, comps = _ROW_COLUMN[k]
            g.grad = self.grad_rows.data_ptr() + 4 * col
            g.grad_row, g.grad_comps, g.zero_grad = 8, comps, 1
            g.store = (self.store[k].data_ptr()
                       if store_dtype != f32 and not (self.fp32_geometry and k in geom) else None)
            g.numel = p.numel()
            g.lr = float(lrs[k])
            g.stride = 2 if k == "means2D" else 1
            g.clamp = int(clamp and k in ("means2D", "radii"))
            if k == "means2D":
                g.lo[0], g.hi[0], g.lo[1], g.hi[1] = 0.0, float(W - 1), 0.0, float(H - 1)
            elif k == "radii":
                g.lo[0], g.hi[0], g.lo[1], g.hi[1] = 2.0, rmax, 2.0, rmax
        self._bm = Handle.bitmaps(self.templates, self.colors_orig, self.co_mode, self.c_blend,
                                  self.tmpl_version, exact_color_map_grad, colors_version=0)

    def _prims(self, reuse: bool):
        s = self.store
        return Handle.prims(s["means2D"], s["radii"], s["rotations"], s["opacities"], s["colors"],
                            self.sel, reuse=reuse)

    def step(self) -> None:
        """One optimisation step, queued on the current stream."""
        if self.t >= self.losses.numel():
            grown = torch.zeros(2 * self.losses.numel(), dtype=torch.float32, device=self.device)
            grown[: self.losses.numel()] = self.losses
            self.losses = grown
        h = self.handle
        lists = h.bin_lazy(self._prims(False), self.ext, tuple(self.templates.shape), h.bin_factor)
        self.lists = lists
        h.forward(self._prims(True), self._bm, lists, self.eps_T, self.color, self.alpha, self.T, self.last)
        # background composite + MSE + its gradient fused into the backward
        # pass, added into the gradient rows the Adam step then reads and clears
        h.backward_mse_rows(self._prims(True), self._bm, lists, self.eps_T, self.target, self.color,
                            self.T, self.last, self.grad_rows,
                            loss_accum=self.losses[self.t:self.t + 1], bg=self.bg, bg_value=1.0)
        self.t += 1
        self.adam.step = self.t
        h.adam_step(self.adam, self.device)

    def fwd_bwd(self) -> None:
        """Binning, forward and backward of the current parameters, without the
        Adam update: the hot path alone.  The gradients are added into
        ``grad_rows`` (call ``grad_rows.zero_()`` before the next ``step``) and
        the loss into a scratch value, not the loss history."""
        h = self.handle
        if not hasattr(self, "_loss_scratch"):
            self._loss_scratch = torch.zeros(1, dtype=torch.float32, device=self.device)
        lists = h.bin_lazy(self._prims(False), self.ext, tuple(self.templates.shape), h.bin_factor)
        h.forward(self._prims(True), self._bm, lists, self.eps_T, self.color, self.alpha, self.T, self.last)
        h.backward_mse_rows(self._prims(True), self._bm, lists, self.eps_T, self.target, self.color,
                            self.T, self.last, self.grad_rows, loss_accum=self._loss_scratch, bg=self.bg,
                            bg_value=1.0)

    def render(self) -> torch.Tensor:
        """Composite image (H,W,3) of the current parameters over the background."""
        h = self.handle
        lists = h.bin_lazy(self._prims(False), self.ext, tuple(self.templates.shape), h.bin_factor)
        h.forward(self._prims(True), self._bm, lists, self.eps_T, self.color, self.alpha, self.T, self.last)
        bg = self.bg if self.bg is not None else torch.ones_like(self.color)
        return self.color + (1.0 - self.alpha).unsqueeze(-1) * bg

    def loss_history(self) -> torch.Tensor:
        """Per-step MSE (synchronises)."""
        return self.losses[: self.t].clone()

    @staticmethod
    def psnr(mse: float) -> float:
        return 10.0 * math.log10(1.0 / max(mse, 1e-30))
