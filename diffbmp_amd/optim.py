"""Adam for the drop-in path: ``torch.optim.Adam``'s update, one HIP kernel.

The reference's fit loop (``simple_tile_renderer.py:961-968, 1125-1131``)
steps ``torch.optim.Adam`` over the five primitive tensors and then clamps
x, y and r (``simple_tile_renderer.py:122-128``).  torch runs that as a
handful of multi-tensor kernels plus one kernel per clamp.  ``FusedAdam``
keeps the ``torch.optim.Optimizer`` interface: parameter groups, ``step()``,
``zero_grad()`` and ``state_dict()``.  It updates up to five parameter tensors
per launch with ``dbmp_adam_step`` and applies an optional per-group clamp in
the same kernel.

Supported: fp32 CUDA parameters with dense fp32 gradients; ``lr``, ``betas``
and ``eps`` per group.  Not supported (they raise): weight decay, amsgrad,
maximize, and sparse or reduced-precision gradients.  The arithmetic is
torch's, in fp32: ``m = b1 m + (1-b1) g``, ``v = b2 v + (1-b2) g^2``,
``p -= lr/(1-b1^t) * m / (sqrt(v)/sqrt(1-b2^t) + eps)``.  Each operation may
round differently from torch's kernels, by an ulp or so.
"""
from __future__ import annotations

import ctypes

import torch

from . import _native
from ._native import Adam, Handle

__all__ = ["FusedAdam"]


class FusedAdam(torch.optim.Optimizer):
    """``torch.optim.Adam`` (no weight decay, no amsgrad) on fp32 device
    tensors, up to five tensors per kernel launch.

    ``clamp`` (per group, extension): ``None``, or ``(lo, hi)`` applied to
    every element after its update.  For an (N, 2) tensor, ``lo`` and ``hi``
    may be pairs, one bound per column.  For example, the reference's means2D
    clamp is ``((0, 0), (W - 1, H - 1))``.

    ``handle``: the rasterizer handle the launches go through (a
    ``TileRasterizer``'s ``.handle``); by default one of the optimizer's own.
    """

    def __init__(self, params, lr: float = 1e-3, betas=(0.9, 0.999), eps: float = 1e-8,
                 clamp=None, handle: Handle | None = None):
        if lr < 0.0 or eps < 0.0 or not (0.0 <= betas[0] < 1.0) or not (0.0 <= betas[1] < 1.0):
            raise ValueError(f"invalid Adam hyper-parameters lr={lr} betas={betas} eps={eps}")
        super().__init__(params, dict(lr=lr, betas=tuple(betas), eps=eps, clamp=clamp))
        self._handle = handle

    def _h(self, dev) -> Handle:
        if self._handle is None:
            self._handle = Handle(16, 16, 16, device=dev)
        return self._handle

    @staticmethod
    def _bounds(clamp, p):
        lo, hi = clamp
        pair = lambda b: (float(b[0]), float(b[1])) if isinstance(b, (tuple, list)) else (float(b),) * 2
        lo, hi = pair(lo), pair(hi)
        stride = 2 if (p.dim() == 2 and p.shape[1] == 2) else 1
        if stride == 1 and (lo[0] != lo[1] or hi[0] != hi[1]):
            raise ValueError("per-column clamp bounds need an (N, 2) parameter")
        return stride, lo, hi

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        # (step count, betas, eps) -> tensors of one launch
        batches: dict = {}
        for group in self.param_groups:
            for p in group["params"]:
                if p.grad is None:
                    continue
                g = p.grad
                if p.dtype != torch.float32 or g.dtype != torch.float32 or not p.is_cuda:
                    raise TypeError("FusedAdam takes fp32 device parameters and gradients")
                if g.is_sparse:
                    raise TypeError("FusedAdam does not take sparse gradients")
                if not p.is_contiguous() or not g.is_contiguous():
                    raise ValueError("FusedAdam takes contiguous parameters and gradients")
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                st["step"] += 1
                key = (st["step"], group["betas"], group["eps"], p.device)
                batches.setdefault(key, []).append((p, g, st, group))
        for (t, (b1, b2), eps, dev), items in batches.items():
            h = self._h(dev)
            for k in range(0, len(items), 5):
                cfg = Adam()
                chunk = items[k:k + 5]
                cfg.num_groups = len(chunk)
                cfg.beta1, cfg.beta2, cfg.eps = float(b1), float(b2), float(eps)
                cfg.step = int(t)
                cfg.store_dtype = _native._DT[torch.float32]
                for i, (p, g, st, group) in enumerate(chunk):
                    c = cfg.groups[i]
                    c.param, c.grad = p.data_ptr(), g.data_ptr()
                    c.exp_avg, c.exp_avg_sq = st["exp_avg"].data_ptr(), st["exp_avg_sq"].data_ptr()
                    c.store = None
                    c.numel = p.numel()
                    c.lr = float(group["lr"])
                    c.grad_row, c.grad_comps, c.zero_grad = 0, 1, 0
                    if group["clamp"] is not None:
                        c.stride, lo, hi = self._bounds(group["clamp"], p)
                        c.clamp = 1
                        c.lo[0], c.lo[1], c.hi[0], c.hi[1] = lo[0], lo[1], hi[0], hi[1]
                    else:
                        c.stride, c.clamp = 1, 0
                with torch.cuda.device(dev):
                    h.adam_step(cfg, dev)
        return loss
