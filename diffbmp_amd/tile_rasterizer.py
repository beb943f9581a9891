"""Reference-compatible autograd boundary of the MI355X tile rasterizer.

Mirrors ``cuda_tile_rasterizer/__init__.py`` (``TileRasterizer`` :60-121,
``TileRasterizerBatchFunction`` :124-252, ``TileRasterizerFunction`` :254-340,
``TileRasterizerFunctionFP16`` :342-436, timing helpers :439-465): same class
and function names, argument order and meaning, 16-argument ``apply`` with a
16-entry backward.  A renderer written against the reference switches by
importing these names from ``diffbmp_amd.tile_rasterizer``.

Differences a caller can observe, all deliberate:

* forward -> backward state (per-pixel ``T_final`` and list position) is
  kept in ``ctx``, not in a process-wide cache, so a second forward before the
  first backward is safe (the reference's batch backward reads the last
  candidate's cache for every candidate, ``tile_rasterize.cu:533-628``);
* ``tile_primitive_mapping=None`` (or empty) bins on the GPU instead of reusing
  the previous mapping (``tile_rasterize.cu:323``);
* ``colors_orig`` is read only when ``c_blend > 0`` (as the kernel does) and is
  not cast or copied otherwise;
* FP16 storage uses fp32 transmittance / colour accumulation and fp32
  gradients (the reference accumulates in half, ``tile_forward_fp16.cu:373-401``);
  outputs are returned as fp16 like the reference;
* a pixel stops compositing once its transmittance is below ``eps_T``
  (``DEFAULT_EPS_T``); the dropped colour is below that bound.  Set it to 0 to
  walk every list entry as the reference does;
* errors raise (``NativeError`` / ``RuntimeError``); there is no silent
  fallback to a different renderer.
"""
from __future__ import annotations

import threading

import torch
from torch.autograd import Function

from . import _native
from ._native import Handle

__all__ = [
    "TileRasterizer", "TileRasterizerFunction", "TileRasterizerFunctionFP16",
    "TileRasterizerBatchFunction", "CUDA_AVAILABLE", "CUDA_FP16_AVAILABLE",
    "print_cuda_timing_stats", "reset_cuda_timing_stats", "print_cuda_timing_stats_fp16",
    "reset_cuda_timing_stats_fp16", "DEFAULT_EPS_T", "bin_primitives", "pack_mapping",
    "unpack_mapping", "CompositeMSELoss", "composite_mse_loss",
]

DEFAULT_EPS_T = 1e-6


def _native_available() -> bool:
    try:
        _native.lib()
        return True
    except Exception:
        return False


# Names kept from the reference: "CUDA" here is torch's device namespace,
# which PyTorch-ROCm maps to HIP.
CUDA_AVAILABLE = _native_available()
CUDA_FP16_AVAILABLE = CUDA_AVAILABLE

# The active rasterizer per precision, set by TileRasterizer.__init__ like the
# reference's init_tile_rasterizer (ext.cpp:5-8).  Only configuration lives
# here; per-call state travels in autograd ctx.
_ACTIVE: dict[str, Handle] = {}
_ACTIVE_LOCK = threading.Lock()


def _active(kind: str) -> Handle:
    h = _ACTIVE.get(kind)
    if h is None:
        raise RuntimeError("TileRasterizer not initialized. Call init_tile_rasterizer first.")
    return h


def pack_mapping(tile_offsets: torch.Tensor, tile_indices: torch.Tensor) -> torch.Tensor:
    """Reference wire format ``[tiles+1, P, offsets..., indices...]`` int32
    (``simple_tile_renderer.py:757-766``)."""
    hdr = torch.tensor([tile_offsets.numel(), tile_indices.numel()], dtype=torch.int32,
                       device=tile_offsets.device)
    return torch.cat([hdr, tile_offsets.to(torch.int32), tile_indices.to(torch.int32)])


def unpack_mapping(mapping: torch.Tensor, tiles: int):
    """Split a packed mapping without reading it back: the offsets block has
    ``tiles + 1`` entries for this canvas and tile size."""
    if mapping.dtype != torch.int32 or mapping.dim() != 1:
        raise TypeError("tile_primitive_mapping must be a 1-D int32 tensor")
    if mapping.numel() < tiles + 3:
        raise ValueError(f"tile_primitive_mapping has {mapping.numel()} entries, "
                         f"expected at least {tiles + 3} for {tiles} tiles")
    return mapping[2:3 + tiles], mapping[3 + tiles:]


def bin_primitives(handle: Handle, means2D, radii, rotations, bmp_sel, templates,
                   extents: torch.Tensor | None = None):
    """GPU binning: ``(tile_offsets, tile_indices)`` for the given primitives,
    identical to the reference's ``_unified_primitive_assignment`` +
    ``_convert_masks_to_mapping`` (``vector_renderer.py:401-444``,
    ``simple_tile_renderer.py:721-766``)."""
    dt = means2D.dtype
    zeros = torch.zeros_like(radii)
    prims = Handle.prims(means2D.contiguous(), radii.contiguous(), rotations.contiguous(),
                         zeros, torch.zeros((radii.shape[0], 3), dtype=dt, device=radii.device),
                         bmp_sel.to(torch.int32).contiguous())
    if extents is None:
        extents = handle.template_extents(templates)
    return handle.bin(prims, extents, tuple(templates.shape))


def _as(t: torch.Tensor, dtype) -> torch.Tensor:
    """``t`` in ``dtype``, contiguous; ``t`` itself when it already is (the
    common case, without a trip through torch's dispatcher)."""
    if t.dtype == dtype and t.is_contiguous():
        return t
    return t.to(dtype=dtype).contiguous()


class _Run:
    """Everything one forward needs, normalised to contiguous device tensors."""

    def __init__(self, handle: Handle, store_dtype, means2D, radii, rotations, opacities, colors,
                 colors_orig, templates, bmp_sel, c_blend, mapping, eps_T,
                 colors_orig_per_template=False):
        for t, nm in ((means2D, "means2D"), (templates, "primitive_templates")):
            _native.require_gpu(t, nm)
        # mixed storage: fp32 positions, radii and rotations, half-precision
        # opacities, colours and bitmaps
        gdt = torch.float32 if getattr(handle, "fp32_geometry", False) else store_dtype
        self.means2D, self.radii, self.rotations = _as(means2D, gdt), _as(radii, gdt), _as(rotations, gdt)
        self.opacities, self.colors = _as(opacities, store_dtype), _as(colors, store_dtype)
        self.templates, self.ext, self.tmpl_version = handle.prepared_templates(templates, store_dtype)
        self.sel = _as(bmp_sel, torch.int32)
        if isinstance(c_blend, torch.Tensor):
            c_blend = float(c_blend.item())
        self.c_blend = float(c_blend)
        self.colors_orig, self.co_mode = None, 0
        if colors_orig is not None and self.c_blend > 0.0 and colors_orig.numel() > 0:
            self.co_mode = 2 if colors_orig_per_template else 1
            need = self.templates.shape[0] if self.co_mode == 2 else self.radii.shape[0]
            if colors_orig.dim() != 4 or colors_orig.shape[0] < need or \
                    tuple(colors_orig.shape[1:]) != (*self.templates.shape[1:], 3):
                raise ValueError(f"colors_orig must be ({need},Ht,Wt,3), got {tuple(colors_orig.shape)}")
            self.colors_orig = _as(colors_orig, store_dtype)
        self.eps_T = float(eps_T)
        self.blend_grad_exact = bool(getattr(handle, "blend_grad_exact", False))
        self.handle = handle
        self.gen = None   # handle.records_gen while the handle holds this call's records
        self._prims = self._bitmaps = None
        if mapping is not None and mapping.numel() > 0:
            self.lists = _native.Lists(*unpack_mapping(mapping, handle.tiles))
        else:
            self.lists = handle.bin_lazy(self.prims(), self.ext, tuple(self.templates.shape),
                                         handle.bin_factor)
            self.gen = handle.records_gen

    def prims(self):
        # validated and packed once per call; the records the binning (or this
        # call's forward) prepared are reused while no other call on the handle
        # has replaced them
        if self._prims is None:
            self._prims = Handle.prims(self.means2D, self.radii, self.rotations, self.opacities,
                                       self.colors, self.sel)
        self._prims.reuse_records = int(self.gen is not None and self.gen == self.handle.records_gen)
        return self._prims

    def bitmaps(self):
        if self._bitmaps is None:
            self._bitmaps = Handle.bitmaps(self.templates, self.colors_orig, self.co_mode,
                                           self.c_blend, self.tmpl_version, self.blend_grad_exact)
        return self._bitmaps

    def forward(self):
        dev, H, W = self.means2D.device, self.handle.H, self.handle.W
        out_color = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
        out_alpha = torch.empty((H, W), dtype=torch.float32, device=dev)
        # per-pixel forward state for the backward
        out_T = torch.empty((H, W), dtype=torch.float32, device=dev)
        out_last = torch.empty((H, W), dtype=torch.int32, device=dev)
        self.handle.forward(self.prims(), self.bitmaps(), self.lists, self.eps_T,
                            out_color, out_alpha, out_T, out_last)
        self.gen = self.handle.records_gen
        return out_color, out_alpha, out_T, out_last

    def backward(self, grad_color, grad_alpha, out_color, out_T, out_last):
        dev, N = self.means2D.device, self.radii.shape[0]
        f32 = torch.float32
        gC = _as(grad_color, f32)
        gA = None if grad_alpha is None else _as(grad_alpha, f32)
        g_means = torch.empty((N, 2), dtype=f32, device=dev)
        g_r = torch.empty(N, dtype=f32, device=dev)
        g_th = torch.empty(N, dtype=f32, device=dev)
        g_v = torch.empty(N, dtype=f32, device=dev)
        g_c = torch.empty((N, 3), dtype=f32, device=dev)
        self.handle.backward(self.prims(), self.bitmaps(), self.lists, self.eps_T,
                             gC, gA, out_color, out_T, out_last, g_means, g_r, g_th, g_v, g_c)
        return g_means, g_r, g_th, g_v, g_c


def _apply_forward(ctx, kind, store_dtype, out_dtype, means2D, radii, rotations, opacities, colors,
                   colors_orig, primitive_templates, global_bmp_sel, c_blend, lr_conf,
                   tile_primitive_mapping, image_height, image_width, tile_size, sigma):
    handle = _active(kind)
    if kind == "fp16":   # reduced-precision storage: fp16 unless the rasterizer chose bf16
        store_dtype = out_dtype = getattr(handle, "store_dtype", store_dtype)
    if (int(image_height), int(image_width), int(tile_size)) != (handle.H, handle.W, handle.ts):
        raise ValueError(f"rasterizer was initialised for {handle.H}x{handle.W} tile {handle.ts}, "
                         f"called with {image_height}x{image_width} tile {tile_size}")
    run = _Run(handle, store_dtype, means2D, radii, rotations, opacities, colors, colors_orig,
               primitive_templates, global_bmp_sel, c_blend, tile_primitive_mapping,
               getattr(handle, "eps_T", DEFAULT_EPS_T))
    out_color, out_alpha, out_T, out_last = run.forward()
    ctx.run = run
    ctx.save_for_backward(out_color, out_T, out_last)
    ctx.mark_non_differentiable(out_T, out_last)
    if out_dtype != torch.float32:
        return out_color.to(out_dtype), out_alpha.to(out_dtype)
    return out_color, out_alpha


def _apply_backward(ctx, grad_out_color, grad_out_alpha):
    out_color, out_T, out_last = ctx.saved_tensors
    run = ctx.run
    if grad_out_color is None:
        grad_out_color = torch.zeros_like(out_color)
    grads = run.backward(grad_out_color, grad_out_alpha, out_color, out_T, out_last)
    return tuple(grads) + (None,) * 11


class TileRasterizerFunction(Function):
    """FP32 forward/backward (``cuda_tile_rasterizer/__init__.py:254-340``)."""

    @staticmethod
    def forward(ctx, means2D, radii, rotations, opacities, colors, colors_orig, primitive_templates,
                global_bmp_sel, c_blend, lr_conf, tile_primitive_mapping, image_height, image_width,
                tile_size, sigma, use_class=False):
        return _apply_forward(ctx, "fp32", torch.float32, torch.float32, means2D, radii, rotations,
                              opacities, colors, colors_orig, primitive_templates, global_bmp_sel,
                              c_blend, lr_conf, tile_primitive_mapping, image_height, image_width,
                              tile_size, sigma)

    @staticmethod
    def backward(ctx, grad_out_color, grad_out_alpha):
        return _apply_backward(ctx, grad_out_color, grad_out_alpha)


class TileRasterizerFunctionFP16(Function):
    """FP16-storage forward/backward (``cuda_tile_rasterizer/__init__.py:342-436``):
    parameters and bitmaps are stored as fp16, arithmetic and accumulation are
    fp32, outputs are returned as fp16."""

    @staticmethod
    def forward(ctx, means2D, radii, rotations, opacities, colors, colors_orig, primitive_templates,
                global_bmp_sel, c_blend, lr_conf, tile_primitive_mapping, image_height, image_width,
                tile_size, sigma, use_class=False):
        return _apply_forward(ctx, "fp16", torch.float16, torch.float16, means2D, radii, rotations,
                              opacities, colors, colors_orig, primitive_templates, global_bmp_sel,
                              c_blend, lr_conf, tile_primitive_mapping, image_height, image_width,
                              tile_size, sigma)

    @staticmethod
    def backward(ctx, grad_out_color, grad_out_alpha):
        return _apply_backward(ctx, grad_out_color, grad_out_alpha)


class TileRasterizerBatchFunction(Function):
    """B candidates with shared templates (``cuda_tile_rasterizer/__init__.py:124-252``):
    ``means2D`` (B,N,2), ``radii`` / ``rotations`` / ``opacities`` (B,N),
    ``colors`` (B,N,3), ``colors_orig`` (B,N,Ht,Wt,3) or one (N,Ht,Wt,3) for
    all.  Each candidate is binned on its own; then one raster launch renders
    every candidate and one launch back-propagates all of them
    (``dbmp_forward_batch`` / ``dbmp_backward_batch``).  Each candidate keeps
    its own tile lists and forward state, so every candidate's gradient is
    its own."""

    @staticmethod
    def forward(ctx, means2D, radii, rotations, opacities, colors, colors_orig, primitive_templates,
                global_bmp_sel, c_blend, lr_conf, tile_primitive_mappings, image_height, image_width,
                tile_size, sigma, use_fp16=False):
        kind = "fp16" if use_fp16 else "fp32"
        handle = _active(kind)
        if (int(image_height), int(image_width), int(tile_size)) != (handle.H, handle.W, handle.ts):
            raise ValueError(f"rasterizer was initialised for {handle.H}x{handle.W} tile {handle.ts}, "
                             f"called with {image_height}x{image_width} tile {tile_size}")
        store = getattr(handle, "store_dtype", torch.float16) if use_fp16 else torch.float32
        B = means2D.shape[0]
        if means2D.dim() != 3 or any(t.shape[0] != B for t in (radii, rotations, opacities, colors)):
            raise ValueError("batch parameters must be means2D (B,N,2), radii/rotations/opacities "
                             "(B,N), colors (B,N,3)")
        if isinstance(c_blend, torch.Tensor):
            c_blend = float(c_blend.item())
        c_blend = float(c_blend)
        maps = list(tile_primitive_mappings) if tile_primitive_mappings else []
        eps_T = getattr(handle, "eps_T", DEFAULT_EPS_T)
        runs = [_Run(handle, store, means2D[b], radii[b], rotations[b], opacities[b], colors[b], None,
                     primitive_templates, global_bmp_sel, 0.0, maps[b] if b < len(maps) else None, eps_T)
                for b in range(B)]
        co_all = None
        tm = runs[0].templates
        if colors_orig is not None and colors_orig.numel() > 0 and c_blend > 0.0:
            N = radii.shape[1]
            want = (N, *tm.shape[1:], 3)
            if colors_orig.dim() == 4 and tuple(colors_orig.shape) == want:
                colors_orig = colors_orig.unsqueeze(0).expand(B, *want)
            if colors_orig.dim() != 5 or tuple(colors_orig.shape) != (B, *want):
                raise ValueError(f"colors_orig must be (B,N,Ht,Wt,3) = {(B, *want)}, "
                                 f"got {tuple(colors_orig.shape)}")
            co_all = colors_orig.to(dtype=store).contiguous().view(B * N, *want[1:])
        bm = Handle.bitmaps(tm, co_all, 1 if co_all is not None else 0, c_blend, runs[0].tmpl_version,
                            bool(getattr(handle, "blend_grad_exact", False)))
        dev, H, W = runs[0].means2D.device, handle.H, handle.W
        out_c = torch.empty((B, H, W, 3), dtype=torch.float32, device=dev)
        out_a = torch.empty((B, H, W), dtype=torch.float32, device=dev)
        out_T = torch.empty((B, H, W), dtype=torch.float32, device=dev)
        out_l = torch.empty((B, H, W), dtype=torch.int32, device=dev)
        handle.forward_batch([r.prims() for r in runs], bm, [r.lists for r in runs], eps_T,
                             out_c, out_a, out_T, out_l)
        ctx.runs, ctx.bm, ctx.co_all, ctx.eps_T = runs, bm, co_all, eps_T
        ctx.save_for_backward(out_c, out_T, out_l)
        ctx.mark_non_differentiable(out_T, out_l)
        if use_fp16:
            return out_c.to(store), out_a.to(store)
        return out_c, out_a

    @staticmethod
    def backward(ctx, grad_out_color, grad_out_alpha):
        out_c, out_T, out_l = ctx.saved_tensors
        runs = ctx.runs
        B, N = len(runs), runs[0].radii.shape[0]
        dev, f32 = out_c.device, torch.float32
        gC = (torch.zeros_like(out_c) if grad_out_color is None
              else grad_out_color.to(f32).contiguous())
        gA = None if grad_out_alpha is None else grad_out_alpha.to(f32).contiguous()
        g = [torch.empty((B, N, 2), dtype=f32, device=dev), torch.empty((B, N), dtype=f32, device=dev),
             torch.empty((B, N), dtype=f32, device=dev), torch.empty((B, N), dtype=f32, device=dev),
             torch.empty((B, N, 3), dtype=f32, device=dev)]
        runs[0].handle.backward_batch([r.prims() for r in runs], ctx.bm, [r.lists for r in runs],
                                      ctx.eps_T, gC, gA, out_c, out_T, out_l, *g)
        return tuple(g) + (None,) * 11


class CompositeMSELoss(Function):
    """``mean((color + (1 - alpha) * bg - target)^2)`` with its gradient, one
    HIP kernel (``dbmp_loss_mse``) in place of the torch composite and
    ``F.mse_loss`` the reference's renderer builds around the rasterizer
    (``simple_tile_renderer.py:703-707``, ``loss_functions.py:20-40``) and their
    autograd backward.  The gradient with respect to color and alpha is formed in
    the forward pass and scaled in place to the incoming gradient in the backward
    (``dbmp_rescale_grads``: nothing to do for the usual incoming 1; the factor
    already applied is tracked, so a second backward over a retained graph is
    scaled correctly); ``target`` and ``bg`` are constants."""

    @staticmethod
    def forward(ctx, color, alpha, target, bg, handle, bg_value):
        H, W = handle.H, handle.W
        for t, nm, sh in ((color, "color", (H, W, 3)), (alpha, "alpha", (H, W)),
                          (target, "target", (H, W, 3)), (bg, "bg", (H, W, 3))):
            if t is not None and tuple(t.shape) != sh:
                raise ValueError(f"{nm} must be {sh}, got {tuple(t.shape)}")
        for t, nm in ((target, "target"), (bg, "bg")):
            if t is not None and t.requires_grad:
                raise ValueError(f"composite_mse_loss: {nm} is not differentiable here")
        f32 = torch.float32
        c, a, t = _as(color, f32), _as(alpha, f32), _as(target, f32)
        b = None if bg is None else _as(bg, f32)
        gC, gA = torch.empty_like(c), torch.empty_like(a)
        # [0]: the loss; [1]: the factor the gradient images carry (0 stands for
        # 1); [2]: the rescale's workgroup counter
        sc = torch.zeros(3, dtype=f32, device=c.device)
        loss = sc[0:1]
        handle.loss_mse(c, a, t, gC, gA, loss, bg=b, bg_value=bg_value)
        ctx.save_for_backward(gC, gA)
        ctx.handle = handle
        ctx.applied = sc[1:3]
        return loss.view(())

    @staticmethod
    def backward(ctx, grad_loss):
        gC, gA = ctx.saved_tensors
        # the incoming gradient of the 0-d loss: one fp32 value on the device
        ctx.handle.rescale_grads(gC, gA, _as(grad_loss, torch.float32), ctx.applied)
        return gC, gA, None, None, None, None


def composite_mse_loss(handle: Handle, color, alpha, target, bg=None, bg_value: float = 1.0):
    """Fused background composite + MSE loss (extension; see ``CompositeMSELoss``).
    ``bg`` is an (H,W,3) image or None for the constant ``bg_value``."""
    return CompositeMSELoss.apply(color, alpha, target, bg, handle, float(bg_value))


class TileRasterizer:
    """Class-based rasterizer (``cuda_tile_rasterizer/__init__.py:60-121``)."""

    def __init__(self, image_height, image_width, tile_size=16, sigma=0.0, alpha_upper_bound=1.0,
                 max_prims_per_pixel=151, num_primitives=None, use_fp16=False, storage_dtype=None,
                 fp32_geometry=False):
        """``storage_dtype`` (extension): torch.float16 or torch.bfloat16 selects the
        reduced-precision storage of the ``use_fp16`` path; torch.float32 with
        ``use_fp16=False`` is the plain path.  ``fp32_geometry`` (extension): in
        the reduced-precision path keep positions, radii and rotations in fp32
        (the reference's fp16 path casts means2D to half,
        ``cuda_tile_rasterizer/__init__.py:348``, which moves a 1024-px
        coordinate by up to 0.25 px)."""
        if storage_dtype is not None and storage_dtype != torch.float32:
            if storage_dtype not in (torch.float16, torch.bfloat16):
                raise ValueError(f"unsupported storage_dtype {storage_dtype}")
            use_fp16 = True
        self.image_height = image_height
        self.image_width = image_width
        self.tile_size = tile_size
        self.sigma = sigma
        self.alpha_upper_bound = alpha_upper_bound
        self.max_prims_per_pixel = max_prims_per_pixel
        self.num_primitives = num_primitives
        self.use_fp16 = use_fp16
        kind = "fp16" if use_fp16 else "fp32"
        self.handle = Handle(image_height, image_width, tile_size, sigma, alpha_upper_bound,
                             max_prims_per_pixel, num_primitives)
        self.handle.eps_T = DEFAULT_EPS_T
        self.handle.blend_grad_exact = False
        self.handle.fp32_geometry = bool(fp32_geometry) and use_fp16
        self.handle.store_dtype = (storage_dtype if storage_dtype is not None and use_fp16
                                   else (torch.float16 if use_fp16 else torch.float32))
        with _ACTIVE_LOCK:
            _ACTIVE[kind] = self.handle
        self.apply_fn = TileRasterizerFunctionFP16.apply if use_fp16 else TileRasterizerFunction.apply

    def __call__(self, means2D, radii, rotations, opacities, colors, colors_orig, primitive_templates,
                 global_bmp_sel, c_blend, lr_conf, tile_primitive_mapping=None):
        with _ACTIVE_LOCK:
            _ACTIVE["fp16" if self.use_fp16 else "fp32"] = self.handle
        return self.apply_fn(means2D, radii, rotations, opacities, colors, colors_orig,
                             primitive_templates, global_bmp_sel, c_blend, lr_conf,
                             tile_primitive_mapping, self.image_height, self.image_width,
                             self.tile_size, self.sigma, True)

    def forward_batch(self, means2D, radii, rotations, opacities, colors, colors_orig,
                      primitive_templates, global_bmp_sel, c_blend, lr_conf,
                      tile_primitive_mappings=None):
        with _ACTIVE_LOCK:
            _ACTIVE["fp16" if self.use_fp16 else "fp32"] = self.handle
        return TileRasterizerBatchFunction.apply(
            means2D, radii, rotations, opacities, colors, colors_orig, primitive_templates,
            global_bmp_sel, c_blend, lr_conf, tile_primitive_mappings, self.image_height,
            self.image_width, self.tile_size, self.sigma, self.use_fp16)

    @property
    def eps_T(self) -> float:
        """Early-termination threshold on transmittance (0 = walk every entry)."""
        return self.handle.eps_T

    @eps_T.setter
    def eps_T(self, v: float) -> None:
        self.handle.eps_T = float(v)

    @property
    def exact_color_map_grad(self) -> bool:
        """Extension, off by default.  Off: under ``c_blend > 0`` the colour
        gradient is the reference's (``tile_backward.cu:253-258``), which leaves
        out the ``(1 - c_blend)`` factor on the colour logits and the colour
        map's dependence on position, radius and rotation.  On: the exact
        gradient of the blended colour."""
        return self.handle.blend_grad_exact

    @exact_color_map_grad.setter
    def exact_color_map_grad(self, v: bool) -> None:
        self.handle.blend_grad_exact = bool(v)

    def enable_timing(self, on: bool = True) -> None:
        self.handle.timing_enable(on)

    def composite_mse_loss(self, color, alpha, target, bg=None, bg_value: float = 1.0):
        """Extension: ``F.mse_loss(color + (1 - alpha)[..., None] * bg, target)``
        and its gradient in one kernel on this rasterizer's canvas."""
        return composite_mse_loss(self.handle, color, alpha, target, bg, bg_value)


def _print_stats(kind: str) -> None:
    h = _ACTIVE.get(kind)
    if h is None:
        print("Rasterizer not initialized: no timing statistics")
        return
    f, fn, b, bn = h.timing_stats()
    print("=== tile rasterizer timing (%s, hipEvent) ===" % kind)
    print("forward : %d calls, total %.3f ms, avg %.4f ms" % (fn, f, f / fn if fn else 0.0))
    print("backward: %d calls, total %.3f ms, avg %.4f ms" % (bn, b, b / bn if bn else 0.0))


def print_cuda_timing_stats():
    """Average forward/backward time (``tile_rasterize.cu:783-806``); timing is
    recorded only after ``TileRasterizer.enable_timing()``."""
    _print_stats("fp32")


def reset_cuda_timing_stats():
    h = _ACTIVE.get("fp32")
    if h is not None:
        h.timing_reset()


def print_cuda_timing_stats_fp16():
    _print_stats("fp16")


def reset_cuda_timing_stats_fp16():
    h = _ACTIVE.get("fp16")
    if h is not None:
        h.timing_reset()
