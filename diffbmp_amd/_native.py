"""ctypes binding of ``libdiffbmp_amd.so`` (C ABI in ``include/diffbmp_amd.h``).

PyTorch is used for device memory and the current HIP stream only.  There is
no fallback: if the library is missing or the device is not a GPU, every call
raises.
"""
from __future__ import annotations

import ctypes
import math
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
# DBMP_LIBRARY selects another build of the same sources (e.g. the lane
# statistics build, libdiffbmp_amd_lanestats.so)
LIB_PATH = os.environ.get("DBMP_LIBRARY") or os.path.join(_HERE, "libdiffbmp_amd.so")

DBMP_F32, DBMP_F16, DBMP_BF16, DBMP_F16_APP, DBMP_BF16_APP = 0, 1, 2, 3, 4
_DT = {torch.float32: DBMP_F32, torch.float16: DBMP_F16, torch.bfloat16: DBMP_BF16}
# fp32 geometry with half-precision opacities and colours
_DT_APP = {torch.float16: DBMP_F16_APP, torch.bfloat16: DBMP_BF16_APP}

_lib = None
_lib_lock = threading.Lock()


class Prims(ctypes.Structure):
    _fields_ = [("means2D", ctypes.c_void_p), ("radii", ctypes.c_void_p),
                ("rotations", ctypes.c_void_p), ("opacities", ctypes.c_void_p),
                ("colors", ctypes.c_void_p), ("bmp_sel", ctypes.c_void_p),
                ("num_primitives", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("reuse_records", ctypes.c_int32)]


class AdamGroup(ctypes.Structure):
    _fields_ = [("param", ctypes.c_void_p), ("exp_avg", ctypes.c_void_p),
                ("exp_avg_sq", ctypes.c_void_p), ("grad", ctypes.c_void_p),
                ("store", ctypes.c_void_p), ("numel", ctypes.c_int64), ("lr", ctypes.c_float),
                ("stride", ctypes.c_int32), ("clamp", ctypes.c_int32),
                ("lo", ctypes.c_float * 2), ("hi", ctypes.c_float * 2),
                ("grad_row", ctypes.c_int32), ("grad_comps", ctypes.c_int32),
                ("zero_grad", ctypes.c_int32)]


class Adam(ctypes.Structure):
    _fields_ = [("groups", AdamGroup * 5), ("num_groups", ctypes.c_int32),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("eps", ctypes.c_float),
                ("step", ctypes.c_int64), ("store_dtype", ctypes.c_int32)]


class Bitmaps(ctypes.Structure):
    _fields_ = [("templates", ctypes.c_void_p), ("colors_orig", ctypes.c_void_p),
                ("num_templates", ctypes.c_int32), ("height", ctypes.c_int32),
                ("width", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("colors_orig_mode", ctypes.c_int32), ("c_blend", ctypes.c_float),
                ("version", ctypes.c_int64), ("blend_grad_exact", ctypes.c_int32),
                ("colors_version", ctypes.c_int64)]


class TileLists(ctypes.Structure):
    _fields_ = [("tile_offsets", ctypes.c_void_p), ("tile_indices", ctypes.c_void_p),
                ("pair_count", ctypes.c_void_p), ("capacity", ctypes.c_int64),
                ("bin_factor", ctypes.c_int32), ("ranges", ctypes.c_void_p),
                ("push_masks", ctypes.c_void_p), ("push_mask_words", ctypes.c_int64)]


class Job(ctypes.Structure):
    """``dbmp_job``: one fit job of a batched step."""
    _fields_ = [("height", ctypes.c_int32), ("width", ctypes.c_int32),
                ("num_primitives", ctypes.c_int32), ("num_templates", ctypes.c_int32),
                ("tmpl_height", ctypes.c_int32), ("tmpl_width", ctypes.c_int32),
                ("quads", ctypes.c_void_p), ("ext", ctypes.c_void_p),
                ("quads_with_color", ctypes.c_int32), ("c_blend", ctypes.c_float),
                ("blend_grad_exact", ctypes.c_int32), ("bmp_sel", ctypes.c_void_p),
                ("params", ctypes.c_void_p), ("exp_avg", ctypes.c_void_p),
                ("exp_avg_sq", ctypes.c_void_p), ("store", ctypes.c_void_p),
                ("target", ctypes.c_void_p), ("background", ctypes.c_void_p),
                ("background_value", ctypes.c_float), ("out_color", ctypes.c_void_p),
                ("out_alpha", ctypes.c_void_p), ("out_T", ctypes.c_void_p),
                ("out_last", ctypes.c_void_p), ("loss_history", ctypes.c_void_p),
                ("loss_capacity", ctypes.c_int32), ("lr", ctypes.c_float * 5),
                ("alpha_upper_bound", ctypes.c_float), ("eps_T", ctypes.c_float),
                ("clamp", ctypes.c_int32), ("steps_done", ctypes.c_int32)]


class LayerPrims(ctypes.Structure):
    """``dbmp_layer_prims``: the layer export's primitive arrays (fp32)."""
    _fields_ = [("means2D", ctypes.c_void_p), ("radii", ctypes.c_void_p),
                ("rotations", ctypes.c_void_p), ("colors", ctypes.c_void_p),
                ("visibility", ctypes.c_void_p), ("bmp_sel", ctypes.c_void_p),
                ("num_primitives", ctypes.c_int32)]


class Lists:
    """Tile lists on the device: CSR offsets, index buffer, and -- for lists
    produced by dbmp_bin -- the device pair count and the capacity used.
    ``factor`` > 1 marks super-tile lists from dbmp_bin_tiles (one list per
    factor x factor tiles) with the binning's per-primitive tile ``ranges``.
    ``masks`` (int64 words, optional) is forward->backward state: the forward
    over these lists records which entries reached each wave's pixels and the
    backward of that forward walks only those."""

    __slots__ = ("offsets", "indices", "count", "capacity", "factor", "ranges", "masks", "_c")

    def __init__(self, offsets, indices, count=None, capacity=0, factor=1, ranges=None, masks=None):
        self.offsets, self.indices, self.count, self.capacity = offsets, indices, count, int(capacity)
        self.factor, self.ranges, self.masks = int(factor), ranges, masks
        self._c = None

    def c(self) -> TileLists:
        """The ``dbmp_tile_lists`` view of these buffers, built on first use
        (the fields are fixed once the lists exist)."""
        if self._c is None:
            idx = self.indices if self.indices.numel() else self.offsets
            m = self.masks
            self._c = TileLists(self.offsets.data_ptr(), idx.data_ptr(),
                                self.count.data_ptr() if self.count is not None else None,
                                self.capacity, self.factor,
                                self.ranges.data_ptr() if self.ranges is not None else None,
                                m.data_ptr() if m is not None else None,
                                m.numel() if m is not None else 0)
        return self._c


class _PackedLists(Lists):
    """Lists that ``Handle.bin_lazy`` lays out in one int32 buffer: pair count
    (int64) | pad | ranges (N,4) | offsets | pad | push masks (int64 words) |
    indices.  The C struct is built from the buffer's address; the tensor
    views of the fields are made only when one is asked for."""

    __slots__ = ("_buf", "_lay", "_views")

    def __init__(self, buf, capacity, factor, lay):
        # lay: (offsets start, offsets length, masks start, indices start) in int32 words
        self._buf, self._lay, self._views = buf, lay, None
        self.capacity, self.factor, self._c = int(capacity), int(factor), None

    def _view(self):
        if self._views is None:
            b = self._buf
            o_offs, n_offs, o_pm, o_idx = self._lay
            self._views = (b[o_offs:o_offs + n_offs], b[o_idx:], b[0:2].view(torch.int64),
                           b[4:o_offs].view(-1, 4), b[o_pm:o_idx].view(torch.int64))
        return self._views

    offsets = property(lambda self: self._view()[0])
    indices = property(lambda self: self._view()[1])
    count = property(lambda self: self._view()[2])
    ranges = property(lambda self: self._view()[3])
    masks = property(lambda self: self._view()[4])

    def c(self) -> TileLists:
        if self._c is None:
            a = self._buf.data_ptr()
            o_offs, _, o_pm, o_idx = self._lay
            self._c = TileLists(a + 4 * o_offs, a + 4 * o_idx, a, self.capacity, self.factor, a + 16,
                                a + 4 * o_pm, (o_idx - o_pm) // 2)
        return self._c


EXPORTS = {
    "dbmp_version": (ctypes.c_int, []),
    "dbmp_last_error": (ctypes.c_char_p, []),
    "dbmp_init": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                 ctypes.c_int]),
    "dbmp_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "dbmp_push_mask_words": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32]),
    "dbmp_template_extents": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Bitmaps),
                                             ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_bin": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Prims), ctypes.c_void_p,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                ctypes.c_void_p]),
    "dbmp_bin_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Prims), ctypes.c_void_p,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int32,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Prims),
                                    ctypes.POINTER(Bitmaps), ctypes.POINTER(TileLists),
                                    ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_forward_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                            ctypes.POINTER(Prims), ctypes.POINTER(Bitmaps),
                                            ctypes.POINTER(TileLists), ctypes.c_float,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_forward_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Prims),
                                          ctypes.POINTER(Bitmaps), ctypes.POINTER(TileLists),
                                          ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_backward_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Prims),
                                           ctypes.POINTER(Bitmaps), ctypes.POINTER(TileLists),
                                           ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Prims),
                                     ctypes.POINTER(Bitmaps), ctypes.POINTER(TileLists),
                                     ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_loss_mse": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]),
    "dbmp_backward_mse": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Prims),
                                         ctypes.POINTER(Bitmaps), ctypes.POINTER(TileLists),
                                         ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_backward_mse_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Prims),
                                              ctypes.POINTER(Bitmaps), ctypes.POINTER(TileLists),
                                              ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]),
    "dbmp_adam_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Adam), ctypes.c_void_p]),
    "dbmp_rescale_grads": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_lane_stats": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int32]),
    "dbmp_sincos": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                   ctypes.c_void_p]),
    "dbmp_timing_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "dbmp_timing_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_int64)]),
    "dbmp_timing_reset": (ctypes.c_int, [ctypes.c_void_p]),
    "dbmp_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32,
                                         ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "dbmp_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "dbmp_batch_bitmap_bytes": (ctypes.c_int64, [ctypes.c_void_p, ctypes.POINTER(Bitmaps)]),
    "dbmp_batch_prepare_bitmaps": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Bitmaps),
                                                  ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_void_p]),
    "dbmp_batch_set_jobs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Job), ctypes.c_int32,
                                           ctypes.c_int64, ctypes.c_void_p]),
    "dbmp_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "dbmp_batch_pair_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "dbmp_batch_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "dbmp_psd_bounds": (ctypes.c_int, [ctypes.POINTER(LayerPrims), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "dbmp_psd_layers": (ctypes.c_int, [ctypes.POINTER(LayerPrims), ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]),
}


class NativeError(RuntimeError):
    """A libdiffbmp_amd call returned an error code."""


def lib() -> ctypes.CDLL:
    """Load the HIP library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built; run diffbmp_amd.build.build() "
                    "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in EXPORTS.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().dbmp_last_error().decode(errors="replace")
        raise NativeError(f"diffbmp_amd error {rc}: {msg}")


def _ptr(t: torch.Tensor | None):
    """Device address of ``t`` (None for None) for a ``void*`` argument: every
    pointer argument in ``EXPORTS`` is declared ``c_void_p``, which takes the
    integer as is."""
    return None if t is None else t.data_ptr()


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream(dev: torch.device):
    """The current torch stream of ``dev`` as a hipStream_t (an integer for a
    ``c_void_p`` argument)."""
    if _raw_stream is not None:
        idx = dev.index if dev.index is not None else torch.cuda.current_device()
        return _raw_stream(idx)
    return torch.cuda.current_stream(dev).cuda_stream


def require_gpu(t: torch.Tensor, name: str) -> None:
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a GPU tensor (got device {t.device}); "
                           "the rasterizer has no CPU path")


class Handle:
    """One native rasterizer (``dbmp_init``): canvas, tile size, alpha bound,
    and device scratch.  Calls are ordered on the current torch stream."""

    READBACK_EVERY = 8   # bin_lazy: calls per asynchronous pair-count read-back

    def __init__(self, image_height: int, image_width: int, tile_size: int = 16,
                 sigma: float = 0.0, alpha_upper_bound: float = 1.0,
                 max_prims_per_pixel: int = 151, num_primitives: int | None = None,
                 device: torch.device | None = None):
        L = lib()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.H, self.W, self.ts = int(image_height), int(image_width), int(tile_size)
        self.alpha_ub = float(alpha_upper_bound)
        h = ctypes.c_void_p()
        # dbmp_init validates the sizes (and raises) before anything is derived from them
        with torch.cuda.device(self.device):
            check(L.dbmp_init(ctypes.byref(h), self.H, self.W, self.ts, float(sigma), self.alpha_ub,
                              int(max_prims_per_pixel), int(num_primitives or 0)))
        self.tilesX = -(-self.W // self.ts)
        self.tilesY = -(-self.H // self.ts)
        self.tiles = self.tilesX * self.tilesY
        self._h = h
        self._L = L
        self.pair_capacity = 0
        # internal binning granularity: 64-px super-tiles for 16-px tiles
        # (same per-tile work, a sixteenth of the pairs to sort)
        self.bin_factor = 4 if self.ts == 16 else 1
        self._pending = None          # event of the last asynchronous pair-count read-back
        self._readback = None         # pinned int64 the read-back lands in (reused)
        self._readback_ev = None
        self._lazy_calls = 0
        # bumped by every call that recomputes the primitive records (a binning,
        # or a raster call without reuse_records): a caller that noted it after
        # preparing its parameters may pass reuse_records while it is unchanged
        self.records_gen = 0
        self._tmpl_cache = None       # (key, converted templates, extents, version)
        self._pm_words = {}           # (capacity, bin factor) -> push-mask words
        self._tmpl_gen = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._L.dbmp_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------------
    @staticmethod
    def prims(means2D, radii, rotations, opacities, colors, bmp_sel, reuse: bool = False) -> Prims:
        dt = _DT[means2D.dtype]
        for t, nm in ((radii, "radii"), (rotations, "rotations")):
            if t.dtype != means2D.dtype:
                raise TypeError(f"{nm} dtype {t.dtype} != means2D dtype {means2D.dtype}")
        if colors.dtype != opacities.dtype:
            raise TypeError(f"colors dtype {colors.dtype} != opacities dtype {opacities.dtype}")
        if opacities.dtype != means2D.dtype:
            # fp32 positions, radii and rotations with fp16 / bf16 opacities and colours
            if means2D.dtype != torch.float32 or opacities.dtype not in _DT_APP:
                raise TypeError(f"opacities dtype {opacities.dtype} with means2D dtype "
                                f"{means2D.dtype}: mixed storage is fp32 geometry with fp16 or bf16 "
                                "appearance")
            dt = _DT_APP[opacities.dtype]
        N = radii.shape[0]
        if means2D.shape != (N, 2) or colors.shape != (N, 3) or rotations.shape != (N,) \
                or opacities.shape != (N,) or bmp_sel.shape != (N,):
            raise ValueError("primitive tensors must be means2D (N,2), radii/rotations/opacities/"
                             "bmp_sel (N,), colors (N,3)")
        if bmp_sel.dtype != torch.int32:
            raise TypeError("bmp_sel must be int32")
        return Prims(means2D.data_ptr(), radii.data_ptr(), rotations.data_ptr(),
                     opacities.data_ptr(), colors.data_ptr(), bmp_sel.data_ptr(), N, dt, int(reuse))

    @staticmethod
    def bitmaps(templates, colors_orig=None, colors_orig_mode=0, c_blend=0.0,
                version: int = -1, blend_grad_exact: bool = False,
                colors_version: int = -1) -> Bitmaps:
        if templates.dim() != 3:
            raise ValueError("templates must be (T,Ht,Wt)")
        T, Ht, Wt = templates.shape
        co = 0
        if colors_orig is not None and colors_orig_mode:
            if colors_orig.dtype != templates.dtype:
                raise TypeError("colors_orig must have the templates' dtype")
            if colors_orig.shape[1:] != (Ht, Wt, 3):
                raise ValueError("colors_orig must be (M,Ht,Wt,3)")
            co = colors_orig.data_ptr()
        return Bitmaps(templates.data_ptr(), co, T, Ht, Wt, _DT[templates.dtype],
                       int(colors_orig_mode if co else 0), float(c_blend), int(version),
                       int(bool(blend_grad_exact)), int(colors_version))

    def template_extents(self, templates: torch.Tensor) -> torch.Tensor:
        require_gpu(templates, "templates")
        templates = templates.contiguous()
        ext = torch.empty((templates.shape[0], 4), dtype=torch.float32, device=templates.device)
        bm = self.bitmaps(templates)
        check(self._L.dbmp_template_extents(self._h, ctypes.byref(bm), _ptr(ext),
                                            _stream(templates.device)))
        return ext

    def list_count(self, factor: int = 1) -> int:
        """Number of tile lists binning with ``factor`` produces."""
        return (-(-self.tilesX // factor)) * (-(-self.tilesY // factor))

    def bin_async(self, prims: Prims, ext: torch.Tensor, tmpl_shape, capacity: int,
                  tile_offsets: torch.Tensor, tile_indices: torch.Tensor,
                  pair_count: torch.Tensor, factor: int = 1,
                  ranges: torch.Tensor | None = None) -> None:
        """Launch binning into caller buffers; no host synchronisation.
        ``factor`` > 1 bins per super-tile of factor x factor tiles.  Each
        primitive's tile range goes into ``ranges`` ((N,4) int32; required for
        ``factor`` > 1, optional for per-tile lists): after a pair overflow the
        raster kernels walk every primitive and use it only inside that range."""
        T, Ht, Wt = tmpl_shape
        if factor > 1 or ranges is not None:
            self._bin_tiles(prims, ext, tmpl_shape, factor, _ptr(tile_offsets), _ptr(tile_indices),
                            capacity, _ptr(pair_count), _ptr(ranges))
            return
        self.records_gen += 1
        check(self._L.dbmp_bin(self._h, ctypes.byref(prims), _ptr(ext), int(T), int(Ht), int(Wt),
                               _ptr(tile_offsets), _ptr(tile_indices), int(capacity),
                               _ptr(pair_count), _stream(ext.device)))

    def _bin_tiles(self, prims: Prims, ext: torch.Tensor, tmpl_shape, factor: int, offsets: int,
                   indices: int, capacity: int, pair_count: int, ranges: int) -> None:
        """dbmp_bin_tiles on device addresses."""
        T, Ht, Wt = tmpl_shape
        self.records_gen += 1
        check(self._L.dbmp_bin_tiles(self._h, ctypes.byref(prims), ext.data_ptr(), int(T), int(Ht),
                                     int(Wt), int(factor), offsets, indices, int(capacity),
                                     pair_count, ranges, _stream(ext.device)))

    @staticmethod
    def _ranges(prims: Prims, dev):
        return torch.empty((max(prims.num_primitives, 1), 4), dtype=torch.int32, device=dev)

    def bin(self, prims: Prims, ext: torch.Tensor, tmpl_shape, factor: int = 1):
        """Tile lists (offsets (lists+1) int32, indices (P) int32; with
        ``factor`` > 1 also the (N,4) tile ranges).  Reads the pair count back
        once (the only host synchronisation of the step) and regrows the index
        buffer if it overflowed."""
        dev = ext.device
        N = prims.num_primitives
        cap = max(self.pair_capacity, 4 * N + 1024)
        offs = torch.empty(self.list_count(factor) + 1, dtype=torch.int32, device=dev)
        cnt = torch.empty(1, dtype=torch.int64, device=dev)
        rg = self._ranges(prims, dev) if factor > 1 else None
        for _ in range(2):
            idx = torch.empty(cap, dtype=torch.int32, device=dev)
            self.bin_async(prims, ext, tmpl_shape, cap, offs, idx, cnt, factor, rg)
            P = int(cnt.item())
            if P <= cap:
                self.pair_capacity = max(self.pair_capacity, cap)
                if factor > 1:
                    return offs, idx[:P], rg
                return offs, idx[:P]
            cap = int(math.ceil(P * 1.25)) + 1024
            if cap >= 2 ** 31:
                raise RuntimeError(f"{P} (tile, primitive) pairs exceed the int32 index range")
        raise RuntimeError("binning did not converge")  # unreachable

    def prepared_templates(self, templates: torch.Tensor, store_dtype):
        """Templates converted to the storage dtype, their extents and a content
        version, cached while the source tensor is unchanged (same storage,
        same torch version counter)."""
        key = (templates.data_ptr(), templates._version, tuple(templates.shape), templates.dtype,
               store_dtype, templates.device)
        c = self._tmpl_cache
        if c is not None and c[0] == key:
            return c[1], c[2], c[3]
        conv = templates.to(dtype=store_dtype).contiguous()
        if conv.data_ptr() == templates.data_ptr():
            conv = conv.clone()        # own the bytes the cache keys on
        ext = self.template_extents(conv)
        self._tmpl_gen += 1
        self._tmpl_cache = (key, conv, ext, self._tmpl_gen)
        return conv, ext, self._tmpl_gen

    def bin_lazy(self, prims: Prims, ext: torch.Tensor, tmpl_shape, factor: int = 1) -> Lists:
        """Tile lists without a host synchronisation in steady state.  The
        capacity follows the pair counts seen so far, read back asynchronously
        every ``READBACK_EVERY`` calls (each read-back is a device-to-host copy
        on the stream); the first call, with no history, reads the count once.
        If a call overflows, the raster kernels walk every primitive for that
        call, each within the tile range the binning gave it (the lists'
        membership, so the result is the same, only slower), until a read-back
        grows the buffer."""
        dev = ext.device
        N = prims.num_primitives
        self._lazy_calls += 1
        if self._pending is not None and self._pending.query():
            seen = int(self._readback[0])
            if seen > 0.9 * self.pair_capacity:
                self.pair_capacity = int(seen * 1.3) + 4096
            self._pending = None
        if self.pair_capacity == 0:
            out = self.bin(prims, ext, tmpl_shape, factor)
            self.pair_capacity = int(out[1].numel() * 1.3) + 4096
            return Lists(out[0], out[1], factor=factor, ranges=out[2] if factor > 1 else None)
        cap = self.pair_capacity
        # one allocation: count (int64) | pad | ranges (N,4) | offsets | pad |
        # push masks (int64 words) | indices
        L = self.list_count(factor) + 1
        nr = 4 * max(N, 1)
        nw = self._pm_words.get((cap, factor))
        if nw is None:
            nw = self._pm_words[(cap, factor)] = int(self._L.dbmp_push_mask_words(self._h, cap,
                                                                                 int(factor)))
        o_offs = 4 + nr
        o_pm = o_offs + ((L + 3) & ~3)
        o_idx = o_pm + 2 * nw
        buf = torch.empty(o_idx + cap, dtype=torch.int32, device=dev)
        a = buf.data_ptr()
        self._bin_tiles(prims, ext, tmpl_shape, factor, a + 4 * o_offs, a + 4 * o_idx, cap, a, a + 16)
        if self._pending is None and self._lazy_calls % self.READBACK_EVERY == 0:
            if self._readback is None:
                self._readback = torch.empty(1, dtype=torch.int64, pin_memory=True)
                self._readback_ev = torch.cuda.Event()
            self._readback.copy_(buf[0:2].view(torch.int64), non_blocking=True)
            self._readback_ev.record(torch.cuda.current_stream(dev))
            self._pending = self._readback_ev
        return _PackedLists(buf, cap, factor, (o_offs, L, o_pm, o_idx))

    def forward(self, prims: Prims, bm: Bitmaps, lists: Lists, eps_T: float, out_color, out_alpha,
                out_T, out_last) -> None:
        L = lists.c()
        if not prims.reuse_records:
            self.records_gen += 1
        check(self._L.dbmp_forward(self._h, ctypes.byref(prims), ctypes.byref(bm), ctypes.byref(L),
                                   float(eps_T), _ptr(out_color), _ptr(out_alpha), _ptr(out_T),
                                   _ptr(out_last), _stream(out_color.device)))

    def forward_batch(self, prims: list, bm: Bitmaps, lists: list, eps_T: float, out_color,
                      out_alpha, out_T, out_last) -> None:
        """B candidates (same N, dtype, bitmaps) in one launch; outputs stacked
        (B,H,W,3) / (B,H,W)."""
        B = len(prims)
        P = (Prims * B)(*prims)
        L = (TileLists * B)(*[x.c() for x in lists])
        self.records_gen += 1
        check(self._L.dbmp_forward_batch(self._h, B, P, ctypes.byref(bm), L, float(eps_T),
                                         _ptr(out_color), _ptr(out_alpha), _ptr(out_T),
                                         _ptr(out_last), _stream(out_color.device)))

    def backward_batch(self, prims: list, bm: Bitmaps, lists: list, eps_T, grad_color, grad_alpha,
                       out_color, out_T, out_last, g_means, g_r, g_th, g_v, g_c) -> None:
        """Backward of forward_batch; gradients stacked (B,N,...)."""
        B = len(prims)
        P = (Prims * B)(*prims)
        L = (TileLists * B)(*[x.c() for x in lists])
        self.records_gen += 1
        check(self._L.dbmp_backward_batch(self._h, B, P, ctypes.byref(bm), L, float(eps_T),
                                          _ptr(grad_color), _ptr(grad_alpha), _ptr(out_color),
                                          _ptr(out_T), _ptr(out_last), _ptr(g_means), _ptr(g_r),
                                          _ptr(g_th), _ptr(g_v), _ptr(g_c),
                                          _stream(grad_color.device)))

    def backward(self, prims: Prims, bm: Bitmaps, lists: Lists, eps_T, grad_color, grad_alpha,
                 out_color, out_T, out_last, g_means, g_r, g_th, g_v, g_c) -> None:
        L = lists.c()
        if not prims.reuse_records:
            self.records_gen += 1
        check(self._L.dbmp_backward(self._h, ctypes.byref(prims), ctypes.byref(bm), ctypes.byref(L),
                                    float(eps_T), _ptr(grad_color), _ptr(grad_alpha), _ptr(out_color),
                                    _ptr(out_T), _ptr(out_last), _ptr(g_means), _ptr(g_r), _ptr(g_th),
                                    _ptr(g_v), _ptr(g_c), _stream(grad_color.device)))

    def backward_mse(self, prims: Prims, bm: Bitmaps, lists: Lists, eps_T, target, out_color, out_T,
                     out_last, g_means, g_r, g_th, g_v, g_c, loss_accum=None, bg=None,
                     bg_value: float = 1.0) -> None:
        """Background composite + MSE + backward in one raster pass."""
        L = lists.c()
        if not prims.reuse_records:
            self.records_gen += 1
        check(self._L.dbmp_backward_mse(self._h, ctypes.byref(prims), ctypes.byref(bm), ctypes.byref(L),
                                        float(eps_T), _ptr(target), _ptr(bg), float(bg_value),
                                        _ptr(loss_accum), _ptr(out_color), _ptr(out_T), _ptr(out_last),
                                        _ptr(g_means), _ptr(g_r), _ptr(g_th), _ptr(g_v), _ptr(g_c),
                                        _stream(target.device)))

    def backward_mse_rows(self, prims: Prims, bm: Bitmaps, lists: Lists, eps_T, target, out_color,
                          out_T, out_last, grad_rows, loss_accum=None, bg=None,
                          bg_value: float = 1.0) -> None:
        """backward_mse that adds the gradients into ``grad_rows`` ((N,8) fp32:
        x, y, r, theta, opacity logit, 3 colour logits)."""
        L = lists.c()
        if not prims.reuse_records:
            self.records_gen += 1
        check(self._L.dbmp_backward_mse_rows(self._h, ctypes.byref(prims), ctypes.byref(bm),
                                             ctypes.byref(L), float(eps_T), _ptr(target), _ptr(bg),
                                             float(bg_value), _ptr(loss_accum), _ptr(out_color),
                                             _ptr(out_T), _ptr(out_last), _ptr(grad_rows),
                                             _stream(target.device)))

    def loss_mse(self, color, alpha, target, grad_color, grad_alpha, loss_accum=None, bg=None,
                 bg_value: float = 1.0) -> None:
        check(self._L.dbmp_loss_mse(self._h, _ptr(color), _ptr(alpha), _ptr(bg), float(bg_value),
                                    _ptr(target), _ptr(grad_color), _ptr(grad_alpha), _ptr(loss_accum),
                                    _stream(color.device)))

    def rescale_grads(self, grad_color, grad_alpha, scale, applied) -> None:
        """Scale the two gradient images in place from the factor in
        ``applied[0]`` to the one in ``scale`` (device fp32 scalars), then store
        it in ``applied[0]``; ``applied[1]`` is the kernel's counter (zero); a
        ratio of 1 touches nothing."""
        check(self._L.dbmp_rescale_grads(_ptr(grad_color), grad_color.numel(), _ptr(grad_alpha),
                                         grad_alpha.numel(), _ptr(scale), _ptr(applied),
                                         _stream(grad_color.device)))

    def adam_step(self, cfg: "Adam", device) -> None:
        check(self._L.dbmp_adam_step(self._h, ctypes.byref(cfg), _stream(device)))

    # timing -------------------------------------------------------------
    def timing_enable(self, on: bool = True) -> None:
        check(self._L.dbmp_timing_enable(self._h, int(on)))

    def timing_stats(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        fn, bn = ctypes.c_int64(), ctypes.c_int64()
        check(self._L.dbmp_timing_stats(self._h, ctypes.byref(f), ctypes.byref(fn), ctypes.byref(b),
                                        ctypes.byref(bn)))
        return f.value, fn.value, b.value, bn.value

    def timing_reset(self) -> None:
        check(self._L.dbmp_timing_reset(self._h))
