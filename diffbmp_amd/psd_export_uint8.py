"""Layer export: per-primitive uint8 RGBA crops on the GPU.

Mirrors ``cuda_tile_rasterizer/psd_export_uint8.py`` of the reference (same
function names, arguments and return values) over the library's
``dbmp_psd_bounds`` / ``dbmp_psd_layers`` kernels.  Differences: the crop
offsets are 64-bit (the reference's int32 offsets overflow past 2 GiB of
layers), a layer's "has some alpha" test reads a per-layer maximum computed on
the GPU, and colour maps per template are sampled in place instead of being
gathered per primitive first.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import numpy as np
import torch

from . import _native

try:
    _native.lib()
    CUDA_AVAILABLE = True
except Exception:   # the HIP library is not built or no GPU: callers see the flag
    CUDA_AVAILABLE = False


def _prims(means2D, r, theta, c, v, sel) -> _native.LayerPrims:
    return _native.LayerPrims(means2D.data_ptr(), r.data_ptr(), theta.data_ptr(), c.data_ptr(),
                              v.data_ptr(), sel.data_ptr(), int(r.shape[0]))


def layer_bounds(means2D: torch.Tensor, r: torch.Tensor, H: int, W: int, scale_factor: float = 1.0):
    """Stage 1 on the device: boxes (N,4) int32, sizes (N,2) int32, byte
    offsets (N+1) int64 and the total byte count (read back)."""
    dev = means2D.device
    N = r.shape[0]
    boxes = torch.empty((N, 4), dtype=torch.int32, device=dev)
    sizes = torch.empty((N, 2), dtype=torch.int32, device=dev)
    offs = torch.empty(N + 1, dtype=torch.int64, device=dev)
    total = torch.empty(1, dtype=torch.int64, device=dev)
    p = _native.LayerPrims(means2D.data_ptr(), r.data_ptr(), None, None, None, None, int(N))
    _native.check(_native.lib().dbmp_psd_bounds(ctypes.byref(p), int(H), int(W), float(scale_factor),
                                                boxes.data_ptr(), sizes.data_ptr(), offs.data_ptr(),
                                                total.data_ptr(), _native._stream(dev)))
    return boxes, sizes, offs, int(total.item())


def export_psd_layers_cuda_uint8(
    renderer_S: torch.Tensor,        # [P, H_t, W_t] (or [H_t, W_t]) primitive templates
    x: torch.Tensor, y: torch.Tensor, r: torch.Tensor, theta: torch.Tensor,
    v: torch.Tensor,                 # [N] visibility (sigmoid applied)
    c: torch.Tensor,                 # [N, 3] colours (sigmoid applied)
    global_bmp_sel: torch.Tensor,    # [N] template indices
    H: int, W: int,
    scale_factor: float = 1.0,
    alpha_upper_bound: float = 1.0,
    c_blend: float = 0.0,
    primitive_colors: Optional[torch.Tensor] = None,   # [P, H_t, W_t, 3] or [P, 3]
) -> Tuple[List[np.ndarray], List[Tuple[int, int, int, int]], List[int]]:
    """RGBA crops of every primitive that has one (size > 1x1 and some alpha),
    their (left, top, right, bottom) boxes and their primitive indices
    (``psd_export_uint8.py:22-150``)."""
    if not CUDA_AVAILABLE:
        raise RuntimeError("the diffbmp_amd HIP library is not available")
    _native.require_gpu(x, "x")
    dev = x.device
    N = int(x.shape[0])
    if N == 0:
        return [], [], []
    means2D = torch.stack([x.float(), y.float()], 1).contiguous()
    rf = r.float().contiguous()
    th = theta.float().contiguous()
    vf = v.float().contiguous()
    cf = c.float().contiguous()
    sel = global_bmp_sel.to(torch.int32).contiguous()
    S = renderer_S.float()
    if S.dim() == 2:
        S = S.unsqueeze(0)
    S = S.contiguous()
    T, Ht, Wt = S.shape
    cmaps, mode = None, 0
    if primitive_colors is not None and c_blend > 0.0:
        pc = primitive_colors.to(dev).float()
        if pc.dim() == 2:   # one colour per template: a constant map, as the reference expands it
            pc = pc[:, None, None, :].expand(pc.shape[0], Ht, Wt, 3)
        if pc.dim() != 4 or tuple(pc.shape[1:]) != (Ht, Wt, 3):
            raise ValueError(f"Unexpected primitive_colors shape: {tuple(primitive_colors.shape)}")
        cmaps, mode = pc.contiguous(), 2
    boxes, sizes, offs, total = layer_bounds(means2D, rf, H, W, scale_factor)
    buf = torch.empty(max(total, 16), dtype=torch.uint8, device=dev)
    amax = torch.empty(N, dtype=torch.int32, device=dev)
    p = _prims(means2D, rf, th, cf, vf, sel)
    _native.check(_native.lib().dbmp_psd_layers(
        ctypes.byref(p), S.data_ptr(), int(T), int(Ht), int(Wt),
        cmaps.data_ptr() if cmaps is not None else None, mode, float(c_blend), int(H), int(W),
        float(scale_factor), float(alpha_upper_bound), boxes.data_ptr(), offs.data_ptr(), total,
        buf.data_ptr(), amax.data_ptr(), _native._stream(dev)))
    h_boxes = boxes.cpu().numpy()
    h_sizes = sizes.cpu().numpy()
    h_offs = offs.cpu().numpy()
    h_amax = amax.cpu().numpy()
    h_buf = buf[:total].cpu().numpy()
    arrays, bounds, indices = [], [], []
    for i in range(N):
        w, h = (int(s) for s in h_sizes[i])
        if w <= 1 or h <= 1 or h_amax[i] == 0:
            continue
        arrays.append(h_buf[h_offs[i]:h_offs[i + 1]].reshape(h, w, 4))
        bounds.append(tuple(int(b) for b in h_boxes[i]))
        indices.append(i)
    return arrays, bounds, indices


def create_pixel_layers_from_numpy(numpy_arrays, bounds_list, valid_indices, names, psd_file):
    """Pixel layers of a layered-image document from the crops
    (``psd_export_uint8.py:152-186``); needs ``psd_tools`` and ``PIL``."""
    from PIL import Image
    from psd_tools.api.layers import PixelLayer
    layers = []
    for rgba, (left, top, _, _), orig in zip(numpy_arrays, bounds_list, valid_indices):
        name = names[orig] if orig < len(names) else f"Layer_{orig}"
        layers.append(PixelLayer.frompil(Image.fromarray(rgba, "RGBA"), psd_file, name, top=top,
                                         left=left))
    return layers
