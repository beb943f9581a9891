"""MI355X-native (gfx950, HIP) implementation of DiffBMP's differentiable tile
rasterizer: GPU binning with a radix sort, forward composite and backward
adjoint behind the reference's ``TileRasterizer`` / ``autograd.Function`` API.
"""
from .tile_rasterizer import (  # noqa: F401
    CUDA_AVAILABLE, CUDA_FP16_AVAILABLE, DEFAULT_EPS_T, TileRasterizer,
    TileRasterizerBatchFunction, TileRasterizerFunction, TileRasterizerFunctionFP16,
    bin_primitives, pack_mapping, print_cuda_timing_stats, print_cuda_timing_stats_fp16,
    reset_cuda_timing_stats, reset_cuda_timing_stats_fp16, unpack_mapping,
)

__version__ = "0.1.0"
