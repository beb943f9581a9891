"""The library's fp32 sine/cosine on the GPU (dbmp_sincos, the function every
kernel uses) equals, bit for bit, the oracle's build of the same source
(dbmp_trig.h compiled by gcc), on the angles binning and raster meet (a turn
and a few), large arguments (the Payne-Hanek path) and special values."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from diffbmp_amd import _native

pytestmark = pytest.mark.gpu


def test_sincos_bit_exact_with_oracle(dev):
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(0, 2 * np.pi, 1 << 20), rng.uniform(-100, 100, 1 << 18),
        rng.uniform(-131072, 131072, 1 << 16), (rng.standard_normal(1 << 16) * 1e6),
        np.exp(rng.uniform(np.log(131072), np.log(3e38), 1 << 16)),
        np.arange(-64, 64) * (np.pi / 4),
        [0.0, -0.0, np.inf, -np.inf, np.nan, 3.4028235e38, 1e-30, -1e-45]]).astype(np.float32)
    xt = torch.from_numpy(x).to(dev)
    s = torch.empty_like(xt)
    c = torch.empty_like(xt)
    _native.check(_native.lib().dbmp_sincos(xt.data_ptr(), s.data_ptr(), c.data_ptr(), x.size,
                                            _native._stream(dev)))
    torch.cuda.synchronize()
    rc, rs = O.cos_sin(x)
    gs, gc = s.cpu().numpy(), c.cpu().numpy()
    # bit patterns (NaN compare as equal patterns of the quiet NaN the sources make)
    same = lambda a, b: (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
    bad_s, bad_c = ~same(gs, rs), ~same(gc, rc)
    print(f"{x.size} angles: sin mismatches {bad_s.sum()}, cos mismatches {bad_c.sum()}")
    assert not bad_s.any(), x[bad_s][:8]
    assert not bad_c.any(), x[bad_c][:8]
