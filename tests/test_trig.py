"""The fp32 sine/cosine shared by the GPU kernels and the oracle
(diffbmp_amd/csrc/dbmp_trig.h), on the CPU through the oracle's build.

* its constants are the ones make_trig_constants.py derives (2/pi bits from
  Machin's formula, the pi/2 split, the fitted polynomials);
* error <= 2 ulp of the fp32 result against the C library's double sin/cos
  over the fp32 angles in [0, 2 pi) (every 16th bit pattern here, every one
  with DBMP_FULL_TRIG_SCAN=1), over [2 pi, 2^17) and over the Payne-Hanek
  range up to FLT_MAX (sampled);
* symmetry, signed zero, inf/NaN.
"""
import ctypes
import importlib.util
import math
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "diffbmp_amd", "csrc")
f32 = np.float32


def _gen():
    spec = importlib.util.spec_from_file_location("mtc", os.path.join(CSRC, "make_trig_constants.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _bits(v):
    return "0x%08X" % int(np.float32(v).view(np.uint32))


def test_constants_are_the_generators():
    m = _gen()
    text = open(os.path.join(CSRC, "dbmp_trig.h")).read()
    words = m.two_over_pi_words()
    for w in words:
        assert ("0x%08Xu" % w) in text
    # 2/pi itself: 0.1010 0010 1111 1001 ... = 0xA2F9836E... (pi from Machin's formula)
    assert abs(words[0] / 2.0 ** 32 - 2 / math.pi) < 2.0 ** -32
    p1, p2, p3 = m.pio2_parts()
    for p in (p1, p2, p3):
        assert _bits(p) + "u" in text
    assert abs(float(p1) + float(p2) + float(p3) - math.pi / 2) < 1e-20
    s, c, es, ec = m.polys()
    assert es < 1e-9 and ec < 1e-9
    for v in list(s) + list(c)[1:]:
        assert _bits(v) + "u" in text, v
    assert float(c[0]) == -0.5


def _scan(lo, hi, stride):
    out = np.zeros(2)
    lb, hb = int(np.float32(lo).view(np.uint32)), int(np.float32(hi).view(np.uint32))
    O.lib("f64").oracle_sincos_scan(ctypes.c_uint(lb), ctypes.c_uint(hb), ctypes.c_uint(stride), O._p(out))
    return out


def test_error_bound_over_a_turn():
    stride = 1 if os.environ.get("DBMP_FULL_TRIG_SCAN") else 16
    es, ec = _scan(0.0, 2 * math.pi, stride)
    print(f"[0, 2pi) every {stride}: max ulp sin {es:.3f} cos {ec:.3f}")
    assert es <= 2.0 and ec <= 2.0


@pytest.mark.parametrize("lo,hi,stride", [(2 * math.pi, 131072.0, 101), (131072.0, 3.4e38, 1009)])
def test_error_bound_large_arguments(lo, hi, stride):
    es, ec = _scan(lo, hi, stride)
    print(f"[{lo}, {hi}) every {stride}: max ulp sin {es:.3f} cos {ec:.3f}")
    assert es <= 2.0 and ec <= 2.0


def test_symmetry_and_special_values():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-50, 50, 10000), rng.uniform(-1e30, 1e30, 1000)]).astype(f32)
    c, s = O.cos_sin(x)
    c2, s2 = O.cos_sin(-x)
    assert np.array_equal(c, c2) and np.array_equal(s, -s2)
    c, s = O.cos_sin(np.array([0.0, -0.0, np.inf, -np.inf, np.nan], f32))
    assert c[0] == 1 and s[0] == 0 and not np.signbit(s[0])
    assert c[1] == 1 and s[1] == 0 and np.signbit(s[1])
    assert np.isnan(c[2:]).all() and np.isnan(s[2:]).all()
    # exact quadrant points of the fp32 grid: sin(fp32(pi)) is pi - fp32(pi)
    c, s = O.cos_sin(np.array([np.pi], f32))
    assert abs(float(s[0]) - (math.pi - float(f32(math.pi)))) <= 2 * abs(float(np.spacing(s[0])))
