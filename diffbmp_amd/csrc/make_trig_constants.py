"""Derives the constants of dbmp_trig.h (fp32 sincos shared by the GPU kernels
and the CPU oracle) and prints them as C.

* 2/pi as bits for the Payne-Hanek reduction: from pi computed with Machin's
  formula (pi/4 = 4 atan(1/5) - atan(1/239)) in integer arithmetic.
* pi/2 split into three fp32 parts P1 + P2 + P3 (Cody-Waite reduction).
* near-minimax polynomials for sin(r) = r + r^3 S(r^2) (S cubic) and
  cos(r) = 1 + r^2 C(r^2) (C quartic) on |r| <= pi/4 (+ slack), fitted on the
  relative error of sin / cos with Lawson's iteratively reweighted least
  squares (float64; the targets from their Taylor series), then rounded to
  fp32.

Run: python diffbmp_amd/csrc/make_trig_constants.py
"""
from __future__ import annotations

import numpy as np

def _atan_inv(n: int, scale: int) -> int:
    """atan(1/n) * scale by its Taylor series, integer arithmetic."""
    total, term, k, sign = 0, scale // n, 1, 1
    n2 = n * n
    while term:
        total += sign * (term // k)
        term //= n2
        k += 2
        sign = -sign
    return total


def pi_scaled(bits: int) -> int:
    """floor(pi * 2^bits) up to a few units in the last place."""
    guard = 32
    scale = 1 << (bits + guard)
    p = 4 * (4 * _atan_inv(5, scale) - _atan_inv(239, scale))
    return p >> guard


def two_over_pi_words(nwords: int = 10) -> list[int]:
    """The first 32*nwords bits after the binary point of 2/pi."""
    nb = 32 * nwords
    pi = pi_scaled(nb + 96)                       # pi * 2^(nb+96)
    v = (1 << (2 * nb + 97)) // pi                # 2/pi * 2^nb (+ guard), floor
    return [(v >> (32 * (nwords - 1 - i))) & 0xFFFFFFFF for i in range(nwords)]


def f32(x) -> np.float32:
    return np.float32(x)


def pio2_parts():
    """pi/2 = P1 + P2 + P3 + tiny, each part the fp32 rounding of what is left."""
    from fractions import Fraction
    pio2 = Fraction(pi_scaled(200), 1 << 201)
    parts = []
    rest = pio2
    for _ in range(3):
        p = f32(float(rest))
        parts.append(p)
        rest -= Fraction(float(p))
    return parts


def _series(x, first, sign_step):
    """sum_k (-1)^k x^(first+2k) / (first+2k)!  (float64, |x| <= 1)."""
    out = np.zeros_like(x)
    term = x ** first / float(np.prod(np.arange(1, first + 1)))
    k = first
    sign = 1.0
    while np.abs(term).max() > 1e-30:
        out += sign * term
        term = term * x * x / ((k + 1) * (k + 2))
        k += 2
        sign = -sign
    return out


def lawson(t, f, w, deg, iters=400):
    """Near-minimax fit of the polynomial P (degree ``deg`` in t) to f on the
    points t, minimising max |w (P - f)|: Lawson's algorithm (iteratively
    reweighted least squares)."""
    V = np.vander(t, deg + 1, increasing=True)
    lw = np.full(len(t), 1.0 / len(t))
    c = None
    for _ in range(iters):
        sw = np.sqrt(lw) * w
        c, *_ = np.linalg.lstsq(V * sw[:, None], f * sw, rcond=None)
        e = np.abs(w * (V @ c - f))
        lw = lw * e
        lw /= lw.sum()
    return c, float(np.abs(w * (V @ c - f)).max())


def polys():
    """sin r = r + r^3 S(r^2), S cubic; cos r = 1 + r^2 C(r^2), C quartic; the
    weights make the fitted error the relative error of sin / cos."""
    r = np.linspace(1e-6, 0.7854 + 2e-3, 40001)
    t = r * r
    sin_r = _series(r, 1, -1)
    cos_r = 1.0 + _series(r, 2, -1) * -1.0 + 0.0
    # _series(r, 2) = r^2/2 - r^4/24 + ...; cos r = 1 - that
    cos_r = 1.0 - _series(r, 2, -1)
    S = (sin_r - r) / r ** 3
    C = (cos_r - 1.0) / r ** 2
    s, es = lawson(t, S, r ** 3 / sin_r, 3)
    c, ec = lawson(t, C, r ** 2 / cos_r, 4)
    return [f32(v) for v in s], [f32(v) for v in c], es, ec


def main():
    w = two_over_pi_words()
    p = pio2_parts()
    s, c, es, ec = polys()
    hx = lambda v: "0x%08X" % np.float32(v).view(np.uint32)
    print("// 2/pi, bits 1..%d after the binary point" % (32 * len(w)))
    print("kTwoOverPiBits[%d] = {%s};" % (len(w), ", ".join("0x%08Xu" % x for x in w)))
    print("// pi/2 = P1 + P2 + P3 (fp32 bit patterns)")
    print("P1 %s (%r)  P2 %s (%r)  P3 %s (%r)" % (hx(p[0]), float(p[0]), hx(p[1]), float(p[1]),
                                                hx(p[2]), float(p[2])))
    print("2/pi fp32 %s (%r)" % (hx(f32(2 / np.pi)), float(f32(2 / np.pi))))
    print("S(t) coefficients t^0..t^3:", [hx(v) for v in s], [float(v) for v in s],
          "relative fit error %.3g" % es)
    print("C(t) coefficients t^0..t^4:", [hx(v) for v in c], [float(v) for v in c],
          "relative fit error %.3g" % ec)


if __name__ == "__main__":
    main()
