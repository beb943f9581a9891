/*
 * dbmp_trig.h -- fp32 sine and cosine of one angle, in fp32 and integer
 * arithmetic only (no fp64).
 *
 * One source for the GPU kernels (hipcc) and the CPU oracle (gcc, C): with
 * every fused multiply-add written as fmaf, FMA contraction off
 * (-ffp-contract=off in both builds) and IEEE round-to-nearest throughout,
 * both produce the same bits, so binning and raster decisions that use cos/sin
 * agree exactly between the GPU and its checker.
 *
 * Method:
 *   - |x| < 2^17: Cody-Waite reduction (W. J. Cody and W. Waite, Software
 *     Manual for the Elementary Functions, 1980) r = x - j pi/2 with
 *     j = rint(x 2/pi) and pi/2 = P1 + P2 + P3 split into three fp32 parts,
 *     each product taken inside a fused multiply-add (the first step is
 *     exact);
 *   - larger |x|: Payne-Hanek reduction (M. Payne and R. Hanek, Radian
 *     reduction for trigonometric functions, SIGNUM Newsletter 18, 1983):
 *     the 24-bit significand times a 96-bit window of the bits of 2/pi in
 *     64-bit integer arithmetic gives the quadrant and a 64-bit fraction;
 *   - sin r = r + r^3 S(r^2), cos r = 1 + r^2 C(r^2) on |r| <= pi/4 with
 *     near-minimax polynomials S (cubic) and C (quartic), Horner with fmaf;
 *   - quadrant (j mod 4) and sign symmetry.
 * The constants (bits of 2/pi, the P1/P2/P3 split, the polynomial
 * coefficients) come from make_trig_constants.py in this directory, which
 * derives them from pi computed by Machin's formula.
 *
 * Accuracy (measured over every fp32 angle in [0, 2 pi) and a sample of large
 * arguments against the C library's double-precision sin/cos): at most 2 ulp
 * of the fp32 result; the test suite checks that bound.
 */
#ifndef DBMP_TRIG_H
#define DBMP_TRIG_H

#include <math.h>
#include <stdint.h>

/* the same function body for a HIP translation unit (host and device) and
 * for the oracle's C */
#if defined(__HIPCC__)
#define DBMP_TRIG_FN static __host__ __device__ inline
#else
#define DBMP_TRIG_FN static inline
#endif

/* bits 1..320 after the binary point of 2/pi */
#define DBMP_TWO_OVER_PI_WORDS 10
#define DBMP_TWO_OVER_PI_BITS                                                         \
  {0xA2F9836Eu, 0x4E441529u, 0xFC2757D1u, 0xF534DDC0u, 0xDB629599u, 0x3C439041u,      \
   0xFE5163ABu, 0xDEBBC561u, 0xB7246E3Au, 0x424DD2E0u}

typedef union {
  float f;
  uint32_t u;
} dbmp_f32_bits;

DBMP_TRIG_FN float dbmp_f32_from_bits(uint32_t u) {
  dbmp_f32_bits b;
  b.u = u;
  return b.f;
}

/* x >= 2^17, finite: r in [-pi/4, pi/4] and the quadrant q with
 * x = (4k + q) pi/2 + r. */
DBMP_TRIG_FN void dbmp_reduce_large(float x, float* r_out, uint32_t* q_out) {
  const uint32_t tbl[DBMP_TWO_OVER_PI_WORDS] = DBMP_TWO_OVER_PI_BITS;
  dbmp_f32_bits v;
  v.f = x;
  const int e = (int)(v.u >> 23) - 150;                 /* x = m 2^e */
  const uint64_t m = (uint64_t)((v.u & 0x7FFFFFu) | 0x800000u);
  /* bits i of 2/pi with e - i >= 2 add multiples of 4 to x 2/pi: the window
   * starts at bit i0 = max(1, e - 1) (1-based) */
  const int i0 = e - 1 > 1 ? e - 1 : 1;
  const int p = i0 - 1, k = p >> 5, o = p & 31;
  uint32_t w[3];
  for (int i = 0; i < 3; ++i)
    w[i] = o ? (tbl[k + i] << o) | (tbl[k + i + 1] >> (32 - o)) : tbl[k + i];
  /* P = m * (w0 2^64 + w1 2^32 + w2) as hi 2^64 + lo */
  const uint64_t a2 = m * (uint64_t)w[2], a1 = m * (uint64_t)w[1], a0 = m * (uint64_t)w[0];
  const uint64_t lo = a2 + (a1 << 32);
  const uint64_t hi = a0 + (a1 >> 32) + (lo < a2 ? 1u : 0u);
  /* x 2/pi (mod 4) = P 2^-sh */
  const int sh = i0 + 95 - e;                           /* in [94, 102] */
  const int hs = sh - 64;
  const uint64_t ip = hi >> hs;
  const uint64_t F = ((hi & ((1ull << hs) - 1ull)) << (64 - hs)) | (lo >> hs);
  /* nearest integer j and the fraction in [-1/2, 1/2) as a signed 64-bit
   * fixed-point value (halved: its conversions to float stay below 2^62) */
  const uint64_t j = ip + (F >> 63);
  const int64_t fs = (int64_t)F >> 1;                   /* fraction * 2^63 */
  const float fh = (float)fs;
  const float fl = (float)(fs - (int64_t)fh);
  const float P1 = dbmp_f32_from_bits(0x3FC90FDBu), P2 = dbmp_f32_from_bits(0xB33BBD2Eu);
  const float r = fmaf(fh, P1, fmaf(fh, P2, fl * P1));
  *r_out = r * dbmp_f32_from_bits(0x20000000u);         /* 2^-63, exact */
  *q_out = (uint32_t)j & 3u;
}

/* *s = sin(x), *c = cos(x); NaN for an infinite or NaN x. */
DBMP_TRIG_FN void dbmp_sincosf(float x, float* s_out, float* c_out) {
  dbmp_f32_bits v;
  v.f = x;
  const uint32_t neg = v.u >> 31;
  v.u &= 0x7FFFFFFFu;
  const float ax = v.f;
  if (!(ax <= 3.4028234663852886e38f)) {               /* inf or NaN */
    *s_out = x - x;
    *c_out = x - x;
    return;
  }
  float r;
  uint32_t q;
  if (ax < 131072.0f) {
    const float j = rintf(ax * dbmp_f32_from_bits(0x3F22F983u));   /* x 2/pi */
    r = fmaf(-j, dbmp_f32_from_bits(0x3FC90FDBu), ax);               /* P1: exact */
    r = fmaf(-j, dbmp_f32_from_bits(0xB33BBD2Eu), r);                /* P2 */
    r = fmaf(-j, dbmp_f32_from_bits(0xA6F72CEDu), r);                /* P3 */
    q = (uint32_t)(int32_t)j & 3u;
  } else {
    dbmp_reduce_large(ax, &r, &q);
  }
  const float t = r * r;
  const float S = fmaf(fmaf(fmaf(dbmp_f32_from_bits(0x363665AFu), t, dbmp_f32_from_bits(0xB95007B2u)),
                            t, dbmp_f32_from_bits(0x3C088884u)),
                       t, dbmp_f32_from_bits(0xBE2AAAABu));
  const float C = fmaf(fmaf(fmaf(fmaf(dbmp_f32_from_bits(0xB491E206u), t, dbmp_f32_from_bits(0x37D0074Eu)),
                                 t, dbmp_f32_from_bits(0xBAB60B5Au)),
                            t, dbmp_f32_from_bits(0x3D2AAAABu)),
                       t, -0.5f);
  const float sr = fmaf(r * t, S, r);
  const float cr = fmaf(t, C, 1.0f);
  float s, c;
  switch (q) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
  }
  *s_out = neg ? -s : s;
  *c_out = c;
}

#endif /* DBMP_TRIG_H */
