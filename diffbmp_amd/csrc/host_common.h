// Host-side helpers shared by the C ABI translation units: error state,
// checked HIP calls, device allocation, fill and texel-quad kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/diffbmp_amd.h"
#include "dbmp_device.h"

namespace dbmp_host {

using dbmp::kF32;

inline thread_local std::string g_err;

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(DBMP_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
  } while (0)

#define LAUNCH_CHECK(name)                                                          \
  do {                                                                              \
    hipError_t e_ = hipGetLastError();                                              \
    if (e_ != hipSuccess)                                                           \
      return fail(DBMP_E_HIP, "launch of %s failed: %s", name, hipGetErrorString(e_)); \
  } while (0)

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

static __global__ void k_fill_f32(float* p, long long n, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Every template's quads / records are Ht + 1 rows of Wt entries: rows
// 0..Ht-1 hold the texels, row Ht is padding (a sample's second row is read
// at entry + Wt, inside its template).
__host__ __device__ inline long long tmpl_entries(int Ht, int Wt) { return (long long)(Ht + 1) * Wt; }

// Entry e of the padded layout -> (template texel index s, x, y); y == Ht on
// the padding row.
__device__ __forceinline__ void padded_entry(long long e, int Ht, int Wt, long long& s, int& x,
                                             int& y) {
  const long long per = (long long)(Ht + 1) * Wt;
  const long long t = e / per;
  const int r = (int)(e - t * per);
  y = r / Wt;
  x = r - y * Wt;
  s = t * Ht * Wt + (long long)y * Wt + x;
}

template <int DT>
static __global__ void k_build_quads(const void* src, int T, int Ht, int Wt, void* dst) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)T * tmpl_entries(Ht, Wt)) return;
  long long i;
  int x, y;
  padded_entry(e, Ht, Wt, i, x, y);
  const bool row = y < Ht;   // the padding row's quads are zero (never pushed)
  const bool rx = x + 1 < Wt, ry = y + 1 < Ht;
  if constexpr (DT == kF32) {
    const float* s = (const float*)src;
    float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
    if (row) {
      q.x = s[i];
      q.y = rx ? s[i + 1] : 0.f;
      q.z = ry ? s[i + Wt] : 0.f;
      q.w = (rx && ry) ? s[i + Wt + 1] : 0.f;
    }
    ((float4*)dst)[e] = q;
  } else {
    const unsigned short* s = (const unsigned short*)src;   // raw fp16 / bf16 bits
    uint2 q = make_uint2(0u, 0u);
    if (row) {
      const uint32_t a = s[i], b = rx ? s[i + 1] : 0u, c = ry ? s[i + Wt] : 0u,
                     d = (rx && ry) ? s[i + Wt + 1] : 0u;
      q = make_uint2(a | (b << 16), c | (d << 16));
    }
    ((uint2*)dst)[e] = q;
  }
}

// RGBA texel records (dbmp::RecT) of T templates (T,Ht,Wt) and their colour
// maps (T,Ht,Wt,3) in the padded layout.  fp32: per texel the alpha quad
// (zero outside the bitmap) and the colour map's clamp-to-edge
// neighbourhood, 16 values; zero on the padding row.  fp16 / bf16: per
// texel one row of the neighbourhood, alpha of (x,y), (x+1,y) and RGB of
// (x,y), (x1,y); the padding row has alpha 0 and row Ht-1's colours.
template <int DT>
static __global__ void k_build_rgba(const void* alpha, const void* color, int T, int Ht, int Wt,
                                    void* dst) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)T * tmpl_entries(Ht, Wt)) return;
  long long i;
  int x, y;
  padded_entry(e, Ht, Wt, i, x, y);
  const bool row = y < Ht;
  const bool rx = x + 1 < Wt, ry = y + 1 < Ht;
  const int yc = row ? y : Ht - 1;   // colour row (clamp-to-edge)
  const long long base = i - ((long long)y * Wt + x);   // the template's first texel
  const int x1 = rx ? x + 1 : Wt - 1, y1 = ry ? y + 1 : Ht - 1;
  if constexpr (DT == kF32) {
    float4* d = (float4*)dst + 4 * e;
    if (!row) {
#pragma unroll
      for (int k = 0; k < 4; ++k) d[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      return;
    }
    const long long t[4] = {base + (long long)y * Wt + x, base + (long long)y * Wt + x1,
                            base + (long long)y1 * Wt + x, base + (long long)y1 * Wt + x1};
    const float* a = (const float*)alpha;
    const float* c = (const float*)color;
    float v[16];
    v[0] = a[i];
    v[1] = rx ? a[i + 1] : 0.f;
    v[2] = ry ? a[i + Wt] : 0.f;
    v[3] = (rx && ry) ? a[i + Wt + 1] : 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) v[4 + 3 * k + ch] = c[3 * t[k] + ch];
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
  } else {
    const unsigned short* a = (const unsigned short*)alpha;   // raw fp16 / bf16 bits
    const unsigned short* c = (const unsigned short*)color;
    const uint32_t a0 = row ? a[i] : 0u, a1 = (row && rx) ? a[i + 1] : 0u;
    const long long c0 = 3 * (base + (long long)yc * Wt + x), c1 = 3 * (base + (long long)yc * Wt + x1);
    const uint32_t r0 = c[c0], g0 = c[c0 + 1], b0 = c[c0 + 2], r1 = c[c1], g1 = c[c1 + 1], b1 = c[c1 + 2];
    ((uint4*)dst)[e] = make_uint4(a0 | (a1 << 16), r0 | (g0 << 16), b0 | (r1 << 16), g1 | (b1 << 16));
  }
}

template <typename T>
inline int grow(T** p, long long n) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (n <= 0) return 0;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(DBMP_E_ALLOC, "hipMalloc of %lld bytes failed: %s", (long long)(sizeof(T) * n),
                hipGetErrorString(e));
  }
  return 0;
}


// Texel quads of a template set: the entry of texel i holds the 2x2
// neighbourhood (i, i+1, i+W, i+W+1), zero outside the bitmap; 16 bytes
// (fp32) or 8 bytes (fp16 / bf16 bit patterns) each, in the padded layout.
inline int launch_build_quads(const dbmp_bitmaps* b, void* dst, hipStream_t st) {
  const long long tot = (long long)b->num_templates * tmpl_entries(b->height, b->width);
  if (tot == 0) return 0;
  const dim3 grid(ceil_div(tot, 256)), block(256);
  if (b->dtype == DBMP_F32)
    hipLaunchKernelGGL(k_build_quads<dbmp::kF32>, grid, block, 0, st, b->templates,
                       b->num_templates, b->height, b->width, dst);
  else
    hipLaunchKernelGGL(k_build_quads<dbmp::kF16>, grid, block, 0, st, b->templates,
                       b->num_templates, b->height, b->width, dst);
  LAUNCH_CHECK("k_build_quads");
  return 0;
}

// RGBA texel records of the templates and per-template colour maps of `b`
// into dst (T*(Ht+1)*Wt records of rgba_bytes(dtype)).
inline int launch_build_rgba(const dbmp_bitmaps* b, void* dst, hipStream_t st) {
  const long long tot = (long long)b->num_templates * tmpl_entries(b->height, b->width);
  if (tot == 0) return 0;
  const dim3 grid(ceil_div(tot, 256)), block(256);
  if (b->dtype == DBMP_F32)
    hipLaunchKernelGGL(k_build_rgba<dbmp::kF32>, grid, block, 0, st, b->templates, b->colors_orig,
                       b->num_templates, b->height, b->width, dst);
  else
    hipLaunchKernelGGL(k_build_rgba<dbmp::kF16>, grid, block, 0, st, b->templates, b->colors_orig,
                       b->num_templates, b->height, b->width, dst);
  LAUNCH_CHECK("k_build_rgba");
  return 0;
}

// Bytes per entry of the alpha quads / RGBA records.
inline size_t quad_bytes(int dtype) { return dtype == DBMP_F32 ? 16 : 8; }
inline size_t rgba_bytes(int dtype) { return dtype == DBMP_F32 ? 64 : 16; }

// Bytes of a template set's quads (rgba: RGBA records).  The raster kernels
// address them with 32-bit byte offsets: callers reject sets of 2^32 bytes
// or more.
inline long long texel_record_bytes(int T, int Ht, int Wt, int dtype, bool rgba) {
  return (long long)T * tmpl_entries(Ht, Wt) * (long long)(rgba ? rgba_bytes(dtype) : quad_bytes(dtype));
}

}  // namespace dbmp_host
