// Device-side building blocks shared by the binning and raster kernels.
//
// Compiled with -ffp-contract=off: the per-sample geometry (push box, texture
// coordinates, floor) must round every fp32 operation exactly as written so
// that its discrete decisions are reproducible on the host.  Value arithmetic
// (mask interpolation, compositing, gradients) uses explicit fmaf.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace dbmp {

enum DType : int { kF32 = 0, kF16 = 1, kBF16 = 2 };

// Per-primitive constants, computed once per call by the prep kernel and
// read by the raster kernels through uniform (scalar) loads.  64 bytes.
// The raster loops read the first three 16-byte quarters (one LDS load each);
// the fourth holds what only the binning and the gradient reduction use.
struct alignas(16) PrimRec {
  float mx, my, hw, hh;      // centre; push-box half extents (tile_forward.cu:195-210)
  float c, s, rinv, abase;   // cos, sin (dbmp_sincosf), 1/r, alpha_ub * sigmoid(v)
  float cr, cg, cb;          // sigmoid(colour logits)
  int tbase;                 // tmpl * (Ht + 1) * Wt: first texel entry of the template
  int tmpl;                  // template index, -1 => never pushed
  float r;                   // radius
  float sop;                 // sigmoid(opacity logit)
  float dav;                 // alpha_ub * s (1 - s): d alpha / d(opacity logit) per unit mask
};
static_assert(sizeof(PrimRec) == 64, "PrimRec is 64 bytes");

// Global-memory access ------------------------------------------------------
// A pointer that a kernel reads out of a descriptor in memory (the batched
// job table) is generic to the compiler, which then emits flat loads and
// stores.  Those count against the LDS wait counter as well as the vector
// memory one, so each wait for an LDS read also waits for them.
// Dereferencing through an address-space-1 (global) pointer gives global
// instructions; for kernel arguments, already known to be global, it
// changes nothing.  Every buffer these kernels touch is device memory.
#define DBMP_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const DBMP_GLOBAL T* gptr(const T* p) { return (const DBMP_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ DBMP_GLOBAL T* gptr(T* p) { return (DBMP_GLOBAL T*)p; }
__device__ __forceinline__ const DBMP_GLOBAL char* gbytes(const void* p) { return (const DBMP_GLOBAL char*)p; }
// 16-byte / 8-byte loads and stores through global pointers.  HIP's float4,
// uint2 and uint4 have user-defined constructors, which cannot copy from an
// address-space-qualified object; clang's native vectors can.
typedef float dbmp_f4 __attribute__((ext_vector_type(4)));
typedef int dbmp_i4 __attribute__((ext_vector_type(4)));
typedef unsigned dbmp_u2 __attribute__((ext_vector_type(2)));
typedef unsigned dbmp_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_f4(const DBMP_GLOBAL char* p) {
  const dbmp_f4 v = *(const DBMP_GLOBAL dbmp_f4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 ld_i4(const DBMP_GLOBAL char* p) {
  const dbmp_i4 v = *(const DBMP_GLOBAL dbmp_i4*)p;
  return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 ld_u2(const DBMP_GLOBAL char* p) {
  const dbmp_u2 v = *(const DBMP_GLOBAL dbmp_u2*)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 ld_u4(const DBMP_GLOBAL char* p) {
  const dbmp_u4 v = *(const DBMP_GLOBAL dbmp_u4*)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_f4(void* p, float4 v) {
  dbmp_f4 w;
  w.x = v.x; w.y = v.y; w.z = v.z; w.w = v.w;
  *(DBMP_GLOBAL dbmp_f4*)p = w;
}

// hardware fp32 atomic add (relaxed, device scope) on global memory
__device__ __forceinline__ void atomic_add_global(float* p, float v) {
  __hip_atomic_fetch_add(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Storage-type loads -------------------------------------------------------
__device__ __forceinline__ float ldf(const float* p, long i) { return gptr(p)[i]; }
__device__ __forceinline__ float ldf(const __half* p, long i) {
  return __half2float(__ushort_as_half(((const DBMP_GLOBAL unsigned short*)p)[i]));
}
__device__ __forceinline__ float ldf(const __hip_bfloat16* p, long i) {
  return __uint_as_float((uint32_t)((const DBMP_GLOBAL unsigned short*)p)[i] << 16);   // exact
}

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ float f16_lo(uint32_t w) {
  return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu)));
}
__device__ __forceinline__ float f16_hi(uint32_t w) {
  return __half2float(__ushort_as_half((unsigned short)(w >> 16)));
}

// Texel quads: entry (y,x) of template t holds its bilinear neighbourhood
// {T[y][x], T[y][x+1], T[y+1][x], T[y+1][x+1]}, zero beyond the last row or
// column (the reference's zero padding, tile_common.h:188-194).  One vector
// load per sample instead of four gathers.
template <int DT> struct QuadT;
template <> struct QuadT<kF32> {
  using type = float4;
  // i < 2^28 (the byte offset fits 32 bits: one base register + a 32-bit offset)
  __device__ static __forceinline__ void get(const float4* q, unsigned i, float& a, float& b, float& c, float& d) {
    const float4 v = ld_f4(gbytes(q) + (i << 4));
    a = v.x; b = v.y; c = v.z; d = v.w;
  }
};
template <> struct QuadT<kF16> {
  using type = uint2;
  __device__ static __forceinline__ void get(const uint2* q, unsigned i, float& a, float& b, float& c, float& d) {
    const uint2 v = ld_u2(gbytes(q) + (i << 3));
    a = f16_lo(v.x); b = f16_hi(v.x); c = f16_lo(v.y); d = f16_hi(v.y);
  }
};
template <> struct QuadT<kBF16> {
  using type = uint2;
  __device__ static __forceinline__ void get(const uint2* q, unsigned i, float& a, float& b, float& c, float& d) {
    const uint2 v = ld_u2(gbytes(q) + (i << 3));
    a = bf16_lo(v.x); b = bf16_hi(v.x); c = bf16_lo(v.y); d = bf16_hi(v.y);
  }
};

// RGBA texel records of a template set with per-template colour maps.  Every
// template holds Ht + 1 rows of Wt entries (as the alpha quads do); `rs` is
// the row stride Wt.
// RGBA texel records of a template set with per-template colour maps.  Every
// template holds Ht + 1 rows of Wt entries (as the alpha quads do); `rs` is
// the row stride Wt.
// fp32: entry (t, y, x) holds the template's bilinear neighbourhood (the alpha
// quad of QuadT, zero outside the bitmap) followed by the colour map's
// clamp-to-edge neighbourhood: RGB of (x,y), (x1,y), (x,y1), (x1,y1),
// x1 = min(x+1, W-1), y1 = min(y+1, H-1) -- 16 values, one aligned 64-byte
// record; the padding row is zero.
// fp16 / bf16: entry (t, y, x) holds one row of the neighbourhood, 8 values in
// one aligned 16-byte record: alpha of (x,y) and (x+1,y) (zero past the
// bitmap), then RGB of (x,y) and (x1,y).  The padding row y = Ht holds alpha
// 0 and the colours of row Ht-1 (clamp-to-edge).  A sample reads the entries
// of rows y and y+1 (two loads, as a 32-byte record of the whole
// neighbourhood would take): each texel is stored twice, not four times, so
// the colour-map workloads keep half the bytes in cache.  (The same layout in
// fp32 made the backward slower: its colour reads became four loads.)
template <int DT> struct RecT;
template <> struct RecT<kF32> {
  // Only the record's address travels with the sample; its colour part is
  // read when the colour is needed (colours()).  Carrying the twelve values
  // instead let the compiler merge the two colour-map sources (this record,
  // or a per-primitive map in global memory) into one load through a
  // selected pointer, which put the record in private (scratch) memory.
  struct Col { const DBMP_GLOBAL char* rec; };
  __device__ static __forceinline__ void get(const void* base, unsigned i, unsigned /*rs*/, float& a,
                                             float& b, float& c, float& d, Col& col) {
    const DBMP_GLOBAL char* q = gbytes(base) + (i << 6);
    const float4 r0 = ld_f4(q);
    a = r0.x; b = r0.y; c = r0.z; d = r0.w;
    col.rec = q;
  }
  // 12 values: RGB of (x,y), (x1,y), (x,y1), (x1,y1)
  __device__ static __forceinline__ void colours(const Col& col, unsigned /*rs*/, float v[12]) {
    const float4 r1 = ld_f4(col.rec + 16), r2 = ld_f4(col.rec + 32), r3 = ld_f4(col.rec + 48);
    v[0] = r1.x; v[1] = r1.y; v[2] = r1.z; v[3] = r1.w;
    v[4] = r2.x; v[5] = r2.y; v[6] = r2.z; v[7] = r2.w;
    v[8] = r3.x; v[9] = r3.y; v[10] = r3.z; v[11] = r3.w;
  }
};
template <int DT>
struct RecHalf {
  struct Col { uint32_t v[6]; };
  __device__ static __forceinline__ float lo(uint32_t v) { return DT == kF16 ? f16_lo(v) : bf16_lo(v); }
  __device__ static __forceinline__ float hi(uint32_t v) { return DT == kF16 ? f16_hi(v) : bf16_hi(v); }
  __device__ static __forceinline__ void get(const void* base, unsigned i, unsigned rs, float& a,
                                             float& b, float& c, float& d, Col& col) {
    const uint4 r0 = ld_u4(gbytes(base) + (i << 4)), r1 = ld_u4(gbytes(base) + ((i + rs) << 4));
    a = lo(r0.x); b = hi(r0.x); c = lo(r1.x); d = hi(r1.x);
    // RGB (x,y), (x1,y) from row y; (x,y1), (x1,y1) from row y + 1
    col.v[0] = r0.y; col.v[1] = r0.z; col.v[2] = r0.w; col.v[3] = r1.y; col.v[4] = r1.z; col.v[5] = r1.w;
  }
  __device__ static __forceinline__ void colours(const Col& col, unsigned /*rs*/, float v[12]) {
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = (k & 1) ? hi(col.v[k >> 1]) : lo(col.v[k >> 1]);
  }
};
template <> struct RecT<kF16> : RecHalf<kF16> {};
template <> struct RecT<kBF16> : RecHalf<kBF16> {};

// sigmoidf_safe (tile_common.h:125-133); evaluated once per primitive
__device__ __forceinline__ float sigmoid_safe(float x) {
  if (x >= 0.f) {
    float z = expf(-x);
    return 1.f / (1.f + z);
  }
  float z = expf(x);
  return z / (1.f + z);
}

// Bilinear value and texel-space gradient from a quad (tile_common.h:201-239)
__device__ __forceinline__ float bilerp(float v00, float v01, float v10, float v11, float wx,
                                        float wy) {
  const float top = fmaf(wx, v01 - v00, v00);
  const float bot = fmaf(wx, v11 - v10, v10);
  return fmaf(wy, bot - top, top);
}

__device__ __forceinline__ float bilerp_grad(float v00, float v01, float v10, float v11, float wx,
                                             float wy, float& gxt, float& gyt) {
  const float dxa = v01 - v00, dxb = v11 - v10;
  const float top = fmaf(wx, dxa, v00);
  const float bot = fmaf(wx, dxb, v10);
  gxt = fmaf(wy, dxb - dxa, dxa);
  gyt = bot - top;
  return fmaf(wy, bot - top, top);
}

// Wave-level helpers (wave64) ---------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// DPP lane moves within 16-lane rows (gfx9 encodings).
constexpr int kDppQuadXor1 = 0xB1;     // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;     // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141;
constexpr int kDppRowRor8 = 0x128;     // row_ror:8 == lane ^ 8 inside a row

template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, true));
}

// Reduce-scatter of eight per-lane values across the wave: afterwards every
// lane of the 8-lane group g = lane >> 3 holds the wave total of value index
// ((lane>>5)&1)*4 + ((lane>>4)&1)*2 + ((lane>>3)&1).  Two halving steps use
// the gfx950 lane-swap instructions (v_permlane32_swap exchanges the upper
// half of one register with the lower half of another, v_permlane16_swap the
// odd rows of one with the even rows of another), so each pair of values
// costs one swap and one add and no LDS traffic; the rest are DPP adds.
__device__ __forceinline__ float wave_reduce_scatter8(const float v[8], int& which) {
  const int l = lane_id();
  float w[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]), __float_as_uint(v[4 + i]),
                                                    false, false);
    w[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
  float x[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(w[i]), __float_as_uint(w[2 + i]),
                                                    false, false);
    x[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
  const bool b3 = (l >> 3) & 1;
  const float s0 = x[0] + dpp<kDppRowRor8>(x[0]);
  const float s1 = x[1] + dpp<kDppRowRor8>(x[1]);
  float y = b3 ? s1 : s0;
  y += dpp<kDppQuadXor1>(y);
  y += dpp<kDppQuadXor2>(y);
  y += dpp<kDppRowHalfMirror>(y);
  which = (((l >> 5) & 1) << 2) | (((l >> 4) & 1) << 1) | (b3 ? 1 : 0);
  return y;
}

}  // namespace dbmp
