// Layer export: per-primitive uint8 RGBA crops, the data of a layered image
// file (one layer per primitive).  Restates the reference's two export
// kernels (psd_export_kernels.cu) with the same arithmetic:
//   k_psd_bounds   one thread per primitive: the crop box of a 1.5 r square
//                  around the scaled centre (:12-55).
//   k_psd_offsets  byte offset of each crop in one packed buffer (the host
//                  prefix sum of psd_export_uint8.py:64-75), one workgroup.
//   k_psd_layers   the crops themselves (:61-198): one thread per 4 output
//                  pixels over the packed buffer, so the work follows the
//                  bytes written (the reference loops every primitive over
//                  every 16x16 canvas tile).
#pragma once
#include "bin_kernels.h"
#include "dbmp_device.h"

namespace dbmp {

struct PsdArgs {
  const float* means2D;     // (N,2)
  const float* radii;       // (N)
  const float* rotations;   // (N)
  const float* colors;      // (N,3), sigmoid applied
  const float* visibility;  // (N), sigmoid applied
  const int* bmp_sel;       // (N)
  int N;
  const float* templates;   // (T,Ht,Wt)
  int T, Ht, Wt;
  const float* cmaps;       // colour maps (M,Ht,Wt,3) or null
  int cmap_mode;            // 0 none, 1 per primitive (M = N), 2 per template (M = T)
  float c_blend;
  int H, W;                 // export canvas
  float scale;              // export scale factor
  float alpha_ub;
};

// float -> int as the reference's (int) casts, for finite values in int
// range; anything else is clamped first so the conversion is defined.
__device__ __forceinline__ int psd_to_int(float v) {
  return (int)fminf(fmaxf(v, -1073741824.f), 1073741824.f);
}

// Crop box (left, top, right, bottom) and size (w, h) of primitive n; an
// empty box becomes (0, 0, 1, 1) with size (1, 1).
static __global__ __launch_bounds__(256) void k_psd_bounds(PsdArgs P, int4* __restrict__ boxes,
                                                           int2* __restrict__ sizes) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= P.N) return;
  const float x = P.means2D[2 * n] * P.scale;
  const float y = P.means2D[2 * n + 1] * P.scale;
  const float r = P.radii[n] * P.scale;
  const float tr = r * 1.5f;
  const int x0 = max(0, psd_to_int(floorf(x - tr)));
  const int y0 = max(0, psd_to_int(floorf(y - tr)));
  const int x1 = min(P.W, psd_to_int(ceilf(x + tr)));
  const int y1 = min(P.H, psd_to_int(ceilf(y + tr)));
  if (x1 <= x0 || y1 <= y0) {
    boxes[n] = make_int4(0, 0, 1, 1);
    sizes[n] = make_int2(1, 1);
  } else {
    boxes[n] = make_int4(x0, y0, x1, y1);
    sizes[n] = make_int2(x1 - x0, y1 - y0);
  }
}

// offs[n] = 4 * sum of w*h over layers < n, offs[N] = *total.  One
// 1024-thread workgroup.
static __global__ __launch_bounds__(kScanThreads) void k_psd_offsets(const int2* __restrict__ sizes, int N,
                                                                     long long* __restrict__ offs,
                                                                     long long* __restrict__ total) {
  __shared__ long long sh[kScanThreads / 64];
  long long carry = 0;
  for (int base = 0; base < N; base += kScanThreads) {
    const int i = base + threadIdx.x;
    long long v = 0;
    if (i < N) {
      const int2 s = sizes[i];
      v = 4LL * s.x * s.y;
    }
    long long tot;
    const long long ex = block_excl_scan<kScanThreads / 64>(v, tot, sh);
    if (i < N) offs[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    offs[N] = carry;
    *total = carry;
  }
}

// Template value at canvas pixel (px, py) (psd_grid_sample_bilinear,
// psd_export_common.h:58-108): zero outside [0, W-1) x [0, H-1) in texel
// coordinates, else the bilinear mix written as the sum of four weighted
// texels, left to right.
__device__ __forceinline__ float psd_template(const PsdArgs& P, const float* __restrict__ tm, int px,
                                              int py, float x, float y, float r, float c, float s) {
  const float un = ((float)px - x) / r;
  const float vn = ((float)py - y) / r;
  const float ur = c * un + s * vn;
  const float vr = -s * un + c * vn;
  const float ut = (ur + 1.0f) * 0.5f * (float)(P.Wt - 1);
  const float vt = (vr + 1.0f) * 0.5f * (float)(P.Ht - 1);
  // NaN coordinates (r = 0) count as outside
  if (!(ut >= 0.0f && ut < (float)(P.Wt - 1) && vt >= 0.0f && vt < (float)(P.Ht - 1))) return 0.0f;
  const int u0 = min(max((int)floorf(ut), 0), P.Wt - 1);
  const int v0 = min(max((int)floorf(vt), 0), P.Ht - 1);
  const int u1 = u0 + 1, v1 = v0 + 1;
  const float wu = ut - (float)u0, wv = vt - (float)v0;
  const float wui = 1.0f - wu, wvi = 1.0f - wv;
  const float a00 = tm[v0 * P.Wt + u0], a01 = tm[v0 * P.Wt + u1];
  const float a10 = tm[v1 * P.Wt + u0], a11 = tm[v1 * P.Wt + u1];
  return wui * wvi * a00 + wu * wvi * a01 + wui * wv * a10 + wu * wv * a11;
}

// Clamp-to-edge bilinear colour-map sample at (dx, dy) from the centre in
// unscaled canvas units (psd_sample_color_map, psd_export_common.h:129-196).
__device__ __forceinline__ void psd_cmap(const PsdArgs& P, const float* __restrict__ cm, float dx,
                                         float dy, float radius, float c, float s, float rgb[3]) {
  const float rinv = radius > 1e-6f ? (1.0f / radius) : 1e6f;
  const float nx = dx * rinv, ny = dy * rinv;
  const float rx = c * nx + s * ny;
  const float ry = -s * nx + c * ny;
  const float mw1 = (float)(P.Wt - 1), mh1 = (float)(P.Ht - 1);
  float cx = (rx + 1.0f) * 0.5f * mw1;
  float cy = (ry + 1.0f) * 0.5f * mh1;
  cx = fmaxf(0.0f, fminf(mw1, cx));
  cy = fmaxf(0.0f, fminf(mh1, cy));
  const int x0 = (int)cx, y0 = (int)cy;   // cx, cy >= 0 (NaN clamps to 0 above)
  const int x1 = min(x0 + 1, P.Wt - 1), y1 = min(y0 + 1, P.Ht - 1);
  const float fx = cx - (float)x0, fy = cy - (float)y0;
  const float* c00 = cm + 3 * (y0 * P.Wt + x0);
  const float* c01 = cm + 3 * (y1 * P.Wt + x0);
  const float* c10 = cm + 3 * (y0 * P.Wt + x1);
  const float* c11 = cm + 3 * (y1 * P.Wt + x1);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float a = c00[k] * (1.0f - fx) + c10[k] * fx;
    const float b = c01[k] * (1.0f - fx) + c11[k] * fx;
    rgb[k] = a * (1.0f - fy) + b * fy;
  }
}

__device__ __forceinline__ unsigned psd_u8(float v) {
  return (unsigned)fmaxf(0.0f, fminf(255.0f, v * 255.0f));
}

// Per-layer constants of primitive n.
struct PsdLayer {
  float x, y, r, c, s, vis, cr, cg, cb;
  const float* tm;   // template, or null (no template: transparent)
  const float* cm;   // colour map, or null
  int left, top, w;
};

__device__ __forceinline__ PsdLayer psd_layer(const PsdArgs& P, const int4* __restrict__ boxes, int n,
                                              float c, float s) {
  PsdLayer L;
  const int4 bx = boxes[n];
  L.left = bx.x;
  L.top = bx.y;
  L.w = bx.z - bx.x;
  L.x = P.means2D[2 * n] * P.scale;
  L.y = P.means2D[2 * n + 1] * P.scale;
  L.r = P.radii[n] * P.scale;
  L.c = c;
  L.s = s;
  L.vis = P.visibility[n];
  L.cr = P.colors[3 * n];
  L.cg = P.colors[3 * n + 1];
  L.cb = P.colors[3 * n + 2];
  const int t = P.bmp_sel[n];
  const bool ok = t >= 0 && t < P.T;
  L.tm = ok ? P.templates + (long)t * P.Ht * P.Wt : nullptr;
  L.cm = (ok && P.cmaps != nullptr && P.c_blend > 0.0f)
             ? P.cmaps + (long)(P.cmap_mode == 1 ? n : t) * P.Ht * P.Wt * 3
             : nullptr;
  return L;
}

// RGBA (little-endian R, G, B, A bytes) of canvas pixel (px, py) of a layer.
__device__ __forceinline__ unsigned psd_shade(const PsdArgs& P, const PsdLayer& L, int px, int py) {
  if (L.tm == nullptr) return 0u;
  const float tv = psd_template(P, L.tm, px, py, L.x, L.y, L.r, L.c, L.s);
  const float alpha = tv * L.vis * P.alpha_ub;
  if (alpha < 0.001f) return 0u;
  float rgb[3] = {L.cr, L.cg, L.cb};
  if (L.cm != nullptr) {
    float co[3];
    psd_cmap(P, L.cm, (float)px / P.scale - L.x / P.scale, (float)py / P.scale - L.y / P.scale,
             L.r / P.scale, L.c, L.s, co);
#pragma unroll
    for (int k = 0; k < 3; ++k) rgb[k] = (1.0f - P.c_blend) * rgb[k] + P.c_blend * co[k];
  }
  const float mask = tv > 0.0f ? 1.0f : 0.0f;
  return psd_u8(mask * rgb[0]) | (psd_u8(mask * rgb[1]) << 8) | (psd_u8(mask * rgb[2]) << 16) |
         (psd_u8(alpha) << 24);
}

// Layer of pixel g: the last n in [lo, hi] with offs[n] / 4 <= g.
__device__ __forceinline__ int psd_layer_of(const long long* __restrict__ offs, long long g, int lo,
                                            int hi) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((offs[mid] >> 2) <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

constexpr int kPsdThreads = 256;
constexpr int kPsdPix = 4 * kPsdThreads;   // pixels per workgroup

// Packed crops: thread t writes pixels 4t..4t+3 as one 16-byte store.  The
// workgroup finds the layers its 1024 pixels span once; cos/sin of their
// angles (dbmp_sincosf, fp32) go to LDS, one layer per
// thread, since a workgroup spans at most one layer per pixel.  Optionally
// records each layer's largest alpha byte (the host's "has some alpha" test,
// psd_export_uint8.py:140-143).
static __global__ __launch_bounds__(kPsdThreads) void k_psd_layers(PsdArgs P, const int4* __restrict__ boxes,
                                                                   const long long* __restrict__ offs,
                                                                   long long total_px,
                                                                   unsigned* __restrict__ out,
                                                                   int* __restrict__ alpha_max) {
  __shared__ int span[2];
  __shared__ float scos[kPsdPix], ssin[kPsdPix];
  const long long g0 = (long long)blockIdx.x * kPsdPix;
  const long long gl = min(g0 + kPsdPix, total_px) - 1;
  if (threadIdx.x == 0) span[0] = psd_layer_of(offs, g0, 0, P.N - 1);
  if (threadIdx.x == 64) span[1] = psd_layer_of(offs, gl, 0, P.N - 1);
  __syncthreads();
  const int nlo = span[0], nhi = span[1];
  for (int k = threadIdx.x; k <= nhi - nlo; k += kPsdThreads) {
    dbmp_sincosf(P.rotations[nlo + k], &ssin[k], &scos[k]);
  }
  __syncthreads();
  const long long g = g0 + 4LL * threadIdx.x;
  const bool active = g <= gl;   // no early return: the wave reduction below needs every lane
  unsigned v[4] = {0u, 0u, 0u, 0u};
  int my_n = -1, my_amax = 0;    // layer and alpha maximum for the wave reduction
  if (active) {
    int n = psd_layer_of(offs, g, nlo, nhi);
    const long long p0 = offs[n] >> 2;
    if (g + 3 <= gl && (offs[n + 1] >> 2) > g + 3) {
      // the common case: all four pixels in one layer; one division for the
      // row, then step along it
      const PsdLayer L = psd_layer(P, boxes, n, scos[n - nlo], ssin[n - nlo]);
      const long long loc = g - p0;
      int ly = (int)(loc / L.w);
      int lx = (int)(loc - (long long)ly * L.w);
      int amax = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        v[q] = psd_shade(P, L, L.left + lx, L.top + ly);
        amax = max(amax, (int)(v[q] >> 24));
        if (++lx == L.w) {
          lx = 0;
          ++ly;
        }
      }
      my_n = n;
      my_amax = amax;
    } else {
      // pixels of several layers (crops of a few pixels) or the buffer's end
      for (int q = 0; q < 4; ++q) {
        const long long gq = g + q;
        if (gq > gl) break;
        while (n < nhi && (offs[n + 1] >> 2) <= gq) ++n;
        const PsdLayer L = psd_layer(P, boxes, n, scos[n - nlo], ssin[n - nlo]);
        const long long loc = gq - (offs[n] >> 2);
        const int ly = (int)(loc / L.w);
        const int lx = (int)(loc - (long long)ly * L.w);
        v[q] = psd_shade(P, L, L.left + lx, L.top + ly);
        if (alpha_max != nullptr && (v[q] >> 24) != 0u) atomicMax(&alpha_max[n], (int)(v[q] >> 24));
      }
    }
    if (g + 3 <= gl) {
      *reinterpret_cast<uint4*>(out + g) = make_uint4(v[0], v[1], v[2], v[3]);
    } else {
      for (int q = 0; g + q <= gl; ++q) out[g + q] = v[q];
    }
  }
  if (alpha_max != nullptr) {
    // one atomic per wave when all its fast-path lanes share one layer (the
    // hot case: large crops), else one per lane
    const int n0 = __builtin_amdgcn_readfirstlane(my_n);
    if (n0 >= 0 && __builtin_amdgcn_ballot_w64(my_n >= 0 && my_n != n0) == 0ull) {
      const int wm = wave_max_i32(my_amax);
      if (lane_id() == 0 && wm > 0) atomicMax(&alpha_max[n0], wm);
    } else if (my_amax > 0) {
      atomicMax(&alpha_max[my_n], my_amax);
    }
  }
}

}  // namespace dbmp
