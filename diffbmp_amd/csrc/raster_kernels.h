// Raster forward / backward over tile lists.
//
// Work decomposition: one wave64 owns a 16x16 pixel block ("sub-tile") and
// walks the tile list of the binning tile that contains it; each lane owns the
// 4 pixels (x0 + lane%16, y0 + lane/16 + 4j), j = 0..3.  The wave's loop over
// list entries is uniform, so each entry's primitive id and 64-byte PrimRec
// come in through scalar loads and every per-sample operand that depends only
// on the primitive sits in SGPRs.  No LDS is needed for the parameter
// batch and no barrier is ever taken: waves are independent.
//
// Semantics follow the reference kernels (tile_forward.cu:96-398,
// tile_backward.cu:101-452) with two documented changes:
//   * early termination: a pixel stops once T < eps_T (eps_T = 0 disables it);
//     the dropped colour is below eps_T.  The forward records per pixel the
//     number of list entries consumed ("last") and T_final;
//   * the backward never stores per-layer state: it replays the list in
//     forward order, recomputing T_k exactly as the forward did, takes the
//     suffix colour as C_final - C_prefix and the suffix transmittance as
//     T_final / T_{k+1}.
#pragma once
#include "dbmp_device.h"

namespace dbmp {

constexpr int kSub = 16;  // pixels per side of a wave's block

struct RasterArgs {
  const PrimRec* recs;
  const int* toffs;
  const int* tidx;
  const void* quads;
  const void* cmap;      // colour maps (M,Ht,Wt,3) or null
  int cmap_per_prim;     // 1: M = N (indexed by primitive), 0: M = T (by template)
  float c_blend;
  int N, Ht, Wt;
  float Wtm1, Htm1, hsx, hsy;
  int H, W, ts, tilesX, tiles, subs;   // subs = (ts/16)^2 sub-tiles per tile
  float eps;
  float alpha_ub;
};

// blockIdx -> (tile, sub-tile).  Consecutive workgroups are dealt round-robin
// to the 8 XCDs; remap so that each XCD gets a contiguous band of tiles and
// shares its L2 over neighbouring tiles' primitives and texels.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  if ((nblocks & 7) != 0) return b;
  return (b & 7) * (nblocks >> 3) + (b >> 3);
}

template <int DT, bool BLEND>
__device__ __forceinline__ void sample_color(const RasterArgs& A, const PrimRec& p, int n,
                                             const Geo& g, float col[3]) {
  col[0] = p.cr; col[1] = p.cg; col[2] = p.cb;
  if constexpr (BLEND) {
    using CT = typename std::conditional<DT == kF32, float,
               typename std::conditional<DT == kF16, __half, __hip_bfloat16>::type>::type;
    const long m = A.cmap_per_prim ? (long)n : (long)p.tmpl;
    const CT* cm = (const CT*)A.cmap + m * (long)A.Ht * A.Wt * 3;
    float co[3];
    color_map_sample(cm, A.Ht, A.Wt, g.tx, g.ty, co);
    const float k = A.c_blend, k1 = 1.f - A.c_blend;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) col[ch] = k1 * col[ch] + k * co[ch];
  }
}

template <int DT, bool BLEND>
__global__ __launch_bounds__(64) void k_raster_fwd(RasterArgs A, float* __restrict__ out_color,
                                                   float* __restrict__ out_alpha,
                                                   float* __restrict__ out_T,
                                                   int* __restrict__ out_last) {
  using Q = QuadT<DT>;
  const int b = xcd_remap(blockIdx.x, A.tiles * A.subs);
  const int tile = b / A.subs, sub = b - tile * A.subs;
  const int spr = A.ts / kSub;
  const int tx = tile % A.tilesX, ty = tile / A.tilesX;
  const int bx = tx * A.ts + (sub % spr) * kSub, by = ty * A.ts + (sub / spr) * kSub;
  const int lane = lane_id();
  const int px = bx + (lane & 15);
  const int py0 = by + (lane >> 4);
  const float fpx = (float)px;

  float T[4], C[4][3];
  int last[4];
  bool alive[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    alive[j] = (px < A.W) && (py0 + 4 * j < A.H);
    T[j] = 1.f;
    C[j][0] = C[j][1] = C[j][2] = 0.f;
    last[j] = -1;
  }
  const int start = A.toffs[tile], end = A.toffs[tile + 1];
  // conservative (1 px slack) block rectangle for the per-entry cull
  const float rx0 = (float)bx - 1.f, rx1 = (float)(bx + kSub), ry0 = (float)by - 1.f,
              ry1 = (float)(by + kSub);
  const typename Q::type* quads = (const typename Q::type*)A.quads;

  for (int k = start; k < end; ++k) {
    if (!__any(alive[0] | alive[1] | alive[2] | alive[3])) break;
    const int n = A.tidx[k];
    if ((unsigned)n >= (unsigned)A.N) continue;
    const PrimRec p = A.recs[n];
    if (p.tmpl < 0) continue;
    if (p.mx - p.hw > rx1 || p.mx + p.hw < rx0 || p.my - p.hh > ry1 || p.my + p.hh < ry0) continue;
    const long tb = (long)p.tmpl * A.Ht * A.Wt;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!alive[j]) continue;
      Geo g;
      if (!sample_geom(p, fpx, (float)(py0 + 4 * j), A.Wtm1, A.Htm1, A.hsx, A.hsy, A.Wt, A.Ht, g))
        continue;
      float v00, v01, v10, v11;
      Q::get(quads, tb + (long)g.y0 * A.Wt + g.x0, v00, v01, v10, v11);
      float m = bilerp(v00, v01, v10, v11, g.wx, g.wy);
      m = fmaxf(0.f, fminf(1.f, m));
      if (m <= 1e-6f) continue;
      const float a = p.abase * m;
      float col[3];
      sample_color<DT, BLEND>(A, p, n, g, col);
      const float w = T[j] * a;
      C[j][0] = fmaf(w, col[0], C[j][0]);
      C[j][1] = fmaf(w, col[1], C[j][1]);
      C[j][2] = fmaf(w, col[2], C[j][2]);
      T[j] = T[j] * fmaxf(1.f - a, 0.f);
      if (T[j] < A.eps) {
        alive[j] = false;
        last[j] = k - start + 1;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int py = py0 + 4 * j;
    if (px < A.W && py < A.H) {
      const long p = (long)py * A.W + px;
      out_color[3 * p + 0] = C[j][0];
      out_color[3 * p + 1] = C[j][1];
      out_color[3 * p + 2] = C[j][2];
      out_alpha[p] = 1.f - T[j];
      out_T[p] = T[j];
      out_last[p] = last[j] < 0 ? end - start : last[j];
    }
  }
}

// grad8[n*8 + i]: 0 mx, 1 my, 2 r, 3 theta, 4 opacity logit, 5..7 colour logits
template <int DT, bool BLEND>
__global__ __launch_bounds__(64) void k_raster_bwd(RasterArgs A, const float* __restrict__ gCin,
                                                   const float* __restrict__ gAin,
                                                   const float* __restrict__ Cfin,
                                                   const float* __restrict__ Tfin,
                                                   const int* __restrict__ lastin,
                                                   float* __restrict__ grad8) {
  using Q = QuadT<DT>;
  const int b = xcd_remap(blockIdx.x, A.tiles * A.subs);
  const int tile = b / A.subs, sub = b - tile * A.subs;
  const int spr = A.ts / kSub;
  const int tx = tile % A.tilesX, ty = tile / A.tilesX;
  const int bx = tx * A.ts + (sub % spr) * kSub, by = ty * A.ts + (sub / spr) * kSub;
  const int lane = lane_id();
  const int px = bx + (lane & 15);
  const int py0 = by + (lane >> 4);
  const float fpx = (float)px;

  float T[4], gc[4][3], ga[4], cf[4][3], cacc[4][3], tf[4];
  int lst[4];
  int kmax = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int py = py0 + 4 * j;
    const bool ok = (px < A.W) && (py < A.H);
    const long p = ok ? (long)py * A.W + px : 0;
    T[j] = 1.f;
    lst[j] = ok ? lastin[p] : 0;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      gc[j][ch] = ok ? gCin[3 * p + ch] : 0.f;
      cf[j][ch] = ok ? Cfin[3 * p + ch] : 0.f;
      cacc[j][ch] = 0.f;
    }
    ga[j] = (ok && gAin) ? gAin[p] : 0.f;
    tf[j] = ok ? Tfin[p] : 0.f;
    kmax = max(kmax, lst[j]);
  }
  const int start = A.toffs[tile];
  kmax = min(wave_max_i32(kmax), A.toffs[tile + 1] - start);
  const float rx0 = (float)bx - 1.f, rx1 = (float)(bx + kSub), ry0 = (float)by - 1.f,
              ry1 = (float)(by + kSub);
  const typename Q::type* quads = (const typename Q::type*)A.quads;

  for (int kk = 0; kk < kmax; ++kk) {
    const int n = A.tidx[start + kk];
    if ((unsigned)n >= (unsigned)A.N) continue;
    const PrimRec p = A.recs[n];
    if (p.tmpl < 0) continue;
    if (p.mx - p.hw > rx1 || p.mx + p.hw < rx0 || p.my - p.hh > ry1 || p.my + p.hh < ry0) continue;
    const long tb = (long)p.tmpl * A.Ht * A.Wt;
    const float dsc0 = p.cr * (1.f - p.cr), dsc1 = p.cg * (1.f - p.cg), dsc2 = p.cb * (1.f - p.cb);
    const float dav = p.abase * (1.f - p.sop);   // d alpha / d v = alpha_ub s (1-s) per unit mask
    float acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0.f;
    bool any = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (kk >= lst[j]) continue;
      Geo g;
      if (!sample_geom(p, fpx, (float)(py0 + 4 * j), A.Wtm1, A.Htm1, A.hsx, A.hsy, A.Wt, A.Ht, g))
        continue;
      float v00, v01, v10, v11, gxt, gyt;
      Q::get(quads, tb + (long)g.y0 * A.Wt + g.x0, v00, v01, v10, v11);
      float m = bilerp_grad(v00, v01, v10, v11, g.wx, g.wy, gxt, gyt);
      m = fmaxf(0.f, fminf(1.f, m));
      if (m <= 1e-6f) continue;
      any = true;
      const float a = p.abase * m;
      float col[3];
      sample_color<DT, BLEND>(A, p, n, g, col);
      const float Tk = T[j];
      const float w = Tk * a;
      // suffix colour S_k = C_final - sum_{m<=k} T_m a_m c_m; the prefix is
      // recomputed with the forward's exact operations, so S is exactly 0
      // after the last composited layer
      float S[3];
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        cacc[j][ch] = fmaf(w, col[ch], cacc[j][ch]);
        S[ch] = cf[j][ch] - cacc[j][ch];
      }
      const float om = 1.f - a;
      const float inv1m = __builtin_amdgcn_rcpf(fmaxf(om, 1e-6f));
      const float Tn = Tk * fmaxf(om, 0.f);
      // suffix transmittance B_k = T_final / T_{k+1}; 1 if nothing follows
      const float B = Tn > 0.f ? fminf(tf[j] * __builtin_amdgcn_rcpf(Tn), 1.f) : 1.f;
      float dLda = ga[j] * (Tk * B);
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) dLda = fmaf(gc[j][ch], fmaf(Tk, col[ch], -S[ch] * inv1m), dLda);
      // colour logits (tile_backward.cu:255-261)
      acc[5] = fmaf(gc[j][0] * w, dsc0, acc[5]);
      acc[6] = fmaf(gc[j][1] * w, dsc1, acc[6]);
      acc[7] = fmaf(gc[j][2] * w, dsc2, acc[7]);
      // opacity logit (:311-312)
      acc[4] = fmaf(dLda, m * dav, acc[4]);
      // mask path (:314-343)
      const float dLdm = dLda * p.abase;
      const float dLdu = dLdm * (gxt * A.hsx);
      const float dLdv = dLdm * (gyt * A.hsy);
      const float pc = p.c * p.rinv, ps = p.s * p.rinv;
      acc[0] = fmaf(dLdu, -pc, fmaf(dLdv, ps, acc[0]));
      acc[1] = fmaf(dLdu, -ps, fmaf(dLdv, -pc, acc[1]));
      acc[2] = fmaf(-(fmaf(dLdu, g.u, dLdv * g.v)), p.rinv, acc[2]);
      const float dudp = fmaf(-p.s, g.ndx, p.c * g.ndy);
      const float dvdp = fmaf(-p.c, g.ndx, -p.s * g.ndy);
      acc[3] = fmaf(dLdu, dudp, fmaf(dLdv, dvdp, acc[3]));
      T[j] = Tn;
    }
    if (__any(any)) {
      int which;
      const float tot = wave_reduce_scatter8(acc, which);
      if ((lane & 7) == 0) unsafeAtomicAdd(&grad8[(long)n * 8 + which], tot);
    }
  }
}

}  // namespace dbmp
