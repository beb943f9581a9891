// Raster forward / backward over tile lists.
//
// Work decomposition: one 256-thread workgroup (4 wave64s) owns a 16x16 pixel
// block ("sub-tile") and walks the tile list of the binning tile containing
// it; wave w owns the 8x8 quadrant (w & 1, w >> 1), one pixel per lane (a
// square footprint meets fewer primitive edges than a 16x4 strip: fewer
// entries per wave pass the cull below).  The list is
// consumed in batches of 64 entries whose 64-byte primitive records are
// staged in LDS, double buffered: while the waves composite batch b, the
// records of batch b+1 are already in flight from L2 into registers.  Each
// wave culls a list entry against its own 8x8 pixel square and stops as
// soon as every lane's transmittance is below eps_T; the workgroup stops when
// all four waves have.
//
// Semantics follow the reference kernels (tile_forward.cu:96-398,
// tile_backward.cu:101-452) with two documented changes:
//   * early termination: a pixel stops once T < eps_T (eps_T = 0 disables it);
//     the dropped colour is below eps_T.  The forward records per pixel the
//     number of list entries consumed ("last") and the final transmittance,
//     kept as the product of the nonzero (1 - a) factors plus a count of
//     exactly opaque layers so that d alpha / d a of an opaque layer (the
//     product of the layers behind it) stays exact;
//   * the backward never stores per-layer state: it replays the list in
//     forward order, recomputing T_k exactly as the forward did, takes the
//     suffix colour as C_final - C_prefix and the suffix transmittance as
//     T_final / T_{k+1}.
#pragma once
#include "bin_kernels.h"
#include "dbmp_device.h"

namespace dbmp {

constexpr int kSub = 16;  // pixels per side of a wave's block

struct RasterArgs {
  const PrimRec* recs;
  const int* toffs;
  const int* tidx;
  const long long* pcount;  // binning pair count (device) or null
  long long pcap;           // capacity the lists were built with
  const void* quads;
  const void* cmap;      // colour maps, or null: no colour-map blend
  const int4* ranges;    // each primitive's binning tile range (tx0, tx1, ty0, ty1), in
                         // tiles of ts >> bin_shift px (16 for super-tile lists, ts for
                         // per-tile lists); an entry is used in a 16x16 block only
                         // inside that range.  Super-tile lists always read it; per-tile
                         // lists only when binning overflowed and every primitive is
                         // walked.  Null: caller-supplied, complete per-tile lists
  int cmap_mode;         // 1: cmap = (N,Ht,Wt,3) by primitive, quads = alpha quads;
                         // 2: quads = RGBA texel records (RecT) of the per-template
                         //    maps (cmap = quads when blending)
  float c_blend;
  int exact_blend;       // 1: gradient of the blended colour w.r.t. logits and geometry;
                         // 0: the reference's (no (1 - c_blend), no colour-map term)
  int N, Ht, Wt;
  float HWm1f;           // Ht*Wt - 1: last texel of one template, as a float
  float Wtf;             // Wt as a float (a scalar operand of the texel-index fma)
  float hsx, hsy;        // (Wt-1)/2, (Ht-1)/2
  int H, W, ts, tilesX, tiles, subs;   // ts: list tile; subs = (ts/16)^2 blocks per tile
  float eps;
  int bin_shift;         // log2(list tile / 16) when ranges is set
  const int* order;      // workgroup -> block, heaviest first (k_block_order), or null
  int* cost;             // forward: each block's walk cost, or null
  const float4* cull;    // per template: content box (k_template_extents<_, true>), or null
  int ntmpl;             // templates in `cull`
  // Push masks, or null: per (list batch, block, wave) one 64-bit word whose
  // bit j is set when entry j of the batch pushed a lane of the wave.  The
  // forward writes them; a backward of that forward (same lists, parameters
  // and per-pixel state) walks exactly those entries instead of culling.
  // Word of batch i of list t: pm_word below.  Unused after a binning
  // overflow (every primitive walked).
  unsigned long long* pmask;
};

// Block a workgroup runs: the heaviest-first order of the last forward's
// costs when there is one, else the XCD-interleaved natural order.
__device__ __forceinline__ int xcd_remap(int b, int nblocks);
__device__ __forceinline__ int block_of(const RasterArgs& A, int wg) {
  return A.order != nullptr ? gptr(A.order)[wg] : xcd_remap(wg, A.tiles * A.subs);
}

// blockIdx -> (tile, sub-tile).  Consecutive workgroups are dealt round-robin
// to the 8 XCDs; remap so that each XCD gets a contiguous band of tiles and
// shares its L2 over neighbouring tiles' primitives and texels.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
  if ((nblocks & 7) != 0) return b;
  return (b & 7) * (nblocks >> 3) + (b >> 3);
}

template <int DT>
using StoreT = typename std::conditional<DT == kF32, float,
               typename std::conditional<DT == kF16, __half, __hip_bfloat16>::type>::type;

// The four clamp-to-edge neighbours of a colour-map sample
// (sample_color_map, tile_forward.cu:17-89) and the bilinear weights.
struct CCorners {
  float fx, fy;
  float c00[3], c10[3], c01[3], c11[3];   // (x0,y0), (x1,y0), (x0,y1), (x1,y1)
};

// Colour corners of a sample at texel coordinates (cx, cy): from the RGBA
// records of the sample's gather (cmap_mode 2; a pushed sample has
// x0 = floor(tx), the records' texel), or gathered from per-primitive maps
// (cmap_mode 1).
template <int DT>
__device__ __forceinline__ CCorners cmap_corners(const RasterArgs& A, const PrimRec& p, int n,
                                                 float cx, float cy,
                                                 const typename RecT<DT>::Col& raw) {
  CCorners k;
  cx = fmaxf(0.f, fminf((float)(A.Wt - 1), cx));
  cy = fmaxf(0.f, fminf((float)(A.Ht - 1), cy));
  const int x0 = (int)cx, y0 = (int)cy;
  k.fx = cx - (float)x0;
  k.fy = cy - (float)y0;
  if (A.cmap_mode == 2) {
    float rc[12];
    RecT<DT>::colours(raw, (unsigned)A.Wt, rc);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      k.c00[ch] = rc[ch];
      k.c10[ch] = rc[3 + ch];
      k.c01[ch] = rc[6 + ch];
      k.c11[ch] = rc[9 + ch];
    }
  } else {
    const StoreT<DT>* cm = (const StoreT<DT>*)A.cmap + (long)n * A.Ht * A.Wt * 3;
    const int x1 = min(x0 + 1, A.Wt - 1), y1 = min(y0 + 1, A.Ht - 1);
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      k.c00[ch] = ldf(cm, 3L * (y0 * A.Wt + x0) + ch);
      k.c01[ch] = ldf(cm, 3L * (y1 * A.Wt + x0) + ch);
      k.c10[ch] = ldf(cm, 3L * (y0 * A.Wt + x1) + ch);
      k.c11[ch] = ldf(cm, 3L * (y1 * A.Wt + x1) + ch);
    }
  }
  return k;
}

// Blended colour (1 - k) sig(c) + k c_o and, with GRAD, dc_o/d(tx, ty).
template <int DT, bool GRAD>
__device__ __forceinline__ void blend_color(const RasterArgs& A, const PrimRec& p, int n, float tx,
                                            float ty, const typename RecT<DT>::Col& raw,
                                            float col[3], float dcx[3], float dcy[3]) {
  const CCorners k = cmap_corners<DT>(A, p, n, tx, ty, raw);
  const float pc[3] = {p.cr, p.cg, p.cb};
  const float kb = A.c_blend, k1 = 1.f - A.c_blend;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float c0 = fmaf(k.fx, k.c10[ch] - k.c00[ch], k.c00[ch]);
    const float c1 = fmaf(k.fx, k.c11[ch] - k.c01[ch], k.c01[ch]);
    const float co = fmaf(k.fy, c1 - c0, c0);
    col[ch] = k1 * pc[ch] + kb * co;
    if constexpr (GRAD) {
      dcx[ch] = fmaf(k.fy, (k.c11[ch] - k.c01[ch]) - (k.c10[ch] - k.c00[ch]), k.c10[ch] - k.c00[ch]);
      dcy[ch] = c1 - c0;
    }
  }
}

// Profiling build (DBMP_LANE_STATS): how many lanes a (wave, list entry)
// evaluation pushes.  Each wave tallies its evaluations in scalar registers
// and adds them to these counters once per block: [0] forward evaluations,
// [1] of them with a pushed lane, [2] pushed lanes; [3..5] the same for the
// backward.  Off in the library proper (the tallies compile to nothing).
#ifdef DBMP_LANE_STATS
static __device__ unsigned long long g_lane_stats[8];
struct LaneTally {
  unsigned long long ev = 0, hit = 0, lanes = 0;
  __device__ __forceinline__ void add(bool pushed) {
    const unsigned long long m = __builtin_amdgcn_ballot_w64(pushed);
    ev += 1;
    hit += m != 0ull ? 1 : 0;
    lanes += __popcll(m);
  }
  __device__ __forceinline__ void flush(int base) const {
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&g_lane_stats[base], ev);
      atomicAdd(&g_lane_stats[base + 1], hit);
      atomicAdd(&g_lane_stats[base + 2], lanes);
    }
  }
};
#else
struct LaneTally {
  __device__ __forceinline__ void add(bool) {}
  __device__ __forceinline__ void flush(int) const {}
};
#endif

constexpr int kBatch = 64;    // list entries staged per batch
constexpr int kThreads = 256; // 4 waves per 16x16 block

struct BlockGeom {
  int tile, sub, bx, by, px, py, w, lane;
  int rtx, rty;   // the block's tile in the units of A.ranges
};

// Geometry of workgroup b (already remapped) of a canvas.
__device__ __forceinline__ BlockGeom block_geom(const RasterArgs& A, int b) {
  BlockGeom g;
  g.tile = b / A.subs;
  g.sub = b - g.tile * A.subs;
  const int spr = A.ts / kSub;
  const int tx = g.tile % A.tilesX, ty = g.tile / A.tilesX;
  g.bx = tx * A.ts + (g.sub % spr) * kSub;
  g.by = ty * A.ts + (g.sub / spr) * kSub;
  // ranges count 16-px tiles for super-tile lists, list tiles otherwise
  g.rtx = A.bin_shift > 0 ? g.bx >> 4 : tx;
  g.rty = A.bin_shift > 0 ? g.by >> 4 : ty;
  // wave-uniform by construction; readfirstlane lets the compiler keep the
  // cull tests and loop counters in scalar registers
  g.w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  g.lane = threadIdx.x & 63;
  g.px = g.bx + 8 * (g.w & 1) + (g.lane & 7);
  g.py = g.by + 8 * (g.w >> 1) + (g.lane >> 3);
  return g;
}

// List range of a tile.  If binning overflowed its capacity the lists were
// not written: walk all primitives in index order (same result, slower).
__device__ __forceinline__ bool list_range(const RasterArgs& A, int tile, int& start, int& end) {
  if (A.pcount != nullptr && *gptr(A.pcount) > A.pcap) {
    start = 0;
    end = A.N;
    return true;
  }
  start = gptr(A.toffs)[tile];
  end = gptr(A.toffs)[tile + 1];
  return false;
}

// Push-mask word of batch i of the block's list (list t = G.tile starting at
// entry `start`) for wave G.w.  List t's batches take words
// floor(start_t / 64) + t + i, i < ceil(len_t / 64); the next list starts at
// floor((start_t + len_t) / 64) + t + 1 or later, so the words of different
// lists never meet, and every word is below floor(P / 64) + lists for P list
// entries in all.  Each word holds the block's sub-blocks and waves.
__device__ __forceinline__ long pm_word(const RasterArgs& A, const BlockGeom& G, int start, int i) {
  return ((long)((start >> 6) + G.tile + i) * A.subs + G.sub) * 4 + G.w;
}

// Push-mask words for lists of at most `capacity` entries in all, `lists`
// lists and `subs` 16x16 blocks per list tile (pm_word stays below
// floor(capacity / 64) + lists per block of a list tile and wave).
inline long long push_mask_words(long long capacity, long long lists, int subs) {
  return ((capacity >> 6) + lists + 1) * (long long)subs * 4;
}

// Load one 16-byte quarter of the record of list entry k (thread t handles
// quarter t&3 of record t>>2); entries past the end get an empty push box.
__device__ __forceinline__ float4 fetch_rec_part(const RasterArgs& A, bool all, int k, int end,
                                                 int part, int& n_out) {
  const int n = (k < end) ? (all ? k : gptr(A.tidx)[k]) : -1;
  const bool ok = (unsigned)n < (unsigned)A.N;
  n_out = ok ? n : -1;
  if (ok) return ld_f4(gbytes(A.recs + n) + 16 * part);
  return part == 0 ? make_float4(0.f, 0.f, -1.f, -1.f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// A wave-uniform float kept in a scalar register.
__device__ __forceinline__ float uniform_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// The wave's pixel square [x0, x0+7] x [y0, y0+7] widened by one pixel and
// to the pixel edges, as the bounds a push box must reach: from the centre
// (cx, cy) = (x0 + 3.5, y0 + 3.5), lo = x0 - 1 and hi = x0 + 8 (exact floats).
struct WaveBox {
  float xlo, xhi, ylo, yhi;
};
__device__ __forceinline__ WaveBox wave_box(float cx, float cy) {
  return WaveBox{uniform_f(cx - 4.5f), uniform_f(cx + 4.5f), uniform_f(cy - 4.5f),
                 uniform_f(cy + 4.5f)};
}

// Whether the push box (quarter 0 of the record: mx, my, hw, hh) can touch
// the wave's pixels.  Conservative w.r.t. the per-pixel fp32 test
// |px - mx| <= hw: the box widened to whole pixels, floor(mx - hw) <= x0 + 7
// and floor(mx + hw) >= x0 - 1, i.e. mx - hw < x0 + 8 and mx + hw >= x0 - 1
// (the same decisions, without float-to-int conversions).  NaN rejects.
__device__ __forceinline__ bool box_hits(const PrimRec& p, const WaveBox& b) {
  return (p.mx - p.hw < b.xhi) & (p.mx + p.hw >= b.xlo) & (p.my - p.hh < b.yhi) &
         (p.my + p.hh >= b.ylo);
}

// Whether the content of a staged entry's template can reach the wave's
// 8x8 pixels centred on (cx, cy): separating-axis test, on the axes of the
// primitive's rotated content box (its template's texels above the mask
// threshold, one texel wider: outside it no sample is pushed), of that box
// against the wave's pixel rectangle, with one pixel of slack.  The push box
// (box_hits) covers the canvas axes.  An entry it rejects has no pushed
// sample in the wave, so skipping it leaves every result unchanged.
__device__ __forceinline__ bool content_hits(const RasterArgs& A, const PrimRec& p, float cx,
                                             float cy) {
  const int t = p.tmpl;
  const float4 b = ld_f4(gbytes(A.cull + min(max(t, 0), A.ntmpl - 1)));
  const float ac = fabsf(p.c), as = fabsf(p.s);
  const float dx = cx - p.mx, dy = cy - p.my;
  const float pu = fmaf(p.c, dx, p.s * dy) - p.r * b.x;
  const float pv = fmaf(p.c, dy, -p.s * dx) - p.r * b.y;
  const float k = fmaf(ac + as, 3.5f, 1.f);   // the square's projection (half side 3.5) + slack
  const float eu = fmaf(p.r, b.z, k);
  const float ev = fmaf(p.r, b.w, k);
  return ((unsigned)t < (unsigned)A.ntmpl) & (fabsf(pu) <= eu) & (fabsf(pv) <= ev);
}

// Entries of staged batch `buf` that can touch this wave's pixels: lane i
// tests entry i (in the list tile's range, push box, content box).
__device__ __forceinline__ bool entry_hits_wave(const RasterArgs& A, const PrimRec* srec,
                                                const int* sin, int lane, const WaveBox& wb,
                                                float cx, float cy) {
  const PrimRec& p = srec[lane];
  bool in = (sin[lane] != 0) & box_hits(p, wb);
  if (A.cull != nullptr) in = in & content_hits(A, p, cx, cy);
  return in;
}

// Clear / set bit j of a wave-uniform 64-bit mask in one scalar instruction
// (wm & (wm - 1) takes three).  The raster loops are bound by the total
// instruction count per list entry, scalar instructions included.
__device__ __forceinline__ void sbit_clear(unsigned long long& m, int j) {
  asm("s_bitset0_b64 %0, %1" : "+s"(m) : "s"(j));
}
__device__ __forceinline__ void sbit_set(unsigned long long& m, int j) {
  asm("s_bitset1_b64 %0, %1" : "+s"(m) : "s"(j));
}

// Lane-mask vote on a bool: the ballot stays a scalar mask (no per-lane 0/1
// value is materialised first, as an int-predicate vote does).
__device__ __forceinline__ unsigned long long wave_ballot(bool b) {
  return __builtin_amdgcn_ballot_w64(b);
}
__device__ __forceinline__ bool wave_any(bool b) { return wave_ballot(b) != 0ull; }

// Stage batch entry ri (one quarter per thread) into LDS slot `buf`.
struct Stage {
  float4 v;
  int n;
  int4 rg;   // 16-px tile range (super-tile lists)
};
__device__ __forceinline__ Stage stage_load(const RasterArgs& A, bool all, int k, int end, int part) {
  Stage st;
  st.v = fetch_rec_part(A, all, k, end, part, st.n);
  st.rg = make_int4(0, 0x7fffffff, 0, 0x7fffffff);
  // complete per-tile lists hold an entry only in tiles of its range, so the
  // range is read where it decides: super-tile lists, and the walk over every
  // primitive after a binning overflow
  if (A.ranges != nullptr && part == 0 && st.n >= 0 && (all || A.bin_shift > 0))
    st.rg = ld_i4(gbytes(A.ranges + st.n));
  return st;
}
__device__ __forceinline__ void stage_store(const Stage& st, PrimRec (*srec)[kBatch], int (*sid)[kBatch],
                                            int (*smask)[kBatch], int buf, int ri, int part,
                                            const BlockGeom& G) {
  reinterpret_cast<float4*>(&srec[buf][ri])[part] = st.v;
  if (part == 0) {
    sid[buf][ri] = st.n;
    // the entry belongs to this block's tile's list only inside the
    // primitive's binning range (the reference's per-tile lists,
    // vector_renderer.py:401-444: a texel fringe outside it is not drawn)
    const int tx = G.rtx, ty = G.rty;
    const bool in = (tx >= st.rg.x) & (tx <= st.rg.y) & (ty >= st.rg.z) & (ty <= st.rg.w);
    smask[buf][ri] = (in & (st.n >= 0)) ? 1 : 0;
  }
}

// One sample, branch-free: geometry with the fp32 decisions of the oracle
// (tile_forward.cu:187-267), one quad gather (a rejected lane gathers quad 0), bilinear
// mask.  `pushed` is the reference's push condition (box, texture bounds,
// mask > 1e-6) and the caller's activity flag.
template <int DT>
struct Sample {
  float m, gxt, gyt, u, v, ndx, ndy, tx, ty;
  bool pushed;
  typename RecT<DT>::Col craw;   // colour part of an RGBA record (BLEND, cmap_mode 2)
};

template <int DT, bool GRAD, bool BLEND>
__device__ __forceinline__ Sample<DT> eval_sample(const RasterArgs& A, const PrimRec& p, float fpx,
                                                  float fpy, bool active) {
  using Q = QuadT<DT>;
  Sample<DT> r;
  const float dx = fpx - p.mx, dy = fpy - p.my;
  // predicates combined with & (no short-circuit branches)
  const bool inbox = (fabsf(dx) <= p.hw) & (fabsf(dy) <= p.hh);
  r.ndx = dx * p.rinv;
  r.ndy = dy * p.rinv;
  // u = c ndx + s ndy and v = -s ndx + c ndy with the first product fused
  // into the add, as nvcc's default contraction (-fmad=true) compiles the
  // reference's expressions (tile_forward.cu:259-260); the oracle uses the
  // same fused form
  r.u = fmaf(p.c, r.ndx, p.s * r.ndy);
  r.v = fmaf(-p.s, r.ndx, p.c * r.ndy);
  r.tx = (r.u + 1.f) * A.hsx;
  r.ty = (r.v + 1.f) * A.hsy;
  const bool inb = (r.tx >= 0.f) & (r.tx <= (float)(A.Wt - 1)) & (r.ty >= 0.f) &
                   (r.ty <= (float)(A.Ht - 1));
  const bool ok = active & inbox & inb;
  const float fx = floorf(r.tx), fy = floorf(r.ty);
  // clamped rather than selected, so the gather is unconditional; a lane
  // that is not `ok` reads some valid quad and its sample is discarded
  // texel index within the template in float (exact below 2^24), clamped so
  // that NaN or out-of-range coordinates still address a texel of the template
  const float fl = __builtin_amdgcn_fmed3f(fmaf(fy, A.Wtf, fx), 0.f, A.HWm1f);
  // with colour maps (small, minified primitives are typical there) a lane
  // that cannot push reads the template's first texel, so that idle lanes do
  // not spread cache lines over the bitmap
  const unsigned idx = (unsigned)(p.tbase + ((!BLEND || ok) ? (int)fl : 0));
  float v00, v01, v10, v11;
  if (BLEND && A.cmap_mode == 2)
    RecT<DT>::get(A.quads, idx, (unsigned)A.Wt, v00, v01, v10, v11, r.craw);
  else
    Q::get((const typename Q::type*)A.quads, idx, v00, v01, v10, v11);
  const float wx = r.tx - fx, wy = r.ty - fy;
  float m;
  if constexpr (GRAD) {
    m = bilerp_grad(v00, v01, v10, v11, wx, wy, r.gxt, r.gyt);
  } else {
    m = bilerp(v00, v01, v10, v11, wx, wy);
    r.gxt = r.gyt = 0.f;
  }
  r.m = __builtin_amdgcn_fmed3f(m, 0.f, 1.f);   // clamp to [0, 1]
  r.pushed = ok & (r.m > 1e-6f);
  return r;
}

// Forward of one 16x16 block b of a canvas; the caller's workgroup runs it.
template <int DT, bool BLEND>
__device__ __forceinline__ void raster_fwd_block(const RasterArgs& A, int b,
                                                 float* __restrict__ out_color,
                                                 float* __restrict__ out_alpha,
                                                 float* __restrict__ out_T,
                                                 int* __restrict__ out_last) {
  __shared__ PrimRec srec[2][kBatch];
  __shared__ int sid[2][kBatch];
  __shared__ int smask[2][kBatch];   // 1: the entry is in the list tile's range
  __shared__ __align__(16) int salive[2][4];
  const BlockGeom G = block_geom(A, b);
  if (G.bx >= A.W || G.by >= A.H) {   // 16-px block of a tile past the canvas edge
    if (A.cost != nullptr && threadIdx.x == 0) gptr(A.cost)[b] = 0;
    return;
  }
  const bool inside = (G.px < A.W) && (G.py < A.H);
  // A lane that is not compositing (past the canvas edge, or stopped) has a
  // NaN pixel coordinate: its push-box test is false, so it pushes nothing,
  // and its liveness is `thr >= 0` -- no lane mask to carry per entry.
  float fpx = inside ? (float)G.px : __builtin_nanf(""), fpy = (float)G.py;
  // Tnz: product of the nonzero (1 - a) factors; nz: count of zero factors
  // (exactly opaque layers), capped at 2.  Until a lane meets an exactly
  // opaque layer Tnz is its transmittance T; from then on T is 0, the lane's
  // colour is the snapshot Cf taken at that layer, and the C0..C2 it keeps
  // accumulating are dropped.  One running product instead of two.
  float Tnz = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f, Cf0 = 0.f, Cf1 = 0.f, Cf2 = 0.f;
  int nz = 0;
  // stop test threshold: eps_T on a live lane, +inf once it has seen two
  // opaque layers, -1 on a stopped lane (or one past the canvas edge)
  float thr = inside ? A.eps : -1.f;
  int last = -1;
  int walked = 0;   // entries this wave evaluated (wave-uniform)
  LaneTally tally;
  int batches = 0;
  int start, end;
  const bool all = list_range(A, G.tile, start, end);
  DBMP_GLOBAL unsigned long long* const pm = (A.pmask != nullptr && !all) ? gptr(A.pmask) : nullptr;
  const int ri = threadIdx.x >> 2, part = threadIdx.x & 3;
  const float ccx = uniform_f((float)(G.bx + 8 * (G.w & 1)) + 3.5f);   // centre of the wave's pixels
  const float ccy = uniform_f((float)(G.by + 8 * (G.w >> 1)) + 3.5f);
  const WaveBox wb = wave_box(ccx, ccy);
  stage_store(stage_load(A, all, start + ri, end, part), srec, sid, smask, 0, ri, part, G);
  __syncthreads();
  int buf = 0;
  for (int base = start; base < end; base += kBatch, buf ^= 1) {
    const bool have_next = base + kBatch < end;
    ++batches;
    // entries of this batch that can touch this wave's pixels (before the
    // next batch's loads are in flight: fewer live registers)
    const bool any_alive = wave_any(thr >= 0.f);
    unsigned long long wm = 0ull;
    if (any_alive)
      wm = wave_ballot(entry_hits_wave(A, srec[buf], smask[buf], G.lane, wb, ccx, ccy));
    Stage nxt;
    if (have_next) nxt = stage_load(A, all, base + kBatch + ri, end, part);
    unsigned long long hits = 0ull;   // entries of this batch that pushed a lane of the wave
    if (any_alive) {
      walked += __popcll(wm);   // less the entries left when the wave stops (below)
      while (wm) {
        const int j = __builtin_ctzll(wm);
        sbit_clear(wm, j);
        const PrimRec p = srec[buf][j];
        const Sample<DT> S = eval_sample<DT, false, BLEND>(A, p, fpx, fpy, true);
        tally.add(S.pushed);
        if constexpr (BLEND) {
          // skip the colour-map gathers when nothing of this entry reaches the wave
          if (!wave_any(S.pushed)) continue;
          sbit_set(hits, j);
        }
        float col[3];
        if (BLEND && A.cmap != nullptr) {
          blend_color<DT, false>(A, p, sid[buf][j], S.tx, S.ty, S.craw, col, nullptr, nullptr);
        } else {
          // a batch mixing canvases with and without colour maps runs BLEND
          // kernels; a canvas without one keeps its plain colours
          col[0] = p.cr; col[1] = p.cg; col[2] = p.cb;
        }
        const float a = S.pushed ? p.abase * S.m : 0.f;   // a = 0: compositing is a no-op
        if constexpr (!BLEND) {
          // a > 0 on every pushed lane except where alpha_ub sigma(v) mask is 0,
          // whose gradients are 0 too (a ballot of `pushed` itself costs two
          // VALU instructions here: the compiler materialises the predicate)
          if (wave_any(a > 0.f)) sbit_set(hits, j);
        }
        const float wgt = Tnz * a;
        C0 = fmaf(wgt, col[0], C0);
        C1 = fmaf(wgt, col[1], C1);
        C2 = fmaf(wgt, col[2], C2);
        // max(1 - a, 0) as one clamped subtraction (a >= 0: alpha_ub >= 0)
        float f = fminf(fmaxf(1.f - a, 0.f), 1.f);
        // a pixel stops once later layers can change neither its colour nor
        // its alpha (or alpha's derivative w.r.t. an opaque layer) by eps_T
        // (Tnz < thr; a live lane that is not pushed keeps Tnz >= eps_T,
        // since eps_T < 1, so the test needs no `pushed`)
        // nz counts exactly opaque layers (capped at 2), rare enough for a
        // wave-uniform branch, which also keeps their zero factor out of Tnz
        const bool opq = f == 0.f;
        if (wave_any(opq)) {
          __asm__ volatile("");   // keeps this rare block a branch, not selects
          const bool first = opq & (nz == 0);
          Cf0 = first ? C0 : Cf0;
          Cf1 = first ? C1 : Cf1;
          Cf2 = first ? C2 : Cf2;
          nz = opq ? min(nz + 1, 2) : nz;
          thr = ((thr >= 0.f) & (nz >= 2)) ? __builtin_inff() : thr;
          f = opq ? 1.f : f;
        }
        Tnz *= f;
        const bool dead = Tnz < thr;
        if (wave_any(dead)) {
          last = dead ? base - start + j + 1 : last;
          thr = dead ? -1.f : thr;
          fpx = dead ? __builtin_nanf("") : fpx;
          if (!wave_any(thr >= 0.f)) {   // the wave has stopped: leave the batch
            walked -= __popcll(wm);
            wm = 0ull;
          }
        }
      }
    }
    if (pm != nullptr && G.lane == 0) pm[pm_word(A, G, start, (base - start) >> 6)] = hits;
    if (have_next) stage_store(nxt, srec, sid, smask, buf ^ 1, ri, part, G);
    // one barrier: it publishes the staged batch and each wave's liveness
    // (one flag per wave and buffer parity, so a flag is rewritten only after
    // every wave has read it; __syncthreads_or takes three barriers)
    const int wave_alive = wave_any(thr >= 0.f) ? 1 : 0;
    if (G.lane == 0) salive[buf][G.w] = wave_alive;
    __syncthreads();
    const int4 fl = *reinterpret_cast<const int4*>(salive[buf]);
    if ((fl.x | fl.y | fl.z | fl.w) == 0) break;
  }
  tally.flush(0);
  if (G.px < A.W && G.py < A.H) {
    const long p = (long)G.py * A.W + G.px;
    gptr(out_color)[3 * p + 0] = nz ? Cf0 : C0;
    gptr(out_color)[3 * p + 1] = nz ? Cf1 : C1;
    gptr(out_color)[3 * p + 2] = nz ? Cf2 : C2;
    gptr(out_alpha)[p] = 1.f - (nz ? 0.f : Tnz);
    gptr(out_T)[p] = Tnz;
    gptr(out_last)[p] = (last < 0 ? end - start : last) | (nz << 30);
  }
  if (A.cost != nullptr) {
    // block cost: the slowest wave's entry count plus the staging of its batches
    __syncthreads();
    if (G.lane == 0) smask[0][G.w] = walked + 8 * batches;
    __syncthreads();
    if (threadIdx.x == 0)
      gptr(A.cost)[b] = max(max(smask[0][0], smask[0][1]), max(smask[0][2], smask[0][3]));
  }
}

// Without colour blending the forward fits 64 VGPRs: 8 waves per SIMD, the
// most a SIMD holds (the attribute keeps it there); with it, 80 VGPRs and 6
// waves (its natural allocation, a few registers above, would give 5).
template <int DT, bool BLEND>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(BLEND ? 6 : 8))) void k_raster_fwd(RasterArgs A, float* __restrict__ out_color,
                                                         float* __restrict__ out_alpha,
                                                         float* __restrict__ out_T,
                                                         int* __restrict__ out_last) {
  raster_fwd_block<DT, BLEND>(A, block_of(A, blockIdx.x), out_color, out_alpha,
                              out_T, out_last);
}

// Background composite + MSE, fused in front of the backward:
//   result = C + (1 - A) bg,  r = result - target,  loss += sum(r^2) / (3 H W)
//   dL/dC = 2 r / (3 H W),    dL/dA = -sum_c dL/dC_c bg_c
// (the same arithmetic as k_loss_mse).
struct LossArgs {
  const float* target;   // (H,W,3)
  const float* bg;       // (H,W,3) or null for the constant bg_value
  float bg_value;
  float inv_count;       // 1 / (3 H W)
  float* loss;           // accumulated (+=)
};

// grad8[n*8 + i]: 0 mx, 1 my, 2 r, 3 theta, 4 opacity logit, 5..7 colour logits
// Backward of one 16x16 block b.  With LOSS the pixel gradients come from the
// fused background composite + MSE (L; gCin / gAin unused), else from gCin /
// gAin (L unused).
// The backward's LDS: staged records (double buffered), the four waves'
// gradient partials of a batch, the block's furthest forward position.
struct BwdLds {
  PrimRec srec[2][kBatch];
  int sid[2][kBatch];
  int smask[2][kBatch];   // 1: the entry is in the list tile's range
  float sred[4][kBatch][8];
  int skmax;
};

// MASKS: compiled for lists that carry push masks (A.pmask): each batch's
// entries come from the forward's mask, every one of them pushed a lane of
// the wave, so there is no cull and no per-entry push test (after a binning
// overflow, which writes no masks, the entries are culled).  Without MASKS
// the batch is culled and entries that push no lane are skipped.
template <int DT, bool BLEND, bool LOSS, bool MASKS>
__device__ __forceinline__ void raster_bwd_body(BwdLds& lds, const RasterArgs& A, int b,
                                                const float* __restrict__ gCin,
                                                const float* __restrict__ gAin,
                                                const LossArgs& L,
                                                const float* __restrict__ Cfin,
                                                const float* __restrict__ Tfin,
                                                const int* __restrict__ lastin,
                                                float* __restrict__ grad8) {
  PrimRec (*srec)[kBatch] = lds.srec;
  int (*sid)[kBatch] = lds.sid;
  int (*smask)[kBatch] = lds.smask;
  float (*sred)[kBatch][8] = lds.sred;
  int& skmax = lds.skmax;
  const BlockGeom G = block_geom(A, b);
  if (G.bx >= A.W || G.by >= A.H) return;   // 16-px block of a tile past the canvas edge
  const float fpx = (float)G.px, fpy = (float)G.py;
  const bool okpix = (G.px < A.W) && (G.py < A.H);
  const long pix = okpix ? (long)G.py * A.W + G.px : 0;
  const int lstw = okpix ? gptr(lastin)[pix] : 0;
  const int lst = lstw & 0x3fffffff;   // list entries consumed by the forward
  const int nzf = (unsigned)lstw >> 30; // exactly opaque layers seen (capped at 2)
  const float cf0 = okpix ? gptr(Cfin)[3 * pix] : 0.f, cf1 = okpix ? gptr(Cfin)[3 * pix + 1] : 0.f,
              cf2 = okpix ? gptr(Cfin)[3 * pix + 2] : 0.f;
  const float tnzf = okpix ? gptr(Tfin)[pix] : 0.f;   // product of the nonzero (1 - a) factors
  const float tf = nzf ? 0.f : tnzf;            // transmittance after the last layer
  float gc0, gc1, gc2, ga;
  float acc_loss = 0.f;   // this lane's squared error (LOSS)
  if constexpr (LOSS) {
    // the forward's alpha is 1 - T_final exactly
    const float om = 1.f - (1.f - tf);
    const float cf[3] = {cf0, cf1, cf2};
    float g[3];
    ga = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float bgc = L.bg ? (okpix ? gptr(L.bg)[3 * pix + c] : 0.f) : L.bg_value;
      const float r = okpix ? fmaf(om, bgc, cf[c]) - gptr(L.target)[3 * pix + c] : 0.f;
      acc_loss = fmaf(r, r, acc_loss);
      g[c] = 2.f * r * L.inv_count;
      ga = fmaf(-g[c], bgc, ga);
    }
    gc0 = g[0]; gc1 = g[1]; gc2 = g[2];
  } else {
    gc0 = okpix ? gptr(gCin)[3 * pix] : 0.f;
    gc1 = okpix ? gptr(gCin)[3 * pix + 1] : 0.f;
    gc2 = okpix ? gptr(gCin)[3 * pix + 2] : 0.f;
    ga = (okpix && gAin) ? gptr(gAin)[pix] : 0.f;
  }
  float T = 1.f, ca0 = 0.f, ca1 = 0.f, ca2 = 0.f;
  const float gcf = fmaf(gc2, cf2, fmaf(gc1, cf1, gc0 * cf0));   // gc . C_final
  const float hyx = A.hsy / A.hsx;
  int start, end;
  const bool all = list_range(A, G.tile, start, end);
  const int wkmax = __builtin_amdgcn_readfirstlane(wave_max_i32(lst));
  if (threadIdx.x == 0) skmax = 0;
  for (int i = threadIdx.x; i < 4 * kBatch * 8; i += kThreads) (&sred[0][0][0])[i] = 0.f;
  __syncthreads();
  if (G.lane == 0) atomicMax(&skmax, wkmax);
  const int ri = threadIdx.x >> 2, part = threadIdx.x & 3;
  const float ccx = uniform_f((float)(G.bx + 8 * (G.w & 1)) + 3.5f);   // centre of the wave's pixels
  const float ccy = uniform_f((float)(G.by + 8 * (G.w >> 1)) + 3.5f);
  const WaveBox wb = wave_box(ccx, ccy);
  stage_store(stage_load(A, all, start + ri, end, part), srec, sid, smask, 0, ri, part, G);
  // the forward's push masks (one word per batch, wave-uniform), loaded a
  // batch ahead
  const DBMP_GLOBAL unsigned long long* const pm =
      (MASKS && A.pmask != nullptr && !all) ? gptr((const unsigned long long*)A.pmask) : nullptr;
  unsigned long long pm_cur = pm != nullptr ? pm[pm_word(A, G, start, 0)] : 0ull;
  __syncthreads();
  const int stop = start + min(skmax, end - start);
  LaneTally tally;
  int buf = 0;
  for (int base = start; base < stop; base += kBatch, buf ^= 1) {
    const bool have_next = base + kBatch < stop;
    const int cnt = __builtin_amdgcn_readfirstlane(min(kBatch, stop - base));
    // entries of this batch that pushed a lane of this wave in the forward,
    // or (no masks) that can touch this wave's pixels, up to the wave's
    // furthest forward position (before the next batch's loads are in
    // flight: fewer live registers)
    unsigned long long wm;
    if (pm != nullptr)
      wm = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(pm_cur >> 32)) << 32) |
           (unsigned)__builtin_amdgcn_readfirstlane((unsigned)pm_cur);
    else
      wm = wave_ballot(entry_hits_wave(A, srec[buf], smask[buf], G.lane, wb, ccx, ccy));
    const int lim = wkmax - (base - start);
    wm = lim >= 64 ? wm : (lim <= 0 ? 0ull : wm & ((1ull << lim) - 1ull));
    Stage nxt;
    if (have_next) {
      nxt = stage_load(A, all, base + kBatch + ri, end, part);
      if (pm != nullptr) pm_cur = pm[pm_word(A, G, start, ((base - start) >> 6) + 1)];
    }
    while (wm) {
      const int j = __builtin_ctzll(wm);
      sbit_clear(wm, j);
      const int kk = base - start + j;
      const PrimRec p = srec[buf][j];
      const Sample<DT> S = eval_sample<DT, true, BLEND>(A, p, fpx, fpy, kk < lst);
      tally.add(S.pushed);
      // every entry of a push mask pushed a lane; without masks skip the
      // culled entries that push none (an entry with no pushed lane adds 0
      // to every gradient either way)
      if (!MASKS && !wave_any(S.pushed)) continue;
      float col[3], dcx[3] = {0.f, 0.f, 0.f}, dcy[3] = {0.f, 0.f, 0.f};
      const bool has_map = BLEND && A.cmap != nullptr;
      if (has_map) {
        blend_color<DT, true>(A, p, sid[buf][j], S.tx, S.ty, S.craw, col, dcx, dcy);
      } else {
        col[0] = p.cr; col[1] = p.cg; col[2] = p.cb;
      }
      const float a = S.pushed ? p.abase * S.m : 0.f;
      const float Tk = T;
      const float wgt = Tk * a;
      // colour prefix C_k = sum_{m<=k} T_m a_m c_m with the forward's exact
      // operations, so the suffix S_k = C_final - C_k is 0 at the end; the
      // loss only needs its projection on gc: gc.S = gc.C_final - gc.C_k,
      // both dot products formed in the same order (equal when C_k = C_final)
      ca0 = fmaf(wgt, col[0], ca0);
      ca1 = fmaf(wgt, col[1], ca1);
      ca2 = fmaf(wgt, col[2], ca2);
      const float gS = gcf - fmaf(gc2, ca2, fmaf(gc1, ca1, gc0 * ca0));
      const float gcol = fmaf(gc2, col[2], fmaf(gc1, col[1], gc0 * col[0]));
      // fk = max(1 - a, 0) as one clamped subtraction (a >= 0: alpha_ub >= 0);
      // 1 / max(1 - a, 1e-6) = 1 / max(fk, 1e-6)
      const float fk = fminf(fmaxf(1.f - a, 0.f), 1.f);
      const float inv1m = __builtin_amdgcn_rcpf(fmaxf(fk, 1e-6f));
      const float Tn = Tk * fk;
      // T_k B_k, with the suffix transmittance B_k = prod_{m>k} (1 - a_m) =
      // T_final / T_{k+1}: T_k B_k = T_final / (1 - a_k) while 1 - a_k >= 1e-6
      // (it is 0 = T_final once T is 0).  Nearer-opaque layers, rare, take
      // the general form: for an exactly opaque layer B_k is the
      // nonzero-factor product after it, tnzf / T_k, if it is the only opaque
      // layer, else 0
      float TkB = tf * inv1m;
      if (wave_any(fk < 1e-6f)) {
        const bool tpos = Tn > 0.f;
        const float rB = fminf((tpos ? tf : tnzf) * __builtin_amdgcn_rcpf(tpos ? Tn : Tk), 1.f);
        const float g = Tk * (tpos ? rB : (fk == 0.f ? (nzf == 1 ? rB : 0.f) : 1.f));
        TkB = fk < 1e-6f ? g : TkB;
      }
      const float dLda = fmaf(ga, TkB, fmaf(Tk, gcol, -gS * inv1m));
      // Per-sample parts of the 8 gradients.  Factors that are constant per
      // primitive (-1/r, dalpha/dv, sigma'(c), the blend weight) are applied
      // once per (workgroup, entry) when the wave partials are summed below.
      float acc[8];
      const bool exact = has_map && A.exact_blend;
      // colour logits (tile_backward.cu:255-261): dC_c/dcol = T a; wgt = 0 for
      // lanes not pushed
      acc[5] = gc0 * wgt;
      acc[6] = gc1 * wgt;
      acc[7] = gc2 * wgt;
      // opacity logit (:311-312): alpha_ub * mask * s (1 - s); mask path
      // (:314-343); both zero on lanes not pushed
      const float dLdap = S.pushed ? dLda : 0.f;
      acc[4] = dLdap * S.m;
      // dL/dmask = dL/da alpha_ub sigma(v).  Without colour maps the factor
      // alpha_ub sigma(v) is constant per primitive and applied once per
      // entry below; the exact colour-map term added to dL/du, dL/dv is not
      // proportional to it, so the blending kernels apply it per lane
      const float dLdm = BLEND ? dLdap * p.abase : dLdap;
      // dL/du and dL/dv divided by hsx = dtx/du (hsx is applied with the
      // per-primitive factors); hyx = hsy / hsx
      float dLdu = dLdm * S.gxt;
      float dLdv = dLdm * (S.gyt * hyx);
      if constexpr (BLEND) {
        if (exact) {
          // colour-map term: dC/dcol_k = T_k a_k, dcol/d(tx,ty) = c_blend dc_o/d(tx,ty)
          const float kw = wgt * A.c_blend;
          dLdu = fmaf(kw, fmaf(gc0, dcx[0], fmaf(gc1, dcx[1], gc2 * dcx[2])), dLdu);
          dLdv = fmaf(kw * hyx, fmaf(gc0, dcy[0], fmaf(gc1, dcy[1], gc2 * dcy[2])), dLdv);
        }
      }
      // (u, v) = R(theta) (dx, dy) / r.  With (gx, gy) = R^T (dL/du, dL/dv):
      //   dL/dmx = -gx/r, dL/dmy = -gy/r, dL/dr = -(ndx gx + ndy gy)/r,
      //   dL/dtheta = ndy gx - ndx gy,
      // where, R being a rotation, ndx gx + ndy gy = u dL/du + v dL/dv and
      // ndy gx - ndx gy = v dL/du - u dL/dv.  The lanes sum dL/du, dL/dv and
      // those two; R^T (constant per primitive) is applied to the sums once
      // per entry below
      acc[0] = dLdu;
      acc[1] = dLdv;
      acc[2] = fmaf(S.u, dLdu, S.v * dLdv);
      acc[3] = fmaf(S.v, dLdu, -S.u * dLdv);
      T = Tn;
      int which;
      const float tot = wave_reduce_scatter8(acc, which);
      // the eight lanes of a group hold the same total: all of them write it
      // (masking the write to one lane costs more instructions than the
      // repeated LDS stores)
      sred[G.w][j][which] = tot;
    }
    if (have_next) stage_store(nxt, srec, sid, smask, buf ^ 1, ri, part, G);
    __syncthreads();
    // one atomic per (tile, entry, value): sum the four waves' partials and
    // apply the per-primitive factor of value q
    const float cwf = (BLEND && A.exact_blend && A.cmap != nullptr) ? 1.f - A.c_blend : 1.f;
    for (int i = threadIdx.x; i < cnt * 8; i += kThreads) {
      const int j = i >> 3, q = i & 7;
      const PrimRec& pr = srec[buf][j];
      const float cc = q == 5 ? pr.cr : (q == 6 ? pr.cg : pr.cb);
      const float fa = BLEND ? 1.f : pr.abase;   // alpha_ub sigma(v) of the mask path
      const float fac = q < 3 ? -pr.rinv * A.hsx * fa
                              : (q == 3 ? A.hsx * fa : (q == 4 ? pr.dav : cc * (1.f - cc) * cwf));
      float s = sred[0][j][q] + sred[1][j][q] + sred[2][j][q] + sred[3][j][q];
      // (gx, gy) = R^T (sum dL/du, sum dL/dv): values 0 and 1 of an entry sit
      // in neighbouring lanes, active together (cnt * 8 items in all)
      const float so = __shfl_xor(s, 1, 64);
      s = q == 0 ? fmaf(pr.c, s, -pr.s * so) : (q == 1 ? fmaf(pr.s, so, pr.c * s) : s);
      s *= fac;
      sred[0][j][q] = 0.f;
      sred[1][j][q] = 0.f;
      sred[2][j][q] = 0.f;
      sred[3][j][q] = 0.f;
      const int n = sid[buf][j];
      if (s != 0.f && n >= 0) atomic_add_global(&grad8[(long)n * 8 + q], s);
    }
    __syncthreads();
  }
  tally.flush(3);
  if constexpr (LOSS) {
    // one loss atomic per workgroup, issued when its list walk is done (the
    // workgroups finish at different times, so the same-address atomics do
    // not arrive together)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc_loss += __shfl_xor(acc_loss, o, 64);
    if (G.lane == 0) sred[0][0][G.w] = acc_loss;
    __syncthreads();
    if (threadIdx.x == 0 && L.loss)
      atomic_add_global(L.loss, (sred[0][0][0] + sred[0][0][1] + sred[0][0][2] + sred[0][0][3]) *
                                  L.inv_count);
  }
}

// Backward of block b.  MAY_MASK: the caller's lists may carry push masks
// (single-canvas calls), so both bodies are compiled and A.pmask picks one;
// the candidate and batched kernels never have them.
template <int DT, bool BLEND, bool LOSS, bool MAY_MASK = true>
__device__ __forceinline__ void raster_bwd_block(const RasterArgs& A, int b,
                                                 const float* __restrict__ gCin,
                                                 const float* __restrict__ gAin,
                                                 const LossArgs& L,
                                                 const float* __restrict__ Cfin,
                                                 const float* __restrict__ Tfin,
                                                 const int* __restrict__ lastin,
                                                 float* __restrict__ grad8) {
  __shared__ BwdLds lds;
  if (MAY_MASK && A.pmask != nullptr)
    raster_bwd_body<DT, BLEND, LOSS, true>(lds, A, b, gCin, gAin, L, Cfin, Tfin, lastin, grad8);
  else
    raster_bwd_body<DT, BLEND, LOSS, false>(lds, A, b, gCin, gAin, L, Cfin, Tfin, lastin, grad8);
}

// Backward with the background composite + MSE as its per-pixel prologue.
// Without colour blending it is held to 80 VGPRs, 6 waves per SIMD instead of
// 5; the few values that then spill are read once per staged batch, not per
// list entry, and the extra wave is worth more (measured).  The blending
// variants keep their natural allocation (they would spill in the loop).
template <int DT, bool BLEND>
static __global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(BLEND ? 1 : 6))) void k_raster_bwd_mse(RasterArgs A, LossArgs L,
                                                                    const float* __restrict__ Cfin,
                                                                    const float* __restrict__ Tfin,
                                                                    const int* __restrict__ lastin,
                                                                    float* __restrict__ grad8) {
  raster_bwd_block<DT, BLEND, true>(A, block_of(A, blockIdx.x), nullptr, nullptr, L,
                                    Cfin, Tfin, lastin, grad8);
}

template <int DT, bool BLEND>
__global__ __launch_bounds__(kThreads) void k_raster_bwd(RasterArgs A, const float* __restrict__ gCin,
                                                         const float* __restrict__ gAin,
                                                         const float* __restrict__ Cfin,
                                                         const float* __restrict__ Tfin,
                                                         const int* __restrict__ lastin,
                                                         float* __restrict__ grad8) {
  const LossArgs none{};
  raster_bwd_block<DT, BLEND, false>(A, block_of(A, blockIdx.x), gCin, gAin,
                                     none, Cfin, Tfin, lastin, grad8);
}

// Candidates of one canvas and template set rasterised by one launch
// (the reference's batch path, tile_rasterize.cu:500-636): workgroup ->
// (candidate, block), with each candidate's records at recs + b N and its own
// lists, pair count, ranges and per-primitive colour maps from this table.
constexpr int kMaxCands = 32;
struct CandTable {
  const int* toffs[kMaxCands];
  const int* tidx[kMaxCands];
  const long long* pcount[kMaxCands];
  long long pcap[kMaxCands];
  const int4* ranges[kMaxCands];
  const void* cmap[kMaxCands];   // cmap_mode 1 only
};

__device__ __forceinline__ RasterArgs candidate_args(const RasterArgs& A, const CandTable& C, int b) {
  RasterArgs Ab = A;
  Ab.recs = A.recs + (long)b * A.N;
  Ab.toffs = C.toffs[b];
  Ab.tidx = C.tidx[b];
  Ab.pcount = C.pcount[b];
  Ab.pcap = C.pcap[b];
  Ab.ranges = C.ranges[b];
  if (A.cmap != nullptr && A.cmap_mode == 1) Ab.cmap = C.cmap[b];
  return Ab;
}

template <int DT, bool BLEND>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(BLEND ? 1 : 8))) void k_raster_fwd_cands(
    RasterArgs A, CandTable C, int nb, float* __restrict__ out_color, float* __restrict__ out_alpha,
    float* __restrict__ out_T, int* __restrict__ out_last) {
  // contiguous bands of (candidate, block) per XCD
  const int r = xcd_remap(blockIdx.x, gridDim.x);
  const int b = r / nb, blk = r - b * nb;
  const long px = (long)A.H * A.W;
  raster_fwd_block<DT, BLEND>(candidate_args(A, C, b), blk, out_color + 3 * px * b,
                              out_alpha + px * b, out_T + px * b, out_last + px * b);
}

template <int DT, bool BLEND>
__global__ __launch_bounds__(kThreads) void k_raster_bwd_cands(
    RasterArgs A, CandTable C, int nb, const float* __restrict__ gCin, const float* __restrict__ gAin,
    const float* __restrict__ Cfin, const float* __restrict__ Tfin, const int* __restrict__ lastin,
    float* __restrict__ grad8) {
  const int r = xcd_remap(blockIdx.x, gridDim.x);
  const int b = r / nb, blk = r - b * nb;
  const long px = (long)A.H * A.W;
  const LossArgs none{};
  raster_bwd_block<DT, BLEND, false, false>(candidate_args(A, C, b), blk, gCin + 3 * px * b,
                                     gAin ? gAin + px * b : nullptr, none, Cfin + 3 * px * b,
                                     Tfin + px * b, lastin + px * b, grad8 + 8L * A.N * b);
}

// Launch order of nb blocks from their forward costs, heaviest first.
// Consecutive workgroups are dealt round-robin to the 8 XCDs (workgroup w to
// XCD w mod 8).  With DBMP_ORDER_BANDED (and nb a multiple of 8) the blocks
// are cut into 8 contiguous bands of natural (row-major) order and band x's
// i-th heaviest block is launched at position 8 i + x: each XCD then works
// through its own band of neighbouring tiles -- sharing its L2 over their
// primitives and texels -- heaviest first.  Otherwise the order is global
// heaviest first: 64 cost buckets (bucket 0 the heaviest), blocks within a
// bucket in any order.  One 1024-thread workgroup (16 waves: the kernel is
// latency bound, so more waves mean fewer dependent rounds).  The costs are
// staged in LDS with every load in flight; lanes of a wave with the same key
// are found with ballots (one per key bit) and counted with a single LDS
// atomic, so a crowded bucket costs one atomic per wave, not one per lane.
#ifndef DBMP_ORDER_BANDED
#define DBMP_ORDER_BANDED 0
#endif
constexpr int kOrderThreads = 1024;
constexpr int kOrderBits = 6;                       // cost buckets
constexpr int kOrderKeyBits = kOrderBits + 3;       // + the band (8 XCDs)
constexpr int kOrderKeys = 1 << kOrderKeyBits;
constexpr int kOrderLds = 16384;   // costs staged in LDS up to this many blocks
static __global__ __launch_bounds__(kOrderThreads) void k_block_order(const int* __restrict__ cost_g,
                                                                    int nb, int* __restrict__ order) {
  __shared__ int cnt[kOrderKeys];
  __shared__ int wsum[kOrderThreads / 64];
  __shared__ int scost[kOrderLds];
  const bool banded = DBMP_ORDER_BANDED && (nb & 7) == 0;
  const int nkeys = banded ? kOrderKeys : (1 << kOrderBits);
  const int kbits = banded ? kOrderKeyBits : kOrderBits;
  for (int i = threadIdx.x; i < kOrderKeys; i += kOrderThreads) cnt[i] = 0;
  const bool staged = nb <= kOrderLds;
  const int* cost = staged ? scost : cost_g;
  int m = 1;
  if (staged) {
    const int n4 = nb >> 2;
    const int4* c4 = reinterpret_cast<const int4*>(cost_g);
#pragma unroll 4
    for (int i = threadIdx.x; i < n4; i += kOrderThreads) {
      const int4 c = c4[i];
      scost[4 * i] = c.x; scost[4 * i + 1] = c.y; scost[4 * i + 2] = c.z; scost[4 * i + 3] = c.w;
      m = max(max(m, max(c.x, c.y)), max(c.z, c.w));
    }
    for (int i = 4 * n4 + threadIdx.x; i < nb; i += kOrderThreads) {
      scost[i] = cost_g[i];
      m = max(m, cost_g[i]);
    }
  } else {
    for (int i = threadIdx.x; i < nb; i += kOrderThreads) m = max(m, cost_g[i]);
  }
  m = wave_max_i32(m);
  if (lane_id() == 0) wsum[threadIdx.x >> 6] = m;
  __syncthreads();
  int mx = wsum[0];
#pragma unroll
  for (int w = 1; w < kOrderThreads / 64; ++w) mx = max(mx, wsum[w]);
  constexpr int kTop = (1 << kOrderBits) - 1;
  const float qs = (float)kTop / (float)mx;
  const int band_len = nb >> 3;
  // key: (band, bucket) when banded, else the bucket (0 for the heaviest)
  auto key_of = [&](int i) {
    const int bucket = kTop - min((int)((float)max(cost[i], 0) * qs), kTop);
    return banded ? ((i / band_len) << kOrderBits) | bucket : bucket;
  };
  const int l = lane_id();
  const unsigned long long lt = (1ull << l) - 1ull;
  // lanes of this wave with the same key as this lane (among valid lanes)
  auto peers_of = [&](bool ok, int key) {
    unsigned long long p = __builtin_amdgcn_ballot_w64(ok);
    for (int k = 0; k < kbits; ++k) {
      const bool bit = (key >> k) & 1;
      const unsigned long long mk = __builtin_amdgcn_ballot_w64(bit);
      p &= bit ? mk : ~mk;
    }
    return p;
  };
  __syncthreads();   // wsum is reused by the scan below
  const int rounds = (nb + kOrderThreads - 1) / kOrderThreads;
#pragma unroll 4
  for (int r = 0; r < rounds; ++r) {
    const int i = r * kOrderThreads + threadIdx.x;
    const bool ok = i < nb;
    const int key = ok ? key_of(i) : 0;
    const unsigned long long p = peers_of(ok, key);
    if (ok && (p & lt) == 0ull) atomicAdd(&cnt[key], __popcll(p));
  }
  __syncthreads();
  // exclusive scan of the key counts, one per thread of the first waves
  static_assert(kOrderKeys <= kOrderThreads, "one key per thread");
  const int t = threadIdx.x;
  const int c = t < nkeys ? cnt[t] : 0;
  int inc = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (l >= o) inc += u;
  }
  if (l == 63) wsum[t >> 6] = inc;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < (t >> 6); ++w) before += wsum[w];
  if (t < nkeys) cnt[t] = before + inc - c;
  __syncthreads();
#pragma unroll 4
  for (int r = 0; r < rounds; ++r) {
    const int i = r * kOrderThreads + threadIdx.x;
    const bool ok = i < nb;
    const int key = ok ? key_of(i) : 0;
    const unsigned long long p = peers_of(ok, key);
    const int leader = ok ? __builtin_ctzll(p) : 0;
    int base = 0;
    if (ok && l == leader) base = atomicAdd(&cnt[key], __popcll(p));
    base = __shfl(base, leader, 64);
    if (ok) {
      int pos = base + __popcll(p & lt);
      if (banded) {   // rank in its band -> launch position 8 rank + band
        const int band = key >> kOrderBits;
        pos = 8 * (pos - band * band_len) + band;
      }
      order[pos] = i;
    }
  }
}

}  // namespace dbmp
