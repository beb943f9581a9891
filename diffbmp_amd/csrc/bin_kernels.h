// Binning: primitive -> 16x16-tile lists, ascending primitive id per tile.
//
// The reference computes the assignment with dense torch ops
// (pydiffbmp/core/renderer/vector_renderer.py:194-252, 401-444) and packs it
// with a Python loop over tiles (simple_tile_renderer.py:721-766).  Here:
//   k_prep        one thread per primitive: rotated template-extent AABB in
//                 exact fp32 (every torch op rounded separately), closed-interval
//                 tile range, pair count, and the raster constants (PrimRec).
//   k_scan_partials / k_scan_apply  exclusive 64-bit scan of the pair counts (one
//                 1024-thread workgroup per chunk of 8192 counts).
//   k_emit        one wave per primitive writes its (tile id, primitive id)
//                 pairs, tiles in row-major order, primitives in id order.
//   k_radix_*     stable LSD radix sort on the tile id; stability keeps the
//                 primitive ids ascending inside each tile.
//   k_tile_bounds CSR offsets from the boundaries of the sorted tile ids.
// For few lists x primitives, the lists are built directly instead:
//   k_tile_build  one workgroup per list: its primitives counted per
//                 sixteenth of the primitive range (one per wave), its offset in
//                 closed form from every primitive's tile range, then each
//                 wave appends its share's primitives in id order (ballot
//                 compaction), so the lists equal the sorted ones.
#pragma once
#include "dbmp_device.h"
#include "dbmp_trig.h"

namespace dbmp {

constexpr int kRadixThreads = 256;          // 4 waves
constexpr int kRadixItems = 4;              // keys per thread (1024-key blocks: many small blocks)
constexpr int kRadixTile = kRadixThreads * kRadixItems;  // keys per block

// Template extents (vector_renderer.py:151-192): min/max column/row of texels
// > 1e-6 mapped through u = (col/(W-1))*2 - 1 in fp32; empty => (-1,1,-1,1).
// With BOX the output is instead each template's content box for the raster
// kernels' culling, float4 (u centre, v centre, u half-width, v half-height),
// widened by one texel on every side: a bilinear sample outside it reads four
// texels <= 1e-6, so its mask is <= 1e-6 and the sample is never pushed.
template <typename TT, bool BOX = false>
static __global__ __launch_bounds__(256) void k_template_extents(const TT* __restrict__ tmpl, int Ht,
                                                          int Wt, float* __restrict__ ext) {
  const int t = blockIdx.x;
  const TT* m = tmpl + (long)t * Ht * Wt;
  int cmin = Wt, cmax = -1, rmin = Ht, rmax = -1;
  for (int i = threadIdx.x; i < Ht * Wt; i += blockDim.x) {
    if (ldf(m, i) > 1e-6f) {
      const int yy = i / Wt, xx = i - yy * Wt;
      cmin = min(cmin, xx); cmax = max(cmax, xx);
      rmin = min(rmin, yy); rmax = max(rmax, yy);
    }
  }
  __shared__ int red[4];
  if (threadIdx.x == 0) { red[0] = Wt; red[1] = -1; red[2] = Ht; red[3] = -1; }
  __syncthreads();
  atomicMin(&red[0], cmin); atomicMax(&red[1], cmax);
  atomicMin(&red[2], rmin); atomicMax(&red[3], rmax);
  __syncthreads();
  if (BOX && threadIdx.x == 0) {
    const float du = 2.f / (float)(Wt - 1), dv = 2.f / (float)(Ht - 1);
    float u0 = -1.f - du, u1 = 1.f + du, v0 = -1.f - dv, v1 = 1.f + dv;
    if (red[1] >= 0) {
      u0 = (float)(red[0] - 1) * du - 1.f;
      u1 = (float)(red[1] + 1) * du - 1.f;
      v0 = (float)(red[2] - 1) * dv - 1.f;
      v1 = (float)(red[3] + 1) * dv - 1.f;
    }
    reinterpret_cast<float4*>(ext)[t] =
        make_float4(0.5f * (u0 + u1), 0.5f * (v0 + v1), 0.5f * (u1 - u0), 0.5f * (v1 - v0));
  } else if (threadIdx.x == 0) {
    float* e = ext + 4 * t;
    if (red[1] < 0) {
      e[0] = -1.f; e[1] = 1.f; e[2] = -1.f; e[3] = 1.f;
    } else {
      e[0] = ((float)red[0] / (float)(Wt - 1)) * 2.f - 1.f;
      e[1] = ((float)red[1] / (float)(Wt - 1)) * 2.f - 1.f;
      e[2] = ((float)red[2] / (float)(Ht - 1)) * 2.f - 1.f;
      e[3] = ((float)red[3] / (float)(Ht - 1)) * 2.f - 1.f;
    }
  }
}

__device__ __forceinline__ float nan_min(float a, float b) {
  return (isnan(a) || isnan(b)) ? __int_as_float(0x7fc00000) : fminf(a, b);
}
__device__ __forceinline__ float nan_max(float a, float b) {
  return (isnan(a) || isnan(b)) ? __int_as_float(0x7fc00000) : fmaxf(a, b);
}

// Tiles t in [0, nt) with !(bmax < t*ts || bmin > min((t+1)*ts, lim)), the
// reference's closed-interval test (vector_renderer.py:437-441).  The set is
// an interval [lo, hi]; estimates are corrected with the exact predicate.
__device__ __forceinline__ void tile_interval(float bmin, float bmax, int nt, int ts, int lim,
                                              int& lo, int& hi) {
  if (isnan(bmin) || isnan(bmax)) {  // NaN compares false: every tile passes
    lo = 0; hi = nt - 1;
    return;
  }
  if (!(bmax >= 0.f)) {
    hi = -1;
  } else {
    const float q = bmax / (float)ts;
    int t = (q >= (float)nt) ? nt - 1 : (int)q;
    while (t >= 0 && bmax < (float)(t * ts)) --t;
    while (t + 1 < nt && !(bmax < (float)((t + 1) * ts))) ++t;
    hi = t;
  }
  if (bmin > (float)lim) {
    lo = nt;
  } else {
    const float q = bmin / (float)ts;
    int t = (q <= 0.f) ? 0 : (q >= (float)nt ? nt - 1 : (int)q);
    while (t > 0 && !(bmin > (float)min(t * ts, lim))) --t;
    while (t < nt && bmin > (float)min((t + 1) * ts, lim)) ++t;
    lo = t;
  }
}

// Raster record and tile range of one primitive (parameters already in fp32).
// ext == nullptr: raster constants only (cnt = 0).  rg is the range of
// ts-px tiles; cnt counts the super-tiles of ts << bin_shift px it covers.
__device__ __forceinline__ void prep_core(float x, float y, float r, float th, float v, float c0,
                                          float c1, float c2, int t, const float* __restrict__ ext,
                                          int T, int Ht, int Wt, int H, int W, int ts, int tilesX,
                                          int tilesY, float alpha_ub, int bin_shift, PrimRec& p,
                                          int& cnt, int4& rg) {
  // fp32 trig shared by binning, raster and the oracle (dbmp_trig.h)
  float c, s;
  dbmp_sincosf(th, &s, &c);
  p.mx = x; p.my = y; p.r = r; p.c = c; p.s = s;
  const float tw = (float)Wt, tht = (float)Ht;
  const float md = fmaxf(tw, tht);
  const float sf = r / (md * 0.5f);
  p.hw = tw * 0.5f * sf;
  p.hh = tht * 0.5f * sf;
  p.rinv = r > 1e-6f ? 1.f / r : 1e6f;
  p.sop = sigmoid_safe(v);
  p.abase = alpha_ub * p.sop;
  p.cr = sigmoid_safe(c0);
  p.cg = sigmoid_safe(c1);
  p.cb = sigmoid_safe(c2);
  p.dav = alpha_ub * p.sop * (1.f - p.sop);
  const bool valid = (t >= 0 && t < T);
  p.tmpl = valid ? t : -1;
  p.tbase = valid ? t * (Ht + 1) * Wt : 0;   // Ht + 1 rows of Wt entries per template
  if (!valid) {   // never drawn: an empty push box rejects every pixel
    p.hw = -1.f;
    p.hh = -1.f;
  }
  cnt = 0;
  rg = make_int4(0, -1, 0, -1);
  if (valid && ext != nullptr) {
    const DBMP_GLOBAL float* e = gptr(ext) + 4 * t;
    const float cu[4] = {e[0], e[1], e[0], e[1]};
    const float cv[4] = {e[2], e[2], e[3], e[3]};
    float xmn = 0.f, xmx = 0.f, ymn = 0.f, ymx = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float a = cu[i] * c, b = cv[i] * s;
      const float d = a - b;
      const float wx = x + r * d;
      const float a2 = cu[i] * s, b2 = cv[i] * c;
      const float d2 = a2 + b2;
      const float wy = y + r * d2;
      if (i == 0) { xmn = xmx = wx; ymn = ymx = wy; }
      else {
        xmn = nan_min(xmn, wx); xmx = nan_max(xmx, wx);
        ymn = nan_min(ymn, wy); ymx = nan_max(ymx, wy);
      }
    }
    int tx0, tx1, ty0, ty1;
    tile_interval(xmn, xmx, tilesX, ts, W, tx0, tx1);
    tile_interval(ymn, ymx, tilesY, ts, H, ty0, ty1);
    if (tx1 >= tx0 && ty1 >= ty0) {
      cnt = ((tx1 >> bin_shift) - (tx0 >> bin_shift) + 1) * ((ty1 >> bin_shift) - (ty0 >> bin_shift) + 1);
      rg = make_int4(tx0, tx1, ty0, ty1);
    }
  }
}

// GT: storage type of positions, radii and rotations; AT: of opacities and
// colours (the same, or fp32 geometry with half-precision appearance).
template <typename GT, typename AT = GT>
static __global__ __launch_bounds__(256) void k_prep(
    int N, const GT* __restrict__ means2D, const GT* __restrict__ radii,
    const GT* __restrict__ rotations, const AT* __restrict__ opacities,
    const AT* __restrict__ colors, const int* __restrict__ sel, const float* __restrict__ ext,
    int T, int Ht, int Wt, int H, int W, int ts, int tilesX, int tilesY, float alpha_ub,
    int bin_shift, PrimRec* __restrict__ recs, int* __restrict__ counts,
    int4* __restrict__ ranges) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  PrimRec p;
  int cnt;
  int4 rg;
  prep_core(ldf(means2D, 2L * n), ldf(means2D, 2L * n + 1), ldf(radii, n), ldf(rotations, n),
            ldf(opacities, n), ldf(colors, 3L * n), ldf(colors, 3L * n + 1),
            ldf(colors, 3L * n + 2), sel[n], ext, T, Ht, Wt, H, W, ts, tilesX, tilesY, alpha_ub,
            bin_shift, p, cnt, rg);
  recs[n] = p;
  if (ext != nullptr) {   // the binning's ranges stay valid across raster-only calls
    counts[n] = cnt;
    ranges[n] = rg;
  }
}

// Inclusive wave scan + block exclusive scan over 256 threads.
__device__ __forceinline__ long long wave_incl_scan_i64(long long x) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(x, o, 64);
    if (l >= o) x += y;
  }
  return x;
}

__device__ __forceinline__ long long block256_excl_scan(long long v, long long& total,
                                                        long long* sh) {
  const int w = threadIdx.x >> 6, l = lane_id();
  const long long inc = wave_incl_scan_i64(v);
  if (l == 63) sh[w] = inc;
  __syncthreads();
  long long before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long long s = sh[i];
    before += (i < w) ? s : 0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return inc - v + before;
}

// Exclusive scan over a block of NW waves (sh: NW slots).
template <int NW>
__device__ __forceinline__ long long block_excl_scan(long long v, long long& total, long long* sh) {
  const int w = threadIdx.x >> 6, l = lane_id();
  const long long inc = wave_incl_scan_i64(v);
  if (l == 63) sh[w] = inc;
  __syncthreads();
  long long before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const long long s = sh[i];
    before += (i < w) ? s : 0;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return inc - v + before;
}

// offs[i] = sum(counts[0..i)), offs[N] = *total = sum(counts), over chunks
// of kScanChunk counts, one 1024-thread workgroup per chunk (8 contiguous
// counts per thread): k_scan_partials writes each chunk's sum, k_scan_apply
// scans each chunk from the sum of the chunks before it.  With one chunk
// only the second kernel runs.  No workgroup waits for another.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanThreads * kScanItems;

__device__ __forceinline__ long long load_chunk_counts(const int* __restrict__ counts, int N, int i0,
                                                       int loc[kScanItems]) {
  if (i0 + kScanItems <= N) {
    const int4 a = ld_i4(gbytes(counts + i0));
    const int4 b = ld_i4(gbytes(counts + i0 + 4));
    loc[0] = a.x; loc[1] = a.y; loc[2] = a.z; loc[3] = a.w;
    loc[4] = b.x; loc[5] = b.y; loc[6] = b.z; loc[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) loc[j] = (i0 + j < N) ? gptr(counts)[i0 + j] : 0;
  }
  long long s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) s += loc[j];
  return s;
}

static __global__ __launch_bounds__(kScanThreads) void k_scan_partials(const int* __restrict__ counts,
                                                                       int N,
                                                                       long long* __restrict__ partials) {
  __shared__ long long sh[kScanThreads / 64];
  int loc[kScanItems];
  const long long s =
      load_chunk_counts(counts, N, blockIdx.x * kScanChunk + threadIdx.x * kScanItems, loc);
  long long tot;
  (void)block_excl_scan<kScanThreads / 64>(s, tot, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

static __global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int* __restrict__ counts,
                                                                    int N,
                                                                    const long long* __restrict__ partials,
                                                                    long long* __restrict__ offs,
                                                                    long long* __restrict__ total) {
  __shared__ long long sh[kScanThreads / 64];
  // sum of the chunks before this one (at most a few hundred)
  long long before = 0;
  for (int i = threadIdx.x; i < (int)blockIdx.x; i += kScanThreads) before += partials[i];
  long long pre;
  (void)block_excl_scan<kScanThreads / 64>(before, pre, sh);
  const int i0 = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  int loc[kScanItems];
  const long long s = load_chunk_counts(counts, N, i0, loc);
  long long tot;
  long long ex = block_excl_scan<kScanThreads / 64>(s, tot, sh) + pre;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (i0 + j < N) gptr(offs)[i0 + j] = ex;
    ex += loc[j];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    gptr(offs)[N] = pre + tot;
    *gptr(total) = pre + tot;
  }
}

// Chunk sums needed by scan_counts for N counts (the buffer after offs).
inline long long scan_partials_len(long long N) { return N / kScanChunk + 1; }

// Queue the exclusive scan of counts[0..N) into offs[0..N] and *total;
// partials holds scan_partials_len(N) values.
inline void scan_counts(const int* counts, int N, long long* offs, long long* total,
                        long long* partials, hipStream_t st) {
  const int chunks = N > 0 ? (N + kScanChunk - 1) / kScanChunk : 1;
  if (chunks > 1)
    hipLaunchKernelGGL(k_scan_partials, dim3(chunks), dim3(kScanThreads), 0, st, counts, N, partials);
  hipLaunchKernelGGL(k_scan_apply, dim3(chunks), dim3(kScanThreads), 0, st, counts, N,
                     (const long long*)partials, offs, total);
}

// One wave per primitive; pairs land at offs[n] in row-major tile order.
// tilesX: tiles per row of the list grid (super-tiles of 2^shift tiles).
static __global__ __launch_bounds__(256) void k_emit(int N, const int* __restrict__ counts,
                                              const long long* __restrict__ offs,
                                              const int4* __restrict__ ranges, int tilesX, int shift,
                                              const long long* __restrict__ total, long long cap,
                                              int* __restrict__ keys, int* __restrict__ vals) {
  if (*total > cap) return;
  const int n = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (n >= N) return;
  const int cnt = counts[n];
  if (cnt == 0) return;
  const long long off = offs[n];
  const int4 r16 = ranges[n];
  const int4 rg = make_int4(r16.x >> shift, r16.y >> shift, r16.z >> shift, r16.w >> shift);
  const int nx = rg.y - rg.x + 1;
  for (int i = lane_id(); i < cnt; i += 64) {
    const int qy = i / nx, qx = i - qy * nx;
    keys[off + i] = (rg.z + qy) * tilesX + rg.x + qx;
    vals[off + i] = n;
  }
}

__device__ __forceinline__ long long pairs_in_range(const long long* total, long long cap) {
  const long long P = *total;
  return P > cap ? 0 : P;
}

// Per-block digit histogram, stored digit-major: hist[d * nblk + b].
static __global__ __launch_bounds__(kRadixThreads) void k_radix_hist(
    const int* __restrict__ keys, const long long* __restrict__ total, long long cap, int shift,
    int bits, int nblk, int* __restrict__ hist) {
  __shared__ int h[256];
  const int radix = 1 << bits, mask = radix - 1;
  const long long P = pairs_in_range(total, cap);
  for (int d = threadIdx.x; d < radix; d += kRadixThreads) h[d] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kRadixTile + (threadIdx.x >> 6) * (64 * kRadixItems) + lane_id();
#pragma unroll 4
  for (int i = 0; i < kRadixItems; ++i) {
    const long long idx = base + i * 64;
    if (idx < P) atomicAdd(&h[(keys[idx] >> shift) & mask], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < radix; d += kRadixThreads) hist[(long)d * nblk + blockIdx.x] = h[d];
}

// One block per digit: exclusive scan of that digit's row over blocks.  The
// rows are sized for the pair capacity; only the blocks that hold pairs are
// scanned (the scatter reads no other, and the rest count nothing).
static __global__ __launch_bounds__(256) void k_radix_rowscan(int* __restrict__ hist, int nblk,
                                                       int* __restrict__ totals,
                                                       const long long* __restrict__ total,
                                                       long long cap) {
  __shared__ long long sh[4];
  int* row = hist + (long)blockIdx.x * nblk;
  const long long P = pairs_in_range(total, cap);
  const int nused = (int)min((long long)nblk, (P + kRadixTile - 1) / kRadixTile);
  long long carry = 0;
  for (int base = 0; base < nused; base += 256) {
    const int i = base + threadIdx.x;
    const long long v = i < nused ? row[i] : 0;
    long long tot;
    const long long ex = block256_excl_scan(v, tot, sh) + carry;
    if (i < nused) row[i] = (int)ex;
    carry += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = (int)carry;
}

// Stable scatter.  Keys of block b are visited as wave w, item i, lane l
// (index b*4096 + w*1024 + i*64 + l), i.e. in input order; each wave ranks its
// items by digit with ballots and a per-wave running count in LDS.  The block
// then reorders its keys by digit in LDS and writes each digit's run
// contiguously, so global writes are coalesced.
static __global__ __launch_bounds__(kRadixThreads) void k_radix_scatter(
    const int* __restrict__ kin, const int* __restrict__ vin, int* __restrict__ kout,
    int* __restrict__ vout, const long long* __restrict__ total, long long cap, int shift,
    int bits, int nblk, const int* __restrict__ hist, const int* __restrict__ totals) {
  __shared__ int gbase[256];     // global position of this block's first key of digit d
  __shared__ int loff[256];      // block-local offset of digit d
  __shared__ int wcnt[4][256];
  __shared__ int skey[kRadixTile];
  __shared__ int sval[kRadixTile];
  __shared__ long long sh[4];
  const int radix = 1 << bits, mask = radix - 1;
  const long long P = pairs_in_range(total, cap);
  const int w = threadIdx.x >> 6, l = lane_id();
  const long long blk0 = (long long)blockIdx.x * kRadixTile;
  if (blk0 >= P) return;   // uniform: nothing of this block is in range
  const int nvalid = (int)min((long long)kRadixTile, P - blk0);
  {
    const int d = threadIdx.x;
    const long long v = d < radix ? totals[d] : 0;
    long long tot;
    const long long ex = block256_excl_scan(v, tot, sh);
    if (d < radix) gbase[d] = (int)ex + hist[(long)d * nblk + blockIdx.x];
  }
  for (int d = l; d < radix; d += 64) wcnt[w][d] = 0;
  __syncthreads();
  const long long base = blk0 + w * (64 * kRadixItems) + l;
  const unsigned long long lt = (1ull << l) - 1ull;
  int kk[kRadixItems], vv[kRadixItems], rk[kRadixItems];
#pragma unroll
  for (int i = 0; i < kRadixItems; ++i) {
    const long long idx = base + i * 64;
    const bool ok = idx < P;
    kk[i] = ok ? kin[idx] : 0;
    vv[i] = ok ? vin[idx] : 0;
  }
#pragma unroll
  for (int i = 0; i < kRadixItems; ++i) {
    const bool ok = base + i * 64 < P;
    const int d = (kk[i] >> shift) & mask;
    unsigned long long peers = __ballot(ok);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (d >> b) & 1;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    int r = 0;
    if (ok) {
      const int cur = wcnt[w][d];
      const unsigned long long below = peers & lt;
      r = cur + __popcll(below);
      if (below == 0ull) wcnt[w][d] = cur + __popcll(peers);
    }
    rk[i] = r;
  }
  __syncthreads();
  // exclusive prefix over waves per digit; block count per digit
  long long cnt_d = 0;
  if (threadIdx.x < radix) {
    const int d = threadIdx.x;
    int s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = wcnt[q][d];
      wcnt[q][d] = s;
      s += t;
    }
    cnt_d = s;
  }
  {
    long long tot;
    const long long ex = block256_excl_scan(cnt_d, tot, sh);
    if (threadIdx.x < radix) loff[threadIdx.x] = (int)ex;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kRadixItems; ++i) {
    if (base + i * 64 < P) {
      const int d = (kk[i] >> shift) & mask;
      const int lp = loff[d] + wcnt[w][d] + rk[i];
      skey[lp] = kk[i];
      sval[lp] = vv[i];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nvalid; i += kRadixThreads) {
    const int k = skey[i];
    const int d = (k >> shift) & mask;
    const long long pos = (long long)gbase[d] + (i - loff[d]);
    kout[pos] = k;
    vout[pos] = sval[i];
  }
}

// tile_offsets[t] = first index whose tile id >= t, tile_offsets[tiles] = P,
// from the boundaries of the sorted ids.  On overflow (P > cap) every offset
// is 0.  Needs max(cap, tiles) + 1 threads.
static __global__ __launch_bounds__(256) void k_tile_bounds(const int* __restrict__ keys,
                                                     const long long* __restrict__ total,
                                                     long long cap, int tiles,
                                                     int* __restrict__ tile_offsets) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long Praw = *total;
  if (Praw > cap) {
    if (i <= tiles) tile_offsets[i] = 0;
    return;
  }
  const long long P = Praw;
  if (i > P) return;
  const int kcur = (i < P) ? keys[i] : tiles;
  const int kprev = (i > 0) ? keys[i - 1] : -1;
  for (int t = kprev + 1; t <= kcur; ++t) tile_offsets[t] = (int)i;
}

// Direct list build.  A primitive n belongs to list t = (tx, ty) (lists of
// 2^shift x 2^shift tiles, tilesX lists per row) iff its tile range rg[n]
// covers it.  Wave w of a workgroup owns primitives [w*q, (w+1)*q), q =
// ceil(N/kBuildWaves); eight range loads per lane are in flight at a time.
constexpr int kDirectBatch = 8;
constexpr int kBuildWaves = 16;   // waves per list workgroup (1024 threads)

__device__ __forceinline__ bool list_covers(int4 r, int tx, int ty, int shift) {
  return ((r.x >> shift) <= tx) & (tx <= (r.y >> shift)) & ((r.z >> shift) <= ty) &
         (ty <= (r.w >> shift));
}

// Pairs of the lists before list (tx, ty) (row-major, tilesX per row) and in
// all lists, of a primitive whose lists are [x0, x1] x [y0, y1] (empty when
// x1 < x0 or y1 < y0).
__device__ __forceinline__ void pairs_before(int4 r, int tx, int ty, int shift, long long& before,
                                             long long& all) {
  const int x0 = r.x >> shift, x1 = r.y >> shift, y0 = r.z >> shift, y1 = r.w >> shift;
  const int nx = x1 - x0 + 1, ny = y1 - y0 + 1;
  if (nx <= 0 || ny <= 0) return;
  all += (long long)nx * ny;
  const int rows = max(0, min(y1, ty - 1) - y0 + 1);   // whole rows above ty
  const int cols = (y0 <= ty && ty <= y1) ? max(0, min(x1, tx - 1) - x0 + 1) : 0;
  before += (long long)rows * nx + cols;
}

// One workgroup per list.  First pass: each wave counts its share's
// primitives in this list and, in closed form per primitive, the pairs of
// all lists before this one and of all lists; second pass: each wave appends
// its share's primitives in id order (ballot compaction), so the lists
// equal the sorted ones.  offs (tiles+1): list offsets; total: pair count.
// Lists are written only when the pairs fit cap (otherwise every offset is 0
// and the raster kernels walk all primitives).
// STAGE (N <= kDirectStage): the first pass also keeps the ranges in LDS for
// the second (each wave reads back only what it wrote).
constexpr int kDirectStage = 8192;
template <bool STAGE>
static __global__ __launch_bounds__(64 * kBuildWaves) void k_tile_build(const int4* __restrict__ rg, int N,
                                                           int tilesX, int tiles, int shift,
                                                           long long cap, int* __restrict__ offs,
                                                           int* __restrict__ idx,
                                                           long long* __restrict__ total) {
  __shared__ long long sred[2][kBuildWaves];
  __shared__ int scnt[kBuildWaves];
  __shared__ int4 srg[STAGE ? kDirectStage : 1];
  const int t = blockIdx.x;
  const int tx = t % tilesX, ty = t / tilesX;
  const int w = threadIdx.x >> 6, l = lane_id();
  const int q = (N + kBuildWaves - 1) / kBuildWaves;
  const int lo = w * q, hi = min(N, lo + q);
  long long before = 0, all = 0;
  int c = 0;
  for (int base = lo; base < hi; base += 64 * kDirectBatch) {
    int4 r[kDirectBatch];
#pragma unroll
    for (int k = 0; k < kDirectBatch; ++k) {
      const int n = base + k * 64 + l;
      r[k] = n < hi ? rg[n] : make_int4(0, -1, 0, -1);
    }
#pragma unroll
    for (int k = 0; k < kDirectBatch; ++k) {
      c += list_covers(r[k], tx, ty, shift) ? 1 : 0;
      pairs_before(r[k], tx, ty, shift, before, all);
      if (STAGE && base + k * 64 + l < hi) srg[base + k * 64 + l] = r[k];
    }
  }
  c = wave_sum_i32(c);
  before = wave_sum_i64(before);
  all = wave_sum_i64(all);
  if (l == 0) {
    sred[0][w] = before;
    sred[1][w] = all;
    scnt[w] = c;
  }
  __syncthreads();
  before = 0;
  all = 0;
#pragma unroll
  for (int k = 0; k < kBuildWaves; ++k) {
    before += sred[0][k];
    all += sred[1][k];
  }
  const bool fits = all <= cap;
  if (threadIdx.x == 0) {
    offs[t] = fits ? (int)before : 0;
    if (t == tiles - 1) {
      offs[tiles] = fits ? (int)all : 0;
      *total = all;
    }
  }
  if (!fits) return;
  long long pos = before;
  for (int k = 0; k < w; ++k) pos += scnt[k];
  const unsigned long long lt = (1ull << l) - 1ull;
  for (int base = lo; base < hi; base += 64 * kDirectBatch) {
    int4 r[kDirectBatch];
#pragma unroll
    for (int k = 0; k < kDirectBatch; ++k) {
      const int n = base + k * 64 + l;
      r[k] = n < hi ? (STAGE ? srg[n] : rg[n]) : make_int4(0, -1, 0, -1);
    }
#pragma unroll
    for (int k = 0; k < kDirectBatch; ++k) {
      const bool in = list_covers(r[k], tx, ty, shift);
      const unsigned long long m = __builtin_amdgcn_ballot_w64(in);
      if (in) idx[pos + __popcll(m & lt)] = base + k * 64 + l;
      pos += __popcll(m);
    }
  }
}

}  // namespace dbmp
