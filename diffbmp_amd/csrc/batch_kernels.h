// Batched optimisation step over many independent fit jobs.
//
// Every active job's primitives, tiles and 16x16 blocks are concatenated:
// job j owns primitives [prim_start[j], prim_start[j+1]), tiles
// [tile_start[j], ...) and raster workgroups [block_start[j], ...).  One
// launch of each stage covers all jobs:
//   k_batch_prep    records + tile ranges, one thread per primitive
//   k_scan_partials / k_scan_apply   (shared with the single-canvas path)
//   k_batch_emit    pairs keyed by the job's tile base + local tile id, valued
//                   by the local primitive id; the stable radix sort then
//                   yields every job's lists, ascending primitive id per tile,
//                   exactly as a job binned alone
//   k_batch_fwd     raster forward, workgroup -> (job, block)
//   k_batch_bwd     background composite + MSE prologue, raster backward
//   k_batch_tick    per-job Adam step counter and bias corrections
//   k_batch_adam    Adam + clamp over the (N,8) parameter rows of every job
// Jobs differ in canvas, primitive count, template set and learning rates;
// they share the parameter/bitmap storage type.  Nothing in a step depends on
// host-side values that change between steps, so a step can be captured once
// and replayed while the job set is unchanged.
#pragma once
#include "bin_kernels.h"
#include "raster_kernels.h"

namespace dbmp {

// Per-job constants, uploaded when the job set changes.
struct BatchJob {
  RasterArgs A;        // canvas, lists (batch-wide buffers at this job's offsets), bitmaps
  LossArgs L;          // target / background; L.loss unused (loss_hist below)
  const void* store;   // (N,8) parameters in the storage type, read by k_batch_prep
  float* p;            // (N,8) fp32 master: mx, my, r, theta, opacity logit, colour logits
  float* m;            // (N,8) Adam first moment
  float* v;            // (N,8) Adam second moment
  void* store_out;     // == store, written by k_batch_adam (null when store == p)
  const int* sel;      // (N) template index
  const float* ext;    // (T,4) template extents
  int T;
  int tilesX16, tilesY16;   // 16-px tiles of the canvas (the binning's ranges)
  float* color;        // (H,W,3) forward output, premultiplied over black
  float* alpha;        // (H,W)
  float* Tfin;         // (H,W) product of the nonzero (1 - a) factors
  int* last;           // (H,W) entries consumed | opaque count << 30
  float* loss_hist;    // (loss_cap) MSE per step
  int loss_cap;
  float alpha_ub;      // alpha upper bound of the job
  float lr[5];         // means2D, radii, rotations, opacities, colors
  float lo[3], hi[3];  // clamp of x, y, r (lo > hi: no clamp)
};

// Per-job state that the step itself advances.
struct BatchState {
  int step;            // Adam steps taken
  float b1t, b2t;      // beta1^step, beta2^step as fp32 running products
  float step_scale[5]; // lr_g / (1 - beta1^t) of the step in flight
  float inv_bc2_sqrt;  // 1 / sqrt(1 - beta2^t)
};

// Largest j with start[j] <= x (start ascending, start[0] = 0, J entries).
__device__ __forceinline__ int find_job(const int* __restrict__ start, int J, int x) {
  int lo = 0, hi = J - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Content boxes of job j's templates for the raster cull, from its extents
// (texels > 1e-6 mapped to u/v; empty templates (-1,1,-1,1)) widened by one
// texel on each side, as k_template_extents<_, true> builds them from the
// texels themselves.  One workgroup per job.
static __global__ __launch_bounds__(64) void k_batch_cull(const BatchJob* __restrict__ jobs) {
  const BatchJob& d = jobs[blockIdx.x];
  float4* out = const_cast<float4*>(d.A.cull);
  const float du = 2.f / (float)(d.A.Wt - 1), dv = 2.f / (float)(d.A.Ht - 1);
  for (int t = threadIdx.x; t < d.T; t += 64) {
    const DBMP_GLOBAL float* e = gptr(d.ext) + 4 * t;
    const float u0 = e[0] - du, u1 = e[1] + du, v0 = e[2] - dv, v1 = e[3] + dv;
    st_f4(out + t, make_float4(0.5f * (u0 + u1), 0.5f * (v0 + v1), 0.5f * (u1 - u0), 0.5f * (v1 - v0)));
  }
}

template <int DT>
static __global__ __launch_bounds__(256) void k_batch_prep(const BatchJob* __restrict__ jobs,
                                                    const int* __restrict__ prim_start, int J,
                                                    PrimRec* __restrict__ recs,
                                                    int* __restrict__ counts,
                                                    int4* __restrict__ ranges) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= prim_start[J]) return;
  const int j = find_job(prim_start, J, g);
  const BatchJob& B = jobs[j];
  const int n = g - prim_start[j];
  float q[8];
  if constexpr (DT == kF32) {
    const DBMP_GLOBAL char* s = gbytes(B.store) + 32L * n;
    const float4 a = ld_f4(s), b = ld_f4(s + 16);
    q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = a.w; q[4] = b.x; q[5] = b.y; q[6] = b.z; q[7] = b.w;
  } else {
    using ST = typename std::conditional<DT == kF16, __half, __hip_bfloat16>::type;
    const ST* s = reinterpret_cast<const ST*>(B.store) + 8L * n;
#pragma unroll
    for (int i = 0; i < 8; ++i) q[i] = ldf(s, i);
  }
  PrimRec p;
  int cnt;
  int4 rg;
  const RasterArgs& A = B.A;
  prep_core(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], gptr(B.sel)[n], B.ext, B.T, A.Ht, A.Wt, A.H,
            A.W, kSub, B.tilesX16, B.tilesY16, B.alpha_ub, A.bin_shift, p, cnt, rg);
  recs[g] = p;
  counts[g] = cnt;
  ranges[g] = rg;
}

// One wave per primitive: key = job list base + row-major local list id
// (lists per super-tile of 16 << bin_shift px), value = local primitive id.
static __global__ __launch_bounds__(256) void k_batch_emit(const BatchJob* __restrict__ jobs,
                                                    const int* __restrict__ prim_start,
                                                    const int* __restrict__ tile_start, int J,
                                                    const int* __restrict__ counts,
                                                    const long long* __restrict__ offs,
                                                    const int4* __restrict__ ranges,
                                                    const long long* __restrict__ total,
                                                    long long cap, int* __restrict__ keys,
                                                    int* __restrict__ vals) {
  if (*total > cap) return;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= prim_start[J]) return;
  const int cnt = counts[g];
  if (cnt == 0) return;
  const int j = find_job(prim_start, J, g);
  const int n = g - prim_start[j];
  const int tilesX = jobs[j].A.tilesX, tb = tile_start[j], sh = jobs[j].A.bin_shift;
  const long long off = offs[g];
  const int4 r16 = ranges[g];
  const int4 rg = make_int4(r16.x >> sh, r16.y >> sh, r16.z >> sh, r16.w >> sh);
  const int nx = rg.y - rg.x + 1;
  for (int i = lane_id(); i < cnt; i += 64) {
    const int qy = i / nx, qx = i - qy * nx;
    keys[off + i] = tb + (rg.z + qy) * tilesX + rg.x + qx;
    vals[off + i] = n;
  }
}

// Persistent raster launches.  A launch has as many workgroups as fit on the
// device at once; each takes (job, block) work items until none is left.
// The items are cut into 8 contiguous bands of the concatenated blocks (jobs
// in order, tiles row-major), one per XCD: workgroup w runs on XCD w mod 8
// (workgroups are dealt round-robin to the XCDs, and all of them are resident
// from the start), takes items of its own band from that band's counter --
// neighbouring tiles, whose primitives and texels share its L2 -- and, once
// its band is drained, items of the other bands.  Every workgroup leaves when
// all eight counters have passed their band's end.  The counters are zeroed
// by k_batch_reset before each step.
constexpr int kXcds = 8;
constexpr int kCounterStride = 32;   // ints: one 128-byte line per counter
struct WorkQueue {
  int* ctr;   // kXcds counters, kCounterStride apart
  int total;  // items
};

__device__ __forceinline__ int next_item(const WorkQueue& q, int* s_item) {
  if (threadIdx.x == 0) {
    const int x = blockIdx.x & (kXcds - 1);
    int item = -1;
    for (int k = 0; k < kXcds && item < 0; ++k) {
      const int band = (x + k) & (kXcds - 1);
      const int b0 = (int)((long long)q.total * band / kXcds);
      const int b1 = (int)((long long)q.total * (band + 1) / kXcds);
      if (b1 <= b0) continue;
      const int i = atomicAdd(q.ctr + band * kCounterStride, 1);
      if (i < b1 - b0) item = b0 + i;
    }
    *s_item = item;
  }
  __syncthreads();
  // uniform across the workgroup: readfirstlane tells the compiler, so the
  // job lookup and the job's descriptor stay in scalar registers
  const int item = __builtin_amdgcn_readfirstlane(*s_item);
  __syncthreads();   // every wave has read it before thread 0 writes the next
  return item;
}

static __global__ void k_batch_reset(int* __restrict__ ctr, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctr[i] = 0;
}

template <int DT, bool BLEND>
static __global__ __launch_bounds__(kThreads) void k_batch_fwd(const BatchJob* __restrict__ jobs,
                                                        const int* __restrict__ block_start,
                                                        int J, WorkQueue q) {
  __shared__ int s_item;
  for (;;) {
    const int b = next_item(q, &s_item);
    if (b < 0) break;
    const int j = find_job(block_start, J, b);
    const BatchJob& B = jobs[j];
    raster_fwd_block<DT, BLEND>(B.A, b - block_start[j], B.color, B.alpha, B.Tfin, B.last);
  }
}

template <int DT, bool BLEND>
static __global__ __launch_bounds__(kThreads) void k_batch_bwd(const BatchJob* __restrict__ jobs,
                                                        const BatchState* __restrict__ state,
                                                        const int* __restrict__ block_start,
                                                        const int* __restrict__ prim_start,
                                                        int J, float* __restrict__ grad8,
                                                        WorkQueue q) {
  __shared__ int s_item;
  for (;;) {
    const int b = next_item(q, &s_item);
    if (b < 0) break;
    const int j = find_job(block_start, J, b);
    const BatchJob& B = jobs[j];
    LossArgs L = B.L;
    L.loss = B.loss_hist + min(state[j].step, B.loss_cap - 1);
    raster_bwd_block<DT, BLEND, true>(B.A, b - block_start[j], nullptr, nullptr, L, B.color, B.Tfin,
                                      B.last, grad8 + 8L * prim_start[j]);
  }
}

// t <- t + 1 and the bias corrections 1 - beta^t of step t
// (torch.optim.Adam), from fp32 running products of beta (torch forms
// beta^t in double on the host; the products agree to a few fp32 ulp over
// the step counts of a fit).
// Thread 0 also copies the step's pair count to host-mapped memory.
static __global__ void k_batch_tick(const BatchJob* __restrict__ jobs, BatchState* __restrict__ state,
                             int J, float beta1, float beta2, const long long* __restrict__ total,
                             long long* __restrict__ host_total) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0 && host_total) *host_total = *total;
  if (j >= J) return;
  BatchState& S = state[j];
  S.step += 1;
  S.b1t *= beta1;
  S.b2t *= beta2;
  const float bc1 = 1.f - S.b1t, bc2 = 1.f - S.b2t;
#pragma unroll
  for (int i = 0; i < 5; ++i) S.step_scale[i] = jobs[j].lr[i] / bc1;
  S.inv_bc2_sqrt = 1.f / sqrtf(bc2);
}

// One thread per parameter value; value q of a row belongs to group
// {0,0,1,2,3,4,4,4}[q].  The update is k_adam's.
template <int DT>
static __global__ __launch_bounds__(256) void k_batch_adam(const BatchJob* __restrict__ jobs,
                                                    const BatchState* __restrict__ state,
                                                    const int* __restrict__ prim_start, int J,
                                                    const float* __restrict__ grad8, float beta1,
                                                    float beta2, float eps) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = (int)(i >> 3), q = (int)(i & 7);
  if (g >= prim_start[J]) return;
  const int j = find_job(prim_start, J, g);
  const BatchJob& B = jobs[j];
  const long long k = 8LL * (g - prim_start[j]) + q;
  const int grp = q < 2 ? 0 : (q < 5 ? q - 1 : 4);
  const float gr = grad8[i];
  const float m = fmaf(beta1, gptr(B.m)[k], (1.f - beta1) * gr);
  const float v = fmaf(beta2, gptr(B.v)[k], (1.f - beta2) * (gr * gr));
  gptr(B.m)[k] = m;
  gptr(B.v)[k] = v;
  float p = gptr(B.p)[k] - state[j].step_scale[grp] * (m / (sqrtf(v) * state[j].inv_bc2_sqrt + eps));
  if (q < 3 && B.lo[q] <= B.hi[q]) p = fminf(fmaxf(p, B.lo[q]), B.hi[q]);
  gptr(B.p)[k] = p;
  if (B.store_out) {
    DBMP_GLOBAL unsigned short* h = (DBMP_GLOBAL unsigned short*)B.store_out;
    if constexpr (DT == kF16) h[k] = __half_as_ushort(__float2half(p));
    else if constexpr (DT == kBF16) h[k] = __bfloat16_as_ushort(__float2bfloat16(p));
    else gptr((float*)B.store_out)[k] = p;
  }
}

}  // namespace dbmp
