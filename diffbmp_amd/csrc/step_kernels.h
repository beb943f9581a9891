// The optimisation step around the rasterizer, on the device.
//
// The reference builds it from torch ops: background composite
// C + (1 - A) * I_bg (pydiffbmp/core/renderer/simple_tile_renderer.py:703-707),
// F.mse_loss (pydiffbmp/util/loss_functions.py:20-40), autograd through both,
// torch.optim.Adam (simple_tile_renderer.py:961-968, 1125) and the parameter
// clamp (simple_tile_renderer.py:122-128, 1131).  Here each half is one kernel.
#pragma once
#include "dbmp_device.h"

namespace dbmp {

// result = C + (1 - A) * bg;  r = result - target;  loss += sum(r^2) * inv_count
// dL/dC = 2 r * inv_count;  dL/dA = -sum_c dL/dC_c * bg_c.
// bg is (H,W,3) or, when null, the constant bg_value.
__device__ __forceinline__ float loss_pixel(float a, const float* c, const float* b, float bg_value,
                                            const float* t, float inv_count, float* g, float& acc) {
  const float om = 1.f - a;
  float ga = 0.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float bk = b ? b[k] : bg_value;
    const float r = fmaf(om, bk, c[k]) - t[k];
    acc = fmaf(r, r, acc);
    g[k] = 2.f * r * inv_count;
    ga = fmaf(-g[k], bk, ga);
  }
  return ga;
}

#ifndef DBMP_LOSS_GROUPS
#define DBMP_LOSS_GROUPS 4
#endif
// float4 groups (4 pixels each) a thread loads before it computes: the loads
// of one iteration are independent, so a thread keeps G x 10 of them in
// flight and the grid is G times smaller, which also divides the number of
// blocks that end in an atomic add on the one loss value (same-address
// atomics serialise)
constexpr int kLossGroups = DBMP_LOSS_GROUPS;

// kVec: every pointer 16-byte aligned; each thread takes kLossGroups float4
// groups per iteration (group q: alpha[4q..4q+3] as one float4, colour /
// target / bg as three), spaced one grid apart so every load is coalesced
// across the wave; the npix % 4 tail pixels go through the scalar path.
template <bool kVec>
static __global__ __launch_bounds__(256) void k_loss_mse(const float* __restrict__ color,
                                                  const float* __restrict__ alpha,
                                                  const float* __restrict__ bg, float bg_value,
                                                  const float* __restrict__ target, long long npix,
                                                  float inv_count, float* __restrict__ gC,
                                                  float* __restrict__ gA, float* __restrict__ loss) {
  float acc = 0.f;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nthr = (long long)gridDim.x * blockDim.x;
  long long p0 = 0;
  if (kVec) {
    constexpr int G = kLossGroups;
    const long long nq = npix >> 2;
    for (long long base = tid; base < nq; base += G * nthr) {
      float4 av[G], cv[G][3], tv[G][3], bv[G][3];
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const long long q = base + k * nthr;
        if (q < nq) {
          av[k] = reinterpret_cast<const float4*>(alpha)[q];
#pragma unroll
          for (int j = 0; j < 3; ++j) {
            cv[k][j] = reinterpret_cast<const float4*>(color)[3 * q + j];
            tv[k][j] = reinterpret_cast<const float4*>(target)[3 * q + j];
            if (bg) bv[k][j] = reinterpret_cast<const float4*>(bg)[3 * q + j];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const long long q = base + k * nthr;
        if (q >= nq) break;
        float c[12], t[12], b[12], g[12], ga[4];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          c[4 * j] = cv[k][j].x; c[4 * j + 1] = cv[k][j].y; c[4 * j + 2] = cv[k][j].z; c[4 * j + 3] = cv[k][j].w;
          t[4 * j] = tv[k][j].x; t[4 * j + 1] = tv[k][j].y; t[4 * j + 2] = tv[k][j].z; t[4 * j + 3] = tv[k][j].w;
          if (bg) {
            b[4 * j] = bv[k][j].x; b[4 * j + 1] = bv[k][j].y; b[4 * j + 2] = bv[k][j].z; b[4 * j + 3] = bv[k][j].w;
          }
        }
        const float a[4] = {av[k].x, av[k].y, av[k].z, av[k].w};
#pragma unroll
        for (int i = 0; i < 4; ++i)
          ga[i] = loss_pixel(a[i], c + 3 * i, bg ? b + 3 * i : nullptr, bg_value, t + 3 * i, inv_count,
                             g + 3 * i, acc);
#pragma unroll
        for (int j = 0; j < 3; ++j)
          reinterpret_cast<float4*>(gC)[3 * q + j] =
              make_float4(g[4 * j], g[4 * j + 1], g[4 * j + 2], g[4 * j + 3]);
        reinterpret_cast<float4*>(gA)[q] = make_float4(ga[0], ga[1], ga[2], ga[3]);
      }
    }
    p0 = nq << 2;
  }
  for (long long p = p0 + tid; p < npix; p += nthr)
    gA[p] = loss_pixel(alpha[p], color + 3 * p, bg ? bg + 3 * p : nullptr, bg_value, target + 3 * p,
                       inv_count, gC + 3 * p, acc);
  // wave then block reduction, one atomic per block
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0 && loss) unsafeAtomicAdd(loss, (part[0] + part[1] + part[2] + part[3]) * inv_count);
}

struct AdamGroup {
  float* p;          // fp32 master parameter
  float* m;          // first moment
  float* v;          // second moment
  float* g;          // gradient (element i at g[i], or at a row column, below)
  void* store;       // optional reduced-precision copy for the rasterizer (or null)
  long long n;       // elements
  int stride;        // 2 for means2D (component c = i % 2 uses lo[c], hi[c]), else 1
  int clamp;         // apply [lo, hi] after the update
  float lo[2], hi[2];
  int grow, gcomp;   // grow > 0: element i's gradient at g[(i / gcomp) * grow + i % gcomp]
  int zero;          // write 0 over each gradient value read
};

struct AdamArgs {
  AdamGroup g[5];
  float step_scale[5];   // lr_group / (1 - beta1^t)
  float beta1, beta2, eps;
  float inv_bc2_sqrt;    // 1 / sqrt(1 - beta2^t)
  int store_dtype;       // dtype of the `store` copies (kF16 / kBF16 / kF32)
};

// torch.optim.Adam (no weight decay, no amsgrad), then the reference's clamp:
//   m <- b1 m + (1-b1) g;  v <- b2 v + (1-b2) g^2
//   p <- p - lr/(1-b1^t) * m / (sqrt(v)/sqrt(1-b2^t) + eps)
// blockIdx.y selects the parameter group.
static __global__ __launch_bounds__(256) void k_adam(AdamArgs A) {
  const AdamGroup& G = A.g[blockIdx.y];
  const float ss = A.step_scale[blockIdx.y];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < G.n;
       i += (long long)gridDim.x * blockDim.x) {
    const long long gi = G.grow > 0 ? (i / G.gcomp) * G.grow + i % G.gcomp : i;
    const float g = G.g[gi];
    if (G.zero) G.g[gi] = 0.f;
    const float m = fmaf(A.beta1, G.m[i], (1.f - A.beta1) * g);
    const float v = fmaf(A.beta2, G.v[i], (1.f - A.beta2) * (g * g));
    G.m[i] = m;
    G.v[i] = v;
    float p = G.p[i] - ss * (m / (sqrtf(v) * A.inv_bc2_sqrt + A.eps));
    if (G.clamp) {
      const int c = G.stride == 2 ? (int)(i & 1) : 0;
      p = fminf(fmaxf(p, G.lo[c]), G.hi[c]);
    }
    G.p[i] = p;
    if (G.store) {
      if (A.store_dtype == kF16) ((__half*)G.store)[i] = __float2half(p);
      else if (A.store_dtype == kBF16) ((__hip_bfloat16*)G.store)[i] = __float2bfloat16(p);
      else ((float*)G.store)[i] = p;
    }
  }
}

// Rescale two gradient images in place from the factor they carry to *scale:
// x <- x * (*scale / applied[0]), applied[0] = 0 standing for 1.  Every thread
// reads both scalars; a ratio of 1 (the usual incoming gradient of a loss)
// leaves the images untouched and the launch is all it costs.  Otherwise the
// last workgroup to finish (a counter in applied[1], zero between calls)
// records the new factor: every workgroup has read applied[0] by then.
static __global__ __launch_bounds__(256) void k_rescale2(float* __restrict__ a, long long na,
                                                         float* __restrict__ b, long long nb,
                                                         const float* __restrict__ scale,
                                                         float* __restrict__ applied) {
  const float ap = applied[0];
  const float sc = *scale;
  const float f = sc / (ap == 0.f ? 1.f : ap);
  if (f == 1.f) return;   // the same on every thread
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nthr = (long long)gridDim.x * blockDim.x;
  for (long long i = tid; i < na; i += nthr) a[i] *= f;
  for (long long i = tid; i < nb; i += nthr) b[i] *= f;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned* cnt = reinterpret_cast<unsigned*>(applied + 1);
    if (atomicAdd(cnt, 1u) == gridDim.x - 1) {
      applied[0] = sc;
      *cnt = 0u;
    }
  }
}

}  // namespace dbmp
