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
static __global__ __launch_bounds__(256) void k_loss_mse(const float* __restrict__ color,
                                                  const float* __restrict__ alpha,
                                                  const float* __restrict__ bg, float bg_value,
                                                  const float* __restrict__ target, long long npix,
                                                  float inv_count, float* __restrict__ gC,
                                                  float* __restrict__ gA, float* __restrict__ loss) {
  float acc = 0.f;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < npix;
       p += (long long)gridDim.x * blockDim.x) {
    const float om = 1.f - alpha[p];
    float ga = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float b = bg ? bg[3 * p + c] : bg_value;
      const float r = fmaf(om, b, color[3 * p + c]) - target[3 * p + c];
      acc = fmaf(r, r, acc);
      const float g = 2.f * r * inv_count;
      gC[3 * p + c] = g;
      ga = fmaf(-g, b, ga);
    }
    gA[p] = ga;
  }
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

}  // namespace dbmp
