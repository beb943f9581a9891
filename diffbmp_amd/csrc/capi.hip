// extern "C" entry points of libdiffbmp_amd.so (declared in include/diffbmp_amd.h).
//
// Host side of the MI355X rasterizer: argument checks, per-handle scratch,
// launch sequencing on the caller's stream.  No host synchronisation, no
// device->host reads; sizes that depend on data (the pair count) stay on the
// device and kernels read them there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/diffbmp_amd.h"
#include "bin_kernels.h"
#include "host_common.h"
#include "raster_kernels.h"
#include "step_kernels.h"

using namespace dbmp;
using namespace dbmp_host;

namespace {

// Unpack the (N,8) gradient rows into the five reference arrays and clear the
// rows it read, so that the handle's rows hold zeros between calls.
__global__ void k_unpack_grads(float* __restrict__ g8, int N, float* __restrict__ gm,
                               float* __restrict__ gr, float* __restrict__ gt,
                               float* __restrict__ gv, float* __restrict__ gc) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const float4 a = ((const float4*)g8)[2 * n];
  const float4 b = ((const float4*)g8)[2 * n + 1];
  ((float4*)g8)[2 * n] = make_float4(0.f, 0.f, 0.f, 0.f);
  ((float4*)g8)[2 * n + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
  gm[2 * n] = a.x;
  gm[2 * n + 1] = a.y;
  gr[n] = a.z;
  gt[n] = a.w;
  gv[n] = b.x;
  gc[3 * n] = b.y;
  gc[3 * n + 1] = b.z;
  gc[3 * n + 2] = b.w;
}

// Scratch of dbmp_forward_noalloc: primitive records | template content
// boxes | texel quads or RGBA records, at fixed offsets.  One copy per GPU,
// held until the process ends; every call on a GPU shares it.
constexpr int kNoAllocMaxPrims = 65536;
constexpr int kNoAllocMaxTemplates = 4096;
constexpr size_t kNoAllocRecBytes = (size_t)kNoAllocMaxPrims * sizeof(dbmp::PrimRec);   // 4 MiB
constexpr size_t kNoAllocCullBytes = (size_t)kNoAllocMaxTemplates * sizeof(float4);     // 64 KiB
constexpr size_t kNoAllocQuadBytes = 16u << 20;                                         // 16 MiB
constexpr size_t kNoAllocBytes = kNoAllocRecBytes + kNoAllocCullBytes + kNoAllocQuadBytes;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (the host-side symbol is
// only a stand-in): looked up once per device with hipGetSymbolAddress.
int scratch_address(unsigned char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(DBMP_E_ARG, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(DBMP_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

static __global__ void k_sincos(const float* __restrict__ x, float* __restrict__ s,
                                float* __restrict__ c, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dbmp_sincosf(x[i], s + i, c + i);
}

}  // namespace

struct dbmp_rasterizer {
  int H = 0, W = 0, ts = 16, tilesX = 0, tilesY = 0, tiles = 0, subs = 1;
  float alpha_ub = 1.f;
  int device = 0;
  // primitive-sized scratch
  int n_cap = 0;
  int recs_n = -1;   // N of the records currently in `recs` (-1: none)
  PrimRec* recs = nullptr;
  int* counts = nullptr;
  long long* offs = nullptr;
  int4* ranges = nullptr;
  // (N,8) gradient rows of the calls that unpack into the reference arrays:
  // zeros between calls (k_unpack_grads clears what it reads); false after an
  // allocation, or a call that failed between its raster launch and its
  // unpack, and the next call zero-fills first
  float* grad8 = nullptr;
  bool grad8_clean = false;
  // pair-sized scratch
  long long pair_cap = 0;
  int* keysA = nullptr;
  int* keysB = nullptr;
  int* valsB = nullptr;
  int hist_blocks = 0;
  int* hist = nullptr;
  int* totals = nullptr;
  // per-block forward costs and the heaviest-first block order derived from
  // them (valid for order_n blocks; 0: none yet)
  int blk_cap = 0;
  int* blk_cost = nullptr;
  int* blk_order = nullptr;
  int order_n = 0;
  int order_age = 0;          // forwards since the order was last rebuilt
  // texel quads cache
  void* quads = nullptr;
  size_t quads_bytes = 0;     // allocated
  size_t quads_used = 0;      // bytes of the current bitmaps' records
  const void* q_src = nullptr;
  long long q_version = -2;
  int q_T = 0, q_Ht = 0, q_Wt = 0, q_dt = -1;
  // the quads hold RGBA records (alpha + per-template colour maps)
  bool q_rgba = false;
  const void* q_csrc = nullptr;
  long long q_cversion = -2;
  // per-template content boxes of the templates the quads hold (culling)
  int cull_cap = 0;
  float4* cull = nullptr;

  // timing
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_fwd, ev_bwd;
  double fwd_ms = 0.0, bwd_ms = 0.0;
  long long fwd_n = 0, bwd_n = 0;
};

namespace {

int ensure_n(dbmp_rasterizer* h, int N) {
  if (N <= h->n_cap) return 0;
  const int cap = std::max(N, h->n_cap + h->n_cap / 4);
  int rc;
  if ((rc = grow(&h->recs, cap))) return rc;
  if ((rc = grow(&h->counts, cap))) return rc;
  // offsets (cap + 1), then the scan's chunk sums
  if ((rc = grow(&h->offs, (long long)cap + 1 + scan_partials_len(cap)))) return rc;
  if ((rc = grow(&h->ranges, cap))) return rc;
  if ((rc = grow(&h->grad8, 8LL * cap))) return rc;
  h->grad8_clean = false;
  h->n_cap = cap;
  return 0;
}

int ensure_pairs(dbmp_rasterizer* h, long long cap) {
  if (cap <= h->pair_cap) return 0;
  int rc;
  if ((rc = grow(&h->keysA, cap))) return rc;
  if ((rc = grow(&h->keysB, cap))) return rc;
  if ((rc = grow(&h->valsB, cap))) return rc;
  const int nblk = ceil_div(cap, kRadixTile);
  if ((rc = grow(&h->hist, 256LL * nblk))) return rc;
  if (!h->totals && (rc = grow(&h->totals, 256))) return rc;
  h->pair_cap = cap;
  h->hist_blocks = nblk;
  return 0;
}

// Per-block cost and order buffers for nb blocks.
int ensure_blocks(dbmp_rasterizer* h, int nb) {
  if (nb <= h->blk_cap) return 0;
  int rc;
  if ((rc = grow(&h->blk_cost, nb))) return rc;
  if ((rc = grow(&h->blk_order, nb))) return rc;
  h->blk_cap = nb;
  h->order_n = 0;
  return 0;
}

// Raster launches of nb blocks run heaviest first once a forward of nb
// blocks has measured the costs -- when the bitmaps fit in one XCD's 4 MiB
// L2.  Larger bitmaps are L2-bound: there the natural order, which runs
// neighbouring tiles (sharing texels) together, is faster.
constexpr size_t kOrderMaxQuadBytes = 4u << 20;
constexpr int kOrderEvery = 8;
bool uses_block_order(const dbmp_rasterizer* h) { return h->quads_used <= kOrderMaxQuadBytes; }
const int* block_order(const dbmp_rasterizer* h, int nb) {
  return (h->order_n == nb && uses_block_order(h)) ? h->blk_order : nullptr;
}

// Before a call accumulates n values into h->grad8: zero them unless the rows
// are known to be zero; the call's k_unpack_grads leaves them zero again.
int grad8_begin(dbmp_rasterizer* h, long long n, hipStream_t st) {
  if (n > 0 && !h->grad8_clean) {
    hipLaunchKernelGGL(k_fill_f32, dim3(ceil_div(n, 256)), dim3(256), 0, st, h->grad8, n, 0.f);
    LAUNCH_CHECK("k_fill_f32");
  }
  h->grad8_clean = false;
  return 0;
}

int check_prims(const dbmp_prims* p) {
  if (!p) return fail(DBMP_E_ARG, "prims is NULL");
  if (p->num_primitives < 0) return fail(DBMP_E_ARG, "num_primitives < 0");
  if (p->dtype < DBMP_F32 || p->dtype > DBMP_BF16_APP)
    return fail(DBMP_E_ARG, "bad prims dtype %d", p->dtype);
  if (p->num_primitives > 0 && (!p->means2D || !p->radii || !p->rotations || !p->opacities ||
                                !p->colors || !p->bmp_sel))
    return fail(DBMP_E_ARG, "a primitive parameter pointer is NULL");
  return 0;
}

int check_bitmaps(const dbmp_bitmaps* b) {
  if (!b) return fail(DBMP_E_ARG, "bitmaps is NULL");
  if (!b->templates) return fail(DBMP_E_ARG, "templates is NULL");
  if (b->num_templates <= 0 || b->height < 2 || b->width < 2)
    return fail(DBMP_E_ARG, "templates must be (T>0, H>=2, W>=2), got (%d,%d,%d)",
                b->num_templates, b->height, b->width);
  if (b->dtype < DBMP_F32 || b->dtype > DBMP_BF16) return fail(DBMP_E_ARG, "bad bitmap dtype %d", b->dtype);
  // the raster kernels address the texel records with 32-bit byte offsets
  const bool rgba = b->colors_orig_mode == 2 && b->colors_orig != nullptr && b->c_blend > 0.f;
  if (texel_record_bytes(b->num_templates, b->height, b->width, b->dtype, rgba) >= (1LL << 32))
    return fail(DBMP_E_ARG, "the texel records of %d templates of %dx%d reach 4 GiB",
                b->num_templates, b->height, b->width);
  if (b->colors_orig_mode < 0 || b->colors_orig_mode > 2)
    return fail(DBMP_E_ARG, "bad colors_orig_mode %d", b->colors_orig_mode);
  if (b->colors_orig_mode != 0 && !b->colors_orig) return fail(DBMP_E_ARG, "colors_orig is NULL");
  return 0;
}

// ranges: where the binning's tile ranges go (h->ranges when null); shift:
// log2 of the super-tile factor the pair counts are for.
// recs_at: write the records there instead of h->recs (candidate b of a
// batch at h->recs + b N); such records are never reused by later calls.
int launch_prep(dbmp_rasterizer* h, const dbmp_prims* p, const float* ext, int T, int Ht, int Wt,
                hipStream_t st, int4* ranges = nullptr, int shift = 0, PrimRec* recs_at = nullptr) {
  const int N = p->num_primitives;
  if (recs_at == nullptr && ext == nullptr && p->reuse_records && h->recs_n == N) return 0;
  h->recs_n = recs_at == nullptr ? N : -1;
  if (N == 0) return 0;
  const dim3 grid(ceil_div(N, 256)), block(256);
#define PREP_ARGS(GT, AT)                                                                      \
  N, (const GT*)p->means2D, (const GT*)p->radii, (const GT*)p->rotations,                      \
      (const AT*)p->opacities, (const AT*)p->colors, p->bmp_sel, ext, T, Ht, Wt, h->H, h->W,   \
      h->ts, h->tilesX, h->tilesY, h->alpha_ub, shift, recs_at ? recs_at : h->recs, h->counts, \
      ranges ? ranges : h->ranges
  switch (p->dtype) {
    case DBMP_F32:
      hipLaunchKernelGGL((k_prep<float, float>), grid, block, 0, st, PREP_ARGS(float, float));
      break;
    case DBMP_F16:
      hipLaunchKernelGGL((k_prep<__half, __half>), grid, block, 0, st, PREP_ARGS(__half, __half));
      break;
    case DBMP_BF16:
      hipLaunchKernelGGL((k_prep<__hip_bfloat16, __hip_bfloat16>), grid, block, 0, st,
                         PREP_ARGS(__hip_bfloat16, __hip_bfloat16));
      break;
    case DBMP_F16_APP:
      hipLaunchKernelGGL((k_prep<float, __half>), grid, block, 0, st, PREP_ARGS(float, __half));
      break;
    default:
      hipLaunchKernelGGL((k_prep<float, __hip_bfloat16>), grid, block, 0, st,
                         PREP_ARGS(float, __hip_bfloat16));
      break;
  }
#undef PREP_ARGS
  LAUNCH_CHECK("k_prep");
  return 0;
}

// Texel quads of the templates, or RGBA records when the call blends
// per-template colour maps (colors_orig_mode 2, c_blend > 0).  Rebuilt when
// the source pointers, versions (-1: every call), shape, type or kind change.
bool wants_rgba(const dbmp_bitmaps* b) {
  return b->colors_orig_mode == 2 && b->colors_orig != nullptr && b->c_blend > 0.f;
}

int ensure_quads(dbmp_rasterizer* h, const dbmp_bitmaps* b, hipStream_t st) {
  const bool rgba = wants_rgba(b);
  const size_t bytes = (size_t)texel_record_bytes(b->num_templates, b->height, b->width, b->dtype, rgba);
  if (b->version >= 0 && h->q_src == b->templates && h->q_version == b->version &&
      h->q_T == b->num_templates && h->q_Ht == b->height && h->q_Wt == b->width &&
      h->q_dt == b->dtype && h->q_rgba == rgba &&
      (!rgba || (b->colors_version >= 0 && h->q_csrc == b->colors_orig &&
                 h->q_cversion == b->colors_version)))
    return 0;
  if (bytes > h->quads_bytes) {
    if (h->quads) (void)hipFree(h->quads);
    h->quads = nullptr;
    h->quads_bytes = 0;
    hipError_t e = hipMalloc(&h->quads, bytes);
    if (e != hipSuccess) return fail(DBMP_E_ALLOC, "hipMalloc quads: %s", hipGetErrorString(e));
    h->quads_bytes = bytes;
  }
  int rc = rgba ? launch_build_rgba(b, h->quads, st) : launch_build_quads(b, h->quads, st);
  if (rc) return rc;
  // content boxes from the same texels the quads hold
  if (b->num_templates > h->cull_cap) {
    if ((rc = grow(&h->cull, b->num_templates))) return rc;
    h->cull_cap = b->num_templates;
  }
  const dim3 cgrid(b->num_templates), cblock(256);
  if (b->dtype == DBMP_F32)
    hipLaunchKernelGGL((k_template_extents<float, true>), cgrid, cblock, 0, st,
                       (const float*)b->templates, b->height, b->width, (float*)h->cull);
  else if (b->dtype == DBMP_F16)
    hipLaunchKernelGGL((k_template_extents<__half, true>), cgrid, cblock, 0, st,
                       (const __half*)b->templates, b->height, b->width, (float*)h->cull);
  else
    hipLaunchKernelGGL((k_template_extents<__hip_bfloat16, true>), cgrid, cblock, 0, st,
                       (const __hip_bfloat16*)b->templates, b->height, b->width, (float*)h->cull);
  LAUNCH_CHECK("k_template_extents");
  h->q_src = b->templates;
  h->q_version = b->version;
  h->q_T = b->num_templates;
  h->q_Ht = b->height;
  h->q_Wt = b->width;
  h->q_dt = b->dtype;
  h->q_rgba = rgba;
  h->q_csrc = b->colors_orig;
  h->q_cversion = b->colors_version;
  h->quads_used = bytes;
  return 0;
}

RasterArgs make_args(dbmp_rasterizer* h, const dbmp_prims* p, const dbmp_bitmaps* b,
                     const dbmp_tile_lists* L, float eps) {
  RasterArgs A;
  A.recs = h->recs;
  A.toffs = L->tile_offsets;
  A.tidx = L->tile_indices;
  A.pcount = (const long long*)L->pair_count;
  A.pcap = L->capacity;
  A.quads = h->quads;
  const bool blend = b->colors_orig_mode != 0 && b->c_blend > 0.f;
  // per-template maps are sampled through their colour-map quads
  // per-template maps are sampled from the RGBA records in `quads`
  A.cmap = !blend ? nullptr : (b->colors_orig_mode == 2 ? h->quads : b->colors_orig);
  A.cmap_mode = (blend && b->colors_orig_mode == 2) ? 2 : 1;
  A.ranges = nullptr;
  A.bin_shift = 0;
  A.order = nullptr;
  A.cost = nullptr;
  A.cull = h->cull;   // built with the quads (ensure_quads runs first)
  A.ntmpl = b->num_templates;
  A.c_blend = b->c_blend;
  A.exact_blend = b->blend_grad_exact != 0;
  A.N = p->num_primitives;
  A.Ht = b->height;
  A.Wt = b->width;
  A.HWm1f = (float)(b->height * b->width - 1);
  A.Wtf = (float)b->width;
  A.hsx = 0.5f * (float)(b->width - 1);
  A.hsy = 0.5f * (float)(b->height - 1);
  A.H = h->H;
  A.W = h->W;
  A.ts = h->ts;
  A.tilesX = h->tilesX;
  A.tiles = h->tiles;
  A.subs = h->subs;
  A.eps = eps;
  const int f = L->bin_factor;
  if (f <= 1 && L->pair_count != nullptr) {
    // lists from dbmp_bin: after an overflow every primitive is walked, and
    // the binning's ranges keep each one to the tiles whose lists hold it
    A.ranges = L->ranges != nullptr ? (const int4*)L->ranges : h->ranges;
  }
  if (f > 1) {   // super-tile lists (checked by check_lists)
    int s = 0;
    while ((1 << s) < f) ++s;
    A.ranges = (const int4*)L->ranges;
    A.bin_shift = s;
    A.ts = h->ts * f;
    A.tilesX = (h->tilesX + f - 1) / f;
    A.tiles = A.tilesX * ((h->tilesY + f - 1) / f);
    A.subs = f * f;
  }
  // push masks: lists with a pair count and capacity bound the mask words
  A.pmask = nullptr;
  if (L->push_masks != nullptr && L->pair_count != nullptr &&
      L->push_mask_words >= push_mask_words(L->capacity, A.tiles, A.subs))
    A.pmask = (unsigned long long*)L->push_masks;
  return A;
}

int check_lists(dbmp_rasterizer* h, const dbmp_tile_lists* L) {
  if (!L || !L->tile_offsets) return fail(DBMP_E_ARG, "tile lists are NULL");
  const int f = L->bin_factor;
  if (f > 1) {
    if (f != 2 && f != 4 && f != 8) return fail(DBMP_E_ARG, "bin_factor must be 1, 2, 4 or 8");
    if (h->ts != kSub) return fail(DBMP_E_ARG, "bin_factor > 1 needs tile size 16");
    if (!L->ranges) return fail(DBMP_E_ARG, "bin_factor > 1 needs the binning's ranges");
  } else if (f < 0) {
    return fail(DBMP_E_ARG, "bad bin_factor %d", f);
  }
  return 0;
}

void timing_begin(dbmp_rasterizer* h, std::vector<std::pair<hipEvent_t, hipEvent_t>>& v,
                  hipStream_t st) {
  if (!h->timing) return;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return;
  if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
  (void)hipEventRecord(a, st);
  v.emplace_back(a, b);
}

void timing_end(dbmp_rasterizer* h, std::vector<std::pair<hipEvent_t, hipEvent_t>>& v,
                hipStream_t st) {
  if (!h->timing || v.empty()) return;
  (void)hipEventRecord(v.back().second, st);
}

void timing_drain(std::vector<std::pair<hipEvent_t, hipEvent_t>>& v, double& ms, long long& n) {
  for (auto& e : v) {
    float t = 0.f;
    if (hipEventSynchronize(e.second) == hipSuccess && hipEventElapsedTime(&t, e.first, e.second) == hipSuccess) {
      ms += t;
      ++n;
    }
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  v.clear();
}

}  // namespace

extern "C" {

int dbmp_version(void) { return 100; }

int64_t dbmp_push_mask_words(const dbmp_rasterizer* h, int64_t capacity, int32_t bin_factor) {
  if (!h || capacity < 0) return 0;
  const int f = bin_factor > 1 ? bin_factor : 1;
  const int lists = ((h->tilesX + f - 1) / f) * ((h->tilesY + f - 1) / f);
  return push_mask_words(capacity, lists, h->subs * f * f);
}

const char* dbmp_last_error(void) { return g_err.c_str(); }

int dbmp_init(dbmp_rasterizer** out, int image_height, int image_width, int tile_size, float sigma,
              float alpha_upper_bound, int max_prims_per_pixel, int num_primitives) {
  (void)sigma;
  (void)max_prims_per_pixel;
  if (!out) return fail(DBMP_E_ARG, "out is NULL");
  *out = nullptr;
  if (image_height <= 0 || image_width <= 0)
    return fail(DBMP_E_ARG, "image size must be positive, got %dx%d", image_height, image_width);
  if (!(alpha_upper_bound >= 0.f))
    return fail(DBMP_E_ARG, "alpha_upper_bound must be >= 0, got %g", (double)alpha_upper_bound);
  if (tile_size <= 0 || tile_size % kSub != 0)
    return fail(DBMP_E_ARG, "tile_size must be a positive multiple of %d, got %d", kSub, tile_size);
  auto* h = new dbmp_rasterizer();
  h->H = image_height;
  h->W = image_width;
  h->ts = tile_size;
  h->tilesX = (image_width + tile_size - 1) / tile_size;
  h->tilesY = (image_height + tile_size - 1) / tile_size;
  h->tiles = h->tilesX * h->tilesY;
  h->subs = (tile_size / kSub) * (tile_size / kSub);
  h->alpha_ub = alpha_upper_bound;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return fail(DBMP_E_HIP, "hipGetDevice failed");
  }
  if (num_primitives > 0) {
    int rc = ensure_n(h, num_primitives);
    if (rc) {
      dbmp_destroy(h);
      return rc;
    }
  }
  *out = h;
  return 0;
}

int dbmp_destroy(dbmp_rasterizer* h) {
  if (!h) return 0;
  double ms = 0;
  long long n = 0;
  timing_drain(h->ev_fwd, ms, n);
  timing_drain(h->ev_bwd, ms, n);
  void* ptrs[] = {h->recs, h->counts, h->offs, h->ranges, h->grad8, h->keysA,
                  h->keysB, h->valsB, h->hist, h->totals, h->quads, h->blk_cost, h->blk_order,
                  h->cull};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete h;
  return 0;
}

int dbmp_template_extents(dbmp_rasterizer* h, const dbmp_bitmaps* bm, float* ext_out, void* stream) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  int rc = check_bitmaps(bm);
  if (rc) return rc;
  if (!ext_out) return fail(DBMP_E_ARG, "ext_out is NULL");
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(bm->num_templates), block(256);
  if (bm->dtype == DBMP_F32)
    hipLaunchKernelGGL(k_template_extents<float>, grid, block, 0, st, (const float*)bm->templates,
                       bm->height, bm->width, ext_out);
  else if (bm->dtype == DBMP_F16)
    hipLaunchKernelGGL(k_template_extents<__half>, grid, block, 0, st, (const __half*)bm->templates,
                       bm->height, bm->width, ext_out);
  else
    hipLaunchKernelGGL(k_template_extents<__hip_bfloat16>, grid, block, 0, st,
                       (const __hip_bfloat16*)bm->templates, bm->height, bm->width, ext_out);
  LAUNCH_CHECK("k_template_extents");
  return 0;
}

// Lists built directly (k_tile_build) when there are at most
// this many lists and list x primitive tests; a radix sort otherwise.
constexpr int kDirectMaxLists = 1024;
constexpr long long kDirectMaxTests = 1LL << 23;

int dbmp_bin_tiles(dbmp_rasterizer* h, const dbmp_prims* prims, const float* ext, int num_templates,
                   int tmpl_h, int tmpl_w, int32_t bin_factor, int32_t* tile_offsets,
                   int32_t* tile_indices, int64_t capacity, int64_t* d_pair_count,
                   int32_t* ranges_out, void* stream) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  int rc = check_prims(prims);
  if (rc) return rc;
  if (!ext || !tile_offsets || !d_pair_count) return fail(DBMP_E_ARG, "NULL output/extents pointer");
  if (num_templates <= 0 || tmpl_h < 2 || tmpl_w < 2) return fail(DBMP_E_ARG, "bad template shape");
  if (capacity < 0 || capacity > (int64_t)0x7fffffff)
    return fail(DBMP_E_ARG, "capacity must be in [0, 2^31), got %lld", (long long)capacity);
  if (capacity > 0 && !tile_indices) return fail(DBMP_E_ARG, "tile_indices is NULL");
  const int f = bin_factor < 1 ? 1 : bin_factor;
  if (f != 1 && f != 2 && f != 4 && f != 8) return fail(DBMP_E_ARG, "bin_factor must be 1, 2, 4 or 8");
  if (f > 1 && h->ts != kSub) return fail(DBMP_E_ARG, "bin_factor > 1 needs tile size 16");
  if (f > 1 && !ranges_out) return fail(DBMP_E_ARG, "bin_factor > 1 needs ranges_out");
  int shift = 0;
  while ((1 << shift) < f) ++shift;
  const int tX = (h->tilesX + f - 1) / f, tY = (h->tilesY + f - 1) / f, tiles = tX * tY;
  hipStream_t st = (hipStream_t)stream;
  const int N = prims->num_primitives;
  if ((rc = ensure_n(h, std::max(N, 1)))) return rc;
  if ((rc = ensure_pairs(h, std::max<long long>(capacity, 1)))) return rc;
  long long* dtot = (long long*)d_pair_count;
  int4* rg = ranges_out ? (int4*)ranges_out : h->ranges;
  if ((rc = launch_prep(h, prims, ext, num_templates, tmpl_h, tmpl_w, st, rg, shift))) return rc;
  if (tiles <= kDirectMaxLists && (long long)tiles * N <= kDirectMaxTests) {
    // few lists x primitives: build each list directly from the ranges
    if (N <= kDirectStage)
      hipLaunchKernelGGL(k_tile_build<true>, dim3(tiles), dim3(64 * kBuildWaves), 0, st, (const int4*)rg, N, tX,
                         tiles, shift, (long long)capacity, tile_offsets, tile_indices, dtot);
    else
      hipLaunchKernelGGL(k_tile_build<false>, dim3(tiles), dim3(64 * kBuildWaves), 0, st, (const int4*)rg, N, tX,
                         tiles, shift, (long long)capacity, tile_offsets, tile_indices, dtot);
    LAUNCH_CHECK("k_tile_build");
    return 0;
  }
  scan_counts(h->counts, N, h->offs, dtot, h->offs + h->n_cap + 1, st);
  LAUNCH_CHECK("k_scan_apply");

  int bits = 0;
  while ((1 << bits) < tiles) ++bits;
  const int passes = (bits + 7) / 8;
  const int dbits = passes ? (bits + passes - 1) / passes : 0;
  const bool even = (passes % 2) == 0;
  int* kcur = h->keysA;
  int* vcur = even ? tile_indices : h->valsB;
  int* kalt = h->keysB;
  int* valt = even ? h->valsB : tile_indices;
  if (N > 0 && capacity > 0) {
    hipLaunchKernelGGL(k_emit, dim3(ceil_div(N, 4)), dim3(256), 0, st, N, h->counts, h->offs, rg, tX,
                       shift, dtot, (long long)capacity, kcur, vcur);
    LAUNCH_CHECK("k_emit");
    const int nblk = ceil_div(capacity, kRadixTile);
    for (int p = 0; p < passes; ++p) {
      const int sh = p * dbits;
      const int nb = std::min(dbits, bits - sh);
      hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(kRadixThreads), 0, st, kcur, dtot,
                         (long long)capacity, sh, nb, nblk, h->hist);
      LAUNCH_CHECK("k_radix_hist");
      hipLaunchKernelGGL(k_radix_rowscan, dim3(1 << nb), dim3(256), 0, st, h->hist, nblk, h->totals,
                         dtot, (long long)capacity);
      LAUNCH_CHECK("k_radix_rowscan");
      hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(kRadixThreads), 0, st, kcur, vcur, kalt,
                         valt, dtot, (long long)capacity, sh, nb, nblk, h->hist, h->totals);
      LAUNCH_CHECK("k_radix_scatter");
      std::swap(kcur, kalt);
      std::swap(vcur, valt);
    }
  }
  hipLaunchKernelGGL(k_tile_bounds, dim3(ceil_div(std::max<long long>(capacity, tiles) + 1, 256)),
                     dim3(256), 0, st, kcur, dtot, (long long)capacity, tiles, tile_offsets);
  LAUNCH_CHECK("k_tile_bounds");
  return 0;
}

int dbmp_bin(dbmp_rasterizer* h, const dbmp_prims* prims, const float* ext, int num_templates,
             int tmpl_h, int tmpl_w, int32_t* tile_offsets, int32_t* tile_indices,
             int64_t capacity, int64_t* d_pair_count, void* stream) {
  return dbmp_bin_tiles(h, prims, ext, num_templates, tmpl_h, tmpl_w, 1, tile_offsets, tile_indices,
                        capacity, d_pair_count, nullptr, stream);
}

#define DISPATCH_DT_BLEND(DT, BLEND, KERNEL, ...)                                            \
  do {                                                                                       \
    if ((DT) == DBMP_F32) {                                                                  \
      if (BLEND) hipLaunchKernelGGL((KERNEL<kF32, true>), __VA_ARGS__);                      \
      else hipLaunchKernelGGL((KERNEL<kF32, false>), __VA_ARGS__);                           \
    } else if ((DT) == DBMP_F16) {                                                           \
      if (BLEND) hipLaunchKernelGGL((KERNEL<kF16, true>), __VA_ARGS__);                      \
      else hipLaunchKernelGGL((KERNEL<kF16, false>), __VA_ARGS__);                           \
    } else {                                                                                 \
      if (BLEND) hipLaunchKernelGGL((KERNEL<kBF16, true>), __VA_ARGS__);                     \
      else hipLaunchKernelGGL((KERNEL<kBF16, false>), __VA_ARGS__);                          \
    }                                                                                        \
  } while (0)

int dbmp_forward(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                 const dbmp_tile_lists* lists, float eps_T, float* out_color, float* out_alpha,
                 float* out_T_final, int32_t* out_last, void* stream) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  int rc = check_prims(prims);
  if (rc) return rc;
  if ((rc = check_bitmaps(bm))) return rc;
  if ((rc = check_lists(h, lists))) return rc;
  if (!out_color || !out_alpha || !out_T_final || !out_last) return fail(DBMP_E_ARG, "NULL output");
  if (!(eps_T >= 0.f && eps_T < 1.f)) return fail(DBMP_E_ARG, "eps_T must be in [0, 1)");
  hipStream_t st = (hipStream_t)stream;
  const int N = prims->num_primitives;
  if ((rc = ensure_n(h, std::max(N, 1)))) return rc;
  timing_begin(h, h->ev_fwd, st);
  if ((rc = ensure_quads(h, bm, st))) return rc;
  // raster constants only (no extents => no tile ranges)
  if ((rc = launch_prep(h, prims, nullptr, bm->num_templates, bm->height, bm->width, st))) return rc;
  RasterArgs A = make_args(h, prims, bm, lists, eps_T);
  const int nb = A.tiles * A.subs;
  if ((rc = ensure_blocks(h, nb))) return rc;
  // the order is rebuilt from this forward's costs every kOrderEvery forwards
  // (costs change slowly from one call to the next; the order only balances
  // the launch, every order gives the same results)
  const bool refresh = uses_block_order(h) && (h->order_n != nb || h->order_age + 1 >= kOrderEvery);
  A.order = block_order(h, nb);
  A.cost = refresh ? h->blk_cost : nullptr;
  const bool blend = A.cmap != nullptr;
  DISPATCH_DT_BLEND(bm->dtype, blend, k_raster_fwd, dim3(nb), dim3(kThreads), 0, st, A,
                    out_color, out_alpha, out_T_final, out_last);
  LAUNCH_CHECK("k_raster_fwd");
  if (refresh) {
    // this forward's costs order the backward and the next forwards
    hipLaunchKernelGGL(k_block_order, dim3(1), dim3(kOrderThreads), 0, st, h->blk_cost, nb,
                       h->blk_order);
    LAUNCH_CHECK("k_block_order");
    h->order_n = nb;
    h->order_age = 0;
  } else {
    ++h->order_age;
  }
  timing_end(h, h->ev_fwd, st);
  return 0;
}

int dbmp_forward_noalloc(int image_height, int image_width, int tile_size, float alpha_upper_bound,
                         const dbmp_prims* prims, const dbmp_bitmaps* bm,
                         const dbmp_tile_lists* lists, float eps_T, float* out_color,
                         float* out_alpha, float* out_T_final, int32_t* out_last, void* stream) {
  int rc = check_prims(prims);
  if (rc) return rc;
  if ((rc = check_bitmaps(bm))) return rc;
  if (image_height <= 0 || image_width <= 0)
    return fail(DBMP_E_ARG, "image size must be positive, got %dx%d", image_height, image_width);
  if (tile_size <= 0 || tile_size % kSub != 0)
    return fail(DBMP_E_ARG, "tile_size must be a positive multiple of %d, got %d", kSub, tile_size);
  if (!(alpha_upper_bound >= 0.f))
    return fail(DBMP_E_ARG, "alpha_upper_bound must be >= 0, got %g", (double)alpha_upper_bound);
  if (!out_color || !out_alpha || !out_T_final || !out_last) return fail(DBMP_E_ARG, "NULL output");
  if (!(eps_T >= 0.f && eps_T < 1.f)) return fail(DBMP_E_ARG, "eps_T must be in [0, 1)");
  const int N = prims->num_primitives;
  if (N > kNoAllocMaxPrims)
    return fail(DBMP_E_ARG, "dbmp_forward_noalloc: %d primitives exceed its maximum of %d", N,
                kNoAllocMaxPrims);
  if (bm->num_templates > kNoAllocMaxTemplates)
    return fail(DBMP_E_ARG, "dbmp_forward_noalloc: %d templates exceed its maximum of %d",
                bm->num_templates, kNoAllocMaxTemplates);
  const bool rgba = wants_rgba(bm);
  const size_t qbytes = (size_t)texel_record_bytes(bm->num_templates, bm->height, bm->width,
                                                   bm->dtype, rgba);
  if (qbytes > kNoAllocQuadBytes)
    return fail(DBMP_E_ARG, "dbmp_forward_noalloc: texel records of %zu bytes exceed its maximum of %zu",
                qbytes, kNoAllocQuadBytes);
  if (!lists || !lists->tile_offsets) return fail(DBMP_E_ARG, "tile lists are NULL");
  const int f = lists->bin_factor;
  if (f > 1 && tile_size != kSub) return fail(DBMP_E_ARG, "bin_factor > 1 needs tile size 16");
  if ((f > 1 || lists->pair_count) && !lists->ranges)
    return fail(DBMP_E_ARG, "dbmp_forward_noalloc: lists with a pair count or bin_factor > 1 need "
                "their binning's ranges");
  if (f < 0 || (f > 1 && f != 2 && f != 4 && f != 8)) return fail(DBMP_E_ARG, "bad bin_factor %d", f);
  unsigned char* base = nullptr;
  if ((rc = scratch_address(&base))) return rc;
  hipStream_t st = (hipStream_t)stream;
  // a handle on the stack whose scratch pointers address the static array
  // (nothing below allocates: no ensure_* call)
  dbmp_rasterizer h;
  h.H = image_height;
  h.W = image_width;
  h.ts = tile_size;
  h.tilesX = (image_width + tile_size - 1) / tile_size;
  h.tilesY = (image_height + tile_size - 1) / tile_size;
  h.tiles = h.tilesX * h.tilesY;
  h.subs = (tile_size / kSub) * (tile_size / kSub);
  h.alpha_ub = alpha_upper_bound;
  h.recs = (PrimRec*)base;
  h.cull = (float4*)(base + kNoAllocRecBytes);
  h.quads = base + kNoAllocRecBytes + kNoAllocCullBytes;
  h.quads_used = qbytes;
  // records, texel records and content boxes of this call (the previous
  // call's contents are never reused)
  dbmp_prims p = *prims;
  p.reuse_records = 0;
  if ((rc = launch_prep(&h, &p, nullptr, bm->num_templates, bm->height, bm->width, st))) return rc;
  if ((rc = rgba ? launch_build_rgba(bm, h.quads, st) : launch_build_quads(bm, h.quads, st))) return rc;
  const dim3 cgrid(bm->num_templates), cblock(256);
  if (bm->dtype == DBMP_F32)
    hipLaunchKernelGGL((k_template_extents<float, true>), cgrid, cblock, 0, st,
                       (const float*)bm->templates, bm->height, bm->width, (float*)h.cull);
  else if (bm->dtype == DBMP_F16)
    hipLaunchKernelGGL((k_template_extents<__half, true>), cgrid, cblock, 0, st,
                       (const __half*)bm->templates, bm->height, bm->width, (float*)h.cull);
  else
    hipLaunchKernelGGL((k_template_extents<__hip_bfloat16, true>), cgrid, cblock, 0, st,
                       (const __hip_bfloat16*)bm->templates, bm->height, bm->width, (float*)h.cull);
  LAUNCH_CHECK("k_template_extents");
  RasterArgs A = make_args(&h, &p, bm, lists, eps_T);
  A.order = nullptr;
  A.cost = nullptr;
  const int nb = A.tiles * A.subs;
  const bool blend = A.cmap != nullptr;
  DISPATCH_DT_BLEND(bm->dtype, blend, k_raster_fwd, dim3(nb), dim3(kThreads), 0, st, A, out_color,
                    out_alpha, out_T_final, out_last);
  LAUNCH_CHECK("k_raster_fwd");
  return 0;
}

// Shared checks of the candidate-batch calls: B >= 1 candidates of one N and
// dtype, lists that carry their ranges wherever an overflow walk could need them.
static int check_candidates(dbmp_rasterizer* h, int32_t B, const dbmp_prims* prims,
                            const dbmp_tile_lists* lists) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  if (B < 1 || !prims || !lists) return fail(DBMP_E_ARG, "need B >= 1 candidates with prims and lists");
  int rc;
  for (int b = 0; b < B; ++b) {
    if ((rc = check_prims(&prims[b]))) return rc;
    if ((rc = check_lists(h, &lists[b]))) return rc;
    if (prims[b].num_primitives != prims[0].num_primitives || prims[b].dtype != prims[0].dtype)
      return fail(DBMP_E_ARG, "candidate %d: every candidate needs the same N and dtype", b);
    if (lists[b].bin_factor != lists[0].bin_factor)
      return fail(DBMP_E_ARG, "candidate %d: every candidate needs the same bin_factor", b);
    if (lists[b].pair_count && !lists[b].ranges)
      return fail(DBMP_E_ARG, "candidate %d: lists with a pair count need their ranges", b);
  }
  if ((long long)B * prims[0].num_primitives >= (1LL << 30))
    return fail(DBMP_E_ARG, "B * N too large");
  return 0;
}

// Per-candidate table of a launch (at most kMaxCands candidates).
static CandTable cand_table(const dbmp_tile_lists* lists, const dbmp_bitmaps* bm, int b0, int Bc,
                            long long cmap_stride) {
  CandTable C;
  for (int i = 0; i < kMaxCands; ++i) {
    const int b = b0 + (i < Bc ? i : 0);
    const dbmp_tile_lists& L = lists[b];
    C.toffs[i] = L.tile_offsets;
    C.tidx[i] = L.tile_indices;
    C.pcount[i] = (const long long*)L.pair_count;
    C.pcap[i] = L.capacity;
    C.ranges[i] = (const int4*)L.ranges;
    C.cmap[i] = bm->colors_orig_mode == 1 && bm->colors_orig
                    ? (const void*)((const char*)bm->colors_orig + cmap_stride * b)
                    : nullptr;
  }
  return C;
}

// Bytes between candidates' per-primitive colour maps (mode 1): each
// candidate's (N,Ht,Wt,3) block in the bitmaps' dtype.
static long long cmap_cand_stride(const dbmp_bitmaps* bm, int N) {
  return (long long)N * bm->height * bm->width * 3 * (bm->dtype == DBMP_F32 ? 4 : 2);
}

int dbmp_forward_batch(dbmp_rasterizer* h, int32_t B, const dbmp_prims* prims,
                       const dbmp_bitmaps* bm, const dbmp_tile_lists* lists, float eps_T,
                       float* out_color, float* out_alpha, float* out_T_final, int32_t* out_last,
                       void* stream) {
  int rc = check_candidates(h, B, prims, lists);
  if (rc) return rc;
  if ((rc = check_bitmaps(bm))) return rc;
  if (!out_color || !out_alpha || !out_T_final || !out_last) return fail(DBMP_E_ARG, "NULL output");
  if (!(eps_T >= 0.f && eps_T < 1.f)) return fail(DBMP_E_ARG, "eps_T must be in [0, 1)");
  hipStream_t st = (hipStream_t)stream;
  const int N = prims[0].num_primitives;
  if ((rc = ensure_n(h, std::max(B * N, 1)))) return rc;
  timing_begin(h, h->ev_fwd, st);
  if ((rc = ensure_quads(h, bm, st))) return rc;
  for (int b = 0; b < B; ++b)
    if ((rc = launch_prep(h, &prims[b], nullptr, bm->num_templates, bm->height, bm->width, st,
                          nullptr, 0, h->recs + (long long)b * N)))
      return rc;
  RasterArgs A = make_args(h, &prims[0], bm, &lists[0], eps_T);
  A.order = nullptr;
  A.cost = nullptr;
  A.pmask = nullptr;   // push masks are per canvas; candidates cull
  const int nb = A.tiles * A.subs;
  const long long px = (long long)h->H * h->W;
  const bool blend = A.cmap != nullptr;
  const long long cst = cmap_cand_stride(bm, N);
  for (int b0 = 0; b0 < B; b0 += kMaxCands) {
    const int Bc = std::min(kMaxCands, B - b0);
    RasterArgs Ab = A;
    Ab.recs = h->recs + (long long)b0 * N;
    const CandTable C = cand_table(lists, bm, b0, Bc, cst);
    DISPATCH_DT_BLEND(bm->dtype, blend, k_raster_fwd_cands, dim3(nb * Bc), dim3(kThreads), 0, st, Ab,
                      C, nb, out_color + 3 * px * b0, out_alpha + px * b0, out_T_final + px * b0,
                      out_last + px * b0);
    LAUNCH_CHECK("k_raster_fwd_cands");
  }
  timing_end(h, h->ev_fwd, st);
  return 0;
}

int dbmp_backward_batch(dbmp_rasterizer* h, int32_t B, const dbmp_prims* prims,
                        const dbmp_bitmaps* bm, const dbmp_tile_lists* lists, float eps_T,
                        const float* grad_color, const float* grad_alpha, const float* out_color,
                        const float* out_T_final, const int32_t* out_last, float* grad_means2D,
                        float* grad_radii, float* grad_rotations, float* grad_opacities,
                        float* grad_colors, void* stream) {
  (void)eps_T;
  int rc = check_candidates(h, B, prims, lists);
  if (rc) return rc;
  if ((rc = check_bitmaps(bm))) return rc;
  if (!grad_color || !out_color || !out_T_final || !out_last) return fail(DBMP_E_ARG, "NULL input");
  // zero primitives: the gradient outputs are empty (and may be NULL)
  if (prims[0].num_primitives > 0 &&
      (!grad_means2D || !grad_radii || !grad_rotations || !grad_opacities || !grad_colors))
    return fail(DBMP_E_ARG, "NULL gradient output");
  hipStream_t st = (hipStream_t)stream;
  const int N = prims[0].num_primitives;
  const long long BN = (long long)B * N;
  if ((rc = ensure_n(h, (int)std::max<long long>(BN, 1)))) return rc;
  timing_begin(h, h->ev_bwd, st);
  if ((rc = ensure_quads(h, bm, st))) return rc;
  for (int b = 0; b < B; ++b)
    if ((rc = launch_prep(h, &prims[b], nullptr, bm->num_templates, bm->height, bm->width, st,
                          nullptr, 0, h->recs + (long long)b * N)))
      return rc;
  if ((rc = grad8_begin(h, 8 * BN, st))) return rc;
  RasterArgs A = make_args(h, &prims[0], bm, &lists[0], 0.f);
  A.order = nullptr;
  A.cost = nullptr;
  A.pmask = nullptr;   // push masks are per canvas; candidates cull
  const int nb = A.tiles * A.subs;
  const long long px = (long long)h->H * h->W;
  const bool blend = A.cmap != nullptr;
  const long long cst = cmap_cand_stride(bm, N);
  for (int b0 = 0; b0 < B; b0 += kMaxCands) {
    const int Bc = std::min(kMaxCands, B - b0);
    RasterArgs Ab = A;
    Ab.recs = h->recs + (long long)b0 * N;
    const CandTable C = cand_table(lists, bm, b0, Bc, cst);
    DISPATCH_DT_BLEND(bm->dtype, blend, k_raster_bwd_cands, dim3(nb * Bc), dim3(kThreads), 0, st, Ab,
                      C, nb, grad_color + 3 * px * b0, grad_alpha ? grad_alpha + px * b0 : nullptr,
                      out_color + 3 * px * b0, out_T_final + px * b0, out_last + px * b0,
                      h->grad8 + 8LL * N * b0);
    LAUNCH_CHECK("k_raster_bwd_cands");
  }
  if (BN > 0) {
    hipLaunchKernelGGL(k_unpack_grads, dim3(ceil_div(BN, 256)), dim3(256), 0, st, h->grad8, (int)BN,
                       grad_means2D, grad_radii, grad_rotations, grad_opacities, grad_colors);
    LAUNCH_CHECK("k_unpack_grads");
  }
  h->grad8_clean = true;
  timing_end(h, h->ev_bwd, st);
  return 0;
}

int dbmp_backward(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                  const dbmp_tile_lists* lists, float eps_T, const float* grad_color,
                  const float* grad_alpha, const float* out_color, const float* out_T_final,
                  const int32_t* out_last, float* grad_means2D, float* grad_radii,
                  float* grad_rotations, float* grad_opacities, float* grad_colors, void* stream) {
  (void)eps_T;
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  int rc = check_prims(prims);
  if (rc) return rc;
  if ((rc = check_bitmaps(bm))) return rc;
  if ((rc = check_lists(h, lists))) return rc;
  if (!grad_color || !out_color || !out_T_final || !out_last) return fail(DBMP_E_ARG, "NULL input");
  // zero primitives: the gradient outputs are empty (and may be NULL)
  if (prims->num_primitives > 0 &&
      (!grad_means2D || !grad_radii || !grad_rotations || !grad_opacities || !grad_colors))
    return fail(DBMP_E_ARG, "NULL gradient output");
  hipStream_t st = (hipStream_t)stream;
  const int N = prims->num_primitives;
  if ((rc = ensure_n(h, std::max(N, 1)))) return rc;
  timing_begin(h, h->ev_bwd, st);
  if ((rc = ensure_quads(h, bm, st))) return rc;
  if ((rc = launch_prep(h, prims, nullptr, bm->num_templates, bm->height, bm->width, st))) return rc;
  if ((rc = grad8_begin(h, 8LL * N, st))) return rc;
  RasterArgs A = make_args(h, prims, bm, lists, 0.f);
  A.order = block_order(h, A.tiles * A.subs);
  const bool blend = A.cmap != nullptr;
  DISPATCH_DT_BLEND(bm->dtype, blend, k_raster_bwd, dim3(A.tiles * A.subs), dim3(kThreads), 0, st, A,
                    grad_color, grad_alpha, out_color, out_T_final, out_last, h->grad8);
  LAUNCH_CHECK("k_raster_bwd");
  if (N > 0) {
    hipLaunchKernelGGL(k_unpack_grads, dim3(ceil_div(N, 256)), dim3(256), 0, st, h->grad8, N,
                       grad_means2D, grad_radii, grad_rotations, grad_opacities, grad_colors);
    LAUNCH_CHECK("k_unpack_grads");
  }
  h->grad8_clean = true;
  timing_end(h, h->ev_bwd, st);
  return 0;
}

// Shared body of dbmp_backward_mse (rows = null: internal rows, zeroed first,
// then unpacked) and dbmp_backward_mse_rows (accumulate into the caller's rows).
static int backward_mse_impl(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                      const dbmp_tile_lists* lists, const float* target, const float* background,
                      float background_value, float* loss_accum, const float* out_color,
                      const float* out_T_final, const int32_t* out_last, float* rows,
                      float* grad_means2D, float* grad_radii, float* grad_rotations,
                      float* grad_opacities, float* grad_colors, hipStream_t st) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  int rc = check_prims(prims);
  if (rc) return rc;
  if ((rc = check_bitmaps(bm))) return rc;
  if ((rc = check_lists(h, lists))) return rc;
  if (!target || !out_color || !out_T_final || !out_last) return fail(DBMP_E_ARG, "NULL input");
  if (prims->num_primitives > 0 && !rows &&
      (!grad_means2D || !grad_radii || !grad_rotations || !grad_opacities || !grad_colors))
    return fail(DBMP_E_ARG, "NULL gradient output");
  const int N = prims->num_primitives;
  if ((rc = ensure_n(h, std::max(N, 1)))) return rc;
  timing_begin(h, h->ev_bwd, st);
  if ((rc = ensure_quads(h, bm, st))) return rc;
  if ((rc = launch_prep(h, prims, nullptr, bm->num_templates, bm->height, bm->width, st))) return rc;
  if (!rows && (rc = grad8_begin(h, 8LL * N, st))) return rc;
  RasterArgs A = make_args(h, prims, bm, lists, 0.f);
  A.order = block_order(h, A.tiles * A.subs);
  LossArgs L;
  L.target = target;
  L.bg = background;
  L.bg_value = background_value;
  L.inv_count = 1.f / (float)(3.0 * h->H * h->W);
  L.loss = loss_accum;
  const bool blend = A.cmap != nullptr;
  DISPATCH_DT_BLEND(bm->dtype, blend, k_raster_bwd_mse, dim3(A.tiles * A.subs), dim3(kThreads), 0, st,
                    A, L, out_color, out_T_final, out_last, rows ? rows : h->grad8);
  LAUNCH_CHECK("k_raster_bwd_mse");
  if (N > 0 && !rows) {
    hipLaunchKernelGGL(k_unpack_grads, dim3(ceil_div(N, 256)), dim3(256), 0, st, h->grad8, N,
                       grad_means2D, grad_radii, grad_rotations, grad_opacities, grad_colors);
    LAUNCH_CHECK("k_unpack_grads");
  }
  if (!rows) h->grad8_clean = true;
  timing_end(h, h->ev_bwd, st);
  return 0;
}

int dbmp_backward_mse(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                      const dbmp_tile_lists* lists, float eps_T, const float* target,
                      const float* background, float background_value, float* loss_accum,
                      const float* out_color, const float* out_T_final, const int32_t* out_last,
                      float* grad_means2D, float* grad_radii, float* grad_rotations,
                      float* grad_opacities, float* grad_colors, void* stream) {
  (void)eps_T;
  return backward_mse_impl(h, prims, bm, lists, target, background, background_value, loss_accum,
                           out_color, out_T_final, out_last, nullptr, grad_means2D, grad_radii,
                           grad_rotations, grad_opacities, grad_colors, (hipStream_t)stream);
}

int dbmp_backward_mse_rows(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                           const dbmp_tile_lists* lists, float eps_T, const float* target,
                           const float* background, float background_value, float* loss_accum,
                           const float* out_color, const float* out_T_final,
                           const int32_t* out_last, float* grad_rows, void* stream) {
  (void)eps_T;
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  if (!grad_rows && prims && prims->num_primitives > 0) return fail(DBMP_E_ARG, "grad_rows is NULL");
  return backward_mse_impl(h, prims, bm, lists, target, background, background_value, loss_accum,
                           out_color, out_T_final, out_last, grad_rows ? grad_rows : h->grad8,
                           nullptr, nullptr, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int dbmp_loss_mse(dbmp_rasterizer* h, const float* color, const float* alpha, const float* bg,
                  float bg_value, const float* target, float* grad_color, float* grad_alpha,
                  float* loss_accum, void* stream) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  if (!color || !alpha || !target || !grad_color || !grad_alpha)
    return fail(DBMP_E_ARG, "NULL image pointer");
  const long long npix = (long long)h->H * h->W;
  const float inv = 1.f / (float)(3 * npix);
  const auto a16 = [](const void* q) { return ((uintptr_t)q & 15u) == 0; };
  const bool vec = a16(color) && a16(alpha) && a16(target) && a16(grad_color) && a16(grad_alpha) &&
                   (!bg || a16(bg));
  // kLossGroups float4 groups (4 pixels each) per thread on the vector path
  const long long work = vec ? ceil_div(std::max(npix >> 2, 1LL), (long long)kLossGroups) : npix;
  const int grid = (int)std::min<long long>(ceil_div(work, 256), 16384);
  if (vec)
    hipLaunchKernelGGL(k_loss_mse<true>, dim3(grid), dim3(256), 0, (hipStream_t)stream, color, alpha,
                       bg, bg_value, target, npix, inv, grad_color, grad_alpha, loss_accum);
  else
    hipLaunchKernelGGL(k_loss_mse<false>, dim3(grid), dim3(256), 0, (hipStream_t)stream, color, alpha,
                       bg, bg_value, target, npix, inv, grad_color, grad_alpha, loss_accum);
  LAUNCH_CHECK("k_loss_mse");
  return 0;
}

int dbmp_rescale_grads(float* grad_color, int64_t n_color, float* grad_alpha, int64_t n_alpha,
                       const float* scale, float* applied, void* stream) {
  if (!scale || !applied) return fail(DBMP_E_ARG, "NULL scale");
  if (n_color < 0 || n_alpha < 0 || (n_color > 0 && !grad_color) || (n_alpha > 0 && !grad_alpha))
    return fail(DBMP_E_ARG, "bad gradient image");
  hipStream_t st = (hipStream_t)stream;
  const long long n = std::max(n_color, n_alpha);
  // a small grid: with a ratio of 1 (the common case) the launch is all it costs
  const int grid = (int)std::max<long long>(1, std::min<long long>(ceil_div(n, 256), 1024));
  hipLaunchKernelGGL(k_rescale2, dim3(grid), dim3(256), 0, st, grad_color, (long long)n_color,
                     grad_alpha, (long long)n_alpha, scale, applied);
  LAUNCH_CHECK("k_rescale2");
  return 0;
}

int dbmp_adam_step(dbmp_rasterizer* h, const dbmp_adam* cfg, void* stream) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  if (!cfg || cfg->num_groups < 1 || cfg->num_groups > 5) return fail(DBMP_E_ARG, "bad Adam config");
  if (cfg->step < 1) return fail(DBMP_E_ARG, "Adam step must be >= 1");
  if (cfg->store_dtype < DBMP_F32 || cfg->store_dtype > DBMP_BF16) return fail(DBMP_E_ARG, "bad store dtype");
  AdamArgs A;
  long long nmax = 0;
  const double bc1 = 1.0 - std::pow((double)cfg->beta1, (double)cfg->step);
  const double bc2 = 1.0 - std::pow((double)cfg->beta2, (double)cfg->step);
  for (int i = 0; i < 5; ++i) {
    AdamGroup& g = A.g[i];
    if (i < cfg->num_groups) {
      const dbmp_adam_group& c = cfg->groups[i];
      if (c.numel > 0 && (!c.param || !c.exp_avg || !c.exp_avg_sq || !c.grad))
        return fail(DBMP_E_ARG, "Adam group %d has a NULL pointer", i);
      g.p = c.param; g.m = c.exp_avg; g.v = c.exp_avg_sq; g.g = c.grad; g.store = c.store;
      g.n = c.numel; g.stride = c.stride; g.clamp = c.clamp;
      if (c.grad_row < 0 || (c.grad_row > 0 && (c.grad_comps < 1 || c.grad_comps > c.grad_row)))
        return fail(DBMP_E_ARG, "Adam group %d: bad grad_row %d / grad_comps %d", i, c.grad_row,
                    c.grad_comps);
      g.grow = c.grad_row; g.gcomp = c.grad_row > 0 ? c.grad_comps : 1; g.zero = c.zero_grad != 0;
      g.lo[0] = c.lo[0]; g.lo[1] = c.lo[1]; g.hi[0] = c.hi[0]; g.hi[1] = c.hi[1];
      A.step_scale[i] = (float)((double)c.lr / bc1);
      nmax = std::max(nmax, (long long)c.numel);
    } else {
      g = AdamGroup{};
      g.n = 0;
      A.step_scale[i] = 0.f;
    }
  }
  A.beta1 = cfg->beta1;
  A.beta2 = cfg->beta2;
  A.eps = cfg->eps;
  A.inv_bc2_sqrt = (float)(1.0 / std::sqrt(bc2));
  A.store_dtype = cfg->store_dtype;
  if (nmax == 0) return 0;
  const int gx = (int)std::min<long long>(ceil_div(nmax, 256), 1024);
  hipLaunchKernelGGL(k_adam, dim3(gx, cfg->num_groups), dim3(256), 0, (hipStream_t)stream, A);
  LAUNCH_CHECK("k_adam");
  return 0;
}

int dbmp_lane_stats(uint64_t* out, int32_t reset) {
#ifdef DBMP_LANE_STATS
  if (!out) return fail(DBMP_E_ARG, "out is NULL");
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_lane_stats), sizeof(uint64_t) * 8));
  if (reset) {
    const unsigned long long zero[8] = {};
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_lane_stats), zero, sizeof(zero)));
  }
  return 0;
#else
  (void)out;
  (void)reset;
  return fail(DBMP_E_ARG, "this library was built without lane statistics (DBMP_LANE_STATS)");
#endif
}

int dbmp_sincos(const float* x, float* s, float* c, int64_t n, void* stream) {
  if (n < 0) return fail(DBMP_E_ARG, "n < 0");
  if (n == 0) return 0;
  if (!x || !s || !c) return fail(DBMP_E_ARG, "NULL pointer");
  hipLaunchKernelGGL(k_sincos, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, x, s, c,
                     (long long)n);
  LAUNCH_CHECK("k_sincos");
  return 0;
}

int dbmp_timing_enable(dbmp_rasterizer* h, int on) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  h->timing = on != 0;
  return 0;
}

int dbmp_timing_stats(dbmp_rasterizer* h, double* fwd_ms_total, int64_t* fwd_count,
                      double* bwd_ms_total, int64_t* bwd_count) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  long long fn = h->fwd_n, bn = h->bwd_n;
  timing_drain(h->ev_fwd, h->fwd_ms, fn);
  timing_drain(h->ev_bwd, h->bwd_ms, bn);
  h->fwd_n = fn;
  h->bwd_n = bn;
  if (fwd_ms_total) *fwd_ms_total = h->fwd_ms;
  if (fwd_count) *fwd_count = h->fwd_n;
  if (bwd_ms_total) *bwd_ms_total = h->bwd_ms;
  if (bwd_count) *bwd_count = h->bwd_n;
  return 0;
}

int dbmp_timing_reset(dbmp_rasterizer* h) {
  if (!h) return fail(DBMP_E_ARG, "handle is NULL");
  double ms = 0;
  long long n = 0;
  timing_drain(h->ev_fwd, ms, n);
  timing_drain(h->ev_bwd, ms, n);
  h->fwd_ms = h->bwd_ms = 0.0;
  h->fwd_n = h->bwd_n = 0;
  return 0;
}

}  // extern "C"
