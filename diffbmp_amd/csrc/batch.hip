// extern "C" batched-step entry points (declared in include/diffbmp_amd.h).
//
// A dbmp_batch owns the buffers that all jobs of a step share — primitive
// records, pair keys, tile lists, the (N,8) gradient rows — sized for the
// current job set, plus the device job table.  Jobs own their parameters,
// Adam moments, bitmaps, target and image scratch.  A step is a fixed
// sequence of launches whose shapes depend only on the job set, so it can be
// captured once into a hipGraph (on the batch's own stream) and replayed on
// the caller's stream until the set changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/diffbmp_amd.h"
#include "batch_kernels.h"
#include "host_common.h"

using namespace dbmp;
using namespace dbmp_host;

struct dbmp_batch {
  int dtype = DBMP_F32;
  float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  int device = 0;
  hipStream_t cap_stream = nullptr;
  // job tables
  int J = 0, job_cap = 0;
  BatchJob* djobs = nullptr;
  BatchState* dstate = nullptr;
  int* dstarts = nullptr;        // prim_start | tile_start | block_start, (job_cap+1) each
  void* staging = nullptr;       // pinned upload buffer
  size_t staging_bytes = 0;
  hipEvent_t upload_done = nullptr;
  bool upload_pending = false;
  // launch shape of the current set
  int Ntot = 0, tiles_tot = 0, blocks_tot = 0;
  bool blend = false;
  // shared buffers
  int n_cap = 0;
  PrimRec* recs = nullptr;
  int* counts = nullptr;
  long long* offs = nullptr;
  int4* ranges = nullptr;
  float* grad8 = nullptr;
  long long pair_cap = 0, pair_alloc = 0;
  int* keysA = nullptr;
  int* keysB = nullptr;
  int* valsB = nullptr;
  int* tidx = nullptr;
  int hist_blocks = 0;
  int* hist = nullptr;
  int* totals = nullptr;
  int tile_alloc = 0;
  int* toffs = nullptr;
  long long* dtot = nullptr;
  long long* host_total = nullptr;      // mapped pinned mirror of *dtot
  long long* host_total_dev = nullptr;  // its device address
  long long cull_cap = 0;
  float4* cull = nullptr;               // per-template content boxes of every job
  long long pm_cap = 0;
  unsigned long long* pmask = nullptr;  // push masks of all jobs' lists (forward -> backward)
  int* counters = nullptr;              // 2 x kXcds work counters (forward, backward)
  int cus = 0;                          // compute units of the device
  int occ_fwd = 0, occ_bwd = 0;         // resident workgroups per CU of the raster kernels
  std::vector<int> h_prim_start;        // host copy of prim_start (J+1)
  // graph of the current set
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// Binning of the batched step: lists per super-tile of 16 << kBatchBinShift
// px (each primitive's 16-px tile range filters it per 16x16 block), so the
// sort handles 16x fewer pairs than per-tile lists at 4.
#ifndef DBMP_BATCH_BIN_SHIFT
#define DBMP_BATCH_BIN_SHIFT 2
#endif
constexpr int kBatchBinShift = DBMP_BATCH_BIN_SHIFT;
constexpr int kBatchBin = 1 << kBatchBinShift;

int* prim_start(dbmp_batch* b) { return b->dstarts; }
int* tile_start(dbmp_batch* b) { return b->dstarts + (b->job_cap + 1); }
int* block_start(dbmp_batch* b) { return b->dstarts + 2 * (b->job_cap + 1); }

void drop_graph(dbmp_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int check_job(const dbmp_job& j, int i, int dtype) {
  if (j.height <= 0 || j.width <= 0) return fail(DBMP_E_ARG, "job %d: canvas must be positive", i);
  if (j.num_primitives < 0) return fail(DBMP_E_ARG, "job %d: num_primitives < 0", i);
  if (!(j.alpha_upper_bound >= 0.f))
    return fail(DBMP_E_ARG, "job %d: alpha_upper_bound must be >= 0", i);
  if (j.num_templates <= 0 || j.tmpl_height < 2 || j.tmpl_width < 2)
    return fail(DBMP_E_ARG, "job %d: bad template shape", i);
  if (texel_record_bytes(j.num_templates, j.tmpl_height, j.tmpl_width, dtype,
                         j.quads_with_color != 0) >= (1LL << 32))
    return fail(DBMP_E_ARG, "job %d: the texel records reach 4 GiB", i);
  if (!j.quads || !j.ext) return fail(DBMP_E_ARG, "job %d: prepared bitmaps missing", i);
  if (j.num_primitives > 0 && (!j.bmp_sel || !j.params || !j.exp_avg || !j.exp_avg_sq || !j.store))
    return fail(DBMP_E_ARG, "job %d: a parameter pointer is NULL", i);
  if (!j.target || !j.out_color || !j.out_alpha || !j.out_T || !j.out_last)
    return fail(DBMP_E_ARG, "job %d: an image pointer is NULL", i);
  if (!j.loss_history || j.loss_capacity < 1) return fail(DBMP_E_ARG, "job %d: no loss history", i);
  if (!(j.eps_T >= 0.f && j.eps_T < 1.f)) return fail(DBMP_E_ARG, "job %d: eps_T must be in [0, 1)", i);
  if (j.steps_done < 0) return fail(DBMP_E_ARG, "job %d: steps_done < 0", i);
  return 0;
}

int ensure_prims(dbmp_batch* b, int N) {
  if (N <= b->n_cap) return 0;
  const int cap = std::max(N, b->n_cap + b->n_cap / 4);
  int rc;
  if ((rc = grow(&b->recs, cap))) return rc;
  if ((rc = grow(&b->counts, cap))) return rc;
  // offsets (cap + 1), then the scan's chunk sums
  if ((rc = grow(&b->offs, (long long)cap + 1 + scan_partials_len(cap)))) return rc;
  if ((rc = grow(&b->ranges, cap))) return rc;
  if ((rc = grow(&b->grad8, 8LL * cap))) return rc;
  b->n_cap = cap;
  return 0;
}

int ensure_pairs(dbmp_batch* b, long long cap) {
  cap = std::max<long long>(cap, 1);
  if (cap <= b->pair_alloc) return 0;
  int rc;
  if ((rc = grow(&b->keysA, cap))) return rc;
  if ((rc = grow(&b->keysB, cap))) return rc;
  if ((rc = grow(&b->valsB, cap))) return rc;
  if ((rc = grow(&b->tidx, cap))) return rc;
  const int nblk = ceil_div(cap, kRadixTile);
  if ((rc = grow(&b->hist, 256LL * nblk))) return rc;
  b->pair_alloc = cap;
  b->hist_blocks = nblk;
  return 0;
}

int ensure_jobs(dbmp_batch* b, int J) {
  if (J <= b->job_cap) return 0;
  const int cap = std::max(J, 2 * b->job_cap);
  int rc;
  if ((rc = grow(&b->djobs, cap))) return rc;
  if ((rc = grow(&b->dstate, cap))) return rc;
  if ((rc = grow(&b->dstarts, 3LL * (cap + 1)))) return rc;
  b->job_cap = cap;
  return 0;
}

// Resident workgroups per CU of the batched raster kernels of storage type DT.
template <int DT>
int raster_occupancy(dbmp_batch* b) {
  if (b->occ_fwd > 0) return 0;
  HIP_TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device));
  int f = 0, g = 0;
  if (b->blend) {
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&f, k_batch_fwd<DT, true>, kThreads, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&g, k_batch_bwd<DT, true>, kThreads, 0));
  } else {
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&f, k_batch_fwd<DT, false>, kThreads, 0));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&g, k_batch_bwd<DT, false>, kThreads, 0));
  }
  b->occ_fwd = std::max(f, 1);
  b->occ_bwd = std::max(g, 1);
  return 0;
}

template <int DT>
int launch_step(dbmp_batch* b, hipStream_t st) {
  const int J = b->J, N = b->Ntot;
  const long long cap = b->pair_cap;
  const int gf = (int)std::min<long long>(b->blocks_tot, (long long)b->cus * b->occ_fwd);
  const int gb = (int)std::min<long long>(b->blocks_tot, (long long)b->cus * b->occ_bwd);
  const WorkQueue qf{b->counters, b->blocks_tot};
  const WorkQueue qb{b->counters + kXcds * kCounterStride, b->blocks_tot};
  hipLaunchKernelGGL(k_batch_reset, dim3(ceil_div(2 * kXcds * kCounterStride, 256)), dim3(256), 0, st,
                     b->counters, 2 * kXcds * kCounterStride);
  LAUNCH_CHECK("k_batch_reset");
  if (N > 0) {
    hipLaunchKernelGGL(k_batch_prep<DT>, dim3(ceil_div(N, 256)), dim3(256), 0, st, b->djobs,
                       prim_start(b), J, b->recs, b->counts, b->ranges);
    LAUNCH_CHECK("k_batch_prep");
  }
  scan_counts(b->counts, N, b->offs, b->dtot, b->offs + b->n_cap + 1, st);
  LAUNCH_CHECK("k_scan_apply");
  int bits = 0;
  while ((1 << bits) < b->tiles_tot) ++bits;
  const int passes = (bits + 7) / 8;
  const int dbits = passes ? (bits + passes - 1) / passes : 0;
  const bool even = (passes % 2) == 0;
  int* kcur = b->keysA;
  int* vcur = even ? b->tidx : b->valsB;
  int* kalt = b->keysB;
  int* valt = even ? b->valsB : b->tidx;
  if (N > 0 && cap > 0) {
    hipLaunchKernelGGL(k_batch_emit, dim3(ceil_div(N, 4)), dim3(256), 0, st, b->djobs,
                       prim_start(b), tile_start(b), J, b->counts, b->offs, b->ranges, b->dtot, cap,
                       kcur, vcur);
    LAUNCH_CHECK("k_batch_emit");
    const int nblk = ceil_div(cap, kRadixTile);
    for (int p = 0; p < passes; ++p) {
      const int shift = p * dbits;
      const int nb = std::min(dbits, bits - shift);
      hipLaunchKernelGGL(k_radix_hist, dim3(nblk), dim3(kRadixThreads), 0, st, kcur, b->dtot, cap,
                         shift, nb, nblk, b->hist);
      LAUNCH_CHECK("k_radix_hist");
      hipLaunchKernelGGL(k_radix_rowscan, dim3(1 << nb), dim3(256), 0, st, b->hist, nblk, b->totals,
                         b->dtot, cap);
      LAUNCH_CHECK("k_radix_rowscan");
      hipLaunchKernelGGL(k_radix_scatter, dim3(nblk), dim3(kRadixThreads), 0, st, kcur, vcur, kalt,
                         valt, b->dtot, cap, shift, nb, nblk, b->hist, b->totals);
      LAUNCH_CHECK("k_radix_scatter");
      std::swap(kcur, kalt);
      std::swap(vcur, valt);
    }
  }
  hipLaunchKernelGGL(k_tile_bounds, dim3(ceil_div(std::max<long long>(cap, b->tiles_tot) + 1, 256)),
                     dim3(256), 0, st, kcur, b->dtot, cap, b->tiles_tot, b->toffs);
  LAUNCH_CHECK("k_tile_bounds");
  if (b->blend)
    hipLaunchKernelGGL((k_batch_fwd<DT, true>), dim3(gf), dim3(kThreads), 0, st, b->djobs,
                       block_start(b), J, qf);
  else
    hipLaunchKernelGGL((k_batch_fwd<DT, false>), dim3(gf), dim3(kThreads), 0, st, b->djobs,
                       block_start(b), J, qf);
  LAUNCH_CHECK("k_batch_fwd");
  if (N > 0) {
    hipLaunchKernelGGL(k_fill_f32, dim3(ceil_div(8LL * N, 256)), dim3(256), 0, st, b->grad8, 8LL * N, 0.f);
    LAUNCH_CHECK("k_fill_f32");
  }
  if (b->blend)
    hipLaunchKernelGGL((k_batch_bwd<DT, true>), dim3(gb), dim3(kThreads), 0, st, b->djobs,
                       b->dstate, block_start(b), prim_start(b), J, b->grad8, qb);
  else
    hipLaunchKernelGGL((k_batch_bwd<DT, false>), dim3(gb), dim3(kThreads), 0, st, b->djobs,
                       b->dstate, block_start(b), prim_start(b), J, b->grad8, qb);
  LAUNCH_CHECK("k_batch_bwd");
  hipLaunchKernelGGL(k_batch_tick, dim3(ceil_div(J, 64)), dim3(64), 0, st, b->djobs, b->dstate, J,
                     b->beta1, b->beta2, (const long long*)b->dtot,
                     b->host_total_dev);
  LAUNCH_CHECK("k_batch_tick");
  if (N > 0) {
    hipLaunchKernelGGL(k_batch_adam<DT>, dim3(ceil_div(8LL * N, 256)), dim3(256), 0, st, b->djobs,
                       b->dstate, prim_start(b), J, b->grad8, b->beta1, b->beta2, b->eps);
    LAUNCH_CHECK("k_batch_adam");
  }
  return 0;
}

int raster_occupancy_dt(dbmp_batch* b) {
  if (b->dtype == DBMP_F32) return raster_occupancy<kF32>(b);
  if (b->dtype == DBMP_F16) return raster_occupancy<kF16>(b);
  return raster_occupancy<kBF16>(b);
}

int launch_step_dt(dbmp_batch* b, hipStream_t st) {
  if (b->dtype == DBMP_F32) return launch_step<kF32>(b, st);
  if (b->dtype == DBMP_F16) return launch_step<kF16>(b, st);
  return launch_step<kBF16>(b, st);
}

}  // namespace

extern "C" {

int dbmp_batch_create(dbmp_batch** out, int32_t storage_dtype, float beta1, float beta2, float eps) {
  if (!out) return fail(DBMP_E_ARG, "out is NULL");
  *out = nullptr;
  if (storage_dtype < DBMP_F32 || storage_dtype > DBMP_BF16)
    return fail(DBMP_E_ARG, "bad storage dtype %d", storage_dtype);
  auto* b = new dbmp_batch();
  b->dtype = storage_dtype;
  b->beta1 = beta1;
  b->beta2 = beta2;
  b->eps = eps;
  int rc = 0;
  if (hipGetDevice(&b->device) != hipSuccess) rc = fail(DBMP_E_HIP, "hipGetDevice failed");
  if (!rc && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(DBMP_E_HIP, "hipStreamCreate failed");
  if (!rc && hipEventCreateWithFlags(&b->upload_done, hipEventDisableTiming) != hipSuccess)
    rc = fail(DBMP_E_HIP, "hipEventCreate failed");
  if (!rc) rc = grow(&b->dtot, 1);
  if (!rc) rc = grow(&b->totals, 256);
  if (!rc) rc = grow(&b->counters, 2 * kXcds * kCounterStride);
  if (!rc && hipHostMalloc((void**)&b->host_total, sizeof(long long), hipHostMallocMapped) != hipSuccess)
    rc = fail(DBMP_E_ALLOC, "hipHostMalloc failed");
  if (!rc && hipHostGetDevicePointer((void**)&b->host_total_dev, b->host_total, 0) != hipSuccess)
    rc = fail(DBMP_E_HIP, "hipHostGetDevicePointer failed");
  if (rc) {
    dbmp_batch_destroy(b);
    return rc;
  }
  *b->host_total = 0;
  *out = b;
  return 0;
}

int dbmp_batch_destroy(dbmp_batch* b) {
  if (!b) return 0;
  if (b->upload_pending) (void)hipEventSynchronize(b->upload_done);
  drop_graph(b);
  if (b->cap_stream) (void)hipStreamSynchronize(b->cap_stream);
  void* ptrs[] = {b->djobs, b->dstate, b->dstarts, b->recs, b->counts, b->offs, b->ranges,
                  b->grad8, b->keysA, b->keysB, b->valsB, b->tidx, b->hist, b->totals, b->toffs,
                  b->dtot, b->cull, b->counters, b->pmask};
  (void)hipDeviceSynchronize();
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b->staging) (void)hipHostFree(b->staging);
  if (b->host_total) (void)hipHostFree(b->host_total);
  if (b->upload_done) (void)hipEventDestroy(b->upload_done);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  delete b;
  return 0;
}

int64_t dbmp_batch_bitmap_bytes(const dbmp_batch* b, const dbmp_bitmaps* bm) {
  if (!b || !bm || bm->num_templates <= 0 || bm->height < 2 || bm->width < 2) return -1;
  return texel_record_bytes(bm->num_templates, bm->height, bm->width, b->dtype,
                            bm->colors_orig_mode == 2 && bm->colors_orig);
}

int dbmp_batch_prepare_bitmaps(dbmp_batch* b, const dbmp_bitmaps* bm, float* ext_out,
                               void* quads_out, void* stream) {
  if (!b) return fail(DBMP_E_ARG, "batch is NULL");
  if (!bm || !bm->templates) return fail(DBMP_E_ARG, "templates are NULL");
  if (!ext_out || !quads_out) return fail(DBMP_E_ARG, "NULL output");
  if (bm->num_templates <= 0 || bm->height < 2 || bm->width < 2)
    return fail(DBMP_E_ARG, "templates must be (T>0, H>=2, W>=2)");
  if (bm->dtype != b->dtype) return fail(DBMP_E_ARG, "template dtype differs from the batch's");
  if (texel_record_bytes(bm->num_templates, bm->height, bm->width, bm->dtype,
                         bm->colors_orig_mode == 2 && bm->colors_orig) >= (1LL << 32))
    return fail(DBMP_E_ARG, "the texel records of %d templates of %dx%d reach 4 GiB",
                bm->num_templates, bm->height, bm->width);
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
  if (bm->colors_orig_mode == 2 && bm->colors_orig) return launch_build_rgba(bm, quads_out, st);
  if (bm->colors_orig_mode != 0) return fail(DBMP_E_ARG, "colour maps must be per template (mode 2)");
  return launch_build_quads(bm, quads_out, st);
}

int dbmp_batch_set_jobs(dbmp_batch* b, const dbmp_job* jobs, int32_t num_jobs,
                        int64_t pair_capacity, void* stream) {
  if (!b) return fail(DBMP_E_ARG, "batch is NULL");
  if (num_jobs < 0 || (num_jobs > 0 && !jobs)) return fail(DBMP_E_ARG, "bad job list");
  if (pair_capacity < 0 || pair_capacity > (int64_t)0x7fffffff)
    return fail(DBMP_E_ARG, "pair_capacity must be in [0, 2^31)");
  int rc;
  for (int i = 0; i < num_jobs; ++i)
    if ((rc = check_job(jobs[i], i, b->dtype))) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int J = num_jobs;
  std::vector<int> ps(J + 1), ts(J + 1), bs(J + 1);
  std::vector<long long> cs(J + 1);
  long long Nt = 0, Tt = 0, Bt = 0, Ct = 0;
  for (int i = 0; i < J; ++i) {
    cs[i] = Ct;
    Ct += jobs[i].num_templates;
    ps[i] = (int)Nt;
    ts[i] = (int)Tt;
    bs[i] = (int)Bt;
    const long long tx = (jobs[i].width + kSub - 1) / kSub, ty = (jobs[i].height + kSub - 1) / kSub;
    const long long lx = (tx + kBatchBin - 1) / kBatchBin, ly = (ty + kBatchBin - 1) / kBatchBin;
    Nt += jobs[i].num_primitives;
    Tt += lx * ly;                          // lists: one per super-tile
    Bt += lx * ly * kBatchBin * kBatchBin;  // 16x16 blocks of the super-tiles
  }
  ps[J] = (int)Nt;
  ts[J] = (int)Tt;
  bs[J] = (int)Bt;
  if (Nt >= (1LL << 28) || Tt >= (1LL << 30))
    return fail(DBMP_E_ARG, "job set too large (%lld primitives, %lld tiles)", Nt, Tt);
  // every buffer the previous steps may still read stays alive until they end
  // push masks: job i's lists are lists ts[i].. of the batch, so its words
  // start ts[i] list rows in (pm_word counts the list index from there)
  const long long pm_words = push_mask_words(pair_capacity, Tt, kBatchBin * kBatchBin);
  const bool grows = Nt > b->n_cap || pair_capacity > b->pair_alloc || Tt + 1 > b->tile_alloc ||
                     J > b->job_cap || Ct > b->cull_cap || pm_words > b->pm_cap;
  if (grows) HIP_TRY(hipStreamSynchronize(st));
  if ((rc = ensure_prims(b, (int)std::max<long long>(Nt, 1)))) return rc;
  if ((rc = ensure_pairs(b, pair_capacity))) return rc;
  if (Tt + 1 > b->tile_alloc) {
    if ((rc = grow(&b->toffs, Tt + 1))) return rc;
    b->tile_alloc = (int)(Tt + 1);
  }
  if ((rc = ensure_jobs(b, std::max(J, 1)))) return rc;
  if (Ct > b->cull_cap) {
    if ((rc = grow(&b->cull, Ct))) return rc;
    b->cull_cap = Ct;
  }
  if (pm_words > b->pm_cap) {
    if ((rc = grow(&b->pmask, pm_words))) {
      b->pm_cap = 0;
      return rc;
    }
    b->pm_cap = pm_words;
  }

  std::vector<BatchJob> hj(J);
  std::vector<BatchState> hs(J);
  bool blend = false;
  for (int i = 0; i < J; ++i) {
    const dbmp_job& s = jobs[i];
    BatchJob& d = hj[i];
    std::memset(&d, 0, sizeof(d));
    RasterArgs& A = d.A;
    A.recs = b->recs + ps[i];
    A.toffs = b->toffs + ts[i];
    A.tidx = b->tidx;
    A.pcount = b->dtot;
    A.pcap = pair_capacity;
    A.quads = s.quads;
    // RGBA records need the blend kernels (their stride), blending or not
    const bool rec = s.quads_with_color != 0;
    blend = blend || rec;
    A.cmap = (rec && s.c_blend > 0.f) ? s.quads : nullptr;
    A.cmap_mode = rec ? 2 : 1;
    A.c_blend = s.c_blend;
    A.exact_blend = s.blend_grad_exact != 0;
    A.N = s.num_primitives;
    A.Ht = s.tmpl_height;
    A.Wt = s.tmpl_width;
    A.HWm1f = (float)(s.tmpl_height * s.tmpl_width - 1);
    A.Wtf = (float)s.tmpl_width;
    A.hsx = 0.5f * (float)(s.tmpl_width - 1);
    A.hsy = 0.5f * (float)(s.tmpl_height - 1);
    A.H = s.height;
    A.W = s.width;
    A.ts = kSub * kBatchBin;
    A.tilesX = ((s.width + kSub - 1) / kSub + kBatchBin - 1) / kBatchBin;
    A.tiles = ts[i + 1] - ts[i];
    A.subs = kBatchBin * kBatchBin;
    A.eps = s.eps_T;
    // the job's binning ranges: after a pair overflow every pixel walks its
    // job's primitives, each kept to the tiles of its range
    A.ranges = b->ranges + ps[i];
    A.order = nullptr;
    A.cost = nullptr;
    A.cull = b->cull + cs[i];   // filled from the job's extents after the upload
    A.pmask = b->pmask + (long long)ts[i] * (kBatchBin * kBatchBin) * 4;
    A.ntmpl = s.num_templates;
    A.bin_shift = kBatchBinShift;
    d.alpha_ub = s.alpha_upper_bound;
    d.L.target = s.target;
    d.L.bg = s.background;
    d.L.bg_value = s.background_value;
    d.L.inv_count = 1.f / (float)(3.0 * s.height * s.width);
    d.L.loss = nullptr;
    d.store = s.store;
    d.p = s.params;
    d.m = s.exp_avg;
    d.v = s.exp_avg_sq;
    d.store_out = (s.store != (void*)s.params) ? s.store : nullptr;
    d.sel = s.bmp_sel;
    d.ext = s.ext;
    d.T = s.num_templates;
    d.tilesX16 = (s.width + kSub - 1) / kSub;
    d.tilesY16 = (s.height + kSub - 1) / kSub;
    d.color = s.out_color;
    d.alpha = s.out_alpha;
    d.Tfin = s.out_T;
    d.last = s.out_last;
    d.loss_hist = s.loss_history;
    d.loss_cap = s.loss_capacity;
    for (int g = 0; g < 5; ++g) d.lr[g] = s.lr[g];
    if (s.clamp) {
      d.lo[0] = 0.f; d.hi[0] = (float)(s.width - 1);
      d.lo[1] = 0.f; d.hi[1] = (float)(s.height - 1);
      d.lo[2] = 2.f; d.hi[2] = (float)(std::min(s.height, s.width) / 4);
    } else {
      for (int c = 0; c < 3; ++c) { d.lo[c] = 1.f; d.hi[c] = 0.f; }
    }
    std::memset(&hs[i], 0, sizeof(BatchState));
    hs[i].step = s.steps_done;
    // beta^steps_done by the same fp32 products k_batch_tick forms, so a job
    // re-admitted mid-fit continues the sequence it had
    float b1t = 1.f, b2t = 1.f;
    for (int t = 0; t < s.steps_done; ++t) {
      b1t *= b->beta1;
      b2t *= b->beta2;
    }
    hs[i].b1t = b1t;
    hs[i].b2t = b2t;
  }
  // one pinned upload: jobs | state | starts (3 x (job_cap + 1) ints)
  const size_t bj = sizeof(BatchJob) * J, bst = sizeof(BatchState) * J;
  const size_t bss = sizeof(int) * 3 * (size_t)(b->job_cap + 1);
  const size_t need = bj + bst + bss;
  if (b->upload_pending) {
    HIP_TRY(hipEventSynchronize(b->upload_done));
    b->upload_pending = false;
  }
  if (need > b->staging_bytes) {
    if (b->staging) (void)hipHostFree(b->staging);
    b->staging = nullptr;
    b->staging_bytes = 0;
    if (hipHostMalloc(&b->staging, need, hipHostMallocDefault) != hipSuccess)
      return fail(DBMP_E_ALLOC, "hipHostMalloc of %zu bytes failed", need);
    b->staging_bytes = need;
  }
  char* stg = (char*)b->staging;
  if (J) {
    std::memcpy(stg, hj.data(), bj);
    std::memcpy(stg + bj, hs.data(), bst);
  }
  int* starts = (int*)(stg + bj + bst);
  std::memset(starts, 0, bss);
  std::copy(ps.begin(), ps.end(), starts);
  std::copy(ts.begin(), ts.end(), starts + (b->job_cap + 1));
  std::copy(bs.begin(), bs.end(), starts + 2 * (b->job_cap + 1));
  if (J) {
    HIP_TRY(hipMemcpyAsync(b->djobs, stg, bj, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->dstate, stg + bj, bst, hipMemcpyHostToDevice, st));
  }
  HIP_TRY(hipMemcpyAsync(b->dstarts, starts, bss, hipMemcpyHostToDevice, st));
  if (J) {
    hipLaunchKernelGGL(k_batch_cull, dim3(J), dim3(64), 0, st, (const BatchJob*)b->djobs);
    LAUNCH_CHECK("k_batch_cull");
  }
  HIP_TRY(hipEventRecord(b->upload_done, st));
  b->upload_pending = true;
  b->h_prim_start = ps;
  b->J = J;
  b->Ntot = (int)Nt;
  b->tiles_tot = (int)Tt;
  b->blocks_tot = (int)Bt;
  b->pair_cap = pair_capacity;
  if (b->blend != blend) b->occ_fwd = b->occ_bwd = 0;   // the other kernels' occupancy
  b->blend = blend;
  if ((rc = raster_occupancy_dt(b))) return rc;   // here, never under stream capture
  drop_graph(b);
  return 0;
}

int dbmp_batch_step(dbmp_batch* b, int32_t use_graph, void* stream) {
  if (!b) return fail(DBMP_E_ARG, "batch is NULL");
  if (b->J == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return launch_step_dt(b, st);
  if (!b->exec) {
    HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step_dt(b, b->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(DBMP_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(b);
      return fail(DBMP_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
  }
  HIP_TRY(hipGraphLaunch(b->exec, st));
  return 0;
}

int dbmp_batch_gradients(dbmp_batch* b, int32_t job, float* rows, void* stream) {
  if (!b || !rows) return fail(DBMP_E_ARG, "NULL argument");
  if (job < 0 || job >= b->J) return fail(DBMP_E_ARG, "job %d not in the set of %d", job, b->J);
  const int n0 = b->h_prim_start[job], n1 = b->h_prim_start[job + 1];
  if (n1 > n0)
    HIP_TRY(hipMemcpyAsync(rows, b->grad8 + 8LL * n0, sizeof(float) * 8LL * (n1 - n0),
                           hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return 0;
}

int dbmp_batch_pair_count(dbmp_batch* b, int64_t* pairs) {
  if (!b || !pairs) return fail(DBMP_E_ARG, "NULL argument");
  *pairs = *(volatile long long*)b->host_total;
  return 0;
}

}  // extern "C"
