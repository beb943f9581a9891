/*
 * diffbmp_amd.h -- C ABI of the MI355X (gfx950) DiffBMP tile rasterizer.
 *
 * The library replaces the reference's CUDA torch extension `_C` / `_C_fp16`
 * (cuda_tile_rasterizer/ext.cpp:5-121, ext_fp16.cpp:5-82) on the
 * binning -> forward -> backward path.  Every entry point takes plain device
 * pointers, sizes and a hipStream_t (as void*), never synchronises the host,
 * never reads tensors back, and returns 0 on success or a nonzero DBMP_E*
 * code, with a message from dbmp_last_error().  State lives in a handle
 * (not a process-wide singleton): forward->backward state is returned to the
 * caller in caller-owned buffers (T_final, last), so several forward passes can
 * be outstanding at once.  A handle's internal scratch (primitive records,
 * sort buffers, gradient accumulator, texel quads) is shared by its calls:
 * use one handle per stream.
 *
 * Reference conventions kept: means2D (N,2) [x,y] in pixels with integer pixel
 * centres; radii (N) half extent in pixels; rotations (N) radians; opacities
 * (N) and colors (N,3) are logits; templates (T,Ht,Wt) alpha in [0,1];
 * bmp_sel (N) int32 template index; colour output premultiplied over black,
 * alpha = 1 - prod(1 - a_k); tile lists row-major by tile, ascending primitive
 * id inside a tile (cuda_tile_rasterizer/__init__.py:84-92,
 * pydiffbmp/core/renderer/simple_tile_renderer.py:721-766).
 */
#ifndef DIFFBMP_AMD_H
#define DIFFBMP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DBMP_F32 0
#define DBMP_F16 1
#define DBMP_BF16 2
/* dbmp_prims only: positions, radii and rotations fp32; opacities and colours
 * fp16 / bf16 (half-precision positions move a 1024-px coordinate by up to
 * 0.25 px / 2 px; the appearance values do not need that range) */
#define DBMP_F16_APP 3
#define DBMP_BF16_APP 4

#define DBMP_OK 0
#define DBMP_E_ARG 1      /* invalid argument / unsupported shape */
#define DBMP_E_HIP 2      /* HIP runtime or launch error */
#define DBMP_E_ALLOC 3    /* device allocation failed */

typedef struct dbmp_rasterizer dbmp_rasterizer;

/* Per-primitive parameters, storage dtype DBMP_F32 / F16 / BF16, or
 * DBMP_F16_APP / DBMP_BF16_APP (fp32 geometry, half-precision appearance).
 * reuse_records: nonzero asserts that these exact parameter values were the
 * input of the previous dbmp_bin / dbmp_forward / dbmp_backward call on this
 * handle, so the per-primitive constants computed there are reused. */
typedef struct {
  const void* means2D;     /* (N,2) */
  const void* radii;       /* (N)   */
  const void* rotations;   /* (N)   */
  const void* opacities;   /* (N) logits */
  const void* colors;      /* (N,3) logits */
  const int32_t* bmp_sel;  /* (N) template index */
  int32_t num_primitives;
  int32_t dtype;
  int32_t reuse_records;
} dbmp_prims;

/* Alpha bitmaps and optional colour maps (same dtype as the bitmaps). */
typedef struct {
  const void* templates;    /* (T,Ht,Wt) */
  const void* colors_orig;  /* (N,Ht,Wt,3) or (T,Ht,Wt,3) in [0,1]; may be NULL */
  int32_t num_templates, height, width;
  int32_t dtype;
  int32_t colors_orig_mode; /* 0: none, 1: per primitive (N,..), 2: per template (T,..) */
  float c_blend;            /* colour-map blend, used when > 0 and colors_orig set */
  int64_t version;          /* caller's content version of `templates`; -1 = unknown */
  int32_t blend_grad_exact; /* 0: the reference's colour gradient under c_blend > 0
                               (tile_backward.cu:253-258: no (1 - c_blend) factor, no
                               dependence of the colour map on geometry); 1: exact */
  int64_t colors_version;   /* caller's content version of per-template `colors_orig`
                               (mode 2); -1 = unknown */
} dbmp_bitmaps;

/* CSR tile lists: tile_offsets (tiles+1), tile_indices (P).  When the lists
 * come from dbmp_bin / dbmp_bin_tiles, pass its device pair count and the
 * capacity it was given: if the count exceeded the capacity (the lists were
 * not written), the raster kernels walk every primitive 0..N-1 for every tile
 * instead and use a primitive in a tile only inside its binning tile range --
 * the membership the lists would have held -- so image and gradients are the
 * same, only slower.  pair_count = NULL means the lists are complete (e.g. a
 * caller-supplied mapping) and ranges are not needed. */
typedef struct {
  const int32_t* tile_offsets;
  const int32_t* tile_indices;
  const int64_t* pair_count;
  int64_t capacity;
  int32_t bin_factor;       /* 0 or 1: one list per tile; f > 1: lists from
                               dbmp_bin_tiles, one per super-tile of f x f tiles */
  const int32_t* ranges;    /* (N,4) tile ranges written by the binning that made the
                               lists (dbmp_bin_tiles' ranges_out); required for f > 1.
                               For f <= 1 with pair_count set, NULL means the ranges
                               the handle's latest dbmp_bin wrote */
  uint64_t* push_masks;     /* optional forward->backward state, like out_last: for
                               each 64-entry batch of a list, 16x16 block and wave,
                               which entries pushed a pixel of the wave.  dbmp_forward
                               writes them and a backward of that forward (same lists,
                               parameters and out_last) walks only those entries
                               instead of culling; results are the same.  Used when
                               pair_count is set and push_mask_words is at least
                               dbmp_push_mask_words(h, capacity, bin_factor); NULL:
                               none.  Not used by the candidate (batch) calls */
  int64_t push_mask_words;  /* uint64 words at push_masks */
} dbmp_tile_lists;

/* Words of dbmp_tile_lists.push_masks for lists of at most `capacity` entries
 * with this bin_factor on this handle's canvas.  No reference counterpart
 * (the reference re-tests every list entry in its backward,
 * tile_backward.cu:127-157). */
int64_t dbmp_push_mask_words(const dbmp_rasterizer* h, int64_t capacity, int32_t bin_factor);

/* Library version (major*10000 + minor*100 + patch). */
int dbmp_version(void);

/* Message of the last failing call on this host thread ("" if none). */
const char* dbmp_last_error(void);

/* Replaces _C.init_tile_rasterizer (ext.cpp:5-8, tile_rasterize.cu:15-207).
 * tile_size must be a positive multiple of 16.  sigma and
 * max_prims_per_pixel are accepted for signature parity and not used: the
 * reference never reads sigma in its kernels, and this implementation keeps
 * no per-pixel layer cache, so it has no per-pixel capacity limit.
 * alpha_upper_bound must be >= 0 (an opacity bound; the kernels rely on
 * alpha >= 0, and a negative bound is rejected).  A
 * handle's scratch (primitive records, sort buffers, texel records, block
 * costs and launch order) is shared by its calls: use a handle from one
 * stream at a time. */
int dbmp_init(dbmp_rasterizer** out, int image_height, int image_width, int tile_size,
              float sigma, float alpha_upper_bound, int max_prims_per_pixel, int num_primitives);

/* Frees the handle and its scratch. */
int dbmp_destroy(dbmp_rasterizer* h);

/* Template extents (T,4) fp32 (min_u,max_u,min_v,max_v) of texels > 1e-6;
 * replaces VectorRenderer._compute_primitive_bboxes
 * (pydiffbmp/core/renderer/vector_renderer.py:151-192). */
int dbmp_template_extents(dbmp_rasterizer* h, const dbmp_bitmaps* bm, float* ext_out, void* stream);

/* Binning: tile lists for the primitives; replaces the reference's torch
 * binning and packing (vector_renderer.py:194-252,401-444;
 * simple_tile_renderer.py:479-516,721-766).  `ext` is (T,4) from
 * dbmp_template_extents.  Writes the pair count P to *d_pair_count (device
 * int64).  If P > capacity the lists are not written (tile_offsets is set to
 * zeros); passing d_pair_count and capacity on to dbmp_forward/backward in
 * dbmp_tile_lists keeps results exact in that case, and the caller reads P
 * whenever convenient to grow its buffer.  tile_indices must hold `capacity`
 * int32.  No host synchronisation. */
int dbmp_bin(dbmp_rasterizer* h, const dbmp_prims* prims, const float* ext, int num_templates,
             int tmpl_h, int tmpl_w, int32_t* tile_offsets, int32_t* tile_indices,
             int64_t capacity, int64_t* d_pair_count, void* stream);

/* Binning into coarser lists for the raster calls of this handle: one list
 * per super-tile of bin_factor x bin_factor tiles (bin_factor 2, 4 or 8; tile
 * size 16), ascending primitive id, plus each primitive's tile range (N,4)
 * int32 in ranges_out (bin_factor 1: the per-tile lists of dbmp_bin, with
 * ranges_out optional; NULL keeps them in the handle).  The raster kernels
 * use an entry in a tile only inside that range, so forward and backward are
 * exactly those of dbmp_bin's per-tile lists, with bin_factor^2 fewer pairs
 * to sort.  tile_offsets holds
 * ceil(tilesX/f)*ceil(tilesY/f)+1 entries.  Pass bin_factor and ranges_out on
 * in dbmp_tile_lists. */
int dbmp_bin_tiles(dbmp_rasterizer* h, const dbmp_prims* prims, const float* ext, int num_templates,
                   int tmpl_h, int tmpl_w, int32_t bin_factor, int32_t* tile_offsets,
                   int32_t* tile_indices, int64_t capacity, int64_t* d_pair_count,
                   int32_t* ranges_out, void* stream);

/* dbmp_backward_mse that ADDS the gradients into grad_rows, (N,8) fp32 rows
 * (x, y, r, theta, opacity logit, 3 colour logits), and skips the per-array
 * unpacking.  With dbmp_adam_step reading the rows (grad_row 8, zero_grad 1)
 * the rows are zero again for the next step.  No reference counterpart: it is
 * the device-resident step's form of rasterize_tiles_backward_class. */
int dbmp_backward_mse_rows(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                           const dbmp_tile_lists* lists, float eps_T, const float* target,
                           const float* background, float background_value, float* loss_accum,
                           const float* out_color, const float* out_T_final,
                           const int32_t* out_last, float* grad_rows, void* stream);

/* Forward; replaces _C.rasterize_tiles_class / _C_fp16.rasterize_tiles_class_fp16
 * (ext.cpp:11-28 -> tile_rasterize.cu:297-498).  Outputs (fp32):
 *   out_color (H,W,3) premultiplied over black, out_alpha (H,W),
 *   out_T_final (H,W) and out_last (H,W) int32: per-pixel state dbmp_backward
 *   needs -- the product of the pixel's nonzero (1 - alpha) factors, and the
 *   number of list entries consumed in bits 0..29 with the count of exactly
 *   opaque layers (0, 1, 2+) in bits 30..31.  eps_T in [0, 1): a pixel stops
 *   compositing once later layers can change it by less than eps_T (0 = walk
 *   the whole list). */
int dbmp_forward(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                 const dbmp_tile_lists* lists, float eps_T, float* out_color, float* out_alpha,
                 float* out_T_final, int32_t* out_last, void* stream);

/* dbmp_forward without a handle and without allocating: the same kernels and
 * outputs, with this call's primitive records, texel quads (or RGBA records)
 * and template content boxes in a file-scope device array of the library
 * (about 20 MiB, one per GPU, held until the process ends).  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Limits: at most 65536 primitives, 4096 templates and 16 MiB
 * of texel records (T*(Ht+1)*Wt*16 bytes for fp32 alpha bitmaps, *8 for
 * fp16 / bf16, *64 / *16 with per-template colour maps); larger calls return
 * DBMP_E_ARG.  Lists as for dbmp_forward, except that lists with a pair count
 * or bin_factor > 1 must carry their ranges.  The array keeps the last call's
 * contents and every call rewrites what it reads; calls that overlap on one
 * GPU (several streams or threads) race on it. */
int dbmp_forward_noalloc(int image_height, int image_width, int tile_size, float alpha_upper_bound,
                         const dbmp_prims* prims, const dbmp_bitmaps* bm,
                         const dbmp_tile_lists* lists, float eps_T, float* out_color,
                         float* out_alpha, float* out_T_final, int32_t* out_last, void* stream);

/* B candidates of one canvas and template set in one raster launch per
 * stage; replaces _C.rasterize_tiles_batch_class and
 * rasterize_tiles_backward_batch_class (ext.cpp:53-107 ->
 * tile_rasterize.cu:500-636).  prims[b] and lists[b] are candidate b's
 * parameters (the same N and dtype for every candidate) and tile lists (the
 * same bin_factor; lists with a pair count must carry their ranges); with
 * per-primitive colour maps (colors_orig_mode 1) bm->colors_orig holds the
 * candidates' (N,Ht,Wt,3) maps one after another.  Outputs and per-pixel
 * state are stacked: out_color (B,H,W,3), out_alpha / out_T_final / out_last
 * (B,H,W).  Each candidate keeps its own state, so candidate b's gradients
 * are its own (the reference's batch backward reads the last candidate's
 * cache for all, tile_rasterize.cu:533-628). */
int dbmp_forward_batch(dbmp_rasterizer* h, int32_t B, const dbmp_prims* prims,
                       const dbmp_bitmaps* bm, const dbmp_tile_lists* lists, float eps_T,
                       float* out_color, float* out_alpha, float* out_T_final, int32_t* out_last,
                       void* stream);

/* Backward of dbmp_forward_batch: grad_color (B,H,W,3), grad_alpha (B,H,W)
 * or NULL; gradients stacked (B,N,2), (B,N), (B,N), (B,N), (B,N,3). */
int dbmp_backward_batch(dbmp_rasterizer* h, int32_t B, const dbmp_prims* prims,
                        const dbmp_bitmaps* bm, const dbmp_tile_lists* lists, float eps_T,
                        const float* grad_color, const float* grad_alpha, const float* out_color,
                        const float* out_T_final, const int32_t* out_last, float* grad_means2D,
                        float* grad_radii, float* grad_rotations, float* grad_opacities,
                        float* grad_colors, void* stream);

/* Backward; replaces _C.rasterize_tiles_backward_class (ext.cpp:31-50 ->
 * tile_rasterize.cu:638-780).  grad_color (H,W,3), grad_alpha (H,W) or NULL;
 * out_color / out_T_final / out_last are dbmp_forward's outputs for the same
 * inputs.  Gradients (fp32) are overwritten: grad_means2D (N,2), grad_radii
 * (N), grad_rotations (N), grad_opacities (N), grad_colors (N,3). */
int dbmp_backward(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                  const dbmp_tile_lists* lists, float eps_T, const float* grad_color,
                  const float* grad_alpha, const float* out_color, const float* out_T_final,
                  const int32_t* out_last, float* grad_means2D, float* grad_radii,
                  float* grad_rotations, float* grad_opacities, float* grad_colors, void* stream);

/* Background composite + mean squared error + gradient, one kernel:
 *   result = color + (1 - alpha) * bg,  loss = mean((result - target)^2),
 *   grad_color = dloss/dcolor (H,W,3), grad_alpha = dloss/dalpha (H,W);
 * loss is added to *loss_accum (device float).  bg is (H,W,3), or NULL for the
 * constant bg_value (1 = white).  Replaces the torch composite and loss of
 * simple_tile_renderer.py:703-707 and loss_functions.py:20-40 together with
 * their autograd backward. */
int dbmp_loss_mse(dbmp_rasterizer* h, const float* color, const float* alpha, const float* bg,
                  float bg_value, const float* target, float* grad_color, float* grad_alpha,
                  float* loss_accum, void* stream);

/* Gradient images from dbmp_loss_mse scaled in place to an incoming gradient:
 * x <- x * (*scale / applied[0]), then applied[0] = *scale.  scale: a device
 * fp32 scalar; applied: two device words, the factor the images carry (fp32,
 * start at 1 or at 0, which stands for 1) and a counter (start at 0; every
 * call leaves it 0).  One kernel launch; a ratio of 1 leaves the images as
 * they are, so the usual incoming gradient of a loss (1) costs no image
 * traffic, and a second backward over the same saved images rescales them
 * correctly.  No
 * reference counterpart: it is the autograd backward of the fused loss
 * (the reference multiplies by torch's incoming gradient,
 * loss_functions.py:20-40 through torch autograd). */
int dbmp_rescale_grads(float* grad_color, int64_t n_color, float* grad_alpha, int64_t n_alpha,
                       const float* scale, float* applied, void* stream);

/* One parameter group of dbmp_adam_step (fp32 master weights). */
typedef struct {
  float* param;
  float* exp_avg;
  float* exp_avg_sq;
  float* grad;
  void* store;          /* optional copy written in store_dtype (or NULL) */
  int64_t numel;
  float lr;
  int32_t stride;       /* 2 for means2D: element i uses lo[i%2], hi[i%2] */
  int32_t clamp;        /* clamp to [lo, hi] after the update */
  float lo[2], hi[2];
  int32_t grad_row;     /* 0: element i's gradient is grad[i]; > 0: it is
                           grad[(i / grad_comps) * grad_row + i % grad_comps]
                           (e.g. a column of dbmp_backward_mse_rows' rows) */
  int32_t grad_comps;
  int32_t zero_grad;    /* 1: write 0 over each gradient value read */
} dbmp_adam_group;

typedef struct {
  dbmp_adam_group groups[5];
  int32_t num_groups;
  float beta1, beta2, eps;
  int64_t step;         /* 1-based count including this update */
  int32_t store_dtype;
} dbmp_adam;

/* dbmp_loss_mse followed by dbmp_backward in one raster pass: the
 * background composite, squared error and its gradient are computed per pixel
 * in the backward's prologue, so no (H,W,3) gradient image is written or
 * read.  Replaces the reference's torch composite + F.mse_loss + autograd
 * feeding rasterize_tiles_backward (simple_tile_renderer.py:703-707,
 * loss_functions.py:20-40, ext.cpp:31-50).  loss_accum (device, may be NULL)
 * receives += the MSE. */
int dbmp_backward_mse(dbmp_rasterizer* h, const dbmp_prims* prims, const dbmp_bitmaps* bm,
                      const dbmp_tile_lists* lists, float eps_T, const float* target,
                      const float* background, float background_value, float* loss_accum,
                      const float* out_color, const float* out_T_final, const int32_t* out_last,
                      float* grad_means2D, float* grad_radii, float* grad_rotations,
                      float* grad_opacities, float* grad_colors, void* stream);

/* torch.optim.Adam update (no weight decay, no amsgrad) of up to five groups
 * plus the reference's clamp x in [0,W-1], y in [0,H-1], r in [2, min(H,W)/4]
 * (simple_tile_renderer.py:122-128, 961-968, 1125-1131), one kernel. */
int dbmp_adam_step(dbmp_rasterizer* h, const dbmp_adam* cfg, void* stream);

/* Timing: hipEvent brackets around each forward/backward launch sequence;
 * replaces print_/reset_cuda_timing_stats (ext.cpp:110-121,
 * tile_rasterize.cu:783-831).  dbmp_timing_stats synchronises the recorded
 * events. */
int dbmp_timing_enable(dbmp_rasterizer* h, int on);
int dbmp_timing_stats(dbmp_rasterizer* h, double* fwd_ms_total, int64_t* fwd_count,
                      double* bwd_ms_total, int64_t* bwd_count);
int dbmp_timing_reset(dbmp_rasterizer* h);

/* ------------------------------------------------------------------------
 * Batched steps over many independent fit jobs (serving).
 *
 * The reference optimises one image per process (simple_tile_renderer.py:
 * 1043-1140: binning, TileRasterizerFunction forward/backward, F.mse_loss,
 * Adam, clamp); these entry points run that loop for a set of jobs at once,
 * each job with its own canvas, primitive count, template set and target.
 * One step launches each stage once for all jobs.  A job's parameters are
 * rows of 8 values (x, y, r, theta, opacity logit, 3 colour logits) — the
 * reference's means2D, radii, rotations, opacities and colors of one
 * primitive side by side.
 * ------------------------------------------------------------------------ */
typedef struct dbmp_batch dbmp_batch;

typedef struct {
  int32_t height, width;        /* canvas */
  int32_t num_primitives;
  int32_t num_templates, tmpl_height, tmpl_width;
  const void* quads;            /* prepared by dbmp_batch_prepare_bitmaps */
  const float* ext;             /* (T,4) template extents, likewise */
  int32_t quads_with_color;     /* 1: `quads` were prepared with per-template colour maps */
  float c_blend;
  int32_t blend_grad_exact;     /* as in dbmp_bitmaps */
  const int32_t* bmp_sel;       /* (N) */
  float* params;                /* (N,8) fp32 master */
  float* exp_avg;               /* (N,8) */
  float* exp_avg_sq;            /* (N,8) */
  void* store;                  /* (N,8) in the batch's storage type (== params for fp32) */
  const float* target;          /* (H,W,3) */
  const float* background;      /* (H,W,3), or NULL for background_value */
  float background_value;
  float* out_color;             /* (H,W,3) scratch of the step */
  float* out_alpha;             /* (H,W) */
  float* out_T;                 /* (H,W) */
  int32_t* out_last;            /* (H,W) */
  float* loss_history;          /* (loss_capacity) MSE of step t at [t], zeroed by the caller */
  int32_t loss_capacity;
  float lr[5];                  /* means2D, radii, rotations, opacities, colors */
  float alpha_upper_bound;      /* >= 0 */
  float eps_T;
  int32_t clamp;                /* x in [0,W-1], y in [0,H-1], r in [2, min(H,W)/4] */
  int32_t steps_done;           /* Adam steps already taken (the next step is steps_done+1) */
} dbmp_job;

/* storage_dtype: DBMP_F32 / F16 / BF16 for parameters, templates and colour
 * maps of every job; arithmetic is fp32. */
int dbmp_batch_create(dbmp_batch** out, int32_t storage_dtype, float beta1, float beta2, float eps);
int dbmp_batch_destroy(dbmp_batch* b);

/* Template extents (T,4) and texel records of one template set, for
 * dbmp_job, into quads_out (dbmp_batch_bitmap_bytes bytes): T*(Ht+1)*Wt alpha
 * quads of 16 bytes (fp32 storage) or 8 bytes, or, when bm carries
 * per-template colour maps (colors_orig_mode 2), T*(Ht+1)*Wt RGBA records of
 * 64 / 16 bytes holding the texels' alpha and colour neighbourhoods (set
 * quads_with_color in the job).  Sets of 4 GiB or more are rejected. */
int64_t dbmp_batch_bitmap_bytes(const dbmp_batch* b, const dbmp_bitmaps* bm);
int dbmp_batch_prepare_bitmaps(dbmp_batch* b, const dbmp_bitmaps* bm, float* ext_out,
                               void* quads_out, void* stream);

/* Replace the job set (copies the descriptors; jobs keep their own device
 * buffers).  pair_capacity bounds the (tile, primitive) pairs of all jobs
 * together; a step whose pairs exceed it is still exact (every pixel walks all
 * of its job's primitives) and reports the count through
 * dbmp_batch_pair_count.  Waits for the batch's previous table upload. */
int dbmp_batch_set_jobs(dbmp_batch* b, const dbmp_job* jobs, int32_t num_jobs,
                        int64_t pair_capacity, void* stream);

/* One optimisation step of every job: binning, forward, background
 * composite + MSE, backward, Adam, clamp.  use_graph: capture the step once
 * per job set into a hipGraph and replay it. */
int dbmp_batch_step(dbmp_batch* b, int32_t use_graph, void* stream);

/* The gradient rows (N_j,8) of job `job` (its index in the last
 * dbmp_batch_set_jobs) from the batch's most recent step, copied on `stream`
 * into `rows`: dL/d(x, y, r, theta, opacity logit, 3 colour logits) of that
 * step's MSE at the parameters it rendered, before its Adam update. */
int dbmp_batch_gradients(dbmp_batch* b, int32_t job, float* rows, void* stream);

/* Pair count of the most recent step that has completed on the device. */
int dbmp_batch_pair_count(dbmp_batch* b, int64_t* pairs);

/* ---- Layer export: per-primitive uint8 RGBA crops ------------------------
 * Replaces the reference's layer-export extension (psd_export_cuda:
 * psd_export_bindings.cpp:38-129 -> psd_export_kernels.cu:12-275) and the
 * host prefix sum between its two stages (psd_export_uint8.py:64-75).
 * All arrays fp32 except bmp_sel; colors and visibility have the sigmoid
 * applied already (psd_exporter.py:80-81). */
typedef struct {
  const float* means2D;     /* (N,2) */
  const float* radii;       /* (N) */
  const float* rotations;   /* (N) */
  const float* colors;      /* (N,3) */
  const float* visibility;  /* (N) */
  const int32_t* bmp_sel;   /* (N) */
  int32_t num_primitives;
} dbmp_layer_prims;

/* Stage 1 (compute_bounding_boxes_kernel, psd_export_kernels.cu:12-55):
 * crop box (left, top, right, bottom) (N,4) int32 of a 1.5 r square around
 * the centre scaled by scale_factor, clipped to the canvas ((0,0,1,1) when
 * empty); its size (w, h) (N,2) int32; byte_offsets (N+1) int64, the offset
 * of each layer's w*h*4 RGBA bytes in one packed buffer, and the total in
 * *d_total_bytes (device).  Reads means2D and radii only.  The reference
 * kept int32 offsets; here they are 64-bit. */
int dbmp_psd_bounds(const dbmp_layer_prims* prims, int canvas_height, int canvas_width,
                    float scale_factor, int32_t* boxes, int32_t* sizes, int64_t* byte_offsets,
                    int64_t* d_total_bytes, void* stream);

/* Stage 2 (generate_cropped_layers_kernel, psd_export_kernels.cu:61-198):
 * the packed RGBA crops into out (total_bytes, 16-byte aligned).  Colour
 * maps (Ht,Wt,3 fp32 each): color_map_mode 0 none, 1 per primitive (N maps,
 * the reference's gathered layout), 2 per template (T maps, indexed by
 * bmp_sel); blended as (1 - c_blend) c + c_blend c_o when c_blend > 0.
 * layer_alpha_max (N) int32 or NULL: each layer's largest alpha byte.
 * Arithmetic is the reference's, in IEEE fp32 with IEEE division and the
 * library's fp32 cos/sin (dbmp_sincos; the reference builds with fast math). */
int dbmp_psd_layers(const dbmp_layer_prims* prims, const float* templates, int num_templates,
                    int tmpl_h, int tmpl_w, const float* color_maps, int color_map_mode, float c_blend,
                    int canvas_height, int canvas_width, float scale_factor, float alpha_upper_bound,
                    const int32_t* boxes, const int64_t* byte_offsets, int64_t total_bytes,
                    uint8_t* out, int32_t* layer_alpha_max, void* stream);

/* The fp32 sine and cosine every kernel of this library uses (binning,
 * raster, layer export): s[i] = sin(x[i]), c[i] = cos(x[i]) for n angles,
 * at most 2 ulp from the exact values, no fp64.  No counterpart in the
 * reference (it uses torch.cos for binning and __cosf in its kernels); it is
 * exposed so that a host-side restatement of the binning can use the same
 * values (the oracle compiles the same source, csrc/dbmp_trig.h). */
int dbmp_sincos(const float* x, float* s, float* c, int64_t n, void* stream);

/* Profiling counters of a library built with DBMP_LANE_STATS (python -m
 * diffbmp_amd.build --lane-stats): out[0..2] forward (wave, list entry)
 * evaluations, those with at least one pushed lane, pushed lanes; out[3..5]
 * the same for the backward kernels of dbmp_forward / dbmp_backward*.
 * Synchronises the device; reset zeroes the counters.  The library proper
 * returns DBMP_E_ARG.  No counterpart in the reference. */
int dbmp_lane_stats(uint64_t* out, int32_t reset);

#ifdef __cplusplus
}
#endif
#endif /* DIFFBMP_AMD_H */
