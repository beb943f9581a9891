/*
 * dbmp_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of DiffBMP's tile rasterizer (binning -> forward composite ->
 * backward adjoint).  It is the checker for the HIP path and nothing else: only
 * tests, the smoke check and the CPU-baseline leg of bench.py may load it.  The
 * product path never calls it.
 *
 * One source, two builds (see oracle/Makefile):
 *   liboracle_f64.so  REAL=double  -- the FP64 oracle.  Per-sample decisions
 *                     (push box, texture bounds, floor) are evaluated either in
 *                     IEEE fp32 exactly as the reference states them
 *                     (geom64 = 0) or fully in fp64 (geom64 = 1, used by the
 *                     finite-difference harness).  Every value (mask,
 *                     compositing, gradients) is fp64.
 *   liboracle_f32.so  REAL=float + OpenMP -- the reference algorithm in plain
 *                     fp32 (per-pixel caches, O(K^2) suffix loop in backward),
 *                     timed as the CPU baseline ("port").
 *
 * Reference files restated (paths relative to the reference checkout):
 *   binning      pydiffbmp/core/renderer/vector_renderer.py:151-192 (extents),
 *                :194-252 (rotated corner AABB), :401-444 (closed tile test),
 *                pydiffbmp/core/renderer/simple_tile_renderer.py:721-766 (CSR,
 *                ascending primitive id per tile)
 *   forward      cuda_tile_rasterizer/cuda_kernels/tile_forward.cu:96-398,
 *                colour map :17-89, helpers tile_common.h:125-198
 *   backward     cuda_tile_rasterizer/cuda_kernels/tile_backward.cu:101-452,
 *                helper tile_common.h:201-239
 *
 * Deliberate, documented differences from the reference's CUDA build:
 *   - cos/sin are dbmp_sincosf (diffbmp_amd/csrc/dbmp_trig.h, fp32, <= 2 ulp),
 *     the same source the GPU kernels compile, so both make the same fp32
 *     decisions; the reference uses torch.cos (binning) and __cosf (kernel).
 *   - 1/r is IEEE division; the reference builds with --use_fast_math.
 *   - no FMA contraction in fp32 decision arithmetic (-ffp-contract=off).
 *   - the per-pixel cache has no capacity limit (reference defect: overflow
 *     past max_prims_per_pixel is undefined, tile_forward.cu:219-223).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../diffbmp_amd/csrc/dbmp_trig.h"

#ifndef REAL
#define REAL double
#endif

typedef struct {
    int H, W, ts;
    float alpha_ub;
    int N, T, Ht, Wt;
    const float *means2D;    /* (N,2) */
    const float *radii;      /* (N)   */
    const float *rotations;  /* (N)   */
    const float *opacities;  /* (N) logits */
    const float *colors;     /* (N,3) logits */
    const float *colors_orig;/* (N,Ht,Wt,3), or (T,Ht,Wt,3) when colors_orig_per_template,
                                in [0,1]; or NULL */
    float c_blend;
    const float *templates;  /* (T,Ht,Wt) */
    const int *bmp_sel;      /* (N) */
    const int *tile_offsets; /* (tiles+1) */
    const int *tile_indices; /* (P) */
    int geom64;              /* 0: fp32 decisions, 1: fp64 throughout */
    const int *tile_subset;  /* NULL: all tiles */
    int n_subset;
    int blend_grad_exact;    /* 0: the reference's colour gradient under c_blend > 0
                                (tile_backward.cu:253-258); 1: the exact gradient of
                                the blended colour (logits and colour-map geometry) */
    int colors_orig_per_template;
} dbmp_oracle_scene;

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */

/* sigmoidf_safe, tile_common.h:125-133 */
static REAL sigm(REAL x) {
    if (x >= 0) {
        REAL z = (REAL)exp(-(double)x);
        return (REAL)1 / ((REAL)1 + z);
    }
    REAL z = (REAL)exp((double)x);
    return z / ((REAL)1 + z);
}

/* fp32 cos / sin shared with the GPU kernels (dbmp_trig.h) */
static float cos_f(float t) { float s, c; dbmp_sincosf(t, &s, &c); return c; }
static float sin_f(float t) { float s, c; dbmp_sincosf(t, &s, &c); return s; }

int oracle_sincos_array(int n, const float *x, float *s, float *c) {
    for (int i = 0; i < n; ++i) dbmp_sincosf(x[i], s + i, c + i);
    return 0;
}

/* Largest error in ulp of the fp32 result of dbmp_sincosf against the C
 * library's double sin / cos, over the fp32 values with bit patterns
 * lo, lo + stride, ... < hi (positive angles).  out[0] sin, out[1] cos. */
static double ulp_err(float got, double want) {
    const float w = (float)want;
    int e;
    frexpf(fabsf(w) > 0.f ? w : 1e-38f, &e);
    double ulp = ldexp(1.0, e - 24);
    if (ulp < ldexp(1.0, -149)) ulp = ldexp(1.0, -149);
    return fabs((double)got - want) / ulp;
}

int oracle_sincos_scan(unsigned lo, unsigned hi, unsigned stride, double *out) {
    double ms = 0, mc = 0;
#pragma omp parallel for reduction(max : ms, mc) schedule(static)
    for (long long b = lo; b < (long long)hi; b += stride) {
        union { unsigned u; float f; } v;
        v.u = (unsigned)b;
        float s, c;
        dbmp_sincosf(v.f, &s, &c);
        const double es = ulp_err(s, sin((double)v.f)), ec = ulp_err(c, cos((double)v.f));
        if (es > ms) ms = es;
        if (ec > mc) mc = ec;
    }
    out[0] = ms;
    out[1] = mc;
    return 0;
}

/* One sample of primitive n at integer pixel (px,py).  Returns 1 if the
 * sample is pushed (passes the template box test and has mask > 1e-6),
 * tile_forward.cu:187-276. */
typedef struct {
    REAL mask, gxt, gyt;      /* mask and its texel-space gradient */
    REAL u, v, ndx, ndy, c, s, rinv;
    REAL co[3];               /* colour-map sample (c_blend > 0) */
    REAL dco_x[3], dco_y[3];  /* its derivative w.r.t. the texel coordinates */
} sample_t;

static REAL tex_at(const float *t, int Ht, int Wt, int yy, int xx) {
    if (xx < 0 || xx >= Wt || yy < 0 || yy >= Ht) return 0;
    return (REAL)t[yy * Wt + xx];
}

/* bilinear value + gradient, zero padded, tile_common.h:175-239 */
static int bilinear(const float *t, int Ht, int Wt, REAL tx, REAL ty,
                    int x0, int y0, REAL wx, REAL wy, sample_t *o) {
    (void)tx; (void)ty;
    REAL v00 = tex_at(t, Ht, Wt, y0, x0);
    REAL v01 = tex_at(t, Ht, Wt, y0, x0 + 1);
    REAL v10 = tex_at(t, Ht, Wt, y0 + 1, x0);
    REAL v11 = tex_at(t, Ht, Wt, y0 + 1, x0 + 1);
    REAL one = 1;
    REAL f = v00 * (one - wx) * (one - wy) + v01 * wx * (one - wy) +
             v10 * (one - wx) * wy + v11 * wx * wy;
    o->gxt = (v01 - v00) * (one - wy) + (v11 - v10) * wy;
    o->gyt = (v10 - v00) * (one - wx) + (v11 - v01) * wx;
    if (f < 0) f = 0;
    if (f > 1) f = 1;
    o->mask = f;
    return 1;
}

/* sample_color_map, tile_forward.cu:17-89 (clamp-to-edge bilinear RGB) */
static void color_map(const float *cm, int Ht, int Wt, REAL cx, REAL cy, REAL co[3],
                      REAL dcx[3], REAL dcy[3]) {
    if (cx < 0) cx = 0;
    if (cx > (REAL)(Wt - 1)) cx = (REAL)(Wt - 1);
    if (cy < 0) cy = 0;
    if (cy > (REAL)(Ht - 1)) cy = (REAL)(Ht - 1);
    int x0 = (int)cx, y0 = (int)cy;
    int x1 = x0 + 1 < Wt - 1 ? x0 + 1 : Wt - 1;
    int y1 = y0 + 1 < Ht - 1 ? y0 + 1 : Ht - 1;
    REAL fx = cx - (REAL)x0, fy = cy - (REAL)y0;
    for (int ch = 0; ch < 3; ++ch) {
        REAL c00 = cm[3 * (y0 * Wt + x0) + ch], c01 = cm[3 * (y1 * Wt + x0) + ch];
        REAL c10 = cm[3 * (y0 * Wt + x1) + ch], c11 = cm[3 * (y1 * Wt + x1) + ch];
        REAL c0 = c00 * (1 - fx) + c10 * fx;
        REAL c1 = c01 * (1 - fx) + c11 * fx;
        co[ch] = c0 * (1 - fy) + c1 * fy;
        /* derivative of the bilinear interpolant inside the cell; a clamped
         * coordinate (x1 == x0) has zero derivative */
        dcx[ch] = (c10 - c00) * (1 - fy) + (c11 - c01) * fy;
        dcy[ch] = c1 - c0;
    }
}

/* The colour-map sample above at texel coordinates (cx, cy) of one (Ht,Wt,3)
 * map, with its derivative; exported so tests pin it with closed forms. */
int oracle_color_map_sample(const float *cm, int Ht, int Wt, double cx, double cy, REAL *co,
                            REAL *dcx, REAL *dcy) {
    color_map(cm, Ht, Wt, (REAL)cx, (REAL)cy, co, dcx, dcy);
    return 0;
}

static int sample_prim(const dbmp_oracle_scene *s, int n, int px, int py, sample_t *o) {
    const int Ht = s->Ht, Wt = s->Wt;
    const int ti = s->bmp_sel[n];
    /* invalid template index: circle test, then mask stays 0 -> never pushed
     * (tile_forward.cu:211-216, 248-276) */
    if (ti < 0 || ti >= s->T) return 0;
    const float *tex = s->templates + (size_t)ti * Ht * Wt;
    int x0, y0;
    REAL wx, wy, tx, ty;
    if (!s->geom64) {
        /* decisions in IEEE fp32, in the reference's operation order */
        const float mx = s->means2D[2 * n], my = s->means2D[2 * n + 1];
        const float r = s->radii[n];
        const float dx = (float)px - mx, dy = (float)py - my;
        const float tw = (float)Wt, th = (float)Ht;
        const float md = tw > th ? tw : th;
        const float sf = r / (md * 0.5f);
        const float hw = tw * 0.5f * sf, hh = th * 0.5f * sf;
        if (fabsf(dx) > hw || fabsf(dy) > hh) return 0;
        const float rinv = r > 1e-6f ? 1.f / r : 1e6f;
        const float phi = s->rotations[n];
        const float c = cos_f(phi), sn = sin_f(phi);
        const float ndx = dx * rinv, ndy = dy * rinv;
        /* the first product fused into the add, as nvcc contracts the
         * reference's expressions by default (tile_forward.cu:259-260) */
        const float u = fmaf(c, ndx, sn * ndy);
        const float v = fmaf(-sn, ndx, c * ndy);
        const float ftx = (u + 1.f) * 0.5f * (float)(Wt - 1);
        const float fty = (v + 1.f) * 0.5f * (float)(Ht - 1);
        if (s->colors_orig && s->c_blend > 0.f)
            color_map(s->colors_orig + (size_t)(s->colors_orig_per_template ? ti : n) * Ht * Wt * 3, Ht, Wt, ftx, fty, o->co,
                      o->dco_x, o->dco_y);
        /* NaN coordinates are rejected too (undefined in the reference) */
        if (!(ftx >= 0.f && ftx <= (float)(Wt - 1) && fty >= 0.f && fty <= (float)(Ht - 1))) return 0;
        x0 = (int)floorf(ftx);
        y0 = (int)floorf(fty);
        wx = (REAL)(ftx - (float)x0);
        wy = (REAL)(fty - (float)y0);
        tx = ftx; ty = fty;
        o->u = u; o->v = v; o->ndx = ndx; o->ndy = ndy; o->c = c; o->s = sn; o->rinv = rinv;
    } else {
        const double mx = s->means2D[2 * n], my = s->means2D[2 * n + 1];
        const double r = s->radii[n];
        const double dx = (double)px - mx, dy = (double)py - my;
        const double tw = Wt, th = Ht;
        const double md = tw > th ? tw : th;
        const double sf = r / (md * 0.5);
        const double hw = tw * 0.5 * sf, hh = th * 0.5 * sf;
        if (fabs(dx) > hw || fabs(dy) > hh) return 0;
        const double rinv = r > 1e-6 ? 1.0 / r : 1e6;
        const double phi = s->rotations[n];
        const double c = cos(phi), sn = sin(phi);
        const double ndx = dx * rinv, ndy = dy * rinv;
        const double u = c * ndx + sn * ndy;
        const double v = -sn * ndx + c * ndy;
        const double dtx = (u + 1.0) * 0.5 * (Wt - 1);
        const double dty = (v + 1.0) * 0.5 * (Ht - 1);
        if (s->colors_orig && s->c_blend > 0.f)
            color_map(s->colors_orig + (size_t)(s->colors_orig_per_template ? ti : n) * Ht * Wt * 3, Ht, Wt, (REAL)dtx, (REAL)dty, o->co,
                      o->dco_x, o->dco_y);
        if (!(dtx >= 0 && dtx <= (double)(Wt - 1) && dty >= 0 && dty <= (double)(Ht - 1))) return 0;
        x0 = (int)floor(dtx);
        y0 = (int)floor(dty);
        wx = (REAL)(dtx - x0);
        wy = (REAL)(dty - y0);
        tx = (REAL)dtx; ty = (REAL)dty;
        o->u = (REAL)u; o->v = (REAL)v; o->ndx = (REAL)ndx; o->ndy = (REAL)ndy;
        o->c = (REAL)c; o->s = (REAL)sn; o->rinv = (REAL)rinv;
    }
    bilinear(tex, Ht, Wt, tx, ty, x0, y0, wx, wy, o);
    if (o->mask <= (REAL)1e-6) return 0;
    return 1;
}

/* ------------------------------------------------------------------ */
/* A1: template extents (vector_renderer.py:151-192)                   */
/* ------------------------------------------------------------------ */
int oracle_template_extents(const float *tmpl, int T, int Ht, int Wt, float *ext) {
    for (int t = 0; t < T; ++t) {
        const float *m = tmpl + (size_t)t * Ht * Wt;
        int any = 0, cmin = Wt, cmax = -1, rmin = Ht, rmax = -1;
        for (int yy = 0; yy < Ht; ++yy)
            for (int xx = 0; xx < Wt; ++xx)
                if (m[yy * Wt + xx] > 1e-6f) {
                    any = 1;
                    if (xx < cmin) cmin = xx;
                    if (xx > cmax) cmax = xx;
                    if (yy < rmin) rmin = yy;
                    if (yy > rmax) rmax = yy;
                }
        if (!any) {
            ext[4 * t + 0] = -1.f; ext[4 * t + 1] = 1.f;
            ext[4 * t + 2] = -1.f; ext[4 * t + 3] = 1.f;
            continue;
        }
        /* u = (col / (W-1)) * 2 - 1 in fp32; min/max commute with this
         * monotone map, so map the extreme indices */
        ext[4 * t + 0] = ((float)cmin / (float)(Wt - 1)) * 2.f - 1.f;
        ext[4 * t + 1] = ((float)cmax / (float)(Wt - 1)) * 2.f - 1.f;
        ext[4 * t + 2] = ((float)rmin / (float)(Ht - 1)) * 2.f - 1.f;
        ext[4 * t + 3] = ((float)rmax / (float)(Ht - 1)) * 2.f - 1.f;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* A2/A3: rotated template AABB and closed tile test                   */
/* ------------------------------------------------------------------ */
static void world_bbox(float x, float y, float r, float th, const float *e,
                       float *xmin, float *xmax, float *ymin, float *ymax) {
    const float c = cos_f(th), s = sin_f(th);
    const float cu[4] = {e[0], e[1], e[0], e[1]};
    const float cv[4] = {e[2], e[2], e[3], e[3]};
    float wx[4], wy[4];
    for (int i = 0; i < 4; ++i) {
        float a = cu[i] * c, b = cv[i] * s;
        float d = a - b;
        wx[i] = x + r * d;
        float a2 = cu[i] * s, b2 = cv[i] * c;
        float d2 = a2 + b2;
        wy[i] = y + r * d2;
    }
    /* torch.min/max propagate NaN */
    float mn = wx[0], mxv = wx[0], mny = wy[0], mxy = wy[0];
    for (int i = 1; i < 4; ++i) {
        if (isnan(wx[i]) || isnan(mn)) mn = NAN; else if (wx[i] < mn) mn = wx[i];
        if (isnan(wx[i]) || isnan(mxv)) mxv = NAN; else if (wx[i] > mxv) mxv = wx[i];
        if (isnan(wy[i]) || isnan(mny)) mny = NAN; else if (wy[i] < mny) mny = wy[i];
        if (isnan(wy[i]) || isnan(mxy)) mxy = NAN; else if (wy[i] > mxy) mxy = wy[i];
    }
    *xmin = mn; *xmax = mxv; *ymin = mny; *ymax = mxy;
}

static int tile_hit(float bmin, float bmax, int t, int ts, int lim) {
    const int s0 = t * ts;
    const int e0 = s0 + ts < lim ? s0 + ts : lim;
    return !((bmax < (float)s0) || (bmin > (float)e0));
}

/* bboxes (N,4) = xmin,xmax,ymin,ymax; exposed so tests can build tie bands */
int oracle_bin_bboxes(int N, const float *means2D, const float *radii, const float *rot,
                      const int *sel, const float *ext, int T, float *bb) {
    for (int n = 0; n < N; ++n) {
        int t = sel[n];
        if (t < 0 || t >= T) t = 0;
        world_bbox(means2D[2 * n], means2D[2 * n + 1], radii[n], rot[n], ext + 4 * t,
                   bb + 4 * n, bb + 4 * n + 1, bb + 4 * n + 2, bb + 4 * n + 3);
    }
    return 0;
}

/* Tile lists: per tile (row-major) ascending primitive ids.  tile_offsets has
 * tiles+1 entries; tile_indices may be NULL to only count.  Returns P. */
long long oracle_bin_lists(int N, const float *means2D, const float *radii, const float *rot,
                           const int *sel, const float *ext, int T, int H, int W, int ts,
                           int *tile_offsets, int *tile_indices) {
    const int tX = (W + ts - 1) / ts, tY = (H + ts - 1) / ts, tiles = tX * tY;
    float *bb = (float *)malloc(sizeof(float) * 4 * (size_t)(N > 0 ? N : 1));
    oracle_bin_bboxes(N, means2D, radii, rot, sel, ext, T, bb);
    memset(tile_offsets, 0, sizeof(int) * (tiles + 1));
    /* a primitive whose template index is out of range is never drawn
     * (tile_forward.cu:211-216,248-276), so it is left out of every list */
    for (int n = 0; n < N; ++n)
        for (int ty = 0; ty < tY && sel[n] >= 0 && sel[n] < T; ++ty) {
            if (!tile_hit(bb[4 * n + 2], bb[4 * n + 3], ty, ts, H)) continue;
            for (int tx = 0; tx < tX; ++tx)
                if (tile_hit(bb[4 * n], bb[4 * n + 1], tx, ts, W)) tile_offsets[ty * tX + tx + 1]++;
        }
    for (int t = 0; t < tiles; ++t) tile_offsets[t + 1] += tile_offsets[t];
    long long P = tile_offsets[tiles];
    if (tile_indices) {
        int *fill = (int *)malloc(sizeof(int) * tiles);
        memcpy(fill, tile_offsets, sizeof(int) * tiles);
        for (int n = 0; n < N; ++n)
            for (int ty = 0; ty < tY && sel[n] >= 0 && sel[n] < T; ++ty) {
                if (!tile_hit(bb[4 * n + 2], bb[4 * n + 3], ty, ts, H)) continue;
                for (int tx = 0; tx < tX; ++tx)
                    if (tile_hit(bb[4 * n], bb[4 * n + 1], tx, ts, W))
                        tile_indices[fill[ty * tX + tx]++] = n;
            }
        free(fill);
    }
    free(bb);
    return P;
}

/* ------------------------------------------------------------------ */
/* per-pixel list walk shared by forward and backward                  */
/* ------------------------------------------------------------------ */
typedef struct {
    int cap;
    int *n;
    REAL *a, *col, *T;   /* alpha, colour (3 per entry), transmittance */
    sample_t *smp;
} pixcache;

static void cache_reserve(pixcache *c, int need) {
    if (need <= c->cap) return;
    int cap = need + 64;
    c->n = (int *)realloc(c->n, sizeof(int) * cap);
    c->a = (REAL *)realloc(c->a, sizeof(REAL) * cap);
    c->col = (REAL *)realloc(c->col, sizeof(REAL) * 3 * cap);
    c->T = (REAL *)realloc(c->T, sizeof(REAL) * cap);
    c->smp = (sample_t *)realloc(c->smp, sizeof(sample_t) * cap);
    c->cap = cap;
}

static void cache_free(pixcache *c) {
    free(c->n); free(c->a); free(c->col); free(c->T); free(c->smp);
}

/* Phase 1 of tile_forward.cu:182-359: fill the per-pixel cache in list order */
static int walk_pixel(const dbmp_oracle_scene *s, const int *ids, int K, int px, int py,
                      pixcache *pc) {
    cache_reserve(pc, K);
    int k = 0;
    for (int kk = 0; kk < K; ++kk) {
        const int n = ids[kk];
        if (n < 0 || n >= s->N) continue;
        sample_t sm;
        if (!sample_prim(s, n, px, py, &sm)) continue;
        const REAL sop = sigm((REAL)s->opacities[n]);
        const REAL a = (REAL)s->alpha_ub * sop * sm.mask;
        for (int ch = 0; ch < 3; ++ch) {
            REAL ci = sigm((REAL)s->colors[3 * n + ch]);
            if (s->colors_orig && s->c_blend > 0.f)
                ci = ((REAL)1 - (REAL)s->c_blend) * ci + (REAL)s->c_blend * sm.co[ch];
            pc->col[3 * k + ch] = ci;
        }
        pc->a[k] = a;
        pc->n[k] = n;
        pc->smp[k] = sm;
        ++k;
    }
    /* Phase 2, tile_forward.cu:364-391 */
    REAL T = 1;
    for (int j = 0; j < k; ++j) {
        pc->T[j] = T;
        REAL om = (REAL)1 - pc->a[j];
        T *= om > 0 ? om : (REAL)0;
    }
    return k;
}

static int n_work_tiles(const dbmp_oracle_scene *s) {
    const int tX = (s->W + s->ts - 1) / s->ts, tY = (s->H + s->ts - 1) / s->ts;
    return s->tile_subset ? s->n_subset : tX * tY;
}

/* out_color (H,W,3) premultiplied over black, out_alpha (H,W), push_count
 * (H,W) = number of cached layers per pixel (may be NULL).  Pixels of tiles
 * outside tile_subset are left untouched. */
int oracle_forward(const dbmp_oracle_scene *s, REAL *out_color, REAL *out_alpha, int *push_count) {
    const int tX = (s->W + s->ts - 1) / s->ts;
    const int nt = n_work_tiles(s);
#pragma omp parallel
    {
        pixcache pc = {0};
#pragma omp for schedule(dynamic, 1)
        for (int it = 0; it < nt; ++it) {
            const int t = s->tile_subset ? s->tile_subset[it] : it;
            const int tx = t % tX, ty = t / tX;
            const int st = s->tile_offsets[t], K = s->tile_offsets[t + 1] - st;
            const int *ids = s->tile_indices + st;
            for (int py = ty * s->ts; py < (ty + 1) * s->ts && py < s->H; ++py)
                for (int px = tx * s->ts; px < (tx + 1) * s->ts && px < s->W; ++px) {
                    const int k = walk_pixel(s, ids, K, px, py, &pc);
                    REAL C[3] = {0, 0, 0}, T = 1;
                    for (int j = 0; j < k; ++j) {
                        for (int ch = 0; ch < 3; ++ch) C[ch] += pc.T[j] * pc.col[3 * j + ch] * pc.a[j];
                        REAL om = (REAL)1 - pc.a[j];
                        T *= om > 0 ? om : (REAL)0;
                    }
                    const size_t p = (size_t)py * s->W + px;
                    for (int ch = 0; ch < 3; ++ch) out_color[3 * p + ch] = C[ch];
                    out_alpha[p] = (REAL)1 - T;
                    if (push_count) push_count[p] = k;
                }
        }
        cache_free(&pc);
    }
    return 0;
}

/* Gradients w.r.t. (means2D (N,2), radii, rotations, opacity logits, colour
 * logits (N,3)), accumulated (+=) into the outputs; tile_backward.cu:101-410
 * with its O(K^2) suffix loop (:214-226). */
int oracle_backward(const dbmp_oracle_scene *s, const REAL *gCin, const REAL *gAin,
                    REAL *g_means, REAL *g_r, REAL *g_th, REAL *g_v, REAL *g_c) {
    const int tX = (s->W + s->ts - 1) / s->ts;
    const int nt = n_work_tiles(s);
    const int N = s->N;
    const REAL hsx = (REAL)0.5 * (REAL)(s->Wt - 1), hsy = (REAL)0.5 * (REAL)(s->Ht - 1);
#pragma omp parallel
    {
        REAL *acc = (REAL *)calloc((size_t)N * 8, sizeof(REAL));
        pixcache pc = {0};
#pragma omp for schedule(dynamic, 1)
        for (int it = 0; it < nt; ++it) {
            const int t = s->tile_subset ? s->tile_subset[it] : it;
            const int tx = t % tX, ty = t / tX;
            const int st = s->tile_offsets[t], K = s->tile_offsets[t + 1] - st;
            const int *ids = s->tile_indices + st;
            for (int py = ty * s->ts; py < (ty + 1) * s->ts && py < s->H; ++py)
                for (int px = tx * s->ts; px < (tx + 1) * s->ts && px < s->W; ++px) {
                    const size_t p = (size_t)py * s->W + px;
                    const REAL gC[3] = {gCin[3 * p], gCin[3 * p + 1], gCin[3 * p + 2]};
                    const REAL gA = gAin ? gAin[p] : 0;
                    const int k = walk_pixel(s, ids, K, px, py, &pc);
                    for (int j = 0; j < k; ++j) {
                        REAL S[3] = {0, 0, 0}, B = 1;
                        for (int m = j + 1; m < k; ++m) {
                            for (int ch = 0; ch < 3; ++ch) S[ch] += pc.col[3 * m + ch] * pc.a[m] * pc.T[m];
                            REAL om = (REAL)1 - pc.a[m];
                            B *= om > 0 ? om : (REAL)0;
                        }
                        const int n = pc.n[j];
                        const REAL a = pc.a[j], Tk = pc.T[j];
                        const sample_t *sm = &pc.smp[j];
                        REAL *g = acc + (size_t)8 * n;
                        /* (1) colour logits, tile_backward.cu:228-262 (no
                         * (1-c_blend) factor in the reference; the exact
                         * gradient has it) */
                        const int blend = s->colors_orig && s->c_blend > 0.f;
                        const int exact = blend && s->blend_grad_exact;
                        const REAL cw = exact ? (REAL)1 - (REAL)s->c_blend : (REAL)1;
                        for (int ch = 0; ch < 3; ++ch) {
                            const REAL si = sigm((REAL)s->colors[3 * n + ch]);
                            g[5 + ch] += gC[ch] * (a * Tk) * cw * (si * ((REAL)1 - si));
                        }
                        /* (2) dL/dalpha, :264-271 */
                        REAL om = (REAL)1 - a;
                        const REAL inv1m = (REAL)1 / (om > (REAL)1e-6 ? om : (REAL)1e-6);
                        REAL dLda = 0;
                        for (int ch = 0; ch < 3; ++ch)
                            dLda += gC[ch] * (Tk * pc.col[3 * j + ch] - S[ch] * inv1m);
                        dLda += gA * (Tk * B);
                        /* (3)-(6), :273-343 */
                        const REAL sop = sigm((REAL)s->opacities[n]);
                        g[4] += dLda * ((REAL)s->alpha_ub * sm->mask * sop * ((REAL)1 - sop));
                        const REAL dLdm = dLda * ((REAL)s->alpha_ub * sop);
                        REAL dLdu = dLdm * (sm->gxt * hsx);
                        REAL dLdv = dLdm * (sm->gyt * hsy);
                        if (exact) {
                            /* C = sum_k T_k a_k col_k, col = (1-k) sig + k c_o(tx, ty) */
                            REAL gx = 0, gy = 0;
                            for (int ch = 0; ch < 3; ++ch) {
                                gx += gC[ch] * sm->dco_x[ch];
                                gy += gC[ch] * sm->dco_y[ch];
                            }
                            const REAL kw = a * Tk * (REAL)s->c_blend;
                            dLdu += kw * gx * hsx;
                            dLdv += kw * gy * hsy;
                        }
                        const REAL c = sm->c, sn = sm->s, ri = sm->rinv;
                        g[0] += dLdu * (-c * ri) + dLdv * (sn * ri);
                        g[1] += dLdu * (-sn * ri) + dLdv * (-c * ri);
                        g[2] += dLdu * (-sm->u * ri) + dLdv * (-sm->v * ri);
                        g[3] += dLdu * (-sn * sm->ndx + c * sm->ndy) + dLdv * (-c * sm->ndx - sn * sm->ndy);
                    }
                }
        }
#pragma omp critical
        {
            for (int n = 0; n < N; ++n) {
                const REAL *g = acc + (size_t)8 * n;
                g_means[2 * n] += g[0];
                g_means[2 * n + 1] += g[1];
                g_r[n] += g[2];
                g_th[n] += g[3];
                g_v[n] += g[4];
                g_c[3 * n] += g[5];
                g_c[3 * n + 1] += g[6];
                g_c[3 * n + 2] += g[7];
            }
        }
        free(acc);
        cache_free(&pc);
    }
    return 0;
}

/* Total samples that pass the push test and mask gate over the work tiles
 * (the "E_full" sample count of the reference's list walk). */
long long oracle_count_samples(const dbmp_oracle_scene *s) {
    const int tX = (s->W + s->ts - 1) / s->ts;
    const int nt = n_work_tiles(s);
    long long tot = 0;
#pragma omp parallel reduction(+ : tot)
    {
        pixcache pc = {0};
#pragma omp for schedule(dynamic, 1)
        for (int it = 0; it < nt; ++it) {
            const int t = s->tile_subset ? s->tile_subset[it] : it;
            const int tx = t % tX, ty = t / tX;
            const int st = s->tile_offsets[t], K = s->tile_offsets[t + 1] - st;
            for (int py = ty * s->ts; py < (ty + 1) * s->ts && py < s->H; ++py)
                for (int px = tx * s->ts; px < (tx + 1) * s->ts && px < s->W; ++px)
                    tot += walk_pixel(s, s->tile_indices + st, K, px, py, &pc);
        }
        cache_free(&pc);
    }
    return tot;
}

int oracle_real_bytes(void) { return (int)sizeof(REAL); }

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
