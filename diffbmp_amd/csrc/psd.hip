// extern "C" layer-export entry points (declared in include/diffbmp_amd.h):
// per-primitive uint8 RGBA crops packed into one device buffer.
#include <hip/hip_runtime.h>

#include "../../include/diffbmp_amd.h"
#include "host_common.h"
#include "psd_kernels.h"

using namespace dbmp;
using namespace dbmp_host;

namespace {

int check_layer_prims(const dbmp_layer_prims* p) {
  if (!p) return fail(DBMP_E_ARG, "prims is NULL");
  if (p->num_primitives < 0) return fail(DBMP_E_ARG, "num_primitives < 0");
  if (p->num_primitives > 0 && (!p->means2D || !p->radii))
    return fail(DBMP_E_ARG, "means2D or radii is NULL");
  return 0;
}

PsdArgs psd_args(const dbmp_layer_prims* p, int H, int W, float scale) {
  PsdArgs P{};
  P.means2D = p->means2D;
  P.radii = p->radii;
  P.rotations = p->rotations;
  P.colors = p->colors;
  P.visibility = p->visibility;
  P.bmp_sel = p->bmp_sel;
  P.N = p->num_primitives;
  P.H = H;
  P.W = W;
  P.scale = scale;
  return P;
}

}  // namespace

extern "C" {

int dbmp_psd_bounds(const dbmp_layer_prims* prims, int canvas_height, int canvas_width,
                    float scale_factor, int32_t* boxes, int32_t* sizes, int64_t* byte_offsets,
                    int64_t* d_total_bytes, void* stream) {
  int rc = check_layer_prims(prims);
  if (rc) return rc;
  if (canvas_height <= 0 || canvas_width <= 0) return fail(DBMP_E_ARG, "bad canvas size");
  if (!(scale_factor > 0.f)) return fail(DBMP_E_ARG, "scale_factor must be > 0");
  if (!boxes || !sizes || !byte_offsets || !d_total_bytes) return fail(DBMP_E_ARG, "NULL output");
  hipStream_t st = (hipStream_t)stream;
  const PsdArgs P = psd_args(prims, canvas_height, canvas_width, scale_factor);
  if (P.N > 0) {
    hipLaunchKernelGGL(k_psd_bounds, dim3(ceil_div(P.N, 256)), dim3(256), 0, st, P, (int4*)boxes,
                       (int2*)sizes);
    LAUNCH_CHECK("k_psd_bounds");
  }
  hipLaunchKernelGGL(k_psd_offsets, dim3(1), dim3(kScanThreads), 0, st, (const int2*)sizes, P.N,
                     (long long*)byte_offsets, (long long*)d_total_bytes);
  LAUNCH_CHECK("k_psd_offsets");
  return 0;
}

int dbmp_psd_layers(const dbmp_layer_prims* prims, const float* templates, int num_templates,
                    int tmpl_h, int tmpl_w, const float* color_maps, int color_map_mode, float c_blend,
                    int canvas_height, int canvas_width, float scale_factor, float alpha_upper_bound,
                    const int32_t* boxes, const int64_t* byte_offsets, int64_t total_bytes,
                    uint8_t* out, int32_t* layer_alpha_max, void* stream) {
  int rc = check_layer_prims(prims);
  if (rc) return rc;
  if (prims->num_primitives > 0 &&
      (!prims->rotations || !prims->colors || !prims->visibility || !prims->bmp_sel))
    return fail(DBMP_E_ARG, "a primitive rotation/colour/visibility/selection pointer is NULL");
  if (!templates || num_templates <= 0 || tmpl_h < 2 || tmpl_w < 2)
    return fail(DBMP_E_ARG, "templates must be (T>0, H>=2, W>=2)");
  if (color_map_mode < 0 || color_map_mode > 2) return fail(DBMP_E_ARG, "bad color_map_mode");
  if (color_map_mode != 0 && !color_maps) return fail(DBMP_E_ARG, "color_maps is NULL");
  if (canvas_height <= 0 || canvas_width <= 0) return fail(DBMP_E_ARG, "bad canvas size");
  if (!(scale_factor > 0.f)) return fail(DBMP_E_ARG, "scale_factor must be > 0");
  if (!boxes || !byte_offsets) return fail(DBMP_E_ARG, "NULL boxes/offsets");
  if (total_bytes < 0 || (total_bytes & 3) != 0) return fail(DBMP_E_ARG, "total_bytes must be a multiple of 4");
  if (total_bytes > 0 && !out) return fail(DBMP_E_ARG, "out is NULL");
  if (((uintptr_t)out & 15) != 0) return fail(DBMP_E_ARG, "out must be 16-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  PsdArgs P = psd_args(prims, canvas_height, canvas_width, scale_factor);
  P.templates = templates;
  P.T = num_templates;
  P.Ht = tmpl_h;
  P.Wt = tmpl_w;
  P.cmaps = color_map_mode != 0 ? color_maps : nullptr;
  P.cmap_mode = color_map_mode;
  P.c_blend = c_blend;
  P.alpha_ub = alpha_upper_bound;
  if (layer_alpha_max && P.N > 0) {
    hipLaunchKernelGGL(k_fill_f32, dim3(ceil_div(P.N, 256)), dim3(256), 0, st, (float*)layer_alpha_max,
                       (long long)P.N, 0.f);   // int 0 and float +0 share their bits
    LAUNCH_CHECK("k_fill_f32");
  }
  const long long total_px = total_bytes / 4;
  if (P.N == 0 || total_px == 0) return 0;
  const long long blocks = (total_px + kPsdPix - 1) / kPsdPix;
  if (blocks > 0x7fffffffLL) return fail(DBMP_E_ARG, "layer buffer too large");
  hipLaunchKernelGGL(k_psd_layers, dim3((unsigned)blocks), dim3(kPsdThreads), 0, st, P,
                     (const int4*)boxes, (const long long*)byte_offsets, total_px, (unsigned*)out,
                     (int*)layer_alpha_max);
  LAUNCH_CHECK("k_psd_layers");
  return 0;
}

}  // extern "C"
