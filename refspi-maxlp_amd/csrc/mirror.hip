// Single-kernel mirrors of the reference's CuPy RawKernels (kernels.cu), FP64.
//
// Same arguments and results as the reference kernels; the launch shapes are gfx950's:
// the fastest-varying thread index runs along a row (y) so that loads and stores of
// size x size images are coalesced (kernels.cu maps thread x to the row index x).
#include <cmath>

#include "refspi_amd.h"
#include "rsp_common.hpp"

namespace rsp {

__global__ void k_slice_gen(const double* __restrict__ model, double ac, double as,
                            double scale, int size, const double* __restrict__ bg,
                            int log_flag, double* __restrict__ view) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = blockIdx.y;
  if (x >= size || y >= size) return;
  const int64_t t = (int64_t)x * size + y;
  double tx, ty, v;
  rot_source(x, y, size / 2, ac, as, tx, ty);
  if (!bilinear(model, size, tx, ty, v)) {
    view[t] = log_flag ? -1000. : 0.;
    return;
  }
  v = v * scale;
  v = v + bg[t];
  view[t] = log_flag ? log_view(v) : v;
}

__global__ void k_slice_gen_holo(const double* __restrict__ model_c, double shiftx,
                                 double shifty, double diameter, double rel_scale,
                                 double scale, int size, const double* __restrict__ bg,
                                 int log_flag, double* __restrict__ view) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = blockIdx.y;
  if (x >= size || y >= size) return;
  const int64_t t = (int64_t)x * size + y;
  double v = holo_pixel(model_c, x, y, size, shiftx, shifty, diameter, rel_scale);
  v = v * scale;
  v = v + bg[t];
  view[t] = log_flag ? log_view(v) : v;
}

// kernels.cu:116-137: one thread per frame, sequential sum in the reference order.
__global__ void k_calc_prob_all(const double* __restrict__ lview,
                                const int32_t* __restrict__ mask, int64_t ndata,
                                const int32_t* __restrict__ ones,
                                const int32_t* __restrict__ multi,
                                const int64_t* __restrict__ o_acc,
                                const int64_t* __restrict__ m_acc,
                                const int32_t* __restrict__ p_o,
                                const int32_t* __restrict__ p_m,
                                const int32_t* __restrict__ c_m, double init,
                                const double* __restrict__ scales,
                                double* __restrict__ prob_r) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndata) return;
  double acc = init * scales[d];
  for (int64_t t = o_acc[d]; t < o_acc[d] + ones[d]; ++t) {
    const int pixel = p_o[t];
    if (mask[pixel] < 1) acc += lview[pixel];
  }
  for (int64_t t = m_acc[d]; t < m_acc[d] + multi[d]; ++t) {
    const int pixel = p_m[t];
    if (mask[pixel] < 1) acc += lview[pixel] * c_m[t];
  }
  prob_r[d] = acc;
}

// kernels.cu:172-183: f_dt[d*npix + t] = fobj_t[t] + fref_d[d] (complex, interleaved).
__global__ void k_get_f_dt(const double* __restrict__ fobj, const double* __restrict__ fref,
                           int64_t ndata, int64_t npix, double* __restrict__ f_dt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndata * npix) return;
  const int64_t d = i / npix, t = i - d * npix;
  f_dt[2 * i] = fobj[2 * t] + fref[2 * d];
  f_dt[2 * i + 1] = fobj[2 * t + 1] + fref[2 * d + 1];
}

}  // namespace rsp

using namespace rsp;

extern "C" int rsp_slice_gen(const double* model, double angle, double scale, int64_t size,
                             const double* bg, int64_t log_flag, double* view,
                             rsp_stream_t stream) {
  if (!model || !bg || !view) return fail(kErrNull, "rsp_slice_gen: null pointer");
  if (size < 2 || size > 65535) return fail(kErrSize, "rsp_slice_gen: bad size");
  const double ac = std::cos(angle), as = std::sin(angle);
  dim3 grid((unsigned)ceil_div(size, 64), (unsigned)size);
  hipLaunchKernelGGL(k_slice_gen, grid, dim3(64), 0, (hipStream_t)stream, model, ac, as,
                     scale, (int)size, bg, (int)(log_flag != 0), view);
  RSP_LAUNCH_CHECK("k_slice_gen");
  return kOk;
}

extern "C" int rsp_slice_gen_holo(const double* model_c, double shiftx, double shifty,
                                  double diameter, double rel_scale, double scale,
                                  int64_t size, const double* bg, int64_t log_flag,
                                  double* view, rsp_stream_t stream) {
  if (!model_c || !bg || !view) return fail(kErrNull, "rsp_slice_gen_holo: null pointer");
  if (size < 1 || size > 65535) return fail(kErrSize, "rsp_slice_gen_holo: bad size");
  dim3 grid((unsigned)ceil_div(size, 64), (unsigned)size);
  hipLaunchKernelGGL(k_slice_gen_holo, grid, dim3(64), 0, (hipStream_t)stream, model_c,
                     shiftx, shifty, diameter, rel_scale, scale, (int)size, bg,
                     (int)(log_flag != 0), view);
  RSP_LAUNCH_CHECK("k_slice_gen_holo");
  return kOk;
}

extern "C" int rsp_calc_prob_all(const double* lview, const int32_t* mask, int64_t ndata,
                                 const int32_t* ones, const int32_t* multi,
                                 const int64_t* o_acc, const int64_t* m_acc,
                                 const int32_t* p_o, const int32_t* p_m,
                                 const int32_t* c_m, double init, const double* scales,
                                 double* prob_r, rsp_stream_t stream) {
  if (ndata < 0) return fail(kErrSize, "rsp_calc_prob_all: ndata < 0");
  if (ndata == 0) return kOk;
  if (!lview || !mask || !ones || !multi || !o_acc || !m_acc || !scales || !prob_r)
    return fail(kErrNull, "rsp_calc_prob_all: null pointer");
  hipLaunchKernelGGL(k_calc_prob_all, dim3((unsigned)ceil_div(ndata, 64)), dim3(64), 0,
                     (hipStream_t)stream, lview, mask, ndata, ones, multi, o_acc, m_acc,
                     p_o, p_m, c_m, init, scales, prob_r);
  RSP_LAUNCH_CHECK("k_calc_prob_all");
  return kOk;
}

extern "C" int rsp_get_f_dt(const double* fobj_t, const double* fref_d, int64_t ndata,
                            int64_t npix, double* f_dt, rsp_stream_t stream) {
  if (ndata < 0 || npix < 0) return fail(kErrSize, "rsp_get_f_dt: negative size");
  if (ndata == 0 || npix == 0) return kOk;
  if (!fobj_t || !fref_d || !f_dt) return fail(kErrNull, "rsp_get_f_dt: null pointer");
  const int64_t n = ndata * npix;
  hipLaunchKernelGGL(k_get_f_dt, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, fobj_t, fref_d, ndata, npix, f_dt);
  RSP_LAUNCH_CHECK("k_get_f_dt");
  return kOk;
}
