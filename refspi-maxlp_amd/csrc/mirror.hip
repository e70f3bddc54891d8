// Single-kernel mirrors of the reference's CuPy RawKernels (kernels.cu), FP64.
//
// Same arguments and results as the reference kernels; the launch shapes are gfx950's:
// the fastest-varying thread index runs along a row (y) so that loads and stores of
// size x size images are coalesced (kernels.cu maps thread x to the row index x).
#include <cstring>

#include <rocprim/rocprim.hpp>

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

// kernels.cu:84-114 slice_merge in gather form: one thread per model pixel m adds the
// terms view[t]*wx*wy (the reference's operand order) of the <= 3 x 3 view pixels t whose
// bilinear sample touches m, then model[m] += their sum, mweights[m] += the weights' sum.
// No atomics; the terms of one pixel are added in view-pixel order.
__global__ void k_slice_merge(const double* __restrict__ view, double ac, double as, int size,
                              double* __restrict__ model, double* __restrict__ mweights) {
  const int my = blockIdx.x * blockDim.x + threadIdx.x;
  const int mx = blockIdx.y;
  if (mx >= size || my >= size) return;
  double acc = 0., wts = 0.;
  for_each_source(size, mx, my, ac, as, (const int32_t*)nullptr,
                  [&](int t, double wx, double wy) {
                    acc += view[t] * wx * wy;
                    wts += wx * wy;
                  });
  const int64_t m = (int64_t)mx * size + my;
  model[m] += acc;
  mweights[m] += wts;
}

// merge_all (kernels.cu:139-152) without atomics, step 1: one thread per frame writes each
// photon's key (pixel << 32 | g) and term (prob_r[d], or prob_r[d]*count for multi), with g
// the photon's frame-major position (ones before multi within a frame: the order in which
// the reference's thread d adds them).  Photons outside [0, npix) are counted in *bad.
__global__ void k_merge_keys(const double* __restrict__ prob_r, int64_t ndata,
                             const int32_t* __restrict__ ones, const int32_t* __restrict__ multi,
                             const int64_t* __restrict__ o_acc, const int64_t* __restrict__ m_acc,
                             const int32_t* __restrict__ p_o, const int32_t* __restrict__ p_m,
                             const int32_t* __restrict__ c_m, int64_t npix,
                             uint64_t* __restrict__ keys, double* __restrict__ vals,
                             int32_t* __restrict__ bad) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndata) return;
  int64_t g = o_acc[d] + m_acc[d];
  int nb = 0;
  for (int64_t t = o_acc[d]; t < o_acc[d] + ones[d]; ++t, ++g) {
    const int32_t p = p_o[t];
    const bool ok = p >= 0 && p < npix;
    nb += !ok;
    keys[g] = ((uint64_t)(ok ? p : 0) << 32) | (uint64_t)g;
    vals[g] = ok ? prob_r[d] : 0.;
  }
  for (int64_t t = m_acc[d]; t < m_acc[d] + multi[d]; ++t, ++g) {
    const int32_t p = p_m[t];
    const bool ok = p >= 0 && p < npix;
    nb += !ok;
    keys[g] = ((uint64_t)(ok ? p : 0) << 32) | (uint64_t)g;
    vals[g] = ok ? prob_r[d] * c_m[t] : 0.;
  }
  if (nb) atomicAdd(bad, nb);
}

// Step 2 (keys sorted): the first entry of each pixel's run adds the run's terms to view[p]
// one after another, in frame-major order -- the sequential order of the reference's adds.
__global__ void k_merge_runs(const uint64_t* __restrict__ keys, const double* __restrict__ vals,
                             int64_t n, double* __restrict__ view) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = (uint32_t)(keys[i] >> 32);
  if (i > 0 && (uint32_t)(keys[i - 1] >> 32) == p) return;
  double acc = view[p];
  for (int64_t j = i; j < n && (uint32_t)(keys[j] >> 32) == p; ++j) acc += vals[j];
  view[p] = acc;
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

extern "C" int rsp_slice_merge(const double* view, double angle, int64_t size, double* model,
                               double* mweights, rsp_stream_t stream) {
  if (!view || !model || !mweights) return fail(kErrNull, "rsp_slice_merge: null pointer");
  if (size < 2 || size > 65535) return fail(kErrSize, "rsp_slice_merge: bad size");
  const double ac = std::cos(angle), as = std::sin(angle);
  dim3 grid((unsigned)ceil_div(size, 64), (unsigned)size);
  hipLaunchKernelGGL(k_slice_merge, grid, dim3(64), 0, (hipStream_t)stream, view, ac, as,
                     (int)size, model, mweights);
  RSP_LAUNCH_CHECK("k_slice_merge");
  return kOk;
}

extern "C" int rsp_merge_all(const double* prob_r, int64_t ndata, const int32_t* ones,
                             const int32_t* multi, const int64_t* o_acc, const int64_t* m_acc,
                             const int32_t* p_o, const int32_t* p_m, const int32_t* c_m,
                             double* view, int64_t npix, rsp_stream_t stream_) {
  if (ndata < 0 || npix < 1 || npix > 0x7fffffff)
    return fail(kErrSize, "rsp_merge_all: ndata < 0 or npix out of range");
  if (ndata == 0) return kOk;
  if (!prob_r || !ones || !multi || !o_acc || !m_acc || !view)
    return fail(kErrNull, "rsp_merge_all: null pointer");
  hipStream_t st = (hipStream_t)stream_;
  // photon totals from the last frame's offsets
  int64_t oa = 0, ma = 0;
  int32_t no = 0, nm = 0;
  RSP_HIP(hipMemcpyAsync(&oa, o_acc + ndata - 1, sizeof(oa), hipMemcpyDeviceToHost, st));
  RSP_HIP(hipMemcpyAsync(&ma, m_acc + ndata - 1, sizeof(ma), hipMemcpyDeviceToHost, st));
  RSP_HIP(hipMemcpyAsync(&no, ones + ndata - 1, sizeof(no), hipMemcpyDeviceToHost, st));
  RSP_HIP(hipMemcpyAsync(&nm, multi + ndata - 1, sizeof(nm), hipMemcpyDeviceToHost, st));
  RSP_HIP(hipStreamSynchronize(st));
  const int64_t n = oa + no + ma + nm;
  if (n <= 0) return kOk;
  if (n > 0xffffffffLL) return fail(kErrSize, "rsp_merge_all: more than 2^32 photons");
  if (!p_o || !p_m || !c_m) return fail(kErrNull, "rsp_merge_all: null photon list");
  size_t tb = 0;
  RSP_HIP(rocprim::radix_sort_pairs(nullptr, tb, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                    (double*)nullptr, (double*)nullptr, (size_t)n, 0, 64, st));
  const size_t kb = sizeof(uint64_t) * (size_t)n, vb = sizeof(double) * (size_t)n;
  const size_t off_k2 = kb, off_v = 2 * kb, off_v2 = 2 * kb + vb, off_bad = 2 * kb + 2 * vb,
               off_tmp = off_bad + 256;
  char* buf = nullptr;
  RSP_HIP(hipMallocAsync((void**)&buf, off_tmp + tb, st));
  uint64_t *k1 = (uint64_t*)buf, *k2 = (uint64_t*)(buf + off_k2);
  double *v1 = (double*)(buf + off_v), *v2 = (double*)(buf + off_v2);
  int32_t* bad = (int32_t*)(buf + off_bad);
  int rc = kOk;
  int32_t nbad = 0;
  hipError_t e = hipMemsetAsync(bad, 0, sizeof(int32_t), st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_merge_keys, dim3((unsigned)ceil_div(ndata, 64)), dim3(64), 0, st, prob_r,
                       ndata, ones, multi, o_acc, m_acc, p_o, p_m, c_m, npix, k1, v1, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = rocprim::radix_sort_pairs(buf + off_tmp, tb, k1, k2, v1, v2, (size_t)n, 0, 64, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess && nbad == 0) {
    hipLaunchKernelGGL(k_merge_runs, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, k2, v2, n,
                       view);
    e = hipGetLastError();
  }
  if (e != hipSuccess) rc = hip_fail(e, "rsp_merge_all");
  else if (nbad) rc = fail(kErrSize, "rsp_merge_all: photon pixel outside [0, npix)");
  (void)hipFreeAsync(buf, st);
  return rc;
}
