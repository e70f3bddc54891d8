// Shared helpers for the RefSPI-MaxLP gfx950 library: status handling, launch checks,
// small device utilities.  All arithmetic that feeds an index or a comparison with the
// FP64 oracle is written without fused multiply-add (the library is compiled with
// -ffp-contract=off) so that the device reproduces the oracle's rounding.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace rsp {

// Status codes of the C-ABI (0 = OK, <0 = argument / state error, >0 = hipError_t).
enum Status : int {
  kOk = 0,
  kErrNull = -1,
  kErrSize = -2,
  kErrState = -3,
  kErrAlloc = -4,
  kErrUnsupported = -5,
};

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define RSP_HIP(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return ::rsp::hip_fail(e_, #call); \
  } while (0)

#define RSP_LAUNCH_CHECK(name)                          \
  do {                                                  \
    hipError_t e_ = hipGetLastError();                  \
    if (e_ != hipSuccess) return ::rsp::hip_fail(e_, name); \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------------

// Rotated source coordinate of detector pixel (x, y): kernels.cu:19-23.
//   tx = (x-c)*ac - (y-c)*as + c ; ty = (x-c)*as + (y-c)*ac + c
__device__ __forceinline__ void rot_source(int x, int y, int cen, double ac, double as,
                                           double& tx, double& ty) {
  const double xc = (double)(x - cen), yc = (double)(y - cen);
  tx = xc * ac - yc * as + (double)cen;
  ty = xc * as + yc * ac + (double)cen;
}

// Bilinear sample of an FP64 image at (tx, ty) with kernels.cu:23-33 bounds rule.
// Returns false when the sample is out of bounds.
__device__ __forceinline__ bool bilinear(const double* __restrict__ img, int size, double tx,
                                         double ty, double& out) {
  const double fxl = floor(tx), fyl = floor(ty);
  const int ix = (int)fxl, iy = (int)fyl;
  if (ix < 0 || ix > size - 2 || iy < 0 || iy > size - 2) return false;
  const double fx = tx - ix, fy = ty - iy;
  const double cx = 1. - fx, cy = 1. - fy;
  const double* p = img + (int64_t)ix * size + iy;
  out = p[0] * cx * cy + p[size] * fx * cy + p[1] * cx * fy + p[size + 1] * fx * fy;
  return true;
}

// log-view value of kernels.cu:34-41 (log_flag=1): <1e-20 -> -1000, else log.
__device__ __forceinline__ double log_view(double v) { return v < 1.e-20 ? -1000. : log(v); }

// Sphere form factor of kernels.cu:66-69 / maxLP.py:240-241 (no guard here).
__device__ __forceinline__ double sphere_ff(double s) {
  return (sin(s) - s * cos(s)) / (s * s * s);
}

// kernels.cu:44-82 (rel_scale multiplies before the division by s^3, as there).
__device__ __forceinline__ double holo_pixel(const double* __restrict__ model_c, int x, int y,
                                             int size, double shiftx, double shifty,
                                             double diameter, double rel_scale) {
  const double cen = floor(size / 2.);
  const double pi = M_PI;
  const double xc = x - cen, yc = y - cen;
  const double phase = 2. * pi * (xc * shiftx + yc * shifty) / size;
  const double ramp_r = cos(phase), ramp_i = sin(phase);
  double s = sqrt(xc * xc + yc * yc) * pi * diameter / size;
  if (s == 0.) s = 1.e-5;
  const double sph = rel_scale * (sin(s) - s * cos(s)) / (s * s * s);
  const int64_t t = (int64_t)x * size + y;
  const double cr = ramp_r * sph + model_c[2 * t];
  const double ci = ramp_i * sph + model_c[2 * t + 1];
  const double a = hypot(cr, ci);
  return a * a;
}

// Visit the view pixels j whose bilinear sample (kernels.cu:19-33) under the rotation
// (ac, as) touches model pixel (mx, my), with the bilinear weight's two factors wx * wy:
// f(j, wx, wy) -- the gather
// form of slice_merge's scatter (kernels.cu:84-114).  compact_of maps a pixel to its index
// j (< 0: skipped); null visits every pixel with j = its flat index.
template <typename F>
__device__ __forceinline__ void for_each_source(int size, int mx, int my, double ac, double as,
                                                const int32_t* __restrict__ compact_of, F f) {
  const int cen = size / 2;
  // inverse rotation of m: the view pixels that can reach m lie within +-1 of it
  const double mxc = (double)(mx - cen), myc = (double)(my - cen);
  const int x0 = (int)rint(mxc * ac + myc * as + cen);
  const int y0 = (int)rint(-mxc * as + myc * ac + cen);
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy) {
      const int x = x0 + dx, y = y0 + dy;
      if (x < 0 || x >= size || y < 0 || y >= size) continue;
      const int j = compact_of ? compact_of[x * size + y] : x * size + y;
      if (j < 0) continue;
      double tx, ty;
      rot_source(x, y, cen, ac, as, tx, ty);
      const int ix = (int)floor(tx), iy = (int)floor(ty);
      if (ix < 0 || ix > size - 2 || iy < 0 || iy > size - 2) continue;
      const int ox = mx - ix, oy = my - iy;
      if (ox < 0 || ox > 1 || oy < 0 || oy > 1) continue;
      const double fx = tx - ix, fy = ty - iy;
      const double cx = 1. - fx, cy = 1. - fy;
      f(j, ox ? fx : cx, oy ? fy : cy);
    }
}

}  // namespace rsp
