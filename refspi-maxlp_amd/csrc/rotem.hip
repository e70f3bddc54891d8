// Rotation-form E-step and merge on one GPU.
//
// Model: per-frame reference parameters (shift sx_d, sy_d, diameter dia_d) and in-plane
// rotation r = j*pi/R as the latent variable.  Over the scored pixels (probmask == 0,
// emc.py:89-98), compact index j:
//   O_r(j)  = bilinear rotation of the complex model (kernels.cu:6-42 geometry, re and
//             im separately, out of bounds -> 0)
//   R_d(j)  = rel * S(pi*|p-c|*dia_d/size) * exp(2*pi*i*((x-c)*sx_d + (y-c)*sy_d)/size)
//             (the sphere term of kernels.cu:44-82, s == 0 -> 1e-5)
//   W_dr(j) = scale_d * |O_r(j) + R_d(j)|^2
//   LL_dr   = sum_{photons} K log W_dr (log clamp of kernels.cu:37-40)
//             - scale_d * (|O_r|^2 + |R_d|^2 + 2 X[d, r]),   X = Re sum_j O_r conj(R_d)
// X is the cross-term GEMM (xterm.hip) in FP32 (rocBLAS), split BF16 (MFMA) or FP64
// (rocBLAS).  P(r|d) and log Z_d come from an FP64 log-sum-exp over r.  The merge is
// merge_all + slice_merge (kernels.cu:139-152, 84-114) for every rotation, normalised
// by the slice_merge weights times sum_d P(r|d) scale_d.
//
// Kernels
//   k_rot_slices   O_r(j) for a 64-pixel x 32-rotation tile: GEMM operand rows (A), the
//                  pixel-major copy OT[j][r] through an LDS transpose, per-tile |O|^2.
//   k_ref_rows     one work-group per frame: S over the distinct radii in LDS, separable
//                  phase ramps, GEMM operand rows (B) and |R_d|^2.
//   k_rot_ll       one wave per frame, lanes over rotations: sparse log term over the
//                  frame's photons (OT rows are coalesced 64-rotation reads), the dense
//                  term, and the FP64 log-sum-exp.
//   k_merge_view   one work-group per pixel (pixel-major photon lists): VT[j][r] =
//                  sum_d K P(r|d), no atomics.
//   k_merge_model  one thread per model pixel: slice_merge in gather form (the 3 x 3
//                  view pixels around the inverse-rotated position), no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <type_traits>
#include <vector>

#include "refspi_amd.h"
#include "rsp_common.hpp"
#include "xterm.hpp"

using namespace rsp;

namespace {

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
constexpr double kPi = 3.14159265358979323846;

enum Mode { kFP32 = 0, kSplit = 1, kFP64 = 2 };

// O_r(j) of the bilinear rotation, weights formed first as the oracle does.
__device__ __forceinline__ void rot_sample(const double* __restrict__ m, int size, int t,
                                           double ac, double as, double& vr, double& vi) {
  const int cen = size / 2;
  const int x = t / size, y = t - (t / size) * size;
  double tx, ty;
  rot_source(x, y, cen, ac, as, tx, ty);
  const double fxl = floor(tx), fyl = floor(ty);
  const int ix = (int)fxl, iy = (int)fyl;
  vr = 0.;
  vi = 0.;
  if (ix < 0 || ix > size - 2 || iy < 0 || iy > size - 2) return;
  const double fx = tx - ix, fy = ty - iy;
  const double cx = 1. - fx, cy = 1. - fy;
  const double w0 = cx * cy, w1 = fx * cy, w2 = cx * fy, w3 = fx * fy;
  const int64_t a = (int64_t)ix * size + iy;
  const double* p0 = m + 2 * a;
  const double* p1 = m + 2 * (a + size);
  vr = p0[0] * w0 + p1[0] * w1 + p0[2] * w2 + p1[2] * w3;
  vi = p0[1] * w0 + p1[1] * w1 + p0[3] * w2 + p1[3] * w3;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_rot_slices(
    const double* __restrict__ model_c, int size, int K, int R, const int32_t* __restrict__ scored,
    const double* __restrict__ rcos, const double* __restrict__ rsin, void* __restrict__ A,
    int64_t Mpad, int64_t Kp, void* __restrict__ OT, int Rs, double* __restrict__ norm_part) {
  __shared__ double2 tile[32][65];
  const int tid = threadIdx.x;
  const int jl = tid & 63, rg = tid >> 6;
  const int j = blockIdx.x * 64 + jl;
  const int r0 = blockIdx.y * 32;
  const int t = j < K ? scored[j] : 0;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int rl = rg + 4 * u, r = r0 + rl;
    double vr = 0., vi = 0.;
    if (j < K && r < R) rot_sample(model_c, size, t, rcos[r], rsin[r], vr, vi);
    tile[rl][jl] = make_double2(vr, vi);
    if (j < K && r < R) {
      const int64_t k = (int64_t)r * Kp + 2 * j;
      if constexpr (MODE == kFP32) {
        *reinterpret_cast<float2*>((float*)A + k) = make_float2((float)vr, (float)vi);
      } else if constexpr (MODE == kFP64) {
        *reinterpret_cast<double2*>((double*)A + k) = make_double2(vr, vi);
      } else {
        __bf16* ah = (__bf16*)A;
        __bf16* al = ah + Mpad * Kp;
        const float fr = (float)vr, fi = (float)vi;
        const __bf16 hr = (__bf16)fr, hi = (__bf16)fi;
        bf16x2 h = {hr, hi};
        bf16x2 l = {(__bf16)(fr - (float)hr), (__bf16)(fi - (float)hi)};
        *reinterpret_cast<bf16x2*>(ah + k) = h;
        *reinterpret_cast<bf16x2*>(al + k) = l;
      }
    }
  }
  __syncthreads();
  // per-rotation partial |O|^2 over this pixel tile (FP64, fixed order)
  if (tid < 32) {
    double s = 0.;
    for (int q = 0; q < 64; ++q) {
      const double2 v = tile[tid][q];
      s += v.x * v.x + v.y * v.y;
    }
    if (r0 + tid < R) norm_part[(int64_t)blockIdx.x * R + r0 + tid] = s;
  }
  // pixel-major copy: 32 consecutive rotations per pixel row
  const int rl = tid & 31;
  for (int q = tid >> 5; q < 64; q += 8) {
    const int jj = blockIdx.x * 64 + q, r = r0 + rl;
    if (jj >= K || r >= Rs) continue;
    const double2 v = tile[rl][q];
    if constexpr (MODE == kFP64)
      ((double2*)OT)[(int64_t)jj * Rs + r] = v;
    else
      ((float2*)OT)[(int64_t)jj * Rs + r] = make_float2((float)v.x, (float)v.y);
  }
}

__global__ void k_norm_reduce(const double* __restrict__ part, int ntiles, int R,
                              double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double s = 0.;
  for (int q = 0; q < ntiles; ++q) s += part[(int64_t)q * R + r];
  out[r] = s;
}

__device__ __forceinline__ double sphere_guarded(double s) {
  if (s == 0.) s = 1.e-5;
  return (sin(s) - s * cos(s)) / (s * s * s);
}

// R_d(j) rows of frames [d0, d0 + gridDim.x): B[(d - d0) * Kp + 2j], normR[d - d0].
template <int MODE>
__global__ __launch_bounds__(256) void k_ref_rows(
    int size, int K, const int32_t* __restrict__ scored, const int32_t* __restrict__ jrad,
    const double* __restrict__ rad, int nrad, double rel, const double* __restrict__ sx,
    const double* __restrict__ sy, const double* __restrict__ dia, int64_t d0, int64_t N,
    void* __restrict__ B, int64_t Bpad, int64_t Kp, double* __restrict__ normR) {
  extern __shared__ double sh[];
  double* S = sh;                                     // nrad
  double2* ex = reinterpret_cast<double2*>(sh + ((nrad + 1) & ~1));   // size, 16 B aligned
  double2* ey = ex + size;                            // size
  __shared__ double red[256];
  const int64_t dl = blockIdx.x, d = d0 + dl;
  if (d >= N) return;
  const double fsx = sx[d], fsy = sy[d], fd = dia[d];
  const int cen = size / 2;
  for (int i = threadIdx.x; i < nrad; i += 256)
    S[i] = rel * sphere_guarded(rad[i] * kPi * fd / size);
  for (int i = threadIdx.x; i < size; i += 256) {
    double s, c;
    sincos(2. * kPi * ((double)(i - cen) * fsx) / size, &s, &c);
    ex[i] = make_double2(c, s);
    sincos(2. * kPi * ((double)(i - cen) * fsy) / size, &s, &c);
    ey[i] = make_double2(c, s);
  }
  __syncthreads();
  double acc = 0.;
  for (int j = threadIdx.x; j < K; j += 256) {
    const int t = scored[j];
    const int x = t / size, y = t - (t / size) * size;
    const double2 a = ex[x], b = ey[y];
    const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
    const double sv = S[jrad[j]];
    const double vr = sv * pr, vi = sv * pi;
    acc += vr * vr + vi * vi;
    const int64_t k = dl * Kp + 2 * j;
    if constexpr (MODE == kFP32) {
      *reinterpret_cast<float2*>((float*)B + k) = make_float2((float)vr, (float)vi);
    } else if constexpr (MODE == kFP64) {
      *reinterpret_cast<double2*>((double*)B + k) = make_double2(vr, vi);
    } else {
      __bf16* bh = (__bf16*)B;
      __bf16* bl = bh + Bpad * Kp;
      const float fr = (float)vr, fi = (float)vi;
      const __bf16 hr = (__bf16)fr, hi = (__bf16)fi;
      bf16x2 h = {hr, hi};
      bf16x2 l = {(__bf16)(fr - (float)hr), (__bf16)(fi - (float)hi)};
      *reinterpret_cast<bf16x2*>(bh + k) = h;
      *reinterpret_cast<bf16x2*>(bl + k) = l;
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) normR[dl] = red[0];
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Reference term of one photon (FP64, the oracle's operation order).
__device__ __forceinline__ void ref_photon(int t, int size, double rel, double fsx, double fsy,
                                          double fd, double& rr, double& ri) {
  const int cen = size / 2;
  const double xc = (double)(t / size - cen), yc = (double)(t - (t / size) * size - cen);
  const double phase = 2. * kPi * (xc * fsx + yc * fsy) / size;
  const double s = sqrt(xc * xc + yc * yc) * kPi * fd / size;
  const double sv = rel * sphere_guarded(s);
  double sn, cs;
  sincos(phase, &sn, &cs);
  rr = sv * cs;
  ri = sv * sn;
}

template <typename T>
__device__ __forceinline__ T log_floor(T w) {
  if constexpr (sizeof(T) == 4)
    return w >= 1.e-20f ? logf(w) : -1000.f;
  else
    return w >= 1.e-20 ? log(w) : -1000.;
}

// One wave per frame of [d0, d1): LL, P, log Z.  NT = ceil(Rs / 64) lanes-iterations.
template <typename T, int NT>
__global__ __launch_bounds__(256) void k_rot_ll(
    int size, int R, int Rs, double rel, const int64_t* __restrict__ fo,
    const int32_t* __restrict__ fj, const float* __restrict__ fk,
    const int32_t* __restrict__ scored, const void* __restrict__ OT_,
    const double* __restrict__ normO, const double* __restrict__ normR,
    const void* __restrict__ X_, int64_t ldx, const double* __restrict__ sx,
    const double* __restrict__ sy, const double* __restrict__ dia,
    const double* __restrict__ scales, int64_t d0, int64_t d1, double* __restrict__ ll,
    double* __restrict__ prob, double* __restrict__ logz) {
  typedef typename std::conditional<sizeof(T) == 4, float2, double2>::type T2;
  __shared__ int32_t s_j[4][64];
  __shared__ T s_k[4][64], s_rr[4][64], s_ri[4][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t d = d0 + (int64_t)blockIdx.x * 4 + w;
  if (d >= d1) return;
  const T2* OT = (const T2*)OT_;
  const double scale = scales[d];
  const T sc = (T)scale;
  const double fsx = sx[d], fsy = sy[d], fd = dia[d];
  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0.;
  const int64_t o0 = fo[d], o1 = fo[d + 1];
  for (int64_t base = o0; base < o1; base += 64) {
    const int64_t i = base + lane;
    if (i < o1) {
      const int j = fj[i];
      double rr, ri;
      ref_photon(scored[j], size, rel, fsx, fsy, fd, rr, ri);
      s_j[w][lane] = j;
      s_k[w][lane] = (T)fk[i];
      s_rr[w][lane] = (T)rr;
      s_ri[w][lane] = (T)ri;
    }
    __builtin_amdgcn_wave_barrier();
    const int n = (int)min((int64_t)64, o1 - base);
    for (int q = 0; q < n; ++q) {
      const T2* row = OT + (int64_t)s_j[w][q] * Rs + lane;
      const T kq = s_k[w][q], rr = s_rr[w][q], ri = s_ri[w][q];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const T2 o = row[64 * t];
        const T zr = o.x + rr, zi = o.y + ri;
        const T wv = sc * (zr * zr + zi * zi);
        acc[t] += (double)(kq * log_floor<T>(wv));
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  const double nr = normR[d - d0];
  const T* X = (const T*)X_;
  double m = -INFINITY;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int r = lane + 64 * t;
    if (r < R) {
      acc[t] -= scale * (normO[r] + nr + 2. * (double)X[(d - d0) * ldx + r]);
      m = fmax(m, acc[t]);
    }
  }
  m = wave_max(m);
  double s = 0.;
#pragma unroll
  for (int t = 0; t < NT; ++t)
    if (lane + 64 * t < R) s += exp(acc[t] - m);
  s = wave_sum(s);
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int r = lane + 64 * t;
    if (r >= R) continue;
    if (ll) ll[d * R + r] = acc[t];
    if (prob) prob[d * R + r] = exp(acc[t] - m) / s;
  }
  if (lane == 0 && logz) logz[d] = m + log(s);
}

// wsum partials: part[b][r] = sum_{d in block b} P[d, r] * scale_d
__global__ void k_wsum_part(const double* __restrict__ P, const double* __restrict__ scales,
                            int64_t N, int R, int64_t per, double* __restrict__ part) {
  const int r = blockIdx.y * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const int64_t a = (int64_t)blockIdx.x * per, b = min(N, a + per);
  double s = 0.;
  for (int64_t d = a; d < b; ++d) s += P[d * R + r] * scales[d];
  part[(int64_t)blockIdx.x * R + r] = s;
}

// VT[j][r] = sum over the photons of pixel j of K * P[d, r] (pixel-major lists).
__global__ __launch_bounds__(256) void k_merge_view(
    const int64_t* __restrict__ po, const int32_t* __restrict__ pd,
    const float* __restrict__ pk, const double* __restrict__ P, int R, int Rs,
    double* __restrict__ VT) {
  const int j = blockIdx.x;
  const int64_t e0 = po[j], e1 = po[j + 1];
  for (int r0 = 0; r0 < Rs; r0 += 256 * 4) {
    double acc[4] = {0., 0., 0., 0.};
    for (int64_t e = e0; e < e1; ++e) {
      const int64_t d = pd[e];
      const double k = (double)pk[e];
      const double* row = P + d * R;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int r = r0 + threadIdx.x + 256 * u;
        if (r < R) acc[u] += k * row[r];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = r0 + threadIdx.x + 256 * u;
      if (r < Rs) VT[(int64_t)j * Rs + r] = r < R ? acc[u] : 0.;
    }
  }
}

// slice_merge of every rotation in gather form; one thread per model pixel m.
__global__ void k_merge_model(int size, int R, int Rs, const double* __restrict__ rcos,
                              const double* __restrict__ rsin,
                              const int32_t* __restrict__ compact_of,
                              const double* __restrict__ VT, const double* __restrict__ wsum,
                              double* __restrict__ model, double* __restrict__ acc_out,
                              double* __restrict__ wts_out) {
  const int npix = size * size;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npix) return;
  const int mx = m / size, my = m - (m / size) * size, cen = size / 2;
  double acc = 0., wts = 0.;
  for (int r = 0; r < R; ++r) {
    const double ac = rcos[r], as = rsin[r];
    // inverse rotation of m: the view pixels that can reach m lie within +-1 of it
    const double mxc = (double)(mx - cen), myc = (double)(my - cen);
    const int x0 = (int)rint(mxc * ac + myc * as + cen);
    const int y0 = (int)rint(-mxc * as + myc * ac + cen);
    double wr = 0.;
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy) {
        const int x = x0 + dx, y = y0 + dy;
        if (x < 0 || x >= size || y < 0 || y >= size) continue;
        const int j = compact_of[x * size + y];
        if (j < 0) continue;
        double tx, ty;
        rot_source(x, y, cen, ac, as, tx, ty);
        const int ix = (int)floor(tx), iy = (int)floor(ty);
        if (ix < 0 || ix > size - 2 || iy < 0 || iy > size - 2) continue;
        const int ox = mx - ix, oy = my - iy;
        if (ox < 0 || ox > 1 || oy < 0 || oy > 1) continue;
        const double fx = tx - ix, fy = ty - iy;
        const double cx = 1. - fx, cy = 1. - fy;
        const double w = (ox ? fx : cx) * (oy ? fy : cy);
        acc += VT[(int64_t)j * Rs + r] * w;
        wr += w;
      }
    wts += wr * wsum[r];
  }
  if (acc_out) acc_out[m] = acc;
  if (wts_out) wts_out[m] = wts;
  model[m] = wts > 0. ? acc / wts : 0.;
}

// ---------------------------------------------------------------------------------
// Adjoint of L = -sum_d log((1/R) sum_r exp LL_dr) (recomputes O_r and R_d; FP64).
//   dL/dW_drj = -P_dr (K_dj / W_drj - 1)        (-P_dr * -1 where W < 1e-20: log clamp)
//   dW/dtheta = 2 scale_d Re(conj(Z) dZ/dtheta),   Z = O_r(j) + R_d(j)
// Model:  g_O_r(j) = -2 sum_{d in photons(j)} P_dr K Z / |Z|^2 + 2 ws_r O_r(j)
//                    + 2 sum_d P_dr scale_d R_d(j)          (ws_r = sum_d P_dr scale_d)
//         g_model  = sum_r Interp_r^T g_O_r               (slice_merge's transpose form)
// Frame:  g_theta_d = -2 sum_r P_dr sum_photons K Re(conj(Z) dR) / |Z|^2
//                    + 2 scale_d sum_j Re(conj(Obar_d(j) + R_d(j)) dR_d(j)),
//         Obar_d = sum_r P_dr O_r                          (sum_r P_dr = 1)
// ---------------------------------------------------------------------------------

// Q[dl * R + r] = P[(d0 + dl) * R + r] * scale[d0 + dl]
__global__ void k_scaled_prob(const double* __restrict__ P, const double* __restrict__ scales,
                              int64_t d0, int64_t n, int R, double* __restrict__ Q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * R) return;
  const int64_t dl = i / R;
  Q[i] = P[d0 * R + i] * scales[d0 + dl];
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Per-frame tables of the reference term: S(s_i), dS/ddia at every distinct radius, and the
// separable phase factors exp(2 pi i (x-c) sx / size), exp(2 pi i (y-c) sy / size).
struct RefTables {
  double* S;
  double* Sd;
  double2* ex;
  double2* ey;
};

__device__ __forceinline__ RefTables ref_tables(double* sh, int nrad, int size) {
  RefTables t;
  t.S = sh;
  t.Sd = sh + nrad;
  t.ex = reinterpret_cast<double2*>(sh + ((2 * nrad + 1) & ~1));
  t.ey = t.ex + size;
  return t;
}

__device__ void fill_ref_tables(RefTables tb, const double* __restrict__ rad, int nrad, int size,
                                double rel, double fsx, double fsy, double fd) {
  const int cen = size / 2;
  for (int i = threadIdx.x; i < nrad; i += blockDim.x) {
    double s = rad[i] * kPi * fd / size;
    if (s == 0.) s = 1.e-5;
    double sn, cs;
    sincos(s, &sn, &cs);
    const double sv = (sn - s * cs) / (s * s * s);
    tb.S[i] = rel * sv;
    // dS/ds = sin(s)/s^2 - 3 S(s)/s ; ds/ddia = pi rad / size
    tb.Sd[i] = rel * (sn / (s * s) - 3. * sv / s) * (kPi * rad[i] / size);
  }
  for (int i = threadIdx.x; i < size; i += blockDim.x) {
    double sn, cs;
    sincos(2. * kPi * ((double)(i - cen) * fsx) / size, &sn, &cs);
    tb.ex[i] = make_double2(cs, sn);
    sincos(2. * kPi * ((double)(i - cen) * fsy) / size, &sn, &cs);
    tb.ey[i] = make_double2(cs, sn);
  }
}

// R_d(j) and its derivatives by sx, sy, dia at scored pixel t (radius index jr).
__device__ __forceinline__ void ref_and_derivs(const RefTables& tb, int t, int jr, int size,
                                               double2& Rv, double2& Dx, double2& Dy,
                                               double2& Dd) {
  const int cen = size / 2;
  const int x = t / size, y = t - (t / size) * size;
  const double2 e = cmul(tb.ex[x], tb.ey[y]);
  const double sv = tb.S[jr], sd = tb.Sd[jr];
  Rv = make_double2(sv * e.x, sv * e.y);
  Dd = make_double2(sd * e.x, sd * e.y);
  const double kx = 2. * kPi * (double)(x - cen) / size;
  const double ky = 2. * kPi * (double)(y - cen) / size;
  Dx = make_double2(-kx * Rv.y, kx * Rv.x);    // i kx R
  Dy = make_double2(-ky * Rv.y, ky * Rv.x);    // i ky R
}

__device__ __forceinline__ double block_sum256(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// One work-group per frame of [d0, d0 + gridDim.x): frame gradients (sx, sy, dia).
// Dense part: threads over the scored pixels (Obar row of the chunk).  Sparse part: wave w
// takes photons w, w+4, ...; lanes over rotations r = lane + 64 t.
template <int NT>
__global__ __launch_bounds__(256) void k_adj_frame(
    int size, int K, int R, int Rs, double rel, const int32_t* __restrict__ scored,
    const int32_t* __restrict__ jrad, const double* __restrict__ rad, int nrad,
    const int64_t* __restrict__ fo, const int32_t* __restrict__ fj,
    const float* __restrict__ fk, const double2* __restrict__ OT, const double* __restrict__ P,
    const double* __restrict__ Obar, int64_t Kp, const double* __restrict__ sx,
    const double* __restrict__ sy, const double* __restrict__ dia,
    const double* __restrict__ scales, int64_t d0, double* __restrict__ gsx,
    double* __restrict__ gsy, double* __restrict__ gdia) {
  extern __shared__ double sh[];
  __shared__ double red[4];
  const int64_t dl = blockIdx.x, d = d0 + dl;
  const double sc = scales[d];
  const RefTables tb = ref_tables(sh, nrad, size);
  fill_ref_tables(tb, rad, nrad, size, rel, sx[d], sy[d], dia[d]);
  __syncthreads();
  // dense: sum_j Re(conj(Obar + R) dR)
  double ex_ = 0., ey_ = 0., ed_ = 0.;
  const double2* ob = reinterpret_cast<const double2*>(Obar + dl * Kp);
  for (int j = threadIdx.x; j < K; j += 256) {
    double2 Rv, Dx, Dy, Dd;
    ref_and_derivs(tb, scored[j], jrad[j], size, Rv, Dx, Dy, Dd);
    const double2 o = ob[j];
    const double zr = o.x + Rv.x, zi = o.y + Rv.y;
    ex_ += zr * Dx.x + zi * Dx.y;
    ey_ += zr * Dy.x + zi * Dy.y;
    ed_ += zr * Dd.x + zi * Dd.y;
  }
  // sparse: per rotation sum_photons K Re(conj(Z) dR) / |Z|^2
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double ax[NT], ay[NT], ad[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) ax[t] = ay[t] = ad[t] = 0.;
  const int64_t o1 = fo[d + 1];
  for (int64_t q = fo[d] + w; q < o1; q += 4) {
    const int j = fj[q];
    const double k = (double)fk[q];
    double2 Rv, Dx, Dy, Dd;
    ref_and_derivs(tb, scored[j], jrad[j], size, Rv, Dx, Dy, Dd);
    const double2* row = OT + (int64_t)j * Rs + lane;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const double2 o = row[64 * t];
      const double zr = o.x + Rv.x, zi = o.y + Rv.y;
      const double n2 = zr * zr + zi * zi;
      if (sc * n2 >= 1.e-20) {
        const double f = k / n2;
        ax[t] += f * (zr * Dx.x + zi * Dx.y);
        ay[t] += f * (zr * Dy.x + zi * Dy.y);
        ad[t] += f * (zr * Dd.x + zi * Dd.y);
      }
    }
  }
  double px = 0., py = 0., pd = 0.;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int r = lane + 64 * t;
    if (r < R) {
      const double pr = P[d * R + r];
      px += pr * ax[t];
      py += pr * ay[t];
      pd += pr * ad[t];
    }
  }
  const double gx = -2. * block_sum256(px, red) + 2. * sc * block_sum256(ex_, red);
  const double gy = -2. * block_sum256(py, red) + 2. * sc * block_sum256(ey_, red);
  const double gd = -2. * block_sum256(pd, red) + 2. * sc * block_sum256(ed_, red);
  if (threadIdx.x == 0) {
    gsx[d] = gx;
    gsy[d] = gy;
    gdia[d] = gd;
  }
}

// One work-group per scored pixel j: g_O_r(j) for every rotation (threads over r), from
// the pixel-major photon lists, O_r(j) (OT), the dense GEMM term GT (column-major R x Kp)
// and ws_r.  GO[j * Rs + r] (complex); rotations r >= R get 0.
template <int NU>
__global__ __launch_bounds__(256) void k_adj_pixel(
    int size, int R, int Rs, double rel, const int32_t* __restrict__ scored,
    const int64_t* __restrict__ po, const int32_t* __restrict__ pd,
    const float* __restrict__ pk, const double2* __restrict__ OT, const double* __restrict__ P,
    const double* __restrict__ GT, const double* __restrict__ wsum,
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ dia, const double* __restrict__ scales,
    double2* __restrict__ GO) {
  __shared__ double s_rr[256], s_ri[256], s_k[256], s_sc[256];
  __shared__ int32_t s_d[256];
  const int j = blockIdx.x, tid = threadIdx.x;
  const int t = scored[j];
  double2 O[NU], acc[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int r = tid + 256 * u;
    O[u] = r < Rs ? OT[(int64_t)j * Rs + r] : make_double2(0., 0.);
    acc[u] = make_double2(0., 0.);
  }
  const int64_t e0 = po[j], e1 = po[j + 1];
  for (int64_t base = e0; base < e1; base += 256) {
    const int64_t e = base + tid;
    if (e < e1) {
      const int d = pd[e];
      double rr, ri;
      ref_photon(t, size, rel, sx[d], sy[d], dia[d], rr, ri);
      s_rr[tid] = rr;
      s_ri[tid] = ri;
      s_k[tid] = (double)pk[e];
      s_sc[tid] = scales[d];
      s_d[tid] = d;
    }
    __syncthreads();
    const int n = (int)min((int64_t)256, e1 - base);
    for (int q = 0; q < n; ++q) {
      const double rr = s_rr[q], ri = s_ri[q], k = s_k[q], sc = s_sc[q];
      const double* prow = P + (int64_t)s_d[q] * R;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int r = tid + 256 * u;
        if (r < R) {
          const double zr = O[u].x + rr, zi = O[u].y + ri;
          const double n2 = zr * zr + zi * zi;
          if (sc * n2 >= 1.e-20) {
            const double f = k * prow[r] / n2;
            acc[u].x += f * zr;
            acc[u].y += f * zi;
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int r = tid + 256 * u;
    if (r >= Rs) continue;
    double2 g = make_double2(0., 0.);
    if (r < R) {
      const double ws = wsum[r];
      g.x = -2. * acc[u].x + 2. * ws * O[u].x + 2. * GT[(int64_t)(2 * j) * R + r];
      g.y = -2. * acc[u].y + 2. * ws * O[u].y + 2. * GT[(int64_t)(2 * j + 1) * R + r];
    }
    GO[(int64_t)j * Rs + r] = g;
  }
}

// g_model(m) = sum_r sum_{scored view pixels j whose bilinear sample touches m} w * GO[j][r]
// (the transpose of k_rot_slices' interpolation; gather form, one thread per model pixel).
__global__ void k_adj_model(int size, int R, int Rs, const double* __restrict__ rcos,
                            const double* __restrict__ rsin,
                            const int32_t* __restrict__ compact_of,
                            const double2* __restrict__ GO, double* __restrict__ g_model) {
  const int npix = size * size;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npix) return;
  const int mx = m / size, my = m - (m / size) * size, cen = size / 2;
  double gr = 0., gi = 0.;
  for (int r = 0; r < R; ++r) {
    const double ac = rcos[r], as = rsin[r];
    const double mxc = (double)(mx - cen), myc = (double)(my - cen);
    const int x0 = (int)rint(mxc * ac + myc * as + cen);
    const int y0 = (int)rint(-mxc * as + myc * ac + cen);
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy) {
        const int x = x0 + dx, y = y0 + dy;
        if (x < 0 || x >= size || y < 0 || y >= size) continue;
        const int j = compact_of[x * size + y];
        if (j < 0) continue;
        double tx, ty;
        rot_source(x, y, cen, ac, as, tx, ty);
        const int ix = (int)floor(tx), iy = (int)floor(ty);
        if (ix < 0 || ix > size - 2 || iy < 0 || iy > size - 2) continue;
        const int ox = mx - ix, oy = my - iy;
        if (ox < 0 || ox > 1 || oy < 0 || oy > 1) continue;
        const double fx = tx - ix, fy = ty - iy;
        const double cx = 1. - fx, cy = 1. - fy;
        const double w = (ox ? fx : cx) * (oy ? fy : cy);
        const double2 g = GO[(int64_t)j * Rs + r];
        gr += g.x * w;
        gi += g.y * w;
      }
  }
  g_model[2 * m] = gr;
  g_model[2 * m + 1] = gi;
}

template <typename T>
int dalloc(T** p, int64_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  return kOk;
}

template <typename T>
int dupload(T** p, const std::vector<T>& v) {
  int rc = dalloc(p, (int64_t)v.size());
  if (rc) return rc;
  if (!v.empty()) RSP_HIP(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return kOk;
}

}  // namespace

struct rsp_rot {
  int device = 0, size = 0, npix = 0, R = 0, Rs = 0, K = 0, nrad = 0;
  double rel = 1.;
  int64_t Kp = 0, Mpad = 0;
  std::vector<int32_t> scored, compact_of;
  double *d_rcos = nullptr, *d_rsin = nullptr, *d_rad = nullptr;
  int32_t *d_scored = nullptr, *d_jrad = nullptr, *d_compact_of = nullptr;
  // frames
  int64_t N = 0, nnz = 0;
  int64_t *d_fo = nullptr, *d_po = nullptr;
  int32_t *d_fj = nullptr, *d_pd = nullptr;
  float *d_fk = nullptr, *d_pk = nullptr;
  // work
  int mode = kSplit, mode_alloc = -1;
  int64_t chunk = 0, chunk_alloc = 0;
  void *d_A = nullptr, *d_OT = nullptr, *d_B = nullptr, *d_X = nullptr;
  double *d_normO = nullptr, *d_part = nullptr, *d_normR = nullptr;
  double *d_VT = nullptr, *d_wsum = nullptr, *d_wpart = nullptr;
  // adjoint (FP64 operands of its own, so that the forward mode's buffers stay)
  int64_t adj_chunk = 0;
  double *d_aA = nullptr, *d_aB = nullptr, *d_aObar = nullptr, *d_aQ = nullptr;
  double *d_aGT = nullptr, *d_apart = nullptr, *d_anormO = nullptr, *d_anormR = nullptr;
  double2 *d_aOT = nullptr, *d_aGO = nullptr;
  // timing
  bool timing = false;
  hipEvent_t ev[8] = {};
  double t_ms[5] = {0., 0., 0., 0., 0.};
};

namespace {

void free_work(rsp_rot* p) {
  void* ptrs[] = {p->d_A, p->d_OT, p->d_B, p->d_X, p->d_normO, p->d_part, p->d_normR};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  p->d_A = p->d_OT = p->d_B = p->d_X = nullptr;
  p->d_normO = p->d_part = p->d_normR = nullptr;
  p->mode_alloc = -1;
  p->chunk_alloc = 0;
}

void free_adj(rsp_rot* p) {
  void* ptrs[] = {p->d_aA, p->d_aB, p->d_aObar, p->d_aQ, p->d_aGT, p->d_apart,
                  p->d_anormO, p->d_anormR, p->d_aOT, p->d_aGO};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  p->d_aA = p->d_aB = p->d_aObar = p->d_aQ = p->d_aGT = p->d_apart = nullptr;
  p->d_anormO = p->d_anormR = nullptr;
  p->d_aOT = p->d_aGO = nullptr;
  p->adj_chunk = 0;
}

int ensure_adj(rsp_rot* p, int64_t chunk) {
  if (p->adj_chunk == chunk) return kOk;
  free_adj(p);
  const int64_t ntiles = ceil_div(p->K, 64);
  int rc;
  if ((rc = dalloc(&p->d_aA, p->Mpad * p->Kp)) || (rc = dalloc(&p->d_aB, chunk * p->Kp)) ||
      (rc = dalloc(&p->d_aObar, chunk * p->Kp)) || (rc = dalloc(&p->d_aQ, chunk * p->R)) ||
      (rc = dalloc(&p->d_aGT, (int64_t)p->R * p->Kp)) ||
      (rc = dalloc(&p->d_apart, ntiles * p->R)) || (rc = dalloc(&p->d_anormO, p->R)) ||
      (rc = dalloc(&p->d_anormR, chunk)) ||
      (rc = dalloc(&p->d_aOT, (int64_t)p->K * p->Rs)) ||
      (rc = dalloc(&p->d_aGO, (int64_t)p->K * p->Rs))) {
    free_adj(p);
    return rc;
  }
  // operand padding (rows >= R of A, columns >= 2K of A and B) stays zero
  RSP_HIP(hipMemset(p->d_aA, 0, sizeof(double) * (size_t)(p->Mpad * p->Kp)));
  RSP_HIP(hipMemset(p->d_aB, 0, sizeof(double) * (size_t)(chunk * p->Kp)));
  RSP_HIP(hipMemset(p->d_aOT, 0, sizeof(double2) * (size_t)((int64_t)p->K * p->Rs)));
  p->adj_chunk = chunk;
  return kOk;
}

template <int NT>
void launch_adj_frame_nt(rsp_rot* p, size_t shmem, const double* prob, const double* sx,
                         const double* sy, const double* dia, const double* scales, int64_t d0,
                         int64_t n, double* gsx, double* gsy, double* gdia, hipStream_t st) {
  hipLaunchKernelGGL((k_adj_frame<NT>), dim3((unsigned)n), dim3(256), shmem, st, p->size, p->K,
                     p->R, p->Rs, p->rel, p->d_scored, p->d_jrad, p->d_rad, p->nrad, p->d_fo,
                     p->d_fj, p->d_fk, p->d_aOT, prob, p->d_aObar, p->Kp, sx, sy, dia, scales,
                     d0, gsx, gsy, gdia);
}

int launch_adj_frame(rsp_rot* p, size_t shmem, const double* prob, const double* sx,
                     const double* sy, const double* dia, const double* scales, int64_t d0,
                     int64_t n, double* gsx, double* gsy, double* gdia, hipStream_t st) {
  switch (p->Rs / 64) {
#define RSP_NT(k)                                                                        \
  case k:                                                                                \
    launch_adj_frame_nt<k>(p, shmem, prob, sx, sy, dia, scales, d0, n, gsx, gsy, gdia, st); \
    break;
    RSP_NT(1) RSP_NT(2) RSP_NT(3) RSP_NT(4) RSP_NT(5) RSP_NT(6) RSP_NT(7) RSP_NT(8)
    RSP_NT(9) RSP_NT(10) RSP_NT(11) RSP_NT(12)
#undef RSP_NT
    default: return fail(kErrUnsupported, "rotation form supports at most 768 rotations");
  }
  RSP_LAUNCH_CHECK("k_adj_frame");
  return kOk;
}

int launch_adj_pixel(rsp_rot* p, const double* prob, const double* sx, const double* sy,
                     const double* dia, const double* scales, hipStream_t st) {
  const dim3 g((unsigned)p->K), b(256);
  switch (ceil_div(p->Rs, 256)) {
    case 1:
      hipLaunchKernelGGL(k_adj_pixel<1>, g, b, 0, st, p->size, p->R, p->Rs, p->rel, p->d_scored,
                         p->d_po, p->d_pd, p->d_pk, p->d_aOT, prob, p->d_aGT, p->d_wsum, sx, sy,
                         dia, scales, p->d_aGO);
      break;
    case 2:
      hipLaunchKernelGGL(k_adj_pixel<2>, g, b, 0, st, p->size, p->R, p->Rs, p->rel, p->d_scored,
                         p->d_po, p->d_pd, p->d_pk, p->d_aOT, prob, p->d_aGT, p->d_wsum, sx, sy,
                         dia, scales, p->d_aGO);
      break;
    case 3:
      hipLaunchKernelGGL(k_adj_pixel<3>, g, b, 0, st, p->size, p->R, p->Rs, p->rel, p->d_scored,
                         p->d_po, p->d_pd, p->d_pk, p->d_aOT, prob, p->d_aGT, p->d_wsum, sx, sy,
                         dia, scales, p->d_aGO);
      break;
    default: return fail(kErrUnsupported, "rotation form supports at most 768 rotations");
  }
  RSP_LAUNCH_CHECK("k_adj_pixel");
  return kOk;
}

// ws_r = sum_d P[d, r] scale_d into p->d_wsum (fixed summation order).
int launch_wsum(rsp_rot* p, const double* prob, const double* scales, hipStream_t st) {
  const int64_t per = 1024;
  const int64_t nb = ceil_div(p->N, per);
  int rc;
  if (!p->d_wsum && (rc = dalloc(&p->d_wsum, p->R))) return rc;
  if ((rc = dalloc(&p->d_wpart, nb * p->R))) return rc;
  hipLaunchKernelGGL(k_wsum_part, dim3((unsigned)nb, (unsigned)ceil_div(p->R, 256)), dim3(256),
                     0, st, prob, scales, p->N, p->R, per, p->d_wpart);
  RSP_LAUNCH_CHECK("k_wsum_part");
  hipLaunchKernelGGL(k_norm_reduce, dim3((unsigned)ceil_div(p->R, 256)), dim3(256), 0, st,
                     p->d_wpart, (int)nb, p->R, p->d_wsum);
  RSP_LAUNCH_CHECK("k_norm_reduce");
  return kOk;
}

int64_t elem_bytes(int mode) { return mode == kFP64 ? 8 : (mode == kSplit ? 4 : 4); }

int ensure_work(rsp_rot* p, int64_t chunk) {
  if (p->mode_alloc == p->mode && p->chunk_alloc == chunk) return kOk;
  free_work(p);
  const int64_t eb = elem_bytes(p->mode);
  const int64_t Bpad = xterm_pad_n(chunk);
  const int64_t ntiles = ceil_div(p->K, 64);
  int rc;
  char** A = reinterpret_cast<char**>(&p->d_A);
  char** B = reinterpret_cast<char**>(&p->d_B);
  char** OT = reinterpret_cast<char**>(&p->d_OT);
  char** X = reinterpret_cast<char**>(&p->d_X);
  if ((rc = dalloc(A, p->Mpad * p->Kp * eb)) || (rc = dalloc(B, Bpad * p->Kp * eb)) ||
      (rc = dalloc(OT, (int64_t)p->K * p->Rs * (p->mode == kFP64 ? 16 : 8))) ||
      (rc = dalloc(X, chunk * p->Mpad * (p->mode == kFP64 ? 8 : 4))) ||
      (rc = dalloc(&p->d_normO, p->R)) || (rc = dalloc(&p->d_part, ntiles * p->R)) ||
      (rc = dalloc(&p->d_normR, chunk))) {
    free_work(p);
    return rc;
  }
  // padding rows / columns of the operands stay zero
  RSP_HIP(hipMemset(p->d_A, 0, (size_t)(p->Mpad * p->Kp * eb)));
  RSP_HIP(hipMemset(p->d_B, 0, (size_t)(Bpad * p->Kp * eb)));
  RSP_HIP(hipMemset(p->d_OT, 0, (size_t)((int64_t)p->K * p->Rs * (p->mode == kFP64 ? 16 : 8))));
  p->mode_alloc = p->mode;
  p->chunk_alloc = chunk;
  return kOk;
}

// O_r rows (A), the pixel-major copy (OT) and |O_r|^2 (normO) into the given buffers.
template <int MODE>
int launch_slices_to(rsp_rot* p, const double* model_c, void* A, void* OT, double* part,
                     double* normO, hipStream_t st) {
  dim3 g((unsigned)ceil_div(p->K, 64), (unsigned)ceil_div(p->R, 32));
  hipLaunchKernelGGL(k_rot_slices<MODE>, g, dim3(256), 0, st, model_c, p->size, p->K, p->R,
                     p->d_scored, p->d_rcos, p->d_rsin, A, p->Mpad, p->Kp, OT, p->Rs, part);
  RSP_LAUNCH_CHECK("k_rot_slices");
  hipLaunchKernelGGL(k_norm_reduce, dim3((unsigned)ceil_div(p->R, 256)), dim3(256), 0, st,
                     part, (int)ceil_div(p->K, 64), p->R, normO);
  RSP_LAUNCH_CHECK("k_norm_reduce");
  return kOk;
}

template <int MODE>
int launch_slices(rsp_rot* p, const double* model_c, hipStream_t st) {
  return launch_slices_to<MODE>(p, model_c, p->d_A, p->d_OT, p->d_part, p->d_normO, st);
}

// R_d rows of frames [d0, d0 + n) into B (Bpad rows per bf16 plane) and normR.
template <int MODE>
int launch_ref_to(rsp_rot* p, const double* sx, const double* sy, const double* dia, int64_t d0,
                  int64_t n, void* B, int64_t Bpad, double* normR, hipStream_t st) {
  const size_t shmem =
      sizeof(double) * (size_t)((p->nrad + 1) & ~1) + 2 * sizeof(double2) * (size_t)p->size;
  hipLaunchKernelGGL(k_ref_rows<MODE>, dim3((unsigned)n), dim3(256), shmem, st, p->size, p->K,
                     p->d_scored, p->d_jrad, p->d_rad, p->nrad, p->rel, sx, sy, dia, d0, p->N, B,
                     Bpad, p->Kp, normR);
  RSP_LAUNCH_CHECK("k_ref_rows");
  return kOk;
}

template <int MODE>
int launch_ref(rsp_rot* p, const double* sx, const double* sy, const double* dia, int64_t d0,
               int64_t n, hipStream_t st) {
  return launch_ref_to<MODE>(p, sx, sy, dia, d0, n, p->d_B, xterm_pad_n(p->chunk_alloc),
                             p->d_normR, st);
}

template <typename T, int NT>
void launch_ll_nt(rsp_rot* p, const double* sx, const double* sy, const double* dia,
                  const double* scales, int64_t d0, int64_t d1, double* ll, double* prob,
                  double* logz, hipStream_t st) {
  hipLaunchKernelGGL((k_rot_ll<T, NT>), dim3((unsigned)ceil_div(d1 - d0, 4)), dim3(256), 0, st,
                     p->size, p->R, p->Rs, p->rel, p->d_fo, p->d_fj, p->d_fk, p->d_scored,
                     p->d_OT, p->d_normO, p->d_normR, p->d_X, p->Mpad, sx, sy, dia, scales, d0,
                     d1, ll, prob, logz);
}

template <typename T>
int launch_ll(rsp_rot* p, const double* sx, const double* sy, const double* dia,
              const double* scales, int64_t d0, int64_t d1, double* ll, double* prob,
              double* logz, hipStream_t st) {
  const int nt = p->Rs / 64;
  switch (nt) {
#define RSP_NT(n) \
  case n: launch_ll_nt<T, n>(p, sx, sy, dia, scales, d0, d1, ll, prob, logz, st); break;
    RSP_NT(1) RSP_NT(2) RSP_NT(3) RSP_NT(4) RSP_NT(5) RSP_NT(6) RSP_NT(7) RSP_NT(8)
    RSP_NT(9) RSP_NT(10) RSP_NT(11) RSP_NT(12)
#undef RSP_NT
    default: return fail(kErrUnsupported, "rotation form supports at most 768 rotations");
  }
  RSP_LAUNCH_CHECK("k_rot_ll");
  return kOk;
}

}  // namespace

extern "C" {

int rsp_rot_create(int device, int64_t size, int64_t num_rot, const double* rot_cos,
                   const double* rot_sin, double rel_scale, rsp_rot** out) {
  if (!out || !rot_cos || !rot_sin) return fail(kErrNull, "rsp_rot_create: null argument");
  *out = nullptr;
  if (size < 9 || size > 4096) return fail(kErrSize, "rsp_rot_create: size out of range");
  if (num_rot < 1 || num_rot > 768) return fail(kErrSize, "rsp_rot_create: 1 <= num_rot <= 768");
  RSP_HIP(hipSetDevice(device));
  rsp_rot* p = new rsp_rot();
  p->device = device;
  p->size = (int)size;
  p->npix = (int)(size * size);
  p->R = (int)num_rot;
  p->Rs = (int)ceil_div(num_rot, 64) * 64;
  p->rel = rel_scale;
  // scored pixels: probmask == 0 of emc.py:89-98
  const int c = (int)size / 2;
  p->compact_of.assign(p->npix, -1);
  std::map<int64_t, int> radix;
  std::vector<int64_t> r2_of;
  for (int x = 0; x < size; ++x)
    for (int y = 0; y < size; ++y) {
      const double xc = x - c, yc = y - c;
      const double rad = std::sqrt(xc * xc + yc * yc);
      int pm = 0;
      if (rad >= (double)c) pm = 2;
      if (rad < (double)(size / 8)) pm = 1;
      if (rad < 4.) pm = 2;
      if (pm != 0) continue;
      const int t = x * (int)size + y;
      p->compact_of[t] = (int)p->scored.size();
      p->scored.push_back(t);
      const int64_t r2 = (int64_t)(x - c) * (x - c) + (int64_t)(y - c) * (y - c);
      r2_of.push_back(r2);
      radix.emplace(r2, 0);
    }
  p->K = (int)p->scored.size();
  std::vector<double> rad;
  for (auto& kv : radix) {
    kv.second = (int)rad.size();
    rad.push_back(std::sqrt((double)kv.first));
  }
  std::vector<int32_t> jrad(p->K);
  for (int j = 0; j < p->K; ++j) jrad[j] = radix[r2_of[j]];
  p->nrad = (int)rad.size();
  p->Kp = xterm_pad_k(2 * (int64_t)p->K);
  p->Mpad = xterm_pad_m(p->R);
  std::vector<double> rc(rot_cos, rot_cos + num_rot), rs(rot_sin, rot_sin + num_rot);
  int rc_ = kOk;
  if ((rc_ = dupload(&p->d_rcos, rc)) || (rc_ = dupload(&p->d_rsin, rs)) ||
      (rc_ = dupload(&p->d_rad, rad)) || (rc_ = dupload(&p->d_scored, p->scored)) ||
      (rc_ = dupload(&p->d_jrad, jrad)) || (rc_ = dupload(&p->d_compact_of, p->compact_of))) {
    rsp_rot_destroy(p);
    return rc_;
  }
  if (sizeof(double) * (size_t)(p->nrad + 1) + 2 * sizeof(double2) * (size_t)size > 60000) {
    rsp_rot_destroy(p);
    return fail(kErrSize, "rsp_rot_create: radius table exceeds LDS");
  }
  *out = p;
  return kOk;
}

int rsp_rot_set_frames(rsp_rot* p, int64_t num_data, const int32_t* ones, const int32_t* multi,
                       const int32_t* place_ones, const int32_t* place_multi,
                       const int32_t* count_multi) {
  if (!p || !ones || !multi) return fail(kErrNull, "rsp_rot_set_frames: null argument");
  if (num_data < 1) return fail(kErrSize, "rsp_rot_set_frames: no frames");
  RSP_HIP(hipSetDevice(p->device));
  const int64_t N = num_data;
  std::vector<int64_t> fo(N + 1, 0);
  int64_t no = 0, nm = 0;
  for (int64_t d = 0; d < N; ++d) {
    if (ones[d] < 0 || multi[d] < 0) return fail(kErrSize, "negative photon count");
    no += ones[d];
    nm += multi[d];
  }
  if ((no && !place_ones) || (nm && (!place_multi || !count_multi)))
    return fail(kErrNull, "rsp_rot_set_frames: null photon list");
  std::vector<int32_t> fj;
  std::vector<float> fk;
  fj.reserve((size_t)(no + nm));
  fk.reserve((size_t)(no + nm));
  int64_t io = 0, im = 0;
  for (int64_t d = 0; d < N; ++d) {
    for (int q = 0; q < ones[d]; ++q, ++io) {
      const int t = place_ones[io];
      if (t < 0 || t >= p->npix) return fail(kErrSize, "photon index out of range");
      const int j = p->compact_of[t];
      if (j >= 0) { fj.push_back(j); fk.push_back(1.f); }
    }
    for (int q = 0; q < multi[d]; ++q, ++im) {
      const int t = place_multi[im];
      if (t < 0 || t >= p->npix) return fail(kErrSize, "photon index out of range");
      const int j = p->compact_of[t];
      if (j >= 0) { fj.push_back(j); fk.push_back((float)count_multi[im]); }
    }
    fo[d + 1] = (int64_t)fj.size();
  }
  const int64_t nnz = (int64_t)fj.size();
  // pixel-major lists (stable in frame order)
  std::vector<int64_t> po(p->K + 1, 0);
  for (int64_t e = 0; e < nnz; ++e) ++po[fj[e] + 1];
  for (int j = 0; j < p->K; ++j) po[j + 1] += po[j];
  std::vector<int64_t> fill(po.begin(), po.end() - 1);
  std::vector<int32_t> pd(nnz);
  std::vector<float> pk(nnz);
  for (int64_t d = 0; d < N; ++d)
    for (int64_t e = fo[d]; e < fo[d + 1]; ++e) {
      const int64_t q = fill[fj[e]]++;
      pd[q] = (int32_t)d;
      pk[q] = fk[e];
    }
  int rc;
  if ((rc = dupload(&p->d_fo, fo)) || (rc = dupload(&p->d_fj, fj)) ||
      (rc = dupload(&p->d_fk, fk)) || (rc = dupload(&p->d_po, po)) ||
      (rc = dupload(&p->d_pd, pd)) || (rc = dupload(&p->d_pk, pk)))
    return rc;
  p->N = N;
  p->nnz = nnz;
  return kOk;
}

int rsp_rot_set_option(rsp_rot* p, int key, int64_t value) {
  if (!p) return fail(kErrNull, "rsp_rot_set_option: null plan");
  switch (key) {
    case 1:
      if (value < 0 || value > 2) return fail(kErrSize, "mode must be 0, 1 or 2");
      p->mode = (int)value;
      return kOk;
    case 2:
      if (value < 0) return fail(kErrSize, "chunk must be >= 0");
      p->chunk = value;
      return kOk;
    case 5:
      p->timing = value != 0;
      return kOk;
    default:
      return fail(kErrUnsupported, "unknown option key");
  }
}

int rsp_rot_estep(rsp_rot* p, const double* model_c, const double* sx, const double* sy,
                  const double* dia, const double* scales, double* ll, double* prob,
                  double* logz, rsp_stream_t stream_) {
  if (!p || !model_c || !sx || !sy || !dia || !scales)
    return fail(kErrNull, "rsp_rot_estep: null argument");
  if (p->N == 0) return fail(kErrState, "rsp_rot_estep: no frames");
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  const int64_t eb = elem_bytes(p->mode);
  int64_t chunk = p->chunk;
  if (chunk <= 0) {
    // operand rows of a chunk within about 8 GB
    chunk = std::max<int64_t>(128, (int64_t)(8e9 / (double)(p->Kp * eb)) / 128 * 128);
  }
  chunk = std::min(chunk, xterm_pad_n(p->N));
  int rc = ensure_work(p, chunk);
  if (rc) return rc;
  if (p->timing) {
    for (auto& e : p->ev)
      if (!e) RSP_HIP(hipEventCreate(&e));
    for (double& t : p->t_ms) t = 0.;
    RSP_HIP(hipEventRecord(p->ev[0], st));
  }
  if (p->mode == kFP32) rc = launch_slices<kFP32>(p, model_c, st);
  else if (p->mode == kSplit) rc = launch_slices<kSplit>(p, model_c, st);
  else rc = launch_slices<kFP64>(p, model_c, st);
  if (rc) return rc;
  if (p->timing) {
    RSP_HIP(hipEventRecord(p->ev[1], st));
    RSP_HIP(hipEventSynchronize(p->ev[1]));
    float ms = 0.f;
    RSP_HIP(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
    p->t_ms[0] += ms;
  }
  const int64_t Bpad = xterm_pad_n(chunk);
  for (int64_t d0 = 0; d0 < p->N; d0 += chunk) {
    const int64_t n = std::min(chunk, p->N - d0);
    if (p->timing) RSP_HIP(hipEventRecord(p->ev[2], st));
    if (p->mode == kFP32) rc = launch_ref<kFP32>(p, sx, sy, dia, d0, n, st);
    else if (p->mode == kSplit) rc = launch_ref<kSplit>(p, sx, sy, dia, d0, n, st);
    else rc = launch_ref<kFP64>(p, sx, sy, dia, d0, n, st);
    if (rc) return rc;
    if (p->timing) RSP_HIP(hipEventRecord(p->ev[3], st));
    if (p->mode == kFP32) {
      rc = xterm_gemm_blas32((const float*)p->d_A, (const float*)p->d_B, (float*)p->d_X, p->Mpad,
                             n, p->Kp, p->Kp, p->Kp, p->Mpad, st);
    } else if (p->mode == kFP64) {
      rc = xterm_gemm_blas64((const double*)p->d_A, (const double*)p->d_B, (double*)p->d_X,
                             p->Mpad, n, p->Kp, p->Kp, p->Kp, p->Mpad, st);
    } else {
      const __bf16* ah = (const __bf16*)p->d_A;
      const __bf16* bh = (const __bf16*)p->d_B;
      rc = xterm_gemm_split(ah, ah + p->Mpad * p->Kp, bh, bh + Bpad * p->Kp, (float*)p->d_X,
                            p->Mpad, n, p->Kp, p->Mpad, st);
    }
    if (rc) return rc;
    if (p->timing) RSP_HIP(hipEventRecord(p->ev[4], st));
    if (p->mode == kFP64)
      rc = launch_ll<double>(p, sx, sy, dia, scales, d0, d0 + n, ll, prob, logz, st);
    else
      rc = launch_ll<float>(p, sx, sy, dia, scales, d0, d0 + n, ll, prob, logz, st);
    if (rc) return rc;
    if (p->timing) {
      RSP_HIP(hipEventRecord(p->ev[5], st));
      RSP_HIP(hipEventSynchronize(p->ev[5]));
      float a = 0.f, b = 0.f, c = 0.f;
      RSP_HIP(hipEventElapsedTime(&a, p->ev[2], p->ev[3]));
      RSP_HIP(hipEventElapsedTime(&b, p->ev[3], p->ev[4]));
      RSP_HIP(hipEventElapsedTime(&c, p->ev[4], p->ev[5]));
      p->t_ms[1] += a;
      p->t_ms[2] += b;
      p->t_ms[3] += c;
      p->t_ms[4] += 1.;
    }
  }
  RSP_HIP(hipStreamSynchronize(st));
  return kOk;
}

int rsp_rot_merge(rsp_rot* p, const double* prob, const double* scales, double* model_out,
                  double* acc_out, double* weights_out, rsp_stream_t stream_) {
  if (!p || !prob || !scales || !model_out) return fail(kErrNull, "rsp_rot_merge: null argument");
  if (p->N == 0) return fail(kErrState, "rsp_rot_merge: no frames");
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  int rc;
  if (!p->d_VT && (rc = dalloc(&p->d_VT, (int64_t)p->K * p->Rs))) return rc;
  if ((rc = launch_wsum(p, prob, scales, st))) return rc;
  hipLaunchKernelGGL(k_merge_view, dim3((unsigned)p->K), dim3(256), 0, st, p->d_po, p->d_pd,
                     p->d_pk, prob, p->R, p->Rs, p->d_VT);
  RSP_LAUNCH_CHECK("k_merge_view");
  hipLaunchKernelGGL(k_merge_model, dim3((unsigned)ceil_div(p->npix, 256)), dim3(256), 0, st,
                     p->size, p->R, p->Rs, p->d_rcos, p->d_rsin, p->d_compact_of, p->d_VT,
                     p->d_wsum, model_out, acc_out, weights_out);
  RSP_LAUNCH_CHECK("k_merge_model");
  RSP_HIP(hipStreamSynchronize(st));
  return kOk;
}

int rsp_rot_adjoint(rsp_rot* p, const double* model_c, const double* sx, const double* sy,
                    const double* dia, const double* scales, const double* prob,
                    double* g_model, double* g_sx, double* g_sy, double* g_dia,
                    rsp_stream_t stream_) {
  if (!p || !model_c || !sx || !sy || !dia || !scales || !prob || !g_model || !g_sx || !g_sy ||
      !g_dia)
    return fail(kErrNull, "rsp_rot_adjoint: null argument");
  if (p->N == 0) return fail(kErrState, "rsp_rot_adjoint: no frames");
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  const size_t shmem =
      sizeof(double) * (size_t)((2 * p->nrad + 1) & ~1) + 2 * sizeof(double2) * (size_t)p->size;
  int lds_max = 0;
  RSP_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device));
  if (shmem + 64 > (size_t)lds_max)
    return fail(kErrSize, "rsp_rot_adjoint: reference tables exceed the LDS of a work-group");
  int64_t chunk = p->chunk > 0 ? p->chunk : 4096;
  chunk = std::min(chunk, p->N);
  int rc = ensure_adj(p, chunk);
  if (rc) return rc;
  if ((rc = launch_slices_to<kFP64>(p, model_c, p->d_aA, p->d_aOT, p->d_apart, p->d_anormO, st)))
    return rc;
  if ((rc = launch_wsum(p, prob, scales, st))) return rc;
  for (int64_t d0 = 0; d0 < p->N; d0 += chunk) {
    const int64_t n = std::min(chunk, p->N - d0);
    if ((rc = launch_ref_to<kFP64>(p, sx, sy, dia, d0, n, p->d_aB, chunk, p->d_anormR, st)))
      return rc;
    hipLaunchKernelGGL(k_scaled_prob, dim3((unsigned)ceil_div(n * p->R, 256)), dim3(256), 0, st,
                       prob, scales, d0, n, p->R, p->d_aQ);
    RSP_LAUNCH_CHECK("k_scaled_prob");
    // GT (R x Kp, column-major) += Q^T B : GT[k * R + r] = sum_d Q[d, r] B[d, k]
    if ((rc = xterm_dgemm(false, true, p->R, p->Kp, n, 1., p->d_aQ, p->R, p->d_aB, p->Kp,
                          d0 == 0 ? 0. : 1., p->d_aGT, p->R, st)))
      return rc;
    // Obar (n x Kp, row-major) = P A : Obar[d * Kp + k] = sum_r P[d, r] A[r, k]
    if ((rc = xterm_dgemm(false, false, p->Kp, n, p->R, 1., p->d_aA, p->Kp, prob + d0 * p->R,
                          p->R, 0., p->d_aObar, p->Kp, st)))
      return rc;
    if ((rc = launch_adj_frame(p, shmem, prob, sx, sy, dia, scales, d0, n, g_sx, g_sy, g_dia,
                               st)))
      return rc;
  }
  if ((rc = launch_adj_pixel(p, prob, sx, sy, dia, scales, st))) return rc;
  hipLaunchKernelGGL(k_adj_model, dim3((unsigned)ceil_div(p->npix, 256)), dim3(256), 0, st,
                     p->size, p->R, p->Rs, p->d_rcos, p->d_rsin, p->d_compact_of, p->d_aGO,
                     g_model);
  RSP_LAUNCH_CHECK("k_adj_model");
  RSP_HIP(hipStreamSynchronize(st));
  return kOk;
}

int rsp_rot_timing(const rsp_rot* p, double* out) {
  if (!p || !out) return fail(kErrNull, "rsp_rot_timing: null argument");
  for (int i = 0; i < 5; ++i) out[i] = p->t_ms[i];
  return kOk;
}

int rsp_rot_info(const rsp_rot* p, int64_t* out) {
  if (!p || !out) return fail(kErrNull, "rsp_rot_info: null argument");
  out[0] = p->N;
  out[1] = p->K;
  out[2] = p->Kp;
  out[3] = p->Mpad;
  out[4] = p->chunk_alloc;
  out[5] = p->nnz;
  out[6] = p->nrad;
  return kOk;
}

int rsp_rot_destroy(rsp_rot* p) {
  if (!p) return kOk;
  (void)hipSetDevice(p->device);
  free_work(p);
  free_adj(p);
  void* ptrs[] = {p->d_rcos, p->d_rsin, p->d_rad, p->d_scored, p->d_jrad, p->d_compact_of,
                  p->d_fo, p->d_po, p->d_fj, p->d_pd, p->d_fk, p->d_pk, p->d_VT, p->d_wsum,
                  p->d_wpart};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (hipEvent_t e : p->ev)
    if (e) (void)hipEventDestroy(e);
  delete p;
  return kOk;
}

}  // extern "C"
