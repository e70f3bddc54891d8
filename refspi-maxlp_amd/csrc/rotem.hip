// Rotation-form E-step, merge and adjoint on one GPU.
//
// Model: per-frame reference parameters (shift sx_d, sy_d, diameter dia_d) and a latent
// hypothesis h = c * R + r: object class c < C (one complex model per class) and in-plane
// rotation r = j*pi/R.  Over the scored pixels (probmask == 0, emc.py:89-98), compact
// index j:
//   O_h(j)  = bilinear rotation of class c's complex model by r (kernels.cu:6-42
//             geometry, re and im separately, out of bounds -> 0)
//   R_d(j)  = rel * S(pi*|p-c|*dia_d/size) * exp(2*pi*i*((x-c)*sx_d + (y-c)*sy_d)/size)
//             (the sphere term of kernels.cu:44-82, s == 0 -> 1e-5)
//   W_dh(j) = scale_d * |O_h(j) + R_d(j)|^2
//   LL_dh   = sum_{photons} K log W_dh (log clamp of kernels.cu:37-40)
//             - scale_d * (|O_h|^2 + |R_d|^2 + 2 X[d, h]),   X = Re sum_j O_h conj(R_d)
// X is the cross-term GEMM (xterm.hip) in FP32 (rocBLAS), split BF16 (MFMA) or FP64
// (rocBLAS).  P(h|d) and log Z_d come from an FP64 log-sum-exp over h.  The merge is
// merge_all + slice_merge (kernels.cu:139-152, 84-114) for every hypothesis, normalised
// per class by the slice_merge weights times sum_d P(h|d) scale_d.
//
// Kernels
//   k_rot_slices   O_h(j) for a 64-pixel x 32-hypothesis tile: GEMM operand rows (A), the
//                  pixel-major copy OT[j][h] through an LDS transpose, per-tile |O_h|^2.
//   k_ref_rows     one work-group per frame: S over the distinct radii in LDS, separable
//                  phase ramps, GEMM operand rows (B) and |R_d|^2.
//   k_rot_lt       one wave per (frame, 64-hypothesis tile), tiles outermost: sparse log
//                  term over the frame's photons (OT row segments are coalesced 64-hypothesis
//                  reads of a cache-resident tile) and the dense term -> LL.
//   k_rot_lse      one wave per frame: FP64 log-sum-exp over LL, first-occurrence argmax.
//   k_merge_view   one work-group per pixel (pixel-major photon lists): VT[j][h] +=
//                  sum_d K P(h|d), no atomics.
//   k_merge_model  one thread per (model pixel, class): slice_merge in gather form (the
//                  3 x 3 view pixels around the inverse-rotated position), no atomics.
//   k_adj_*        the adjoint (below).
//
// Frames come either resident (rsp_rot_set_frames) or as chunks of the caller's (packed by
// rsp_rot_pack, e.g. into pinned host memory and copied in on a side stream) through the
// pass API: rsp_rot_pass_begin / rsp_rot_pass_chunk / rsp_rot_pass_end, which accumulates
// the merge and the model gradient over chunks.
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
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef double f64x2v __attribute__((ext_vector_type(2)));
constexpr double kPi = 3.14159265358979323846;
constexpr double kLn2 = 0.69314718055994530942;
// buffer resource word 3 for raw dword loads (32-bit data format)
constexpr int kBufWord3 = 0x00020000;

enum Mode { kFP32 = 0, kSplit = 1, kFP64 = 2 };
constexpr int kGenSizeMax = 513;  // generated reference rows: |x - c| <= 256 exact in bf16
enum PassFlags { kPassMerge = 1, kPassFrameGrad = 2, kPassModelGrad = 4 };

// Lane-iteration counts of the hypothesis kernels; the hypothesis stride is 64 * NT.
constexpr int kNTs[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24};
constexpr int kMaxH = 64 * 24;

int nt_for(int64_t H) {
  for (int nt : kNTs)
    if (64 * (int64_t)nt >= H) return nt;
  return 0;
}

// O_h(j) of the bilinear rotation, weights formed first as the oracle does.
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
    const double* __restrict__ model_c, int size, int K, int R, int H,
    const int32_t* __restrict__ scored, const double* __restrict__ rcos,
    const double* __restrict__ rsin, void* __restrict__ A, int64_t Mpad, int64_t Kp,
    void* __restrict__ OT, int Hs, double* __restrict__ norm_part) {
  __shared__ double2 tile[32][65];
  const int tid = threadIdx.x;
  const int jl = tid & 63, rg = tid >> 6;
  const int j = blockIdx.x * 64 + jl;
  const int h0 = blockIdx.y * 32;
  const int t = j < K ? scored[j] : 0;
  const int64_t npix2 = 2 * (int64_t)size * size;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int hl = rg + 4 * u, h = h0 + hl;
    double vr = 0., vi = 0.;
    if (j < K && h < H) {
      const int c = h / R, r = h - c * R;
      rot_sample(model_c + c * npix2, size, t, rcos[r], rsin[r], vr, vi);
    }
    tile[hl][jl] = make_double2(vr, vi);
    if (j < K && h < H) {
      const int64_t k = (int64_t)h * Kp + 2 * j;
      if constexpr (MODE == kFP32) {
        *reinterpret_cast<float2*>((float*)A + k) = make_float2((float)vr, (float)vi);
      } else if constexpr (MODE == kFP64) {
        *reinterpret_cast<double2*>((double*)A + k) = make_double2(vr, vi);
      } else {
        __bf16* ah = (__bf16*)A;
        __bf16* al = ah + Mpad * Kp;
        const float fr = (float)vr, fi = (float)vi;
        const __bf16 hr = (__bf16)fr, hi = (__bf16)fi;
        bf16x2 hv = {hr, hi};
        bf16x2 lv = {(__bf16)(fr - (float)hr), (__bf16)(fi - (float)hi)};
        *reinterpret_cast<bf16x2*>(ah + k) = hv;
        *reinterpret_cast<bf16x2*>(al + k) = lv;
      }
    }
  }
  __syncthreads();
  // per-hypothesis partial |O|^2 over this pixel tile (FP64, fixed order)
  if (tid < 32) {
    double s = 0.;
    for (int q = 0; q < 64; ++q) {
      const double2 v = tile[tid][q];
      s += v.x * v.x + v.y * v.y;
    }
    if (h0 + tid < H) norm_part[(int64_t)blockIdx.x * H + h0 + tid] = s;
  }
  // pixel-major copy: 32 consecutive hypotheses per pixel row
  const int hl = tid & 31;
  for (int q = tid >> 5; q < 64; q += 8) {
    const int jj = blockIdx.x * 64 + q, h = h0 + hl;
    if (jj >= K || h >= Hs) continue;
    const double2 v = tile[hl][q];
    if constexpr (MODE == kFP64) {
      ((double2*)OT)[(int64_t)jj * Hs + h] = v;
    } else {
      // FP32 hi and the remainders as bf16 (12 bytes): O = hi + lo to about 2^-33 relative
      // (see k_rot_lt)
      const float hr = (float)v.x, hi = (float)v.y;
      const bf16x2 lo = {(__bf16)(float)(v.x - (double)hr), (__bf16)(float)(v.y - (double)hi)};
      ((uint3*)OT)[(int64_t)jj * Hs + h] =
          make_uint3(__float_as_uint(hr), __float_as_uint(hi), __builtin_bit_cast(uint32_t, lo));
    }
  }
}

// out[i] (+)= sum_q part[q * n + i] in a fixed order.
__global__ void k_norm_reduce(const double* __restrict__ part, int ntiles, int n,
                              double* __restrict__ out, int accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.;
  for (int q = 0; q < ntiles; ++q) s += part[(int64_t)q * n + i];
  out[i] = accumulate ? out[i] + s : s;
}

__device__ __forceinline__ double sphere_guarded(double s) {
  if (s == 0.) s = 1.e-5;
  return (sin(s) - s * cos(s)) / (s * s * s);
}

// R_d(j) rows of frames [d0, d0 + gridDim.x): B[(d - d0) * Kp + 2j], normR[d - d0].
template <int MODE>
__global__ __launch_bounds__(256) void k_ref_rows(
    int size, int K, const uint32_t* __restrict__ pxy, const int32_t* __restrict__ jrad,
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
  // four consecutive pixels per thread (one 16 B bf16 store per plane, or 32 B FP32); the
  // row padding j in [K, Kp/2) is written as zeros
  double acc = 0.;
  const int jpad = (int)(Kp / 2);
  for (int j0 = 4 * threadIdx.x; j0 < jpad; j0 += 4 * 256) {
    double vr[4], vi[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = j0 + u;
      vr[u] = 0.;
      vi[u] = 0.;
      if (j < K) {
        const uint32_t q = pxy[j];
        const double2 a = ex[q & 0xffffu], b = ey[q >> 16];
        const double pr = a.x * b.x - a.y * b.y, pi = a.x * b.y + a.y * b.x;
        const double sv = S[jrad[j]];
        vr[u] = sv * pr;
        vi[u] = sv * pi;
        acc += vr[u] * vr[u] + vi[u] * vi[u];
      }
    }
    const int64_t k = dl * Kp + 2 * j0;
    // streaming stores for the FP32 and bf16 rows: they are read once, by the GEMM, after
    // the whole chunk is written (far more than the caches hold).  The FP64 rows keep
    // ordinary stores: their four 16-byte streaming stores per thread measured 4.6x slower
    // (2.8 against 0.6 ms per 4096-frame chunk at 185^2)
    if constexpr (MODE == kFP32) {
      f32x4v* o = reinterpret_cast<f32x4v*>((float*)B + k);
      const f32x4v a = {(float)vr[0], (float)vi[0], (float)vr[1], (float)vi[1]};
      const f32x4v b = {(float)vr[2], (float)vi[2], (float)vr[3], (float)vi[3]};
      __builtin_nontemporal_store(a, o);
      __builtin_nontemporal_store(b, o + 1);
    } else if constexpr (MODE == kFP64) {
      f64x2v* o = reinterpret_cast<f64x2v*>((double*)B + k);
#pragma unroll
      for (int u = 0; u < 4; ++u) o[u] = f64x2v{vr[u], vi[u]};
    } else {
      __bf16* bh = (__bf16*)B;
      __bf16* bl = bh + Bpad * Kp;
      bf16x8 hv, lv;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float fr = (float)vr[u], fi = (float)vi[u];
        const __bf16 hr = (__bf16)fr, hi = (__bf16)fi;
        hv[2 * u] = hr;
        hv[2 * u + 1] = hi;
        lv[2 * u] = (__bf16)(fr - (float)hr);
        lv[2 * u + 1] = (__bf16)(fi - (float)hi);
      }
      __builtin_nontemporal_store(hv, reinterpret_cast<bf16x8*>(bh + k));
      __builtin_nontemporal_store(lv, reinterpret_cast<bf16x8*>(bl + k));
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

// |R_d|^2 of frames [d0, d0 + gridDim.x) over the radius histogram: the sum over the
// distinct radii of cnt * (rel S)^2 (the shift ramp has modulus 1), FP64, fixed reduction
// order.  Used with the reference rows generated in the split GEMM (no rows are written).
__global__ __launch_bounds__(256) void k_ref_norm(const double* __restrict__ rad,
                                                  const int32_t* __restrict__ cnt, int nrad,
                                                  double rel, int size,
                                                  const double* __restrict__ dia, int64_t d0,
                                                  double* __restrict__ normR) {
  __shared__ double red[256];
  const double fd = dia[d0 + blockIdx.x];
  double acc = 0.;
  for (int i = threadIdx.x; i < nrad; i += 256) {
    // sphere_guarded with one sincos
    double s = rad[i] * kPi * fd / size;
    if (s == 0.) s = 1.e-5;
    double sn, cs;
    sincos(s, &sn, &cs);
    const double sv = rel * ((sn - s * cs) / (s * s * s));
    acc += (double)cnt[i] * (sv * sv);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) normR[blockIdx.x] = red[0];
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

// LL of hypothesis tile blockIdx.y (h = 64 * tile + lane) for the frames [d0, d1), one
// wave per frame: the sparse log term over the frame's photons (OT row segments of the tile,
// coalesced) and the dense term -scale (|O_h|^2 + |R_d|^2 + 2 X), X being the sum of nxs
// K-slice partial products xstride elements apart (added in FP64).  Tile-major launch order
// (blockIdx.y outermost) keeps one tile's OT columns (K x 64 entries) cache-resident while
// every frame visits them.  Output llb[(d - d0) * H + h] (FP64).
template <typename T>
__global__ __launch_bounds__(256) void k_rot_lt(
    int size, int H, int Hs, double rel, const int64_t* __restrict__ fo,
    const int32_t* __restrict__ fj, const float* __restrict__ fk,
    const int32_t* __restrict__ scored, const void* __restrict__ OT_,
    const double* __restrict__ normO, const double* __restrict__ normR,
    const void* __restrict__ X_, int64_t ldx, int nxs, int64_t xstride,
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ dia, const double* __restrict__ scales, int64_t d0, int64_t d1,
    double* __restrict__ llb) {
  // OT entries: FP64 (re, im), or FP32 (re_hi, im_hi) with the remainders as bf16 (re_lo in
  // the low half of the third word, im_lo in the high half)
  typedef typename std::conditional<sizeof(T) == 4, uint3, double2>::type T2;
  __shared__ int32_t s_j[4][64];
  __shared__ T s_k[4][64], s_rr[4][64], s_ri[4][64], s_rrl[4][64], s_ril[4][64];
  // the wave index as a known-uniform value: the frame, its photon range and the photon
  // loop's trip count are then scalar
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t d = d0 + (int64_t)blockIdx.x * 4 + w;
  if (d >= d1) return;
  const int h = (int)blockIdx.y * 64 + lane;
  const T2* OT = (const T2*)OT_ + h;
  const double scale = scales[d];
  const T sc = (T)scale;
  const double fsx = sx[d], fsy = sy[d], fd = dia[d];
  double acc = 0., acc2 = 0.;   // acc2: sum of K log2 W (FP32 form)
  float kf = 0.f;               // FP32 form: sum of K over photons below the log floor
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
      s_rrl[w][lane] = (T)(rr - (double)(T)rr);
      s_ril[w][lane] = (T)(ri - (double)(T)ri);
    }
    __builtin_amdgcn_wave_barrier();
    const int n = (int)min((int64_t)64, o1 - base);
    if constexpr (sizeof(T) == 4) {
      // The photon's pixel index comes by a scalar load (uniform address) and its OT row
      // is a buffer resource formed on the scalar unit (the lane's 32-bit offset h * 12 is
      // the only vector part).  The resource covers the H hypotheses only: lanes of the
      // last tile with h >= H (padding, not written) read 0 and fetch nothing.  The log is the hardware base-2 log (inputs >= 1e-20 are normal numbers), its
      // K-weighted sum taken in FP64 and multiplied by ln 2 once; photons below the floor
      // are counted apart (kf) so that each adds exactly -1000 K.
      const char* otb = (const char*)OT_;
      const uint32_t lofs = (uint32_t)h * 12u;
#pragma unroll 8
      for (int q = 0; q < n; ++q) {
        const char* row = otb + (int64_t)fj[base + q] * Hs * 12;
        const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)row, 0, H * 12, kBufWord3);
        const auto ov = __builtin_amdgcn_raw_buffer_load_b96(rs, lofs, 0, 0);
        const uint3 o = make_uint3(ov[0], ov[1], ov[2]);
        const float kq = s_k[w][q], rr = s_rr[w][q], ri = s_ri[w][q];
        // the hi parts cancel exactly when O ~ -R (Sterbenz); the lo parts restore the
        // digits the FP32 roundings of O and R dropped
        const float zr = (__uint_as_float(o.x) + rr) + (__uint_as_float(o.z << 16) + s_rrl[w][q]);
        const float zi =
            (__uint_as_float(o.y) + ri) + (__uint_as_float(o.z & 0xffff0000u) + s_ril[w][q]);
        const float wv = sc * (zr * zr + zi * zi);
        const bool ok = wv >= 1.e-20f;
        acc2 += (double)(ok ? kq * __builtin_amdgcn_logf(wv) : 0.f);
        kf += ok ? 0.f : kq;
      }
    } else {
#pragma unroll 8
      for (int q = 0; q < n; ++q) {
        const T2 o = OT[(int64_t)s_j[w][q] * Hs];
        const T kq = s_k[w][q];
        const T zr = o.x + s_rr[w][q], zi = o.y + s_ri[w][q];
        acc += kq * log_floor<T>(sc * (zr * zr + zi * zi));
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (h >= H) return;
  if constexpr (sizeof(T) == 4) acc = acc2 * kLn2 - 1000. * (double)kf;
  const T* xp = (const T*)X_ + (d - d0) * ldx + h;
  double x = 0.;
  for (int z = 0; z < nxs; ++z) x += (double)xp[z * xstride];
  llb[(d - d0) * H + h] = acc - scale * (normO[h] + normR[d - d0] + 2. * x);
}

// One wave per frame of [d0, d1): the FP64 log-sum-exp over the frame's LL row
// (llb[(d - d0) * H + h]) -> P, log Z and the first-occurrence argmax.
__global__ __launch_bounds__(256) void k_rot_lse(int H, int64_t d0, int64_t d1,
                                                 const double* __restrict__ llb,
                                                 double* __restrict__ prob,
                                                 double* __restrict__ logz,
                                                 int32_t* __restrict__ best) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t d = d0 + (int64_t)blockIdx.x * 4 + w;
  if (d >= d1) return;
  const double* row = llb + (d - d0) * H;
  double m = -INFINITY;
  int hbest = 0x7fffffff;
  for (int h = lane; h < H; h += 64) {
    const double v = row[h];
    if (v > m) {                  // first occurrence within the lane (h increases)
      m = v;
      hbest = h;
    }
  }
  const double lm = m;
  m = wave_max(m);
  // smallest hypothesis index attaining the wave maximum
  int hb = (lm == m) ? hbest : 0x7fffffff;
  for (int o = 32; o > 0; o >>= 1) hb = min(hb, __shfl_xor(hb, o));
  double s = 0.;
  for (int h = lane; h < H; h += 64) s += exp(row[h] - m);
  s = wave_sum(s);
  if (prob)
    for (int h = lane; h < H; h += 64) prob[d * H + h] = exp(row[h] - m) / s;
  if (lane == 0) {
    if (logz) logz[d] = m + log(s);
    if (best) best[d] = hb;
  }
}

// wsum partials: part[b][h] = sum_{d in block b} P[d, h] * scale_d
__global__ void k_wsum_part(const double* __restrict__ P, const double* __restrict__ scales,
                            int64_t N, int H, int64_t per, double* __restrict__ part) {
  const int h = blockIdx.y * blockDim.x + threadIdx.x;
  if (h >= H) return;
  const int64_t a = (int64_t)blockIdx.x * per, b = min(N, a + per);
  double s = 0.;
  for (int64_t d = a; d < b; ++d) s += P[d * H + h] * scales[d];
  part[(int64_t)blockIdx.x * H + h] = s;
}

// VT[j][h] (+)= sum over the photons of pixel j of K * P[d, h] (pixel-major lists).
__global__ __launch_bounds__(256) void k_merge_view(
    const int64_t* __restrict__ po, const int32_t* __restrict__ pd,
    const float* __restrict__ pk, const double* __restrict__ P, int H, int Hs,
    double* __restrict__ VT, int accumulate) {
  const int j = blockIdx.x;
  const int64_t e0 = po[j], e1 = po[j + 1];
  for (int h0 = 0; h0 < Hs; h0 += 256 * 4) {
    double acc[4] = {0., 0., 0., 0.};
    for (int64_t e = e0; e < e1; ++e) {
      const int64_t d = pd[e];
      const double k = (double)pk[e];
      const double* row = P + d * H;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int h = h0 + threadIdx.x + 256 * u;
        if (h < H) acc[u] += k * row[h];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int h = h0 + threadIdx.x + 256 * u;
      if (h >= Hs) continue;
      double* v = VT + (int64_t)j * Hs + h;
      const double a = h < H ? acc[u] : 0.;
      *v = accumulate ? *v + a : a;
    }
  }
}

// slice_merge of every rotation of class c = blockIdx.y in gather form; one thread per
// model pixel m.  Outputs at c * npix + m.
__global__ void k_merge_model(int size, int R, int Hs, const double* __restrict__ rcos,
                              const double* __restrict__ rsin,
                              const int32_t* __restrict__ compact_of,
                              const double* __restrict__ VT, const double* __restrict__ wsum,
                              double* __restrict__ model, double* __restrict__ acc_out,
                              double* __restrict__ wts_out) {
  const int npix = size * size;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npix) return;
  const int c = blockIdx.y;
  const int mx = m / size, my = m - (m / size) * size;
  double acc = 0., wts = 0.;
  for (int r = 0; r < R; ++r) {
    const int h = c * R + r;
    double wr = 0.;
    for_each_source(size, mx, my, rcos[r], rsin[r], compact_of, [&](int j, double wx, double wy) {
      const double w = wx * wy;
      acc += VT[(int64_t)j * Hs + h] * w;
      wr += w;
    });
    wts += wr * wsum[h];
  }
  const int64_t o = (int64_t)c * npix + m;
  if (acc_out) acc_out[o] = acc;
  if (wts_out) wts_out[o] = wts;
  model[o] = wts > 0. ? acc / wts : 0.;
}

// ---------------------------------------------------------------------------------
// Adjoint of L = -sum_d log((1/H) sum_h exp LL_dh) (recomputes O_h and R_d; FP64).
//   dL/dW_dhj = -P_dh (K_dj / W_dhj - 1)        (-P_dh * -1 where W < 1e-20: log clamp)
//   dW/dtheta = 2 scale_d Re(conj(Z) dZ/dtheta),   Z = O_h(j) + R_d(j)
// Model:  g_O_h(j) = -2 sum_{d in photons(j)} P_dh K Z / |Z|^2 + 2 ws_h O_h(j)
//                    + 2 sum_d P_dh scale_d R_d(j)          (ws_h = sum_d P_dh scale_d)
//         g_model_c = sum_r Interp_r^T g_O_{cR+r}          (slice_merge's transpose form)
// Frame:  g_theta_d = -2 sum_h P_dh sum_photons K Re(conj(Z) dR) / |Z|^2
//                    + 2 scale_d sum_j Re(conj(Obar_d(j) + R_d(j)) dR_d(j)),
//         Obar_d = sum_h P_dh O_h                          (sum_h P_dh = 1)
// ---------------------------------------------------------------------------------

// Q[dl * H + h] = P[(d0 + dl) * H + h] * scale[d0 + dl]
__global__ void k_scaled_prob(const double* __restrict__ P, const double* __restrict__ scales,
                              int64_t d0, int64_t n, int H, double* __restrict__ Q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * H) return;
  const int64_t dl = i / H;
  Q[i] = P[d0 * H + i] * scales[d0 + dl];
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

// A planes of the split-BF16 frame-gradient GEMM Obar = P A: row 2j (2j+1) holds Re (Im)
// O_h(j) over h = 0 .. Hp-1 (zero for h >= H), from the pixel-major FP64 copy OT; each
// value rounded to FP32, then split into bf16 hi + lo.  Rows >= 2K stay as allocated (zero).
__global__ void k_ot_planes(const double2* __restrict__ OT, int K, int H, int Hs, int64_t Hp,
                            __bf16* __restrict__ hi, __bf16* __restrict__ lo) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)K * Hp) return;
  const int64_t j = t / Hp;
  const int h = (int)(t - j * Hp);
  const double2 o = h < H ? OT[j * Hs + h] : make_double2(0., 0.);
  const float re = (float)o.x, im = (float)o.y;
  const __bf16 rh = (__bf16)re, ih = (__bf16)im;
  hi[(2 * j) * Hp + h] = rh;
  lo[(2 * j) * Hp + h] = (__bf16)(re - (float)rh);
  hi[(2 * j + 1) * Hp + h] = ih;
  lo[(2 * j + 1) * Hp + h] = (__bf16)(im - (float)ih);
}

// A planes of the split-BF16 model-gradient GEMM GT = Q^T B: row h holds Q[d, h] over the
// sub-chunk's frames d = 0 .. Kc-1 (zero for h >= H or d >= n), rounded to FP32 and split
// into bf16 hi + lo.  Every element of the Mp x Kc planes is written.
__global__ void k_qt_planes(const double* __restrict__ Q, int64_t n, int H, int64_t Mp,
                            int64_t Kc, __bf16* __restrict__ hi, __bf16* __restrict__ lo) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= Mp * Kc) return;
  const int64_t h = t / Kc, d = t - h * Kc;
  const float v = (h < H && d < n) ? (float)Q[d * H + h] : 0.f;
  const __bf16 vh = (__bf16)v;
  hi[t] = vh;
  lo[t] = (__bf16)(v - (float)vh);
}

// B planes of GT = Q^T B: the transpose of the sub-chunk's FP64 reference rows B [n x ldb]
// (columns >= kval are zero) into rows k = 0 .. Np-1 over frames d = 0 .. Kc-1, through a
// 64 x 64 LDS tile (reads along k and writes along d both coalesced); every element of the
// Np x Kc planes is written (zero outside k < kval, d < n).
__global__ __launch_bounds__(256) void k_bt_planes(const double* __restrict__ B, int64_t n,
                                                   int64_t ldb, int64_t kval, int64_t Np,
                                                   int64_t Kc, __bf16* __restrict__ hi,
                                                   __bf16* __restrict__ lo) {
  __shared__ double tile[64][65];
  const int64_t k0 = (int64_t)blockIdx.x * 64, d0 = (int64_t)blockIdx.y * 64;
  const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
  for (int i = 0; i < 16; ++i) {
    const int r = r0 + 4 * i;
    const int64_t d = d0 + r, k = k0 + c;
    tile[r][c] = (d < n && k < kval) ? B[d * ldb + k] : 0.;
  }
  __syncthreads();
  for (int i = 0; i < 16; ++i) {
    const int r = r0 + 4 * i;
    const int64_t k = k0 + r, d = d0 + c;
    if (k >= Np || d >= Kc) continue;
    const float v = (float)tile[c][r];
    const __bf16 vh = (__bf16)v;
    hi[k * Kc + d] = vh;
    lo[k * Kc + d] = (__bf16)(v - (float)vh);
  }
}

// GT[k * H + h] = (accumulate ? GT : 0) + C[k * ldc + h] in FP64, k < Kp, h < H.
__global__ void k_gt_accum(const float* __restrict__ C, int64_t ldc, int64_t Kp, int H,
                           int accumulate, double* __restrict__ GT) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= Kp * H) return;
  const int64_t k = t / H, h = t - k * H;
  const double v = (double)C[k * ldc + h];
  GT[t] = accumulate ? GT[t] + v : v;
}

// One work-group per frame of [d0, d0 + gridDim.x): frame gradients (sx, sy, dia).
// Dense part: threads over the scored pixels (Obar row dl of the sub-chunk).  Sparse part:
// wave w takes photons w, w+4, ...; lanes over hypotheses h = lane + 64 t.
// Obar rows: (re, im) pairs of TO (double: the FP64 GEMM; float: the split-BF16 one),
// row stride ldo elements.
template <int NT, typename TO>
__global__ __launch_bounds__(256) void k_adj_frame(
    int size, int K, int H, int Hs, double rel, const int32_t* __restrict__ scored,
    const int32_t* __restrict__ jrad, const double* __restrict__ rad, int nrad,
    const int64_t* __restrict__ fo, const int32_t* __restrict__ fj,
    const float* __restrict__ fk, const double2* __restrict__ OT, const double* __restrict__ P,
    const TO* __restrict__ Obar, int64_t ldo, const double* __restrict__ sx,
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
  const TO* ob = Obar + dl * ldo;
  for (int j = threadIdx.x; j < K; j += 256) {
    double2 Rv, Dx, Dy, Dd;
    ref_and_derivs(tb, scored[j], jrad[j], size, Rv, Dx, Dy, Dd);
    const double2 o = make_double2((double)ob[2 * j], (double)ob[2 * j + 1]);
    const double zr = o.x + Rv.x, zi = o.y + Rv.y;
    ex_ += zr * Dx.x + zi * Dx.y;
    ey_ += zr * Dy.x + zi * Dy.y;
    ed_ += zr * Dd.x + zi * Dd.y;
  }
  // sparse: per hypothesis sum_photons K Re(conj(Z) dR) / |Z|^2
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
    const double2* row = OT + (int64_t)j * Hs + lane;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const double2 o = row[64 * t];
      const double zr = o.x + Rv.x, zi = o.y + Rv.y;
      const double n2 = zr * zr + zi * zi;
      if (sc * n2 >= 1.e-20) {
        // k / n2 through the hardware reciprocal and two Newton steps (within about an ulp
        // of the division, in fewer instructions: this loop runs per photon x hypothesis)
        double rn = __builtin_amdgcn_rcp(n2);
        rn = fma(rn, fma(-n2, rn, 1.), rn);
        rn = fma(rn, fma(-n2, rn, 1.), rn);
        const double f = k * rn;
        ax[t] += f * (zr * Dx.x + zi * Dx.y);
        ay[t] += f * (zr * Dy.x + zi * Dy.y);
        ad[t] += f * (zr * Dd.x + zi * Dd.y);
      }
    }
  }
  double px = 0., py = 0., pd = 0.;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int h = lane + 64 * t;
    if (h < H) {
      const double pr = P[d * H + h];
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

// One work-group per scored pixel j: GS[j * Hs + h] (+)= sum over the pixel's photons
// (d, K) of K P_dh Z / |Z|^2 (threads over hypotheses).  Pixel-major lists po/pd/pk.
template <int NU>
__global__ __launch_bounds__(256) void k_adj_pixel_acc(
    int size, int H, int Hs, double rel, const int32_t* __restrict__ scored,
    const int64_t* __restrict__ po, const int32_t* __restrict__ pd,
    const float* __restrict__ pk, const double2* __restrict__ OT, const double* __restrict__ P,
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ dia, const double* __restrict__ scales,
    double2* __restrict__ GS, int accumulate) {
  __shared__ double s_rr[256], s_ri[256], s_k[256], s_sc[256];
  __shared__ int32_t s_d[256];
  const int j = blockIdx.x, tid = threadIdx.x;
  const int t = scored[j];
  double2 O[NU], acc[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int h = tid + 256 * u;
    O[u] = h < Hs ? OT[(int64_t)j * Hs + h] : make_double2(0., 0.);
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
      const double* prow = P + (int64_t)s_d[q] * H;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int h = tid + 256 * u;
        if (h < H) {
          const double zr = O[u].x + rr, zi = O[u].y + ri;
          const double n2 = zr * zr + zi * zi;
          if (sc * n2 >= 1.e-20) {
            const double f = k * prow[h] / n2;
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
    const int h = tid + 256 * u;
    if (h >= Hs) continue;
    double2* g = GS + (int64_t)j * Hs + h;
    const double2 a = h < H ? acc[u] : make_double2(0., 0.);
    *g = accumulate ? make_double2(g->x + a.x, g->y + a.y) : a;
  }
}

// In place: G[j][h] = -2 GS + 2 ws_h O_h(j) + 2 GT(h, j) (GT column-major H x Kp, re/im at
// columns 2j, 2j+1); 0 for h >= H.  One thread per (j, h).
__global__ void k_adj_combine(int K, int H, int Hs, const double2* __restrict__ OT,
                              const double* __restrict__ GT, const double* __restrict__ wsum,
                              double2* __restrict__ G) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)K * Hs) return;
  const int64_t j = i / Hs;
  const int h = (int)(i - j * Hs);
  double2 g = make_double2(0., 0.);
  if (h < H) {
    const double2 s = G[i], o = OT[i];
    const double ws = wsum[h];
    g.x = -2. * s.x + 2. * ws * o.x + 2. * GT[(2 * j) * H + h];
    g.y = -2. * s.y + 2. * ws * o.y + 2. * GT[(2 * j + 1) * H + h];
  }
  G[i] = g;
}

// g_model[c](m) = sum_r sum_{scored view pixels j whose bilinear sample touches m} w *
// GO[j][c R + r] (the transpose of k_rot_slices' interpolation; gather form).
__global__ void k_adj_model(int size, int R, int Hs, const double* __restrict__ rcos,
                            const double* __restrict__ rsin,
                            const int32_t* __restrict__ compact_of,
                            const double2* __restrict__ GO, double* __restrict__ g_model) {
  const int npix = size * size;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= npix) return;
  const int c = blockIdx.y;
  const int mx = m / size, my = m - (m / size) * size;
  double gr = 0., gi = 0.;
  for (int r = 0; r < R; ++r) {
    const int h = c * R + r;
    for_each_source(size, mx, my, rcos[r], rsin[r], compact_of, [&](int j, double wx, double wy) {
      const double w = wx * wy;
      const double2 g = GO[(int64_t)j * Hs + h];
      gr += g.x * w;
      gi += g.y * w;
    });
  }
  const int64_t o = (int64_t)c * npix + m;
  g_model[2 * o] = gr;
  g_model[2 * o + 1] = gi;
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

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Scored frames as CSR in both orders (device pointers).
struct Csr {
  const int64_t* fo = nullptr;   // n + 1 frame offsets
  const int32_t* fj = nullptr;   // compact pixel of each entry
  const float* fk = nullptr;     // count of each entry
  const int64_t* po = nullptr;   // K + 1 pixel offsets
  const int32_t* pd = nullptr;   // frame (local to the CSR) of each entry
  const float* pk = nullptr;
  int64_t n = 0;
};

}  // namespace

struct rsp_rot {
  int device = 0, size = 0, npix = 0, R = 0, C = 1, H = 0, Hs = 0, NT = 0, K = 0, nrad = 0;
  double rel = 1.;
  int64_t Kp = 0, Mpad = 0;
  std::vector<int32_t> scored, compact_of;
  double *d_rcos = nullptr, *d_rsin = nullptr, *d_rad = nullptr;
  int32_t *d_scored = nullptr, *d_jrad = nullptr, *d_compact_of = nullptr;
  uint32_t* d_pxy = nullptr;  // scored pixel j as x | y << 16 (no division per pixel)
  // split form with the reference rows generated in the GEMM (option 8): pixel j of the K
  // order as bf16 (x - c) | bf16 (y - c) << 16 and its squared radius (Kp/2 entries), the
  // sphere table, the number of scored pixels at each distinct radius, the largest squared
  // radius
  int gen_ref = 1;
  uint32_t* d_gpix = nullptr;
  float* d_gr2 = nullptr;
  float4* d_gtab = nullptr;
  int32_t* d_radcnt = nullptr;
  double r2max = 0.;
  // resident frames
  int64_t N = 0, nnz = 0;
  int64_t *d_fo = nullptr, *d_po = nullptr;
  int32_t *d_fj = nullptr, *d_pd = nullptr;
  float *d_fk = nullptr, *d_pk = nullptr;
  // forward work
  int mode = kSplit, mode_alloc = -1, gen_alloc = -1;
  int64_t chunk = 0, chunk_alloc = 0;
  int xs = 1;                                   // K slices of the cross term (split form)
  void *d_A = nullptr, *d_OT = nullptr, *d_B = nullptr, *d_X = nullptr;
  double* d_LLc = nullptr;                      // LL of a GEMM chunk (when not asked for)
  double *d_normO = nullptr, *d_part = nullptr, *d_normR = nullptr;
  // merge
  double *d_VT = nullptr, *d_wsum = nullptr, *d_wpart = nullptr;
  int64_t wpart_alloc = 0;
  // adjoint (FP64 operands of its own, so that the forward mode's buffers stay)
  int64_t adj_chunk = 0;
  double *d_aA = nullptr, *d_aB = nullptr, *d_aObar = nullptr, *d_aQ = nullptr;
  double *d_aGT = nullptr, *d_apart = nullptr, *d_anormO = nullptr, *d_anormR = nullptr;
  double2 *d_aOT = nullptr, *d_aGO = nullptr;
  // frame gradients with a split-BF16 Obar (option 6): A planes [2 x fg_Mp x fg_Hp] of the
  // slices (r = pixel component, k = hypothesis), B planes [2 x fg_Np x fg_Hp] of P
  int fg_split = 0;
  int64_t fg_Mp = 0, fg_Np = 0, fg_Hp = 0;
  __bf16 *d_fgA = nullptr, *d_fgB = nullptr;
  // model gradient with a split-BF16 GT = Q^T B (option 7): A planes [2 x gt_Mp x gt_Kc]
  // of Q^T (r = hypothesis, k = frame of the sub-chunk), B planes [2 x gt_Np x gt_Kc] of
  // B^T (r = pixel component), FP32 product [Kp x gt_Mp]
  int gt_split = 0;
  int64_t gt_Mp = 0, gt_Np = 0, gt_Kc = 0;
  __bf16 *d_gtA = nullptr, *d_gtB = nullptr;
  float* d_gtC = nullptr;
  // streamed pass
  int pass_flags = -1;
  int64_t pass_max = 0, pc_alloc = 0;
  bool pass_first = true;
  double* d_Pc = nullptr;
  // rocBLAS handle of this plan (FP32 / FP64 forms and the adjoint's FP64 products)
  rocblas_handle blas = nullptr;
  // timing
  bool timing = false;
  hipEvent_t ev[8] = {};
  // side stream (default priority) for |R_d|^2 beside the split GEMM with generated rows,
  // joined by events: [0] chunk start on the caller's stream, [1] norms done
  hipStream_t side = nullptr;
  hipEvent_t sev[2] = {};
  double t_ms[5] = {0., 0., 0., 0., 0.};
};

namespace {

void free_work(rsp_rot* p) {
  dfree(p->d_A);
  dfree(p->d_OT);
  dfree(p->d_B);
  dfree(p->d_X);
  dfree(p->d_LLc);
  dfree(p->d_normO);
  dfree(p->d_part);
  dfree(p->d_normR);
  p->mode_alloc = -1;
  p->gen_alloc = -1;
  p->chunk_alloc = 0;
}

void free_adj(rsp_rot* p) {
  dfree(p->d_aA);
  dfree(p->d_aB);
  dfree(p->d_aObar);
  dfree(p->d_aQ);
  dfree(p->d_aGT);
  dfree(p->d_apart);
  dfree(p->d_anormO);
  dfree(p->d_anormR);
  dfree(p->d_aOT);
  dfree(p->d_aGO);
  dfree(p->d_fgA);
  dfree(p->d_fgB);
  dfree(p->d_gtA);
  dfree(p->d_gtB);
  dfree(p->d_gtC);
  p->adj_chunk = 0;
}

void free_merge(rsp_rot* p) {
  dfree(p->d_VT);
  dfree(p->d_wsum);
  dfree(p->d_wpart);
  p->wpart_alloc = 0;
}

int64_t elem_bytes(int mode) { return mode == kFP64 ? 8 : 4; }

// the B rows are stored (not generated in the GEMM)
bool stores_b(const rsp_rot* p) { return !(p->mode == kSplit && p->gen_ref); }

int ensure_work(rsp_rot* p, int64_t chunk) {
  if (p->mode_alloc == p->mode && p->chunk_alloc == chunk && p->gen_alloc == p->gen_ref)
    return kOk;
  free_work(p);
  const int64_t eb = elem_bytes(p->mode);
  // no B buffer when the split GEMM generates its reference rows
  const int64_t Bpad = stores_b(p) ? xterm_pad_n(chunk) : 1;
  const int64_t ntiles = ceil_div(p->K, 64);
  // OT entries: FP64 re/im (16 B), or FP32 hi re/im + bf16 lo re/im (12 B)
  const int64_t ot_bytes = (int64_t)p->K * p->Hs * (p->mode == kFP64 ? 16 : 12);
  const int xs = p->mode == kSplit ? xterm_split_k(p->Kp) : 1;
  int rc;
  char** A = reinterpret_cast<char**>(&p->d_A);
  char** B = reinterpret_cast<char**>(&p->d_B);
  char** OT = reinterpret_cast<char**>(&p->d_OT);
  char** X = reinterpret_cast<char**>(&p->d_X);
  if ((rc = dalloc(A, p->Mpad * p->Kp * eb)) || (rc = dalloc(B, Bpad * p->Kp * eb)) ||
      (rc = dalloc(OT, ot_bytes)) || (rc = dalloc(X, xs * chunk * p->Mpad * eb)) ||
      (rc = dalloc(&p->d_LLc, chunk * p->H)) ||
      (rc = dalloc(&p->d_normO, p->H)) || (rc = dalloc(&p->d_part, ntiles * p->H)) ||
      (rc = dalloc(&p->d_normR, chunk))) {
    free_work(p);
    return rc;
  }
  // padding rows / columns of the operands stay zero
  RSP_HIP(hipMemset(p->d_A, 0, (size_t)(p->Mpad * p->Kp * eb)));
  RSP_HIP(hipMemset(p->d_B, 0, (size_t)(Bpad * p->Kp * eb)));
  RSP_HIP(hipMemset(p->d_OT, 0, (size_t)ot_bytes));
  p->mode_alloc = p->mode;
  p->gen_alloc = p->gen_ref;
  p->chunk_alloc = chunk;
  p->xs = xs;
  return kOk;
}

int ensure_adj(rsp_rot* p, int64_t chunk) {
  if (p->adj_chunk == chunk) return kOk;
  free_adj(p);
  const int64_t ntiles = ceil_div(p->K, 64);
  int rc;
  if ((rc = dalloc(&p->d_aA, p->Mpad * p->Kp)) || (rc = dalloc(&p->d_aB, chunk * p->Kp)) ||
      (rc = dalloc(&p->d_aObar, chunk * p->Kp)) || (rc = dalloc(&p->d_aQ, chunk * p->H)) ||
      (rc = dalloc(&p->d_aGT, (int64_t)p->H * p->Kp)) ||
      (rc = dalloc(&p->d_apart, ntiles * p->H)) || (rc = dalloc(&p->d_anormO, p->H)) ||
      (rc = dalloc(&p->d_anormR, chunk)) ||
      (rc = dalloc(&p->d_aOT, (int64_t)p->K * p->Hs)) ||
      (rc = dalloc(&p->d_aGO, (int64_t)p->K * p->Hs))) {
    free_adj(p);
    return rc;
  }
  // split-BF16 Obar operands (rows >= 2K of the A planes stay zero)
  p->fg_Mp = xterm_pad_m(2 * (int64_t)p->K);
  p->fg_Hp = xterm_pad_k(p->H);
  p->fg_Np = xterm_pad_n(chunk);
  if ((rc = dalloc(&p->d_fgA, 2 * p->fg_Mp * p->fg_Hp)) ||
      (rc = dalloc(&p->d_fgB, 2 * p->fg_Np * p->fg_Hp))) {
    free_adj(p);
    return rc;
  }
  RSP_HIP(hipMemset(p->d_fgA, 0, sizeof(__bf16) * (size_t)(2 * p->fg_Mp * p->fg_Hp)));
  if (p->gt_split) {
    p->gt_Mp = xterm_pad_m(p->H);
    p->gt_Np = xterm_pad_n(p->Kp);
    p->gt_Kc = xterm_pad_k(chunk);
    if (xterm_split_k(p->gt_Kc) != 1) {
      free_adj(p);
      return fail(kErrSize, "split model-gradient GEMM: adjoint chunk above 12288 frames");
    }
    if ((rc = dalloc(&p->d_gtA, 2 * p->gt_Mp * p->gt_Kc)) ||
        (rc = dalloc(&p->d_gtB, 2 * p->gt_Np * p->gt_Kc)) ||
        (rc = dalloc(&p->d_gtC, p->Kp * p->gt_Mp))) {
      free_adj(p);
      return rc;
    }
  }
  // operand padding (rows >= H of A, columns >= 2K of A and B) stays zero
  RSP_HIP(hipMemset(p->d_aA, 0, sizeof(double) * (size_t)(p->Mpad * p->Kp)));
  RSP_HIP(hipMemset(p->d_aB, 0, sizeof(double) * (size_t)(chunk * p->Kp)));
  RSP_HIP(hipMemset(p->d_aOT, 0, sizeof(double2) * (size_t)((int64_t)p->K * p->Hs)));
  p->adj_chunk = chunk;
  return kOk;
}

int ensure_merge(rsp_rot* p, int64_t nblocks) {
  int rc;
  if (!p->d_VT && (rc = dalloc(&p->d_VT, (int64_t)p->K * p->Hs))) return rc;
  if (!p->d_wsum && (rc = dalloc(&p->d_wsum, p->H))) return rc;
  if (p->wpart_alloc < nblocks) {
    if ((rc = dalloc(&p->d_wpart, nblocks * p->H))) return rc;
    p->wpart_alloc = nblocks;
  }
  return kOk;
}

// O_h rows (A), the pixel-major copy (OT) and |O_h|^2 (normO) into the given buffers.
template <int MODE>
int launch_slices_to(rsp_rot* p, const double* model_c, void* A, void* OT, double* part,
                     double* normO, hipStream_t st) {
  dim3 g((unsigned)ceil_div(p->K, 64), (unsigned)ceil_div(p->H, 32));
  hipLaunchKernelGGL(k_rot_slices<MODE>, g, dim3(256), 0, st, model_c, p->size, p->K, p->R,
                     p->H, p->d_scored, p->d_rcos, p->d_rsin, A, p->Mpad, p->Kp, OT, p->Hs,
                     part);
  RSP_LAUNCH_CHECK("k_rot_slices");
  hipLaunchKernelGGL(k_norm_reduce, dim3((unsigned)ceil_div(p->H, 256)), dim3(256), 0, st,
                     part, (int)ceil_div(p->K, 64), p->H, normO, 0);
  RSP_LAUNCH_CHECK("k_norm_reduce");
  return kOk;
}

int launch_slices(rsp_rot* p, const double* model_c, hipStream_t st) {
  if (p->mode == kFP32)
    return launch_slices_to<kFP32>(p, model_c, p->d_A, p->d_OT, p->d_part, p->d_normO, st);
  if (p->mode == kSplit)
    return launch_slices_to<kSplit>(p, model_c, p->d_A, p->d_OT, p->d_part, p->d_normO, st);
  return launch_slices_to<kFP64>(p, model_c, p->d_A, p->d_OT, p->d_part, p->d_normO, st);
}

// R_d rows of frames [d0, d0 + n) of arrays of N frames into B (Bpad rows per bf16 plane).
template <int MODE>
int launch_ref_to(rsp_rot* p, const double* sx, const double* sy, const double* dia, int64_t d0,
                  int64_t n, int64_t N, void* B, int64_t Bpad, double* normR, hipStream_t st) {
  const size_t shmem =
      sizeof(double) * (size_t)((p->nrad + 1) & ~1) + 2 * sizeof(double2) * (size_t)p->size;
  hipLaunchKernelGGL(k_ref_rows<MODE>, dim3((unsigned)n), dim3(256), shmem, st, p->size, p->K,
                     p->d_pxy, p->d_jrad, p->d_rad, p->nrad, p->rel, sx, sy, dia, d0, N, B,
                     Bpad, p->Kp, normR);
  RSP_LAUNCH_CHECK("k_ref_rows");
  return kOk;
}

// LL of frames [d0, d1) into ll (rows d of the CSR's frames) or the chunk buffer, then the
// log-sum-exp.
template <typename T>
int launch_ll(rsp_rot* p, const Csr& f, const double* sx, const double* sy, const double* dia,
              const double* scales, int64_t d0, int64_t d1, double* ll, double* prob,
              double* logz, int32_t* best, hipStream_t st) {
  double* llb = ll ? ll + d0 * p->H : p->d_LLc;
  const unsigned fb = (unsigned)ceil_div(d1 - d0, 4);
  hipLaunchKernelGGL((k_rot_lt<T>), dim3(fb, (unsigned)ceil_div(p->H, 64)), dim3(256), 0, st,
                     p->size, p->H, p->Hs, p->rel, f.fo, f.fj, f.fk, p->d_scored, p->d_OT,
                     p->d_normO, p->d_normR, p->d_X, p->Mpad, p->xs,
                     p->chunk_alloc * p->Mpad, sx, sy, dia, scales, d0, d1, llb);
  RSP_LAUNCH_CHECK("k_rot_lt");
  if (prob || logz || best) {
    hipLaunchKernelGGL(k_rot_lse, dim3(fb), dim3(256), 0, st, p->H, d0, d1, llb, prob, logz,
                       best);
    RSP_LAUNCH_CHECK("k_rot_lse");
  }
  return kOk;
}

int ensure_events(rsp_rot* p) {
  for (auto& e : p->ev)
    if (!e) RSP_HIP(hipEventCreate(&e));
  return kOk;
}

int ensure_side(rsp_rot* p) {
  if (!p->side) RSP_HIP(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
  for (auto& e : p->sev)
    if (!e) RSP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return kOk;
}

// E-step over the frames of `f` (per-frame arrays indexed like f's frames), in GEMM chunks of
// p->chunk_alloc frames.  The slices of the forward mode must be current.
int run_estep(rsp_rot* p, const Csr& f, const double* sx, const double* sy, const double* dia,
              const double* scales, double* ll, double* prob, double* logz, int32_t* best,
              hipStream_t st) {
  const int64_t chunk = p->chunk_alloc;
  const int64_t Bpad = xterm_pad_n(chunk);
  const bool gen = !stores_b(p);
  int rc = kOk;
  if (p->timing && (rc = ensure_events(p))) return rc;
  if (gen && (rc = ensure_side(p))) return rc;
  for (int64_t d0 = 0; d0 < f.n; d0 += chunk) {
    const int64_t n = std::min(chunk, f.n - d0);
    if (p->timing) RSP_HIP(hipEventRecord(p->ev[2], st));
    if (gen) {
      // |R_d|^2 on the side stream, beside the GEMM (it needs only the frames' diameters);
      // it starts after everything queued on st so far (the previous chunk's log term has
      // read normR) and the log term below waits for it
      RSP_HIP(hipEventRecord(p->sev[0], st));
      RSP_HIP(hipStreamWaitEvent(p->side, p->sev[0], 0));
      hipLaunchKernelGGL(k_ref_norm, dim3((unsigned)n), dim3(256), 0, p->side, p->d_rad,
                         p->d_radcnt, p->nrad, p->rel, p->size, dia, d0, p->d_normR);
      RSP_LAUNCH_CHECK("k_ref_norm");
      RSP_HIP(hipEventRecord(p->sev[1], p->side));
    } else if (p->mode == kFP32)
      rc = launch_ref_to<kFP32>(p, sx, sy, dia, d0, n, f.n, p->d_B, Bpad, p->d_normR, st);
    else if (p->mode == kSplit)
      rc = launch_ref_to<kSplit>(p, sx, sy, dia, d0, n, f.n, p->d_B, Bpad, p->d_normR, st);
    else
      rc = launch_ref_to<kFP64>(p, sx, sy, dia, d0, n, f.n, p->d_B, Bpad, p->d_normR, st);
    if (rc) return rc;
    if (p->timing) RSP_HIP(hipEventRecord(p->ev[3], st));
    if (p->mode == kFP32) {
      rc = xterm_gemm_blas32(p->blas, (const float*)p->d_A, (const float*)p->d_B, (float*)p->d_X, p->Mpad,
                             n, p->Kp, p->Kp, p->Kp, p->Mpad, st);
    } else if (p->mode == kFP64) {
      rc = xterm_gemm_blas64(p->blas, (const double*)p->d_A, (const double*)p->d_B, (double*)p->d_X,
                             p->Mpad, n, p->Kp, p->Kp, p->Kp, p->Mpad, st);
    } else if (gen) {
      const __bf16* ah = (const __bf16*)p->d_A;
      rc = xterm_gemm_split_gen(ah, ah + p->Mpad * p->Kp, p->d_gpix, p->d_gr2, p->d_gtab, p->r2max,
                                p->rel, sx, sy, dia, d0, p->size, (float*)p->d_X, p->Mpad, n,
                                p->Kp, p->Mpad, chunk * p->Mpad, st);
    } else {
      const __bf16* ah = (const __bf16*)p->d_A;
      const __bf16* bh = (const __bf16*)p->d_B;
      rc = xterm_gemm_split(ah, ah + p->Mpad * p->Kp, bh, bh + Bpad * p->Kp, (float*)p->d_X,
                            p->Mpad, n, p->Kp, p->Mpad, chunk * p->Mpad, st);
    }
    if (rc) return rc;
    if (p->timing) RSP_HIP(hipEventRecord(p->ev[4], st));
    if (gen) RSP_HIP(hipStreamWaitEvent(st, p->sev[1], 0));
    if (p->mode == kFP64)
      rc = launch_ll<double>(p, f, sx, sy, dia, scales, d0, d0 + n, ll, prob, logz, best, st);
    else
      rc = launch_ll<float>(p, f, sx, sy, dia, scales, d0, d0 + n, ll, prob, logz, best, st);
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
  return kOk;
}

// ws_h (+)= sum_d P[d, h] scale_d over n frames into p->d_wsum (fixed summation order).
int run_wsum(rsp_rot* p, const double* prob, const double* scales, int64_t n, bool accumulate,
             hipStream_t st) {
  const int64_t per = 1024;
  const int64_t nb = ceil_div(n, per);
  int rc = ensure_merge(p, nb);
  if (rc) return rc;
  hipLaunchKernelGGL(k_wsum_part, dim3((unsigned)nb, (unsigned)ceil_div(p->H, 256)), dim3(256),
                     0, st, prob, scales, n, p->H, per, p->d_wpart);
  RSP_LAUNCH_CHECK("k_wsum_part");
  hipLaunchKernelGGL(k_norm_reduce, dim3((unsigned)ceil_div(p->H, 256)), dim3(256), 0, st,
                     p->d_wpart, (int)nb, p->H, p->d_wsum, accumulate ? 1 : 0);
  RSP_LAUNCH_CHECK("k_norm_reduce");
  return kOk;
}

int run_merge_view(rsp_rot* p, const Csr& f, const double* prob, bool accumulate,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_merge_view, dim3((unsigned)p->K), dim3(256), 0, st, f.po, f.pd, f.pk,
                     prob, p->H, p->Hs, p->d_VT, accumulate ? 1 : 0);
  RSP_LAUNCH_CHECK("k_merge_view");
  return kOk;
}

int run_merge_model(rsp_rot* p, double* model_out, double* acc_out, double* wts_out,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_merge_model, dim3((unsigned)ceil_div(p->npix, 256), (unsigned)p->C),
                     dim3(256), 0, st, p->size, p->R, p->Hs, p->d_rcos, p->d_rsin,
                     p->d_compact_of, p->d_VT, p->d_wsum, model_out, acc_out, wts_out);
  RSP_LAUNCH_CHECK("k_merge_model");
  return kOk;
}

size_t adj_frame_shmem(const rsp_rot* p) {
  return sizeof(double) * (size_t)((2 * p->nrad + 1) & ~1) +
         2 * sizeof(double2) * (size_t)p->size;
}

int check_adj_lds(const rsp_rot* p) {
  int lds_max = 0;
  RSP_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device));
  if (adj_frame_shmem(p) + 64 > (size_t)lds_max)
    return fail(kErrSize, "adjoint: reference tables exceed the LDS of a work-group");
  return kOk;
}

template <int NT>
void launch_adj_frame_nt(rsp_rot* p, const Csr& f, const double* prob, const double* sx,
                         const double* sy, const double* dia, const double* scales, int64_t d0,
                         int64_t n, double* gsx, double* gsy, double* gdia, hipStream_t st) {
  if (p->fg_split)   // Obar in FP32 from the split GEMM, row stride fg_Mp
    hipLaunchKernelGGL((k_adj_frame<NT, float>), dim3((unsigned)n), dim3(256),
                       adj_frame_shmem(p), st, p->size, p->K, p->H, p->Hs, p->rel, p->d_scored,
                       p->d_jrad, p->d_rad, p->nrad, f.fo, f.fj, f.fk, p->d_aOT, prob,
                       reinterpret_cast<const float*>(p->d_aObar), p->fg_Mp, sx, sy, dia,
                       scales, d0, gsx, gsy, gdia);
  else
    hipLaunchKernelGGL((k_adj_frame<NT, double>), dim3((unsigned)n), dim3(256),
                       adj_frame_shmem(p), st, p->size, p->K, p->H, p->Hs, p->rel, p->d_scored,
                       p->d_jrad, p->d_rad, p->nrad, f.fo, f.fj, f.fk, p->d_aOT, prob,
                       p->d_aObar, p->Kp, sx, sy, dia, scales, d0, gsx, gsy, gdia);
}

int launch_adj_frame(rsp_rot* p, const Csr& f, const double* prob, const double* sx,
                     const double* sy, const double* dia, const double* scales, int64_t d0,
                     int64_t n, double* gsx, double* gsy, double* gdia, hipStream_t st) {
  switch (p->NT) {
#define RSP_NT(k)                                                                          \
  case k:                                                                                  \
    launch_adj_frame_nt<k>(p, f, prob, sx, sy, dia, scales, d0, n, gsx, gsy, gdia, st);   \
    break;
    RSP_NT(1) RSP_NT(2) RSP_NT(3) RSP_NT(4) RSP_NT(5) RSP_NT(6) RSP_NT(8) RSP_NT(10)
    RSP_NT(12) RSP_NT(16) RSP_NT(20) RSP_NT(24)
#undef RSP_NT
    default: return fail(kErrUnsupported, "hypothesis count not supported");
  }
  RSP_LAUNCH_CHECK("k_adj_frame");
  return kOk;
}

// Adjoint sub-chunks over the frames of `f`: the dense model term GT (+)= Q^T B and/or the
// frame gradients.  The FP64 slices (d_aA, d_aOT) must be current.
int run_adj_frames(rsp_rot* p, const Csr& f, const double* prob, const double* sx,
                   const double* sy, const double* dia, const double* scales, double* gsx,
                   double* gsy, double* gdia, bool want_frame, bool want_model, bool gt_first,
                   hipStream_t st) {
  const int64_t chunk = p->adj_chunk;
  int rc;
  if (want_frame && p->fg_split) {
    const int64_t nt = (int64_t)p->K * p->fg_Hp;
    hipLaunchKernelGGL(k_ot_planes, dim3((unsigned)ceil_div(nt, 256)), dim3(256), 0, st,
                       p->d_aOT, p->K, p->H, p->Hs, p->fg_Hp, p->d_fgA,
                       p->d_fgA + p->fg_Mp * p->fg_Hp);
    RSP_LAUNCH_CHECK("k_ot_planes");
  }
  for (int64_t d0 = 0; d0 < f.n; d0 += chunk) {
    const int64_t n = std::min(chunk, f.n - d0);
    if ((rc = launch_ref_to<kFP64>(p, sx, sy, dia, d0, n, f.n, p->d_aB, chunk, p->d_anormR, st)))
      return rc;
    if (want_model) {
      hipLaunchKernelGGL(k_scaled_prob, dim3((unsigned)ceil_div(n * p->H, 256)), dim3(256), 0,
                         st, prob, scales, d0, n, p->H, p->d_aQ);
      RSP_LAUNCH_CHECK("k_scaled_prob");
      // GT (H x Kp, column-major) (+)= Q^T B : GT[k * H + h] = sum_d Q[d, h] B[d, k]
      const bool first = gt_first && d0 == 0;
      if (p->gt_split) {
        // on MFMA in split BF16: r = h, d = k (pixel component), reduction over frames
        const int64_t nq = p->gt_Mp * p->gt_Kc;
        hipLaunchKernelGGL(k_qt_planes, dim3((unsigned)ceil_div(nq, 256)), dim3(256), 0, st,
                           p->d_aQ, n, p->H, p->gt_Mp, p->gt_Kc, p->d_gtA, p->d_gtA + nq);
        RSP_LAUNCH_CHECK("k_qt_planes");
        const int64_t nb = p->gt_Np * p->gt_Kc;
        hipLaunchKernelGGL(k_bt_planes, dim3((unsigned)(p->gt_Np / 64), (unsigned)ceil_div(p->gt_Kc, 64)),
                           dim3(256), 0, st, p->d_aB, n, p->Kp, 2 * (int64_t)p->K, p->gt_Np,
                           p->gt_Kc, p->d_gtB, p->d_gtB + nb);
        RSP_LAUNCH_CHECK("k_bt_planes");
        if ((rc = xterm_gemm_split(p->d_gtA, p->d_gtA + nq, p->d_gtB, p->d_gtB + nb, p->d_gtC,
                                   p->gt_Mp, p->Kp, p->gt_Kc, p->gt_Mp, 0, st)))
          return rc;
        const int64_t ng = p->Kp * p->H;
        hipLaunchKernelGGL(k_gt_accum, dim3((unsigned)ceil_div(ng, 256)), dim3(256), 0, st,
                           p->d_gtC, p->gt_Mp, p->Kp, p->H, first ? 0 : 1, p->d_aGT);
        RSP_LAUNCH_CHECK("k_gt_accum");
      } else if ((rc = xterm_dgemm(p->blas, false, true, p->H, p->Kp, n, 1., p->d_aQ, p->H,
                                   p->d_aB, p->Kp, first ? 0. : 1., p->d_aGT, p->H, st))) {
        return rc;
      }
    }
    if (want_frame) {
      if (p->fg_split) {
        // Obar (n x fg_Mp FP32, row-major) = P A on MFMA: r = k (2K pixel components),
        // d = frame, reduction over h (one K slice: H <= 1536 < 12288)
        if ((rc = xterm_split64(prob + d0 * p->H, n, p->H, p->H, p->d_fgB, p->d_fgB + p->fg_Np * p->fg_Hp,
                                p->fg_Np, p->fg_Hp, st)) ||
            (rc = xterm_gemm_split(p->d_fgA, p->d_fgA + p->fg_Mp * p->fg_Hp, p->d_fgB,
                                   p->d_fgB + p->fg_Np * p->fg_Hp,
                                   reinterpret_cast<float*>(p->d_aObar), p->fg_Mp, n, p->fg_Hp,
                                   p->fg_Mp, 0, st)))
          return rc;
      } else if ((rc = xterm_dgemm(p->blas, false, false, p->Kp, n, p->H, 1., p->d_aA, p->Kp,
                                   prob + d0 * p->H, p->H, 0., p->d_aObar, p->Kp, st))) {
        // Obar (n x Kp, row-major) = P A : Obar[d * Kp + k] = sum_h P[d, h] A[h, k]
        return rc;
      }
      if ((rc = launch_adj_frame(p, f, prob, sx, sy, dia, scales, d0, n, gsx, gsy, gdia, st)))
        return rc;
    }
  }
  return kOk;
}

int run_adj_pixel(rsp_rot* p, const Csr& f, const double* prob, const double* sx,
                  const double* sy, const double* dia, const double* scales, bool accumulate,
                  hipStream_t st) {
  const dim3 g((unsigned)p->K), b(256);
  const int acc = accumulate ? 1 : 0;
  switch (ceil_div(p->Hs, 256)) {
#define RSP_NU(u)                                                                          \
  case u:                                                                                  \
    hipLaunchKernelGGL(k_adj_pixel_acc<u>, g, b, 0, st, p->size, p->H, p->Hs, p->rel,       \
                       p->d_scored, f.po, f.pd, f.pk, p->d_aOT, prob, sx, sy, dia, scales,  \
                       p->d_aGO, acc);                                                      \
    break;
    RSP_NU(1) RSP_NU(2) RSP_NU(3) RSP_NU(4) RSP_NU(5) RSP_NU(6)
#undef RSP_NU
    default: return fail(kErrUnsupported, "hypothesis count not supported");
  }
  RSP_LAUNCH_CHECK("k_adj_pixel_acc");
  return kOk;
}

int run_adj_finish(rsp_rot* p, double* g_model, hipStream_t st) {
  const int64_t n = (int64_t)p->K * p->Hs;
  hipLaunchKernelGGL(k_adj_combine, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, p->K,
                     p->H, p->Hs, p->d_aOT, p->d_aGT, p->d_wsum, p->d_aGO);
  RSP_LAUNCH_CHECK("k_adj_combine");
  hipLaunchKernelGGL(k_adj_model, dim3((unsigned)ceil_div(p->npix, 256), (unsigned)p->C),
                     dim3(256), 0, st, p->size, p->R, p->Hs, p->d_rcos, p->d_rsin,
                     p->d_compact_of, p->d_aGO, g_model);
  RSP_LAUNCH_CHECK("k_adj_model");
  return kOk;
}

int64_t default_chunk(const rsp_rot* p, int64_t nframes) {
  int64_t chunk = p->chunk;
  if (chunk <= 0) {
    // operand rows of a chunk within about 8 GB; with the reference rows generated in the
    // GEMM, its K-slice partial products and log-likelihoods within about 8 GB (one chunk
    // for 200k frames at 256^2: fewer, larger launches; the last wave of work-groups is a
    // smaller share of each)
    const double per_frame =
        stores_b(p) ? (double)(p->Kp * elem_bytes(p->mode))
                    : (double)(xterm_split_k(p->Kp) * p->Mpad * 4 + (int64_t)p->H * 8);
    chunk = std::max<int64_t>(128, (int64_t)(8e9 / per_frame) / 128 * 128);
  }
  return std::min(chunk, xterm_pad_n(nframes));
}

Csr resident(const rsp_rot* p) {
  Csr f;
  f.fo = p->d_fo;
  f.fj = p->d_fj;
  f.fk = p->d_fk;
  f.po = p->d_po;
  f.pd = p->d_pd;
  f.pk = p->d_pk;
  f.n = p->N;
  return f;
}

// Host packing of frames (dset layout) into the scored CSR of both orders.
int pack_frames(const rsp_rot* p, int64_t N, const int32_t* ones, const int32_t* multi,
                const int32_t* place_ones, const int32_t* place_multi,
                const int32_t* count_multi, std::vector<int64_t>& fo, std::vector<int32_t>& fj,
                std::vector<float>& fk, std::vector<int64_t>& po, std::vector<int32_t>& pd,
                std::vector<float>& pk) {
  fo.assign(N + 1, 0);
  int64_t no = 0, nm = 0;
  for (int64_t d = 0; d < N; ++d) {
    if (ones[d] < 0 || multi[d] < 0) return fail(kErrSize, "negative photon count");
    no += ones[d];
    nm += multi[d];
  }
  if ((no && !place_ones) || (nm && (!place_multi || !count_multi)))
    return fail(kErrNull, "null photon list");
  fj.clear();
  fk.clear();
  fj.reserve((size_t)(no + nm));
  fk.reserve((size_t)(no + nm));
  int64_t io = 0, im = 0;
  for (int64_t d = 0; d < N; ++d) {
    for (int q = 0; q < ones[d]; ++q, ++io) {
      const int t = place_ones[io];
      if (t < 0 || t >= p->npix) return fail(kErrSize, "photon index out of range");
      const int j = p->compact_of[t];
      if (j >= 0) {
        fj.push_back(j);
        fk.push_back(1.f);
      }
    }
    for (int q = 0; q < multi[d]; ++q, ++im) {
      const int t = place_multi[im];
      if (t < 0 || t >= p->npix) return fail(kErrSize, "photon index out of range");
      if (count_multi[im] < 1 || count_multi[im] >= (1 << 22))
        return fail(kErrSize, "count_multi out of range (1 <= count < 2^22)");
      const int j = p->compact_of[t];
      if (j >= 0) {
        fj.push_back(j);
        fk.push_back((float)count_multi[im]);
      }
    }
    fo[d + 1] = (int64_t)fj.size();
  }
  const int64_t nnz = (int64_t)fj.size();
  // pixel-major lists (stable in frame order)
  po.assign(p->K + 1, 0);
  for (int64_t e = 0; e < nnz; ++e) ++po[fj[e] + 1];
  for (int j = 0; j < p->K; ++j) po[j + 1] += po[j];
  std::vector<int64_t> fill(po.begin(), po.end() - 1);
  pd.assign(nnz, 0);
  pk.assign(nnz, 0.f);
  for (int64_t d = 0; d < N; ++d)
    for (int64_t e = fo[d]; e < fo[d + 1]; ++e) {
      const int64_t q = fill[fj[e]]++;
      pd[q] = (int32_t)d;
      pk[q] = fk[e];
    }
  return kOk;
}

void set_hypotheses(rsp_rot* p, int C) {
  p->C = C;
  p->H = C * p->R;
  p->NT = nt_for(p->H);
  p->Hs = 64 * p->NT;
  p->Mpad = xterm_pad_m(p->H);
}

}  // namespace

extern "C" {

int rsp_rot_create(int device, int64_t size, int64_t num_rot, const double* rot_cos,
                   const double* rot_sin, double rel_scale, rsp_rot** out) {
  if (!out || !rot_cos || !rot_sin) return fail(kErrNull, "rsp_rot_create: null argument");
  *out = nullptr;
  if (size < 9 || size > 4096) return fail(kErrSize, "rsp_rot_create: size out of range");
  if (num_rot < 1 || num_rot > kMaxH)
    return fail(kErrSize, "rsp_rot_create: 1 <= num_rot <= 1536");
  RSP_HIP(hipSetDevice(device));
  rsp_rot* p = new rsp_rot();
  p->device = device;
  p->size = (int)size;
  p->npix = (int)(size * size);
  p->R = (int)num_rot;
  set_hypotheses(p, 1);
  p->rel = rel_scale;
  // the generated reference rows pack |x - c| <= 256 exactly in bf16: sizes up to 513
  p->gen_ref = size <= kGenSizeMax ? 1 : 0;
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
  std::vector<uint32_t> pxy(p->K);
  for (int j = 0; j < p->K; ++j)
    pxy[j] = (uint32_t)(p->scored[j] / size) | ((uint32_t)(p->scored[j] % size) << 16);
  p->nrad = (int)rad.size();
  p->Kp = xterm_pad_k(2 * (int64_t)p->K);
  // generated reference rows: packed pixel coordinates of the K order (exact in bf16 for
  // |x - c| <= 256), pixels per distinct radius, the largest squared radius
  std::vector<uint32_t> gpix((size_t)(p->Kp / 2), 0u);
  std::vector<float> gr2((size_t)(p->Kp / 2), 0.f);
  std::vector<int32_t> radcnt(p->nrad, 0);
  auto bf16_bits = [](float v) {  // exact for the small integers stored here
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return u >> 16;
  };
  for (int j = 0; j < p->K; ++j) {
    const int xc = p->scored[j] / (int)size - c, yc = p->scored[j] % (int)size - c;
    gpix[j] = bf16_bits((float)xc) | (bf16_bits((float)yc) << 16);
    gr2[j] = (float)r2_of[j];  // exact: below 2^24
    ++radcnt[jrad[j]];
    p->r2max = std::max(p->r2max, (double)r2_of[j]);
  }
  std::vector<double> rc(rot_cos, rot_cos + num_rot), rs(rot_sin, rot_sin + num_rot);
  int rc_ = kOk;
  if ((rc_ = dupload(&p->d_rcos, rc)) || (rc_ = dupload(&p->d_rsin, rs)) ||
      (rc_ = dupload(&p->d_rad, rad)) || (rc_ = dupload(&p->d_scored, p->scored)) ||
      (rc_ = dupload(&p->d_jrad, jrad)) || (rc_ = dupload(&p->d_compact_of, p->compact_of)) ||
      (rc_ = dupload(&p->d_pxy, pxy)) || (rc_ = dupload(&p->d_gpix, gpix)) ||
      (rc_ = dupload(&p->d_gr2, gr2)) ||
      (rc_ = dupload(&p->d_radcnt, radcnt)) ||
      (rc_ = dalloc(&p->d_gtab, (int64_t)xterm_gen_table_len())) ||
      (rc_ = xterm_sphere_table(p->rel, p->d_gtab, nullptr))) {
    rsp_rot_destroy(p);
    return rc_;
  }
  {
    const hipError_t e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
      rsp_rot_destroy(p);
      return hip_fail(e, "rsp_rot_create: sphere table");
    }
  }
  if (sizeof(double) * (size_t)(p->nrad + 1) + 2 * sizeof(double2) * (size_t)size > 60000) {
    rsp_rot_destroy(p);
    return fail(kErrSize, "rsp_rot_create: radius table exceeds LDS");
  }
  if (rocblas_create_handle(&p->blas) != rocblas_status_success) {
    p->blas = nullptr;
    rsp_rot_destroy(p);
    return fail(kErrState, "rsp_rot_create: rocblas_create_handle failed");
  }
  *out = p;
  return kOk;
}

int rsp_rot_set_frames(rsp_rot* p, int64_t num_data, const int32_t* ones, const int32_t* multi,
                       const int32_t* place_ones, const int32_t* place_multi,
                       const int32_t* count_multi) {
  if (!p || !ones || !multi) return fail(kErrNull, "rsp_rot_set_frames: null argument");
  if (num_data < 1) return fail(kErrSize, "rsp_rot_set_frames: no frames");
  if (num_data > 0x7fffffff) return fail(kErrSize, "rsp_rot_set_frames: too many frames");
  RSP_HIP(hipSetDevice(p->device));
  std::vector<int64_t> fo, po;
  std::vector<int32_t> fj, pd;
  std::vector<float> fk, pk;
  int rc = pack_frames(p, num_data, ones, multi, place_ones, place_multi, count_multi, fo, fj,
                       fk, po, pd, pk);
  if (rc) return rc;
  if ((rc = dupload(&p->d_fo, fo)) || (rc = dupload(&p->d_fj, fj)) ||
      (rc = dupload(&p->d_fk, fk)) || (rc = dupload(&p->d_po, po)) ||
      (rc = dupload(&p->d_pd, pd)) || (rc = dupload(&p->d_pk, pk)))
    return rc;
  p->N = num_data;
  p->nnz = (int64_t)fj.size();
  return kOk;
}

int rsp_rot_pack(const rsp_rot* p, int64_t num_data, const int32_t* ones, const int32_t* multi,
                 const int32_t* place_ones, const int32_t* place_multi,
                 const int32_t* count_multi, int64_t* fo, int32_t* fj, float* fk, int64_t* po,
                 int32_t* pd, float* pk, int64_t* nnz_out) {
  if (!p || !ones || !multi || !nnz_out) return fail(kErrNull, "rsp_rot_pack: null argument");
  if (num_data < 1 || num_data > 0x7fffffff) return fail(kErrSize, "rsp_rot_pack: frame count");
  std::vector<int64_t> vfo, vpo;
  std::vector<int32_t> vfj, vpd;
  std::vector<float> vfk, vpk;
  int rc = pack_frames(p, num_data, ones, multi, place_ones, place_multi, count_multi, vfo, vfj,
                       vfk, vpo, vpd, vpk);
  if (rc) return rc;
  *nnz_out = (int64_t)vfj.size();
  if (!fo && !fj && !fk && !po && !pd && !pk) return kOk;   // size query
  if (!fo || !fj || !fk || !po || !pd || !pk) return fail(kErrNull, "rsp_rot_pack: null output");
  std::memcpy(fo, vfo.data(), sizeof(int64_t) * vfo.size());
  std::memcpy(po, vpo.data(), sizeof(int64_t) * vpo.size());
  if (!vfj.empty()) {
    std::memcpy(fj, vfj.data(), sizeof(int32_t) * vfj.size());
    std::memcpy(fk, vfk.data(), sizeof(float) * vfk.size());
    std::memcpy(pd, vpd.data(), sizeof(int32_t) * vpd.size());
    std::memcpy(pk, vpk.data(), sizeof(float) * vpk.size());
  }
  return kOk;
}

int rsp_rot_set_option(rsp_rot* p, int key, int64_t value) {
  if (!p) return fail(kErrNull, "rsp_rot_set_option: null plan");
  if (p->pass_flags >= 0) return fail(kErrState, "rsp_rot_set_option: a pass is open");
  switch (key) {
    case 1:
      if (value < 0 || value > 2) return fail(kErrSize, "mode must be 0, 1 or 2");
      p->mode = (int)value;
      return kOk;
    case 2:
      if (value < 0) return fail(kErrSize, "chunk must be >= 0");
      p->chunk = value;
      return kOk;
    case 3: {
      if (value < 1 || value * p->R > kMaxH)
        return fail(kErrSize, "classes must be >= 1 with classes * num_rot <= 1536");
      if (value == p->C) return kOk;
      RSP_HIP(hipSetDevice(p->device));
      free_work(p);
      free_adj(p);
      free_merge(p);
      set_hypotheses(p, (int)value);
      return kOk;
    }
    case 5:
      p->timing = value != 0;
      return kOk;
    case 6:
      if (value < 0 || value > 1) return fail(kErrSize, "frame-gradient GEMM must be 0 or 1");
      p->fg_split = (int)value;
      return kOk;
    case 8:
      if (value < 0 || value > 1) return fail(kErrSize, "generated reference rows must be 0 or 1");
      if (value && p->size > kGenSizeMax)
        return fail(kErrUnsupported, "generated reference rows need size <= 513");
      p->gen_ref = (int)value;  // the forward buffers follow at the next E-step
      return kOk;
    case 7:
      if (value < 0 || value > 1) return fail(kErrSize, "model-gradient GEMM must be 0 or 1");
      if (value != p->gt_split) {
        RSP_HIP(hipSetDevice(p->device));
        free_adj(p);            // its buffers are (re)allocated on the next adjoint
      }
      p->gt_split = (int)value;
      return kOk;
    default:
      return fail(kErrUnsupported, "unknown option key");
  }
}

int rsp_rot_estep(rsp_rot* p, const double* model_c, const double* sx, const double* sy,
                  const double* dia, const double* scales, double* ll, double* prob,
                  double* logz, rsp_stream_t stream_) {
  return rsp_rot_estep_best(p, model_c, sx, sy, dia, scales, ll, prob, logz, nullptr, stream_);
}

int rsp_rot_estep_best(rsp_rot* p, const double* model_c, const double* sx, const double* sy,
                       const double* dia, const double* scales, double* ll, double* prob,
                       double* logz, int32_t* best, rsp_stream_t stream_) {
  if (!p || !model_c || !sx || !sy || !dia || !scales)
    return fail(kErrNull, "rsp_rot_estep: null argument");
  if (p->N == 0) return fail(kErrState, "rsp_rot_estep: no frames");
  if (p->pass_flags >= 0) return fail(kErrState, "rsp_rot_estep: a pass is open");
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  int rc = ensure_work(p, default_chunk(p, p->N));
  if (rc) return rc;
  if (p->timing) {
    if ((rc = ensure_events(p))) return rc;
    for (double& t : p->t_ms) t = 0.;
    RSP_HIP(hipEventRecord(p->ev[0], st));
  }
  if ((rc = launch_slices(p, model_c, st))) return rc;
  if (p->timing) {
    RSP_HIP(hipEventRecord(p->ev[1], st));
    RSP_HIP(hipEventSynchronize(p->ev[1]));
    float ms = 0.f;
    RSP_HIP(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
    p->t_ms[0] += ms;
  }
  if ((rc = run_estep(p, resident(p), sx, sy, dia, scales, ll, prob, logz, best, st))) return rc;
  RSP_HIP(hipStreamSynchronize(st));
  return kOk;
}

int rsp_rot_merge(rsp_rot* p, const double* prob, const double* scales, double* model_out,
                  double* acc_out, double* weights_out, rsp_stream_t stream_) {
  if (!p || !prob || !scales || !model_out) return fail(kErrNull, "rsp_rot_merge: null argument");
  if (p->N == 0) return fail(kErrState, "rsp_rot_merge: no frames");
  if (p->pass_flags >= 0) return fail(kErrState, "rsp_rot_merge: a pass is open");
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  int rc;
  const Csr f = resident(p);
  if ((rc = run_wsum(p, prob, scales, p->N, false, st)) ||
      (rc = run_merge_view(p, f, prob, false, st)) ||
      (rc = run_merge_model(p, model_out, acc_out, weights_out, st)))
    return rc;
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
  if (p->pass_flags >= 0) return fail(kErrState, "rsp_rot_adjoint: a pass is open");
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  int rc = check_adj_lds(p);
  if (rc) return rc;
  const int64_t chunk = std::min(p->chunk > 0 ? p->chunk : 4096, p->N);
  if ((rc = ensure_adj(p, chunk))) return rc;
  if ((rc = launch_slices_to<kFP64>(p, model_c, p->d_aA, p->d_aOT, p->d_apart, p->d_anormO, st)))
    return rc;
  const Csr f = resident(p);
  if ((rc = run_wsum(p, prob, scales, p->N, false, st)) ||
      (rc = run_adj_frames(p, f, prob, sx, sy, dia, scales, g_sx, g_sy, g_dia, true, true, true,
                           st)) ||
      (rc = run_adj_pixel(p, f, prob, sx, sy, dia, scales, false, st)) ||
      (rc = run_adj_finish(p, g_model, st)))
    return rc;
  RSP_HIP(hipStreamSynchronize(st));
  return kOk;
}

int rsp_rot_pass_begin(rsp_rot* p, const double* model_c, int flags, int64_t max_chunk,
                       rsp_stream_t stream_) {
  if (!p || !model_c) return fail(kErrNull, "rsp_rot_pass_begin: null argument");
  if (p->pass_flags >= 0) return fail(kErrState, "rsp_rot_pass_begin: a pass is open");
  if (flags < 0 || flags > 7) return fail(kErrSize, "rsp_rot_pass_begin: flags");
  if (max_chunk < 1 || max_chunk > 0x7fffffff)
    return fail(kErrSize, "rsp_rot_pass_begin: max_chunk");
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  int rc = ensure_work(p, default_chunk(p, max_chunk));
  if (rc) return rc;
  if ((rc = launch_slices(p, model_c, st))) return rc;
  if (flags & (kPassFrameGrad | kPassModelGrad)) {
    if ((rc = check_adj_lds(p))) return rc;
    if ((rc = ensure_adj(p, std::min<int64_t>(p->chunk > 0 ? p->chunk : 4096, max_chunk))))
      return rc;
    if ((rc = launch_slices_to<kFP64>(p, model_c, p->d_aA, p->d_aOT, p->d_apart, p->d_anormO,
                                      st)))
      return rc;
  }
  if (flags & (kPassMerge | kPassModelGrad)) {
    if ((rc = ensure_merge(p, ceil_div(max_chunk, 1024)))) return rc;
    if (p->pc_alloc < max_chunk) {
      if ((rc = dalloc(&p->d_Pc, max_chunk * p->H))) return rc;
      p->pc_alloc = max_chunk;
    }
  } else if (flags & kPassFrameGrad) {
    if (p->pc_alloc < max_chunk) {
      if ((rc = dalloc(&p->d_Pc, max_chunk * p->H))) return rc;
      p->pc_alloc = max_chunk;
    }
  }
  if (p->timing)
    for (double& t : p->t_ms) t = 0.;
  p->pass_flags = flags;
  p->pass_max = max_chunk;
  p->pass_first = true;
  return kOk;
}

int rsp_rot_pass_chunk(rsp_rot* p, int64_t n, const int64_t* fo, const int32_t* fj,
                       const float* fk, const int64_t* po, const int32_t* pd, const float* pk,
                       const double* sx, const double* sy, const double* dia,
                       const double* scales, double* ll, double* prob, double* logz,
                       int32_t* best, double* g_sx, double* g_sy, double* g_dia,
                       rsp_stream_t stream_) {
  if (!p) return fail(kErrNull, "rsp_rot_pass_chunk: null plan");
  if (p->pass_flags < 0) return fail(kErrState, "rsp_rot_pass_chunk: no open pass");
  if (n < 1 || n > p->pass_max) return fail(kErrSize, "rsp_rot_pass_chunk: n > max_chunk");
  const int flags = p->pass_flags;
  if (!fo || !fj || !fk || !sx || !sy || !dia || !scales)
    return fail(kErrNull, "rsp_rot_pass_chunk: null argument");
  if ((flags & (kPassMerge | kPassModelGrad)) && (!po || !pd || !pk))
    return fail(kErrNull, "rsp_rot_pass_chunk: pixel-major lists needed");
  if ((flags & kPassFrameGrad) && (!g_sx || !g_sy || !g_dia))
    return fail(kErrNull, "rsp_rot_pass_chunk: frame gradient outputs needed");
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  Csr f;
  f.fo = fo;
  f.fj = fj;
  f.fk = fk;
  f.po = po;
  f.pd = pd;
  f.pk = pk;
  f.n = n;
  double* P = prob ? prob : (flags ? p->d_Pc : nullptr);
  int rc;
  if ((rc = run_estep(p, f, sx, sy, dia, scales, ll, P, logz, best, st))) return rc;
  const bool acc = !p->pass_first;
  if (flags & (kPassMerge | kPassModelGrad))
    if ((rc = run_wsum(p, P, scales, n, acc, st))) return rc;
  if (flags & kPassMerge)
    if ((rc = run_merge_view(p, f, P, acc, st))) return rc;
  if (flags & (kPassFrameGrad | kPassModelGrad))
    if ((rc = run_adj_frames(p, f, P, sx, sy, dia, scales, g_sx, g_sy, g_dia,
                             (flags & kPassFrameGrad) != 0, (flags & kPassModelGrad) != 0,
                             p->pass_first, st)))
      return rc;
  if (flags & kPassModelGrad)
    if ((rc = run_adj_pixel(p, f, P, sx, sy, dia, scales, acc, st))) return rc;
  p->pass_first = false;
  return kOk;
}

int rsp_rot_pass_end(rsp_rot* p, double* model_out, double* acc_out, double* weights_out,
                     double* g_model, rsp_stream_t stream_) {
  if (!p) return fail(kErrNull, "rsp_rot_pass_end: null plan");
  if (p->pass_flags < 0) return fail(kErrState, "rsp_rot_pass_end: no open pass");
  const int flags = p->pass_flags;
  p->pass_flags = -1;
  hipStream_t st = (hipStream_t)stream_;
  RSP_HIP(hipSetDevice(p->device));
  if (p->pass_first) {
    RSP_HIP(hipStreamSynchronize(st));
    if (flags & (kPassMerge | kPassModelGrad))
      return fail(kErrState, "rsp_rot_pass_end: the pass had no chunks");
    return kOk;
  }
  int rc;
  if ((flags & kPassMerge) && model_out)
    if ((rc = run_merge_model(p, model_out, acc_out, weights_out, st))) return rc;
  if ((flags & kPassModelGrad) && g_model)
    if ((rc = run_adj_finish(p, g_model, st))) return rc;
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
  out[7] = p->C;
  out[8] = p->H;
  out[9] = p->Hs;
  return kOk;
}

int rsp_rot_destroy(rsp_rot* p) {
  if (!p) return kOk;
  (void)hipSetDevice(p->device);
  free_work(p);
  free_adj(p);
  free_merge(p);
  dfree(p->d_Pc);
  dfree(p->d_rcos);
  dfree(p->d_rsin);
  dfree(p->d_rad);
  dfree(p->d_scored);
  dfree(p->d_pxy);
  dfree(p->d_gpix);
  dfree(p->d_gr2);
  dfree(p->d_gtab);
  dfree(p->d_radcnt);
  dfree(p->d_jrad);
  dfree(p->d_compact_of);
  dfree(p->d_fo);
  dfree(p->d_po);
  dfree(p->d_fj);
  dfree(p->d_pd);
  dfree(p->d_fk);
  dfree(p->d_pk);
  for (hipEvent_t e : p->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : p->sev)
    if (e) (void)hipEventDestroy(e);
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->blas) rocblas_destroy_handle(p->blas);
  delete p;
  return kOk;
}

}  // extern "C"
