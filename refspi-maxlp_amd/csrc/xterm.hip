// Dense cross term of the rotation form on MFMA:
//   X[d, r] = sum_k A[r, k] * B[d, k]
// with A = the rotated model slices [Re O_r(p), Im O_r(p)] and B = the per-frame
// reference terms [Re R_d(p), Im R_d(p)] over the scored pixels, so that
// X = Re sum_p O_r(p) conj(R_d(p)).
//
// Split-BF16: each FP32 operand x is carried as hi = bf16(x), lo = bf16(x - hi); the
// product uses hi*hi + hi*lo + lo*hi (the lo*lo term, below 2^-16 relative, is dropped)
// with FP32 accumulation in v_mfma_f32_32x32x16_bf16 over K slices of 12288 (split-K); the
// slices' partial products are added in FP64 by their consumer.
//
// Kernel layout (k_gemm_split): 256 threads = 4 waves, macro tile 256 (r) x 256 (d),
// each wave 128 x 128 = 4 x 4 MFMA tiles of 32 x 32 (accumulators in AGPRs).  K advances
// in stages of 32 bf16 through two static LDS buffers (64 KB per stage) filled by direct
// global->LDS loads one stage ahead; within a stage the fragment reads of the second
// 16-k step overlap the MFMAs of the first.  LDS rows are 64 B (4 slots of 16 B) with the
// slot XOR-swizzled by f(row) = ((row >> 1) ^ (row >> 3)) & 3, which makes the
// fragment reads (ds_read_b128, lane -> row lane%32, slot 2*kk + lane/32) hit distinct
// banks for every 8-, 16- and 32-lane group.  Work-groups are mapped XCD-major so the
// r-tiles of one d-tile run on one XCD and share its L2.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "refspi_amd.h"
#include "rsp_common.hpp"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTM = 256, kTN = 256, kBK = 32;
constexpr int kRowB = kBK * 2;                 // bytes per tile row (64)
constexpr int kTileA = kTM * kRowB;            // 16 KB per A plane tile
constexpr int kTileB = kTN * kRowB;            // 16 KB per B plane tile
constexpr int kLA = kTileA / 1024 / 4;         // DMA instructions per wave per A plane (4)
constexpr int kLB = kTileB / 1024 / 4;         // per B plane (4)
// K stages per split-K slice: 384 stages = 12288 K elements accumulated in FP32 per slice
// (the slices are added in FP64 by the consumer).
constexpr int kSliceStages = 384;

__shared__ __attribute__((aligned(16))) char s_ah0[kTileA], s_ah1[kTileA];
__shared__ __attribute__((aligned(16))) char s_al0[kTileA], s_al1[kTileA];
__shared__ __attribute__((aligned(16))) char s_bh0[kTileB], s_bh1[kTileB];
__shared__ __attribute__((aligned(16))) char s_bl0[kTileB], s_bl1[kTileB];

template <int S> __device__ __forceinline__ char* t_ah() {
  if constexpr (S == 0) return s_ah0; else return s_ah1;
}
template <int S> __device__ __forceinline__ char* t_al() {
  if constexpr (S == 0) return s_al0; else return s_al1;
}
template <int S> __device__ __forceinline__ char* t_bh() {
  if constexpr (S == 0) return s_bh0; else return s_bh1;
}
template <int S> __device__ __forceinline__ char* t_bl() {
  if constexpr (S == 0) return s_bl0; else return s_bl1;
}

__device__ __forceinline__ int swz(int row) { return ((row >> 1) ^ (row >> 3)) & 3; }

__device__ __forceinline__ void dma16(const void* g, char* l) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)g,
                                   (__attribute__((address_space(3))) void*)l, 16, 0, 0);
}

// The fragments of one 16-wide k step of a wave's 128 x 128 tile: rows of A (hi, lo) and
// of B (hi, lo), four 32-row MFMA operands each.
struct Frags {
  bf16x8 ah[4], al[4], bh[4], bl[4];
};

struct SplitGemm {
  const __bf16 *ah, *al, *bh, *bl;   // planes: A [Mpad x Kp], B [Npad x Kp]
  int64_t aoff[kLA], boff[kLB];      // per-lane element offsets of this tile's slots
  int nk, wave, lane;

  template <int S>
  __device__ __forceinline__ void issue(int q) const {
    const int64_t k0 = (int64_t)min(q, nk - 1) * kBK;
#pragma unroll
    for (int u = 0; u < kLA; ++u) {
      dma16(ah + aoff[u] + k0, t_ah<S>() + (wave * kLA + u) * 1024);
      dma16(al + aoff[u] + k0, t_al<S>() + (wave * kLA + u) * 1024);
    }
#pragma unroll
    for (int u = 0; u < kLB; ++u) {
      dma16(bh + boff[u] + k0, t_bh<S>() + (wave * kLB + u) * 1024);
      dma16(bl + boff[u] + k0, t_bl<S>() + (wave * kLB + u) * 1024);
    }
  }

  template <int S>
  __device__ __forceinline__ void load(Frags& f, int kk, int wm, int wn) const {
    const int lr = lane & 31;
    const int kc = ((2 * kk + (lane >> 5)) ^ swz(lr)) * 16;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = (wm * 128 + i * 32 + lr) * kRowB + kc;
      f.ah[i] = *reinterpret_cast<const bf16x8*>(t_ah<S>() + o);
      f.al[i] = *reinterpret_cast<const bf16x8*>(t_al<S>() + o);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = (wn * 128 + j * 32 + lr) * kRowB + kc;
      f.bh[j] = *reinterpret_cast<const bf16x8*>(t_bh<S>() + o);
      f.bl[j] = *reinterpret_cast<const bf16x8*>(t_bl<S>() + o);
    }
  }

  // One of the three products (0: lo*hi, 1: hi*lo, 2: hi*hi, applied in this order) over
  // the 16 tiles.
  template <int PR>
  __device__ __forceinline__ static void mma(f32x16 (&acc)[4][4], const Frags& f) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if constexpr (PR == 0)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.al[i], f.bh[j], acc[i][j], 0, 0, 0);
        else if constexpr (PR == 1)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.ah[i], f.bl[j], acc[i][j], 0, 0, 0);
        else
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.ah[i], f.bh[j], acc[i][j], 0, 0, 0);
      }
  }

  // One stage (32 k) of the wave's tile.  The products run as three groups of 16 MFMAs
  // kept apart by scheduling barriers, so that the three MFMAs on one accumulator are 16
  // apart instead of back to back (each waiting for the previous result); the fragment
  // reads of the second 16-k step are interleaved with the first group.
  template <int S>
  __device__ __forceinline__ void compute(f32x16 (&acc)[4][4], int wm, int wn) const {
    Frags f0, f1;
    load<S>(f0, 0, wm, wn);
    // all of f0 has landed here (one wait), and no read of f1 moves above this point: the
    // MFMAs on f0 then need no waits, and only the f1 reads interleave with them
    asm volatile(""
                 : "+v"(f0.ah[0]), "+v"(f0.ah[1]), "+v"(f0.ah[2]), "+v"(f0.ah[3]),
                   "+v"(f0.al[0]), "+v"(f0.al[1]), "+v"(f0.al[2]), "+v"(f0.al[3]),
                   "+v"(f0.bh[0]), "+v"(f0.bh[1]), "+v"(f0.bh[2]), "+v"(f0.bh[3]),
                   "+v"(f0.bl[0]), "+v"(f0.bl[1]), "+v"(f0.bl[2]), "+v"(f0.bl[3])
                 :
                 : "memory");
    load<S>(f1, 1, wm, wn);
    mma<0>(acc, f0);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
    }
    __builtin_amdgcn_sched_barrier(0);
    mma<1>(acc, f0);
    __builtin_amdgcn_sched_barrier(0);
    mma<2>(acc, f0);
    __builtin_amdgcn_sched_barrier(0);
    mma<0>(acc, f1);
    __builtin_amdgcn_sched_barrier(0);
    mma<1>(acc, f1);
    __builtin_amdgcn_sched_barrier(0);
    mma<2>(acc, f1);
  }

  template <int S>
  __device__ __forceinline__ void step(int q, f32x16 (&acc)[4][4], int wm, int wn) const {
    // vmcnt = 0: this wave's loads of stage q (the only ones in flight) have landed;
    // lgkmcnt = 0: its reads of stage q-1 are done -- the loads issued after the barrier
    // overwrite the buffer stage q-1 read
    __builtin_amdgcn_s_waitcnt(0x070);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    issue<(S + 1) % 2>(q + 1);
    if (q < nk) compute<S>(acc, wm, wn);
  }
};

// K is cut into slices of kst stages (blockIdx.y = slice); slice z accumulates its stages
// in FP32 and writes its partial product to C + z * slice_stride.  256 threads = 2 x 2
// waves of 128 x 128 (4 x 4 MFMA tiles of 32 x 32, accumulators in AGPRs).
__global__ __launch_bounds__(256, 1) void k_gemm_split(
    const __bf16* __restrict__ ah, const __bf16* __restrict__ al,
    const __bf16* __restrict__ bh, const __bf16* __restrict__ bl, float* __restrict__ C,
    int64_t N, int64_t Kp, int64_t ldc, int nM, int nN, int kst, int64_t slice_stride) {
  const int T = nM * nN;
  const int T8 = (T + 7) / 8;
  const int task = (int)(blockIdx.x & 7) * T8 + (int)(blockIdx.x >> 3);
  if (task >= T) return;
  const int mt = task % nM, nt = task / nM;
  const int64_t m0 = (int64_t)mt * kTM, n0 = (int64_t)nt * kTN;
  const int z = blockIdx.y;
  const int nkt = (int)(Kp / kBK);
  const int64_t kz = (int64_t)z * kst * kBK;          // first K element of this slice
  C += (int64_t)z * slice_stride;
  SplitGemm g;
  g.ah = ah + kz; g.al = al + kz; g.bh = bh + kz; g.bl = bl + kz;
  g.lane = threadIdx.x & 63;
  g.wave = threadIdx.x >> 6;
  g.nk = min(kst, nkt - z * kst);
#pragma unroll
  for (int u = 0; u < kLA; ++u) {
    const int s = (g.wave * kLA + u) * 64 + g.lane;
    const int row = s >> 2, kcl = (s & 3) ^ swz(row);
    g.aoff[u] = (m0 + row) * Kp + kcl * 8;
  }
#pragma unroll
  for (int u = 0; u < kLB; ++u) {
    const int s = (g.wave * kLB + u) * 64 + g.lane;
    const int row = s >> 2, kcl = (s & 3) ^ swz(row);
    g.boff[u] = (n0 + row) * Kp + kcl * 8;
  }
  const int wm = g.wave >> 1, wn = g.wave & 1;
  f32x16 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;
  g.issue<0>(0);
  for (int q = 0; q < g.nk; q += 2) {
    g.step<0>(q, acc, wm, wn);
    g.step<1>(q + 1, acc, wm, wn);
  }
  __builtin_amdgcn_s_waitcnt(0xF70);   // the trailing (clamped) loads
  // C[d, r]: lane -> d = lane%32 of the 32-wide tile, rows 8*(e/4) + 4*(lane/32) + e%4
  const int lr = g.lane & 31, hi = g.lane >> 5;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t d = n0 + wn * 128 + j * 32 + lr;
    if (d >= N) continue;
    float* crow = C + d * ldc + m0 + wm * 128;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e4 = 0; e4 < 4; ++e4) {
        f32x4 v = {acc[i][j][4 * e4], acc[i][j][4 * e4 + 1], acc[i][j][4 * e4 + 2],
                   acc[i][j][4 * e4 + 3]};
        *reinterpret_cast<f32x4*>(crow + i * 32 + 8 * e4 + 4 * hi) = v;
      }
  }
}

// FP32 rows [rows x K] (leading dimension ld) -> bf16 hi/lo planes [rows_pad x Kp],
// zero padded.  One thread per 8 consecutive k.
__global__ void k_split_rows(const float* __restrict__ X, int64_t rows, int64_t K, int64_t ld,
                             __bf16* __restrict__ hi, __bf16* __restrict__ lo,
                             int64_t rows_pad, int64_t Kp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_row = Kp / 8;
  if (t >= rows_pad * per_row) return;
  const int64_t r = t / per_row, k0 = (t - r * per_row) * 8;
  bf16x8 h, l;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float x = (r < rows && k0 + e < K) ? X[r * ld + k0 + e] : 0.f;
    const __bf16 xh = (__bf16)x;
    h[e] = xh;
    l[e] = (__bf16)(x - (float)xh);
  }
  *reinterpret_cast<bf16x8*>(hi + r * Kp + k0) = h;
  *reinterpret_cast<bf16x8*>(lo + r * Kp + k0) = l;
}

// As k_split_rows for FP64 rows (each value rounded to FP32 first).
__global__ void k_split_rows64(const double* __restrict__ X, int64_t rows, int64_t K,
                               int64_t ld, __bf16* __restrict__ hi, __bf16* __restrict__ lo,
                               int64_t rows_pad, int64_t Kp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_row = Kp / 8;
  if (t >= rows_pad * per_row) return;
  const int64_t r = t / per_row, k0 = (t - r * per_row) * 8;
  bf16x8 h, l;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float x = (r < rows && k0 + e < K) ? (float)X[r * ld + k0 + e] : 0.f;
    const __bf16 xh = (__bf16)x;
    h[e] = xh;
    l[e] = (__bf16)(x - (float)xh);
  }
  *reinterpret_cast<bf16x8*>(hi + r * Kp + k0) = h;
  *reinterpret_cast<bf16x8*>(lo + r * Kp + k0) = l;
}

// C[d * M + r] = sum over the S slices (FP64, fixed order) of P[z * stride + d * ldp + r].
__global__ void k_slice_sum(const float* __restrict__ P, int64_t stride, int S, int64_t M,
                            int64_t ldp, int64_t n, float* __restrict__ C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t d = i / M, r = i - d * M;
  double s = 0.;
  for (int z = 0; z < S; ++z) s += (double)P[z * stride + d * ldp + r];
  C[i] = (float)s;
}

}  // namespace

namespace rsp {

int64_t xterm_pad_m(int64_t M) { return ceil_div(M, kTM) * kTM; }
int64_t xterm_pad_n(int64_t N) { return ceil_div(N, kTN) * kTN; }
int64_t xterm_pad_k(int64_t K) { return ceil_div(K, kBK) * kBK; }

int xterm_split(const float* X, int64_t rows, int64_t K, int64_t ld, __bf16* hi, __bf16* lo,
                int64_t rows_pad, int64_t Kp, hipStream_t stream) {
  const int64_t n = rows_pad * (Kp / 8);
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_split_rows, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, X,
                     rows, K, ld, hi, lo, rows_pad, Kp);
  RSP_LAUNCH_CHECK("k_split_rows");
  return kOk;
}

int xterm_split64(const double* X, int64_t rows, int64_t K, int64_t ld, __bf16* hi,
                  __bf16* lo, int64_t rows_pad, int64_t Kp, hipStream_t stream) {
  const int64_t n = rows_pad * (Kp / 8);
  if (n == 0) return kOk;
  hipLaunchKernelGGL(k_split_rows64, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, X,
                     rows, K, ld, hi, lo, rows_pad, Kp);
  RSP_LAUNCH_CHECK("k_split_rows64");
  return kOk;
}

int xterm_split_k(int64_t Kp) { return (int)ceil_div(Kp / kBK, kSliceStages); }

// Split-BF16 GEMM on padded planes: partial products of the K slices at
// C + z * slice_stride, [d * ldc + r] for d < N, r < Mpad, z < xterm_split_k(Kp).
int xterm_gemm_split(const __bf16* ah, const __bf16* al, const __bf16* bh, const __bf16* bl,
                     float* C, int64_t Mpad, int64_t N, int64_t Kp, int64_t ldc,
                     int64_t slice_stride, hipStream_t stream) {
  if (Mpad % kTM || Kp % kBK || ldc < Mpad || ldc % 4) return fail(kErrSize, "gemm padding");
  const int S = xterm_split_k(Kp);
  if (S > 1 && slice_stride < (N - 1) * ldc + Mpad) return fail(kErrSize, "gemm slice stride");
  if (N == 0 || Mpad == 0) return kOk;
  const int nM = (int)(Mpad / kTM), nN = (int)ceil_div(N, kTN);
  const int T8 = (nM * nN + 7) / 8;
  hipLaunchKernelGGL(k_gemm_split, dim3(8 * T8, S), dim3(256), 0, stream, ah, al, bh, bl, C, N,
                     Kp, ldc, nM, nN, kSliceStages, slice_stride);
  RSP_LAUNCH_CHECK("k_gemm_split");
  return kOk;
}

// FP32 / FP64 through rocBLAS: C[d * ldc + r] = sum_k A[r * lda + k] B[d * ldb + k].
int xterm_gemm_blas32(rocblas_handle h, const float* A, const float* B, float* C, int64_t M,
                      int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc,
                      hipStream_t stream) {
  if (rocblas_set_stream(h, stream) != rocblas_status_success)
    return fail(kErrState, "rocblas_set_stream failed");
  const float one = 1.f, zero = 0.f;
  rocblas_status s = rocblas_sgemm(h, rocblas_operation_transpose, rocblas_operation_none,
                                   (rocblas_int)M, (rocblas_int)N, (rocblas_int)K, &one, A,
                                   (rocblas_int)lda, B, (rocblas_int)ldb, &zero, C,
                                   (rocblas_int)ldc);
  if (s != rocblas_status_success) return fail(kErrState, "rocblas_sgemm failed");
  return kOk;
}

int xterm_gemm_blas64(rocblas_handle h, const double* A, const double* B, double* C,
                      int64_t M, int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc,
                      hipStream_t stream) {
  if (rocblas_set_stream(h, stream) != rocblas_status_success)
    return fail(kErrState, "rocblas_set_stream failed");
  const double one = 1., zero = 0.;
  rocblas_status s = rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none,
                                   (rocblas_int)M, (rocblas_int)N, (rocblas_int)K, &one, A,
                                   (rocblas_int)lda, B, (rocblas_int)ldb, &zero, C,
                                   (rocblas_int)ldc);
  if (s != rocblas_status_success) return fail(kErrState, "rocblas_dgemm failed");
  return kOk;
}

int xterm_dgemm(rocblas_handle h, bool transA, bool transB, int64_t m, int64_t n, int64_t k,
                double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                double beta, double* C, int64_t ldc, hipStream_t stream) {
  if (m > 0x7fffffff || n > 0x7fffffff || k > 0x7fffffff || lda > 0x7fffffff ||
      ldb > 0x7fffffff || ldc > 0x7fffffff)
    return fail(kErrSize, "dgemm size too large");
  if (m == 0 || n == 0) return kOk;
  if (rocblas_set_stream(h, stream) != rocblas_status_success)
    return fail(kErrState, "rocblas_set_stream failed");
  rocblas_status s = rocblas_dgemm(
      h, transA ? rocblas_operation_transpose : rocblas_operation_none,
      transB ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)m,
      (rocblas_int)n, (rocblas_int)k, &alpha, A, (rocblas_int)lda, B, (rocblas_int)ldb, &beta,
      C, (rocblas_int)ldc);
  if (s != rocblas_status_success) return fail(kErrState, "rocblas_dgemm failed");
  return kOk;
}

}  // namespace rsp

namespace {

// Split-BF16 GEMM of FP32 row-major A [M x K], B [N x K] into C [N x M] with the bf16 planes
// (2 * (Mp + Np) * Kp elements) and the K-slice partial products (S * N * Mp floats) in
// caller-provided device memory; every element of both is written before it is read.
int split_gemm_into(int64_t M, int64_t N, int64_t K, const float* A, const float* B, float* C,
                    __bf16* planes, float* Cp, hipStream_t stream) {
  using namespace rsp;
  const int64_t Mp = xterm_pad_m(M), Np = xterm_pad_n(N), Kp = xterm_pad_k(K);
  __bf16 *ah = planes, *al = ah + Mp * Kp, *bh = al + Mp * Kp, *bl = bh + Np * Kp;
  const int S = xterm_split_k(Kp);
  int rc = xterm_split(A, M, K, K, ah, al, Mp, Kp, stream);
  if (!rc) rc = xterm_split(B, N, K, K, bh, bl, Np, Kp, stream);
  if (!rc) rc = xterm_gemm_split(ah, al, bh, bl, Cp, Mp, N, Kp, Mp, N * Mp, stream);
  if (!rc) {
    const int64_t n = N * M;
    hipLaunchKernelGGL(k_slice_sum, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, stream, Cp,
                       N * Mp, S, M, Mp, n, C);
    hipError_t ce = hipGetLastError();
    if (ce != hipSuccess) rc = hip_fail(ce, "k_slice_sum");
  }
  return rc;
}

// Bytes of plane and partial-product memory split_gemm_into needs.
void split_gemm_bytes(int64_t M, int64_t N, int64_t K, size_t* plane_bytes, size_t* part_bytes) {
  using namespace rsp;
  const int64_t Mp = xterm_pad_m(M), Np = xterm_pad_n(N), Kp = xterm_pad_k(K);
  *plane_bytes = (size_t)(2 * Mp + 2 * Np) * Kp * sizeof(__bf16);
  *part_bytes = (size_t)xterm_split_k(Kp) * N * Mp * sizeof(float);
}

int gemm_args(int64_t M, int64_t N, int64_t K, const void* A, const void* B, void* C) {
  using namespace rsp;
  if (M < 0 || N < 0 || K < 0) return fail(kErrSize, "negative size");
  if ((M && N && K) && (!A || !B || !C)) return fail(kErrNull, "null buffer");
  if (M > 0x7fffffff || N > 0x7fffffff || K > 0x7fffffff) return fail(kErrSize, "size too large");
  return kOk;
}

// Scratch of rsp_xterm_gemm_static: the bf16 planes and partial products of one call.  Each
// GPU holds its own copy from the library's load until the process ends; its device address
// comes from hipGetSymbolAddress (the array's name is a host-side stand-in in host code).
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
int scratch_ptr(void** out) {
  using namespace rsp;
  int dev = 0;
  RSP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(kErrUnsupported, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(kErrState, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace

extern "C" {

int rsp_xterm_gemm(int mode, int64_t M, int64_t N, int64_t K, const void* A, const void* B,
                   void* C, rsp_stream_t stream_) {
  using namespace rsp;
  hipStream_t stream = (hipStream_t)stream_;
  int rc = gemm_args(M, N, K, A, B, C);
  if (rc) return rc;
  if (M == 0 || N == 0) return kOk;
  if (mode == 0 || mode == 2) {
    // a handle of the call's own: concurrent callers never share one
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      return fail(kErrState, "rocblas_create_handle failed");
    if (mode == 0)
      rc = xterm_gemm_blas32(h, (const float*)A, (const float*)B, (float*)C, M, N, K, K, K, M,
                             stream);
    else
      rc = xterm_gemm_blas64(h, (const double*)A, (const double*)B, (double*)C, M, N, K, K, K,
                             M, stream);
    rocblas_destroy_handle(h);
    return rc;
  }
  if (mode != 1) return fail(kErrUnsupported, "mode must be 0 (FP32), 1 (split BF16), 2 (FP64)");
  size_t plane_bytes = 0, part_bytes = 0;
  split_gemm_bytes(M, N, K, &plane_bytes, &part_bytes);
  __bf16* planes = nullptr;
  float* Cp = nullptr;
  RSP_HIP(hipMallocAsync((void**)&planes, plane_bytes, stream));
  hipError_t e = hipMallocAsync((void**)&Cp, part_bytes, stream);
  if (e != hipSuccess) {
    (void)hipFreeAsync(planes, stream);
    return hip_fail(e, "hipMallocAsync");
  }
  rc = split_gemm_into(M, N, K, (const float*)A, (const float*)B, (float*)C, planes, Cp, stream);
  (void)hipFreeAsync(Cp, stream);
  (void)hipFreeAsync(planes, stream);
  return rc;
}

int rsp_xterm_gemm_static(int64_t M, int64_t N, int64_t K, const float* A, const float* B,
                          float* C, rsp_stream_t stream_) {
  using namespace rsp;
  hipStream_t stream = (hipStream_t)stream_;
  int rc = gemm_args(M, N, K, A, B, C);
  if (rc) return rc;
  if (M == 0 || N == 0) return kOk;
  size_t plane_bytes = 0, part_bytes = 0;
  split_gemm_bytes(M, N, K, &plane_bytes, &part_bytes);
  if (plane_bytes + part_bytes > kStaticScratchBytes)
    return fail(kErrSize, "rsp_xterm_gemm_static: needs " +
                              std::to_string(plane_bytes + part_bytes) +
                              " bytes of scratch, more than its 33554432");
  void* base = nullptr;
  if ((rc = scratch_ptr(&base))) return rc;
  __bf16* planes = (__bf16*)base;
  float* Cp = (float*)((unsigned char*)base + plane_bytes);
  return split_gemm_into(M, N, K, A, B, C, planes, Cp, stream);
}

}  // extern "C"
