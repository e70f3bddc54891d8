// Cross-term GEMM entry points shared by the rotation-form plan (xterm.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>

namespace rsp {

// Padding the split kernel needs: rows of A and of B to 256, K to 32.
int64_t xterm_pad_m(int64_t M);
int64_t xterm_pad_n(int64_t N);
int64_t xterm_pad_k(int64_t K);

// FP32 rows -> bf16 hi/lo planes (zero padded to rows_pad x Kp).
int xterm_split(const float* X, int64_t rows, int64_t K, int64_t ld, __bf16* hi, __bf16* lo,
                int64_t rows_pad, int64_t Kp, hipStream_t stream);

// FP64 rows -> bf16 hi/lo planes (each value rounded to FP32 first; zero padded).
int xterm_split64(const double* X, int64_t rows, int64_t K, int64_t ld, __bf16* hi,
                  __bf16* lo, int64_t rows_pad, int64_t Kp, hipStream_t stream);

// Number of K slices of the split GEMM for padded K = Kp (each slice accumulates 12288 K
// elements in FP32).
int xterm_split_k(int64_t Kp);

// Partial products sum_{k in slice z} A[r, k] B[d, k] on padded bf16 hi/lo planes (split
// products) at C + z * slice_stride + d * ldc + r, for z < xterm_split_k(Kp); the caller
// adds the slices (in FP64).
int xterm_gemm_split(const __bf16* ah, const __bf16* al, const __bf16* bh, const __bf16* bl,
                     float* C, int64_t Mpad, int64_t N, int64_t Kp, int64_t ldc,
                     int64_t slice_stride, hipStream_t stream);

// The rotation form's split GEMM with B = the reference rows R_d(j) generated in the kernel
// (B is never stored): frames d0 .. d0 + N - 1 of the arrays sx, sy, dia; pix[j] packs scored
// pixel j of the K order as bf16 (x - c) | bf16 (y - c) << 16 and pr2[j] its (x - c)^2 +
// (y - c)^2, with Kp / 2 entries (zero past the last pixel); r2max the largest of them; table
// from
// xterm_sphere_table (xterm_gen_table_len() float4, built once per rel).
int xterm_gen_table_len();
int xterm_sphere_table(double rel, float4* table, hipStream_t stream);
int xterm_gemm_split_gen(const __bf16* ah, const __bf16* al, const uint32_t* pix,
                         const float* pr2, const float4* table, double r2max, double rel,
                         const double* sx,
                         const double* sy, const double* dia, int64_t d0, int size, float* C,
                         int64_t Mpad, int64_t N, int64_t Kp, int64_t ldc, int64_t slice_stride,
                         hipStream_t stream);

// The same through rocBLAS in FP32 / FP64 (row-major A [M x lda], B [N x ldb]) on the
// caller's handle (set to `stream` for the call; a handle is never shared between plans).
int xterm_gemm_blas32(rocblas_handle h, const float* A, const float* B, float* C, int64_t M,
                      int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc,
                      hipStream_t stream);
int xterm_gemm_blas64(rocblas_handle h, const double* A, const double* B, double* C,
                      int64_t M, int64_t N, int64_t K, int64_t lda, int64_t ldb, int64_t ldc,
                      hipStream_t stream);

// General FP64 GEMM in rocBLAS's column-major convention:
//   C (m x n, ldc) = alpha * op(A) op(B) + beta * C, op = transpose when trans* is true.
int xterm_dgemm(rocblas_handle h, bool transA, bool transB, int64_t m, int64_t n, int64_t k,
                double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
                double beta, double* C, int64_t ldc, hipStream_t stream);

}  // namespace rsp
