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
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

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

  // WITH_B = false: only the A planes (the B tile is generated in LDS by the caller)
  template <int S, bool WITH_B = true>
  __device__ __forceinline__ void issue(int q) const {
    const int64_t k0 = (int64_t)min(q, nk - 1) * kBK;
#pragma unroll
    for (int u = 0; u < kLA; ++u) {
      dma16(ah + aoff[u] + k0, t_ah<S>() + (wave * kLA + u) * 1024);
      dma16(al + aoff[u] + k0, t_al<S>() + (wave * kLA + u) * 1024);
    }
    if constexpr (WITH_B) {
#pragma unroll
      for (int u = 0; u < kLB; ++u) {
        dma16(bh + boff[u] + k0, t_bh<S>() + (wave * kLB + u) * 1024);
        dma16(bl + boff[u] + k0, t_bl<S>() + (wave * kLB + u) * 1024);
      }
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

  // Product PR on tile rows I0, I0 + 1 (8 MFMAs).
  template <int PR, int I0>
  __device__ __forceinline__ static void mma_rows(f32x16 (&acc)[4][4], const Frags& f) {
#pragma unroll
    for (int i = I0; i < I0 + 2; ++i)
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

// ---- the split GEMM with its B operand generated in LDS -------------------------------
// The rotation form's B rows are the reference terms R_d(j) = rel S(s) e^{2 pi i t},
// s = pi r_j dia_d / size, t = ((x_j - c) sx_d + (y_j - c) sy_d) / size, which depend on three
// numbers per frame and two per pixel.  k_gemm_split_gen computes each stage's B tile (256
// frames x 16 pixels, bf16 hi/lo) in the registers of the thread that owns the frame's row
// and writes it to the LDS buffers the A tile's loads do not use, in the layout the direct
// loads would give, instead of reading 8 bytes per frame and pixel from HBM.
//  * S is even in s, so rel S(s) = g(v), v = s^2 = r_j^2 (pi dia_d / size)^2 with r_j^2 an
//    exact integer: g is tabulated once per plan (k_sphere_table, FP64) as kGTab cubic
//    Hermite pieces on [0, kGVmax) (h = 0.5: interpolation error <= 1e-9, below the FP32
//    evaluation's ~3e-8 near S = 1/3), and the table sits in LDS: one ds_read_b128 and
//    three FMAs per element.  s < sqrt(kGVmax) = 22.6 covers diameters up to 14.4 px at the
//    outermost scored pixel (r < size/2) of any grid; a wave holding a frame beyond it
//    evaluates S directly (FP32 library sin/cos) where v leaves the table.
//  * t in FP32 (|t| < 8 revolutions), reduced to [0, 1), cos and sin by v_cos_f32 /
//    v_sin_f32 (their argument is in revolutions).
//  * Pixels are packed one dword each: bf16 (x - c) | bf16 (y - c) << 16 (exact: |x - c| <= 256),
//    read with scalar loads (uniform over the workgroup).  Padding pixels (j >= K) read 0:
//    their A columns are zero, so their finite B values add nothing; rows of frames past the
//    chunk give finite values in C rows that are never stored.
constexpr int kGTab = 1024;
constexpr double kGVmax = 512.;
__shared__ __attribute__((aligned(16))) float4 s_gtab[kGTab];

// DIRECT: the wave holds a frame whose pixels leave the table (a separate instantiation of
// the stage loop, so that the table path stays straight-line).
template <bool DIRECT>
struct GenB {
  const uint32_t* __restrict__ pix;  // this slice's pixels (16 per stage)
  const float* __restrict__ pr2;     // their squared radii (exact integers)
  float a, b, kdx;                   // the row's frame: sx/size, sy/size, (pi dia/size)^2 / h
  float kd2, rel;                    // (pi dia/size)^2 and rel, for the direct evaluation
  int row, nk;

  // The two elements of half hf of k-group kg (pixels 4kg + 2hf, +1: 8 bytes of each
  // plane of the row) in two parts, so that the table reads of one half are in flight
  // while the MFMAs of a region run: prep issues them, finish forms the elements.
  struct Prep {
    float f[2], x[2], r2[2];
    float4 c[2];
  };

  __device__ __forceinline__ static void coords(uint32_t w, float& xc, float& yc) {
    xc = __uint_as_float(w << 16);
    yc = __uint_as_float(w & 0xffff0000u);
  }

  __device__ __forceinline__ void prep(const float (&r2)[16], int kg, int hf, Prep& p) const {
    prep2(r2[4 * kg + 2 * hf], r2[4 * kg + 2 * hf + 1], p);
  }

  // the first half of stage q (its squared radii by two scalar loads), issued one stage ahead
  __device__ __forceinline__ void prep_first(int q, Prep& p) const {
    const int64_t o = (int64_t)min(q, nk - 1) * (kBK / 2);
    prep2(pr2[o], pr2[o + 1], p);
  }

  __device__ __forceinline__ void prep2(float r2a, float r2b, Prep& p) const {
    const float r2[2] = {r2a, r2b};
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const float x = r2[u] * kdx;  // table coordinate v / h (>= 0)
      // x < kGTab - 1/2 outside DIRECT waves: the piece is x's integer part, f its fraction
      const int i = DIRECT ? min((int)x, kGTab - 1) : (int)x;
      p.f[u] = DIRECT ? x - (float)i : __builtin_amdgcn_fractf(x);
      p.c[u] = s_gtab[i];
      if constexpr (DIRECT) {
        p.x[u] = x;
        p.r2[u] = r2[u];
      }
    }
  }

  // g(v): the cubic piece, or (DIRECT, past the table) rel S(s) = rel (sin s - s cos s) / s^3
  // with s = sqrt(v) > 22.6 (no cancellation), by the hardware sine and cosine (argument in
  // revolutions)
  template <int S>
  __device__ __forceinline__ void finish(const uint32_t (&w)[16], int kg, int hf,
                                         const Prep& p) const {
    u32x2 hv, lv;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      float xc, yc;
      coords(w[4 * kg + 2 * hf + u], xc, yc);
      const float4 c = p.c[u];
      const float f = p.f[u];
      float g = fmaf(fmaf(fmaf(c.w, f, c.z), f, c.y), f, c.x);
      if constexpr (DIRECT) {
        const float sv = sqrtf(p.r2[u] * kd2);
        const float t = sv * 0.15915494309189535f;  // s / (2 pi)
        const float gd = rel * (__builtin_amdgcn_sinf(t) - sv * __builtin_amdgcn_cosf(t)) /
                         (sv * sv * sv);
        g = p.x[u] > (float)kGTab ? gd : g;
      }
      const float t = __builtin_amdgcn_fractf(fmaf(xc, a, yc * b));
      const f32x2 v = {g * __builtin_amdgcn_cosf(t), g * __builtin_amdgcn_sinf(t)};
      // bf16 hi = (re, im) rounded, lo = the rounded remainders (one conversion per pair)
      const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
      const f32x2 r = {v[0] - __uint_as_float(h << 16), v[1] - __uint_as_float(h & 0xffff0000u)};
      hv[u] = h;
      lv[u] = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
    }
    // the row's swizzle as an opaque value: the 16 write offsets (2 buffers x 8 halves) are
    // then formed at each write instead of held in registers across the loop
    int sw = swz(row);
    asm volatile("" : "+v"(sw));
    const int o = row * kRowB + ((kg ^ sw) << 4) + 8 * hf;
    *reinterpret_cast<u32x2*>(t_bh<S>() + o) = hv;
    *reinterpret_cast<u32x2*>(t_bl<S>() + o) = lv;
  }

  // The 16 pixel words and squared radii of stage q (uniform: scalar loads, read once per
  // stage).  Stages past the slice (q >= nk) recompute the last one into a buffer no stage
  // reads.
  __device__ __forceinline__ void words(int q, uint32_t (&w)[16], float (&r2)[16]) const {
    const int64_t o = (int64_t)min(q, nk - 1) * (kBK / 2);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      w[u] = pix[o + u];
      r2[u] = pr2[o + u];
    }
  }

  // k-group kg (pixels 4kg .. 4kg+3 of a stage: 8 k values, re/im interleaved) of this
  // thread's row goes to slot kg ^ swz(row) of the row's 64 bytes in the stage's B buffers.
  template <int S>
  __device__ __forceinline__ void half(const uint32_t (&w)[16], const float (&r2)[16], int kg,
                                       int hf) const {
    Prep p;
    prep(r2, kg, hf, p);
    finish<S>(w, kg, hf, p);
  }
};

// The A tile's direct-to-LDS loads with a wave-uniform base and one 32-bit byte offset per
// lane and slot (no 64-bit address registers).
struct AStream {
  const char *ah, *al;   // this tile's first row at the slice's first k
  uint32_t off[kLA];     // byte offsets of this lane's 16-byte slots within the tile
  int nk, wave;

  template <int S>
  __device__ __forceinline__ void issue(int q) const {
    const int64_t kb = (int64_t)min(q, nk - 1) * kBK * 2;
#pragma unroll
    for (int u = 0; u < kLA; ++u) {
      dma16(ah + kb + off[u], t_ah<S>() + (wave * kLA + u) * 1024);
      dma16(al + kb + off[u], t_al<S>() + (wave * kLA + u) * 1024);
    }
  }
};

// One stage of the wave's tile (the products of SplitGemm::compute) with the generation of
// the next stage's B tile (buffer SB).  The generation runs only while one fragment set is
// live: regions of 8 MFMAs with two generated elements each over the first two products of
// the first 16-k step and of the second; the second step's fragment reads are issued during
// the first step's third product, which carries no generation.
template <int S, int SB, bool DIRECT>
__device__ __forceinline__ void compute_gen(const SplitGemm& g, const GenB<DIRECT>& gb, int qn,
                                            f32x16 (&acc)[4][4], int wm, int wn,
                                            typename GenB<DIRECT>::Prep& pa) {
  uint32_t w[16];
  float r2[16];
  gb.words(qn, w, r2);
  Frags f0, f1;
  g.load<S>(f0, 0, wm, wn);
  asm volatile(""
               : "+v"(f0.ah[0]), "+v"(f0.ah[1]), "+v"(f0.ah[2]), "+v"(f0.ah[3]),
                 "+v"(f0.al[0]), "+v"(f0.al[1]), "+v"(f0.al[2]), "+v"(f0.al[3]),
                 "+v"(f0.bh[0]), "+v"(f0.bh[1]), "+v"(f0.bh[2]), "+v"(f0.bh[3]),
                 "+v"(f0.bl[0]), "+v"(f0.bl[1]), "+v"(f0.bl[2]), "+v"(f0.bl[3])
               :
               : "memory");
  // pa holds the first half of stage qn (its table reads issued during the previous stage)
  typename GenB<DIRECT>::Prep pb;
  // a region: 8 MFMAs, the table reads of the next half and the VALU work of this one
  // between them
#define RSP_GEN_REGION(PR, I0, F, KG, HF, PN, KGN, HFN, PC)                  \
  SplitGemm::mma_rows<PR, I0>(acc, F);                                      \
  if (KGN >= 0) gb.prep(r2, KGN, HFN, PN);                                  \
  else gb.prep_first(qn + 1, PN);                                           \
  gb.template finish<SB>(w, KG, HF, PC);                                    \
  for (int u_ = 0; u_ < 8; ++u_) {                                          \
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); /* one MFMA */       \
    __builtin_amdgcn_sched_group_barrier(0x002, 5, 0); /* five VALU */      \
  }                                                                         \
  __builtin_amdgcn_sched_barrier(0);
  RSP_GEN_REGION(0, 0, f0, 0, 0, pb, 0, 1, pa)
  RSP_GEN_REGION(0, 2, f0, 0, 1, pa, 1, 0, pb)
  RSP_GEN_REGION(1, 0, f0, 1, 0, pb, 1, 1, pa)
  RSP_GEN_REGION(1, 2, f0, 1, 1, pa, 2, 0, pb)
  g.load<S>(f1, 1, wm, wn);
  SplitGemm::mma<2>(acc, f0);
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // one LDS read
  }
  __builtin_amdgcn_sched_barrier(0);
  RSP_GEN_REGION(0, 0, f1, 2, 0, pb, 2, 1, pa)
  RSP_GEN_REGION(0, 2, f1, 2, 1, pa, 3, 0, pb)
  RSP_GEN_REGION(1, 0, f1, 3, 0, pb, 3, 1, pa)
  RSP_GEN_REGION(1, 2, f1, 3, 1, pa, -1, 0, pb)  // the next stage's first half into pa
#undef RSP_GEN_REGION
  SplitGemm::mma<2>(acc, f1);
}

// The stage loop of k_gemm_split_gen (accumulators zeroed here).
template <bool DIRECT>
__device__ __forceinline__ void gen_loop(const SplitGemm& g, const AStream& as,
                                         const GenB<DIRECT>& gb, f32x16 (&acc)[4][4], int wm,
                                         int wn) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;
  as.issue<0>(0);
  {
    uint32_t w[16];
    float r2[16];
    gb.words(0, w, r2);
#pragma unroll
    for (int kg = 0; kg < 4; ++kg) {
      gb.template half<0>(w, r2, kg, 0);
      gb.template half<0>(w, r2, kg, 1);
    }
  }
  typename GenB<DIRECT>::Prep pa;
  gb.prep_first(1, pa);
  for (int q = 0; q < g.nk; q += 2) {
    // vmcnt = 0: this wave's A loads of stage q have landed; lgkmcnt = 0: its B-tile writes
    // of stage q and its fragment reads of stage q-1 are done (the buffers of stage q+1
    // are rewritten after the barrier)
    __builtin_amdgcn_s_waitcnt(0x070);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    as.issue<1>(q + 1);
    compute_gen<0, 1>(g, gb, q + 1, acc, wm, wn, pa);
    __builtin_amdgcn_s_waitcnt(0x070);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    as.issue<0>(q + 2);
    if (q + 1 < g.nk) compute_gen<1, 0>(g, gb, q + 2, acc, wm, wn, pa);
  }
}

// X partial products of the K slices (as k_gemm_split) with B = the reference rows of frames
// d0 .. d0 + N - 1 (sx, sy, dia indexed by frame), generated per stage.  gtab: kGTab cubic
// pieces of g and, at gtab[kGTab], {1 / h, v_max, 0, 0} (k_sphere_table); r2max: the largest
// squared radius of the scored pixels.
__global__ __launch_bounds__(256, 1) void k_gemm_split_gen(
    const __bf16* __restrict__ ah, const __bf16* __restrict__ al,
    const uint32_t* __restrict__ pix, const float* __restrict__ pr2,
    const float4* __restrict__ gtab, float r2max, float rel,
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ dia, int64_t d0, int size, float* __restrict__ C, int64_t N,
    int64_t Kp, int64_t ldc, int nM, int nN, int kst, int64_t slice_stride) {
  const int T = nM * nN;
  const int T8 = (T + 7) / 8;
  const int task = (int)(blockIdx.x & 7) * T8 + (int)(blockIdx.x >> 3);
  if (task >= T) return;
  const int mt = task % nM, nt = task / nM;
  const int64_t m0 = (int64_t)mt * kTM, n0 = (int64_t)nt * kTN;
  const int z = blockIdx.y;
  const int nkt = (int)(Kp / kBK);
  const int64_t kz = (int64_t)z * kst * kBK;
  C += (int64_t)z * slice_stride;
  SplitGemm g;   // fragment reads and products (its load addresses are unused)
  g.lane = threadIdx.x & 63;
  g.wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  g.nk = min(kst, nkt - z * kst);
  AStream as;
  as.ah = reinterpret_cast<const char*>(ah + m0 * Kp + kz);
  as.al = reinterpret_cast<const char*>(al + m0 * Kp + kz);
  as.nk = g.nk;
  as.wave = g.wave;
#pragma unroll
  for (int u = 0; u < kLA; ++u) {
    const int s = (g.wave * kLA + u) * 64 + g.lane;
    const int row = s >> 2, kcl = (s & 3) ^ swz(row);
    as.off[u] = (uint32_t)(((int64_t)row * Kp + kcl * 8) * 2);
  }
  // the sphere table into LDS
  for (int i = threadIdx.x; i < kGTab; i += 256) s_gtab[i] = gtab[i];
  const float inv_h = gtab[kGTab].x;
  GenB<false> gb;
  gb.pix = pix + kz / 2;
  gb.pr2 = pr2 + kz / 2;
  gb.row = threadIdx.x;
  gb.nk = g.nk;
  {
    const int64_t r = n0 + threadIdx.x;  // the row's frame within the chunk
    double fa = 0., fb = 0., fk = 0.;
    if (r < N) {
      const double kd = 3.14159265358979323846 * dia[d0 + r] / size;
      fa = sx[d0 + r] / size;
      fb = sy[d0 + r] / size;
      fk = kd * kd * inv_h;
    }
    gb.a = (float)fa;
    gb.b = (float)fb;
    gb.kdx = (float)fk;
    gb.kd2 = (float)(fk / inv_h);
    gb.rel = rel;
  }
  __syncthreads();
  const int wm = g.wave >> 1, wn = g.wave & 1;
  f32x16 acc[4][4];
  // a frame whose outermost pixel comes within half a piece of the table's end sends its
  // wave to the direct path
  if (__any(r2max * gb.kdx >= (float)kGTab - 0.5f)) {
    const GenB<true> gd{gb.pix, gb.pr2, gb.a, gb.b, gb.kdx, gb.kd2, gb.rel, gb.row, gb.nk};
    gen_loop(g, as, gd, acc, wm, wn);
  } else {
    gen_loop(g, as, gb, acc, wm, wn);
  }
  __builtin_amdgcn_s_waitcnt(0xF70);   // the trailing (clamped) loads
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

// rel S(s) and dS/ds in FP64 (the s == 0 limit 1/3 by its series below s = 0.5)
__device__ __forceinline__ void sphere_g(double v, double& g, double& dg) {
  const double s = sqrt(v);
  if (s < 0.5) {
    // S = 1/3 - v/30 + v^2/840 - v^3/45360 + v^4/3991680 - v^5/518918400
    g = 1. / 3. + v * (-1. / 30. + v * (1. / 840. + v * (-1. / 45360. +
                 v * (1. / 3991680. - v / 518918400.))));
    dg = -1. / 30. + v * (2. / 840. + v * (-3. / 45360. + v * (4. / 3991680. -
         5. * v / 518918400.)));
  } else {
    double sn, cs;
    sincos(s, &sn, &cs);
    g = (sn - s * cs) / (s * s * s);
    // dS/ds = sin s / s^2 - 3 S / s;  dg/dv = dS/ds / (2 s)
    dg = (sn / (s * s) - 3. * g / s) / (2. * s);
  }
}

// The table of g(v) = rel S(sqrt v) for k_gemm_split_gen: h = kGVmax / kGTab, piece i on
// [i h, (i+1) h) as c0 + c1 f + c2 f^2 + c3 f^3, f = v/h - i (cubic Hermite from g and dg/dv
// at both ends, FP64, rounded to FP32); out[kGTab] = {1/h, kGVmax, 0, 0}.
__global__ __launch_bounds__(256) void k_sphere_table(double rel, float4* __restrict__ out) {
  const double h = kGVmax / kGTab;
  for (int i = threadIdx.x; i < kGTab; i += 256) {
    double g0, d0, g1, d1;
    sphere_g(i * h, g0, d0);
    sphere_g((i + 1) * h, g1, d1);
    g0 *= rel; g1 *= rel; d0 *= rel * h; d1 *= rel * h;
    out[i] = make_float4((float)g0, (float)d0, (float)(3. * (g1 - g0) - 2. * d0 - d1),
                         (float)(2. * (g0 - g1) + d0 + d1));
  }
  if (threadIdx.x == 0) out[kGTab] = make_float4((float)(1. / h), (float)kGVmax, 0.f, 0.f);
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

int xterm_gen_table_len() { return kGTab + 1; }

int xterm_sphere_table(double rel, float4* table, hipStream_t stream) {
  hipLaunchKernelGGL(k_sphere_table, dim3(1), dim3(256), 0, stream, rel, table);
  RSP_LAUNCH_CHECK("k_sphere_table");
  return kOk;
}

// Split-BF16 GEMM whose B rows are the reference terms of frames d0 .. d0 + N - 1, generated
// in the kernel (pixels packed by the caller, table from xterm_sphere_table).
int xterm_gemm_split_gen(const __bf16* ah, const __bf16* al, const uint32_t* pix,
                         const float* pr2, const float4* table, double r2max, double rel,
                         const double* sx,
                         const double* sy, const double* dia, int64_t d0, int size, float* C,
                         int64_t Mpad, int64_t N, int64_t Kp, int64_t ldc, int64_t slice_stride,
                         hipStream_t stream) {
  if (Mpad % kTM || Kp % kBK || ldc < Mpad || ldc % 4) return fail(kErrSize, "gemm padding");
  const int S = xterm_split_k(Kp);
  if (S > 1 && slice_stride < (N - 1) * ldc + Mpad) return fail(kErrSize, "gemm slice stride");
  if (N == 0 || Mpad == 0) return kOk;
  const int nM = (int)(Mpad / kTM), nN = (int)ceil_div(N, kTN);
  const int T8 = (nM * nN + 7) / 8;
  hipLaunchKernelGGL(k_gemm_split_gen, dim3(8 * T8, S), dim3(256), 0, stream, ah, al, pix,
                     pr2, table, (float)r2max, (float)rel, sx, sy, dia, d0, size, C, N, Kp, ldc, nM,
                     nN, kSliceStages, slice_stride);
  RSP_LAUNCH_CHECK("k_gemm_split_gen");
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
