// EMC E-step plan for gfx950: the fused replacement of emc.py:159-216.
//
// Reference path (emc.py:171-199, kernels.cu:44-137): for every state s and rotation r
// build the log of the rotated hologram and sum it over each frame's scored photons,
// then take the first-occurrence argmax over h = s*R + r.
//
// Device pipeline (one GPU, no state x rotation array ever reaches HBM):
//   1. k_holo_states   I_s = |O + S_s*ramp_s|^2 + bg for all states (FP64) + msums_s.
//   2. k_vscale_init   vscale (emc.py:181) and init_s = -(msums_s*vscale).
//   3. per batch of hypotheses:
//      k_tables        FP32 log-view tables over the scored annulus (kernels.cu:6-42 on
//                      I_s), plus the per-pixel maximum FP32 rounding error.
//      k_gather<P,NW>  one workgroup holds 64*NW frames whose scored photons live in
//                      VGPRs as LDS byte addresses, bank-scheduled on the host so that
//                      each wave-wide ds_read hits distinct banks; per hypothesis the
//                      table is copied to LDS and every lane gathers its photons,
//                      accumulating in FP64 and keeping a running top-3.
//      k_merge_top3    fold the per-segment top-3 into the running per-frame state.
//   4. k_resolve       per frame: a rigorous bound E_d on the FP32-table error decides
//                      whether the FP32 ranking is certain; otherwise the top-2 are
//                      re-evaluated in FP64, or (rare) the frame is re-searched over all
//                      hypotheses in FP64 (k_full_search).  maxll is always the FP64
//                      log-likelihood of the chosen hypothesis, summed in the reference
//                      order (ones then multi, kernels.cu:126-136).
#include <algorithm>
#include <bitset>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "refspi_amd.h"
#include "rsp_common.hpp"

namespace rsp {

constexpr int kNPMax = 4;        // table parts (LDS-sized pieces of the scored annulus)
// The scored annulus is cut into parts of kPartMax entries; a part buffer in LDS holds
// kChunks 1 KB chunks (kPStride floats: the entries, then >= 32 zero pad entries), and a
// workgroup keeps kNBuf part buffers (147 KB of the 160 KB LDS) fed by asynchronous
// direct-to-LDS loads two steps ahead of the gather.  The first two buffers lie within
// the 16-bit immediate offset of ds_read (at 0 and 50176 B), so the slot registers
// address them without per-slot address arithmetic; reads of the third (at 100352 B) add
// kRebase to the slot register and take the rest of its base as the immediate offset.
constexpr int kChunks = 49;
constexpr int kPStride = kChunks * 256;
constexpr int kPartMax = kPStride - 32;
constexpr int kNBuf = 3;
constexpr uint32_t kRebase = 2 * kPStride * 4 - 64512;  // third buffer's base - offset
// word 3 of a raw buffer resource on gfx9-family GPUs: 32-bit data format, no swizzle
// (offsets are plain byte offsets checked against the record count)
constexpr int kBufWord3 = 0x00020000;

struct Top3 {
  double v1, v2, v3;
  int h1, h2;
};

__device__ __forceinline__ bool better(double a, int ha, double b, int hb) {
  return a > b || (a == b && ha < hb);
}

__device__ __forceinline__ void top3_insert(Top3& t, double v, int h) {
  if (better(v, h, t.v1, t.h1)) {
    t.v3 = t.v2;
    t.v2 = t.v1;
    t.h2 = t.h1;
    t.v1 = v;
    t.h1 = h;
  } else if (better(v, h, t.v2, t.h2)) {
    t.v3 = t.v2;
    t.v2 = v;
    t.h2 = h;
  } else if (v > t.v3) {
    t.v3 = v;
  }
}

__device__ __forceinline__ Top3 top3_empty() {
  Top3 t;
  t.v1 = t.v2 = t.v3 = -INFINITY;
  t.h1 = t.h2 = 0x7fffffff;
  return t;
}

// --------------------------------------------------------------------------------
// 1-2: holograms, msums, vscale
// --------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_holo_states(
    const double* __restrict__ model_c, const double* __restrict__ sx,
    const double* __restrict__ sy, const double* __restrict__ dia,
    const double* __restrict__ bg, const int32_t* __restrict__ probmask, int size,
    double* __restrict__ I, double* __restrict__ msums) {
  __shared__ double red[256];
  const int s = blockIdx.x;
  const int npix = size * size;
  const double ssx = sx[s], ssy = sy[s], sdia = dia[s];
  double* Is = I + (int64_t)s * npix;
  double acc = 0.;
  for (int t = threadIdx.x; t < npix; t += 256) {
    const int x = t / size, y = t - x * size;
    double v = holo_pixel(model_c, x, y, size, ssx, ssy, sdia, 1.);
    v = v * 1.;
    v = v + bg[t];
    Is[t] = v;
    if (probmask[t] < 1) acc += v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) msums[s] = red[0];
}

__global__ __launch_bounds__(256) void k_vscale_init(const double* __restrict__ msums, int S,
                                                     double powder_sel_sum,
                                                     double* __restrict__ init,
                                                     double* __restrict__ vscale_out) {
  __shared__ double red[256];
  double acc = 0.;
  for (int s = threadIdx.x; s < S; s += 256) acc += msums[s];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double vscale = powder_sel_sum / red[0] * (double)S;
  if (threadIdx.x == 0) *vscale_out = vscale;
  for (int s = threadIdx.x; s < S; s += 256) init[s] = -(msums[s] * vscale);
}

// --------------------------------------------------------------------------------
// exact FP64 log-view value of detector pixel t for hypothesis (s, r)
// (kernels.cu:6-42 with log_flag = 1, scale = 1, applied to I_s which already holds bg)
// --------------------------------------------------------------------------------
__device__ __forceinline__ double logview_exact(const double* __restrict__ Is,
                                                const double* __restrict__ bg, int size,
                                                int t, double ac, double as) {
  const int x = t / size, y = t - x * size;
  double tx, ty, v;
  rot_source(x, y, size / 2, ac, as, tx, ty);
  if (!bilinear(Is, size, tx, ty, v)) return -1000.;
  v = v * 1.;
  v = v + bg[t];
  return log_view(v);
}

// --------------------------------------------------------------------------------
// The FP64 natural log of the tables: Cody-Waite argument reduction v = 2^e m with
// m in [sqrt(1/2), sqrt(2)) and ln 2 split into hi (32 significant bits: e * hi is exact)
// + lo, and log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716, by its series up to
// s^11 (truncation < 1.8e-11), s through the hardware reciprocal and one Newton step.
// Measured against an 80-bit log (with the reciprocal start 2^-23 off either way):
// |error| <= 1.8e-11 for v in [1e-20, 1e300].  About a fifth of the instructions of the
// library log, which the exact path keeps; kTabLogErr, added to every table entry's
// error bound, covers the difference of the two (by a factor of 50) and is itself far
// below the FP32 rounding of the entries (~5e-8 relative).
// --------------------------------------------------------------------------------
constexpr double kTabLogErr = 1e-9;
// coefficients 1/(2k+1) of the atanh series, highest first
__constant__ double kAtanhC[6] = {1. / 11., 1. / 9., 1. / 7., 1. / 5., 1. / 3., 1.};

__device__ __forceinline__ double log_table(double v) {
  int e;
  double m = frexp(v, &e);  // m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m *= 2.;
    --e;
  }
  // s = (m - 1)/(m + 1) through the hardware reciprocal and one Newton step (a start
  // 2^-23 off ends within 2^-45)
  const double dn = m + 1.;
  double rc = __builtin_amdgcn_rcp(dn);
  rc = fma(rc, fma(-dn, rc, 1.), rc);
  const double s = (m - 1.) * rc;
  const double z = s * s;
  // Horner steps as fma(p, z, c) with the c in scalar registers: one instruction each
  double p = kAtanhC[0];
#pragma unroll
  for (int k = 1; k < 6; ++k) p = fma(p, z, kAtanhC[k]);
  const double ln2_hi = 0x1.62e42feep-1, ln2_lo = 0x1.a39ef35793c76p-33;
  return fma((double)e, ln2_hi, fma((double)e, ln2_lo, 2. * s * p));
}

// logview_exact with log_table, for detector pixel (x, y) with background bgt: the same
// sample v (bit for bit), its log within kTabLogErr of the exact path's
__device__ __forceinline__ double logview_table(const double* __restrict__ Is, double bgt,
                                                int size, int x, int y, double ac,
                                                double as) {
  double tx, ty, v;
  rot_source(x, y, size / 2, ac, as, tx, ty);
  if (!bilinear(Is, size, tx, ty, v)) return -1000.;
  v = v * 1.;
  v = v + bgt;
  return v < 1.e-20 ? -1000. : log_table(v);
}

// One hypothesis of the table loop, split so that its image reads can be issued a
// hypothesis ahead of their use: tab_fetch computes the rotated source position and loads
// the bilinear cell (from the image's first cell when the position is out of bounds, so
// that every read stays in bounds and no branch is taken); tab_value forms the sample of
// bilinear() with the same operations in the same order, adds the background and takes
// log_table.
struct TabSample {
  double p00, p01, p10, p11, fx, fy;
  bool in;
};

__device__ __forceinline__ void tab_fetch(const double* __restrict__ Is, int size, int x, int y,
                                          double ac, double as, TabSample& q) {
  double tx, ty;
  rot_source(x, y, size / 2, ac, as, tx, ty);
  const double fxl = floor(tx), fyl = floor(ty);
  const int ix = (int)fxl, iy = (int)fyl;
  q.in = !(ix < 0 || ix > size - 2 || iy < 0 || iy > size - 2);
  q.fx = tx - ix;
  q.fy = ty - iy;
  const double* p = Is + (q.in ? (int64_t)ix * size + iy : 0);
  q.p00 = p[0];
  q.p01 = p[1];
  q.p10 = p[size];
  q.p11 = p[size + 1];
}

__device__ __forceinline__ double tab_value(const TabSample& q, double bgt) {
  if (!q.in) return -1000.;
  const double cx = 1. - q.fx, cy = 1. - q.fy;
  double v = q.p00 * cx * cy + q.p10 * q.fx * cy + q.p01 * cx * q.fy + q.p11 * q.fx * q.fy;
  v = v * 1.;
  v = v + bgt;
  return v < 1.e-20 ? -1000. : log_table(v);
}

// --------------------------------------------------------------------------------
// 3a: FP32 tables for hypotheses [hb, hb+nb)
// tables[(h-hb)*TS + part*PSTRIDE + (i - part*kPartMax)] = fl32(logview)
// --------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tables(
    const double* __restrict__ I, const double* __restrict__ bg,
    const int32_t* __restrict__ scored_t, int n_scored, int size, int R,
    const double* __restrict__ rcos, const double* __restrict__ rsin, int hb, int nb,
    int HC, int PSTRIDE, int TS, float* __restrict__ tables,
    unsigned* __restrict__ errmax, unsigned* __restrict__ absmax) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_scored) return;
  const int t = scored_t[i];
  const int npix = size * size;
  const int part = i / kPartMax;
  const int64_t col = (int64_t)part * PSTRIDE + (i - part * kPartMax);
  const int h0 = hb + blockIdx.y * HC;
  const int h1 = min(hb + nb, h0 + HC);  // > h0: the grid covers [hb, hb + nb) in HC steps
  const int x = t / size, y = t - x * size;
  const double bgt = bg[t];
  // Software pipeline along h: the cell of hypothesis h+1 is loaded before the log of h,
  // the cos/sin of h+2 before the cell of h+1.  (s1, r1): the (state, rotation) of h+1;
  // r < R always, so the cos/sin loads past the batch stay in bounds; no cell is loaded
  // past the batch's last hypothesis (its state may not exist).
  // Two samples in turn (q0, q1: no copies between iterations, so the wait for one
  // hypothesis's cell leaves the next one's reads in flight).
  int s1 = h0 / R, r1 = h0 - s1 * R;
  TabSample q0, q1;
  tab_fetch(I + (int64_t)s1 * npix, size, x, y, rcos[r1], rsin[r1], q0);
  if (++r1 == R) {
    r1 = 0;
    ++s1;
  }
  double ac1 = rcos[r1], as1 = rsin[r1];
  double emax_d = 0.;
  float amax = 0.f;
  auto step = [&](int h, const TabSample& cur, TabSample& nxt) {
    int r2 = r1 + 1, s2 = s1;
    if (r2 == R) {
      r2 = 0;
      ++s2;
    }
    const double ac2 = rcos[r2], as2 = rsin[r2];
    if (h + 1 < h1) tab_fetch(I + (int64_t)s1 * npix, size, x, y, ac1, as1, nxt);
    const double lv = tab_value(cur, bgt);
    const float f = (float)lv;
    tables[(int64_t)(h - hb) * TS + col] = f;
    emax_d = fmax(emax_d, fabs((double)f - lv));
    amax = fmaxf(amax, fabsf(f));
    s1 = s2;
    r1 = r2;
    ac1 = ac2;
    as1 = as2;
  };
  for (int h = h0; h < h1; h += 2) {
    step(h, q0, q1);
    if (h + 1 < h1) step(h + 1, q1, q0);
  }
  // the error against the exact path's log-view, rounded up to float: the FP32 rounding
  // plus the two logs' difference.  x -> roundup_f32(x + kTabLogErr) is monotone, so the
  // largest rounded bound is the rounded bound of the largest rounding error.
  const double d = emax_d + kTabLogErr;
  float emax = (float)d;
  if ((double)emax < d) emax = nextafterf(emax, INFINITY);
  // max of non-negative floats == max of their bit patterns: order independent
  atomicMax(errmax + i, __float_as_uint(emax));
  atomicMax(absmax + i, __float_as_uint(amax));
}

// Straight-line walk over the slot registers [K0, K1) in blocks of 8: every a[k] is a
// statically named VGPR and, with no branches between blocks, the LDS reads of one block
// overlap the arithmetic of the previous one.  Each block (8 slots, or 4 at the end of a
// part of 8m+4 slots) is summed pairwise in FP32; pairs of blocks are added in FP32 (a
// 16-term pairwise sum: error <= gamma_4 * sum|v|, gamma_4 = 4u/(1-4u), u = 2^-24,
// accounted for in the per-frame bound of k_resolve) and then to an FP64 accumulator.
// rb == kRebase for reads of the third buffer (one add per read: its base exceeds the
// 16-bit immediate offset), else 0 (unused)
template <int K, int N, int P, bool REB>
__device__ __forceinline__ void load_block(const char* lds, const uint32_t (&a)[P],
                                           float (&v)[8], uint32_t rb) {
  if constexpr (REB) {
    const char* base = lds - kRebase;  // within the immediate offset
#pragma unroll
    for (int u = 0; u < N; ++u) v[u] = *reinterpret_cast<const float*>(base + (a[K + u] + rb));
  } else {
#pragma unroll
    for (int u = 0; u < N; ++u) v[u] = *reinterpret_cast<const float*>(lds + a[K + u]);
  }
}

// Holds the block's values until here: an empty asm that reads and rewrites them, placed
// after the next block's reads are issued, so that the wait for this block's reads comes
// after that issue (lgkmcnt(N), the next block still in flight) and the compiler cannot
// start this block's sums -- and so wait for it -- before the next block is issued.
// HOLD = false (slot classes near the register limit, where holding both blocks' values
// would spill): the compiler places the waits, per value.
template <int N, bool HOLD>
__device__ __forceinline__ void hold_block(float (&w)[8]) {
  if constexpr (!HOLD)
    return;
  else if constexpr (N == 8)
    asm volatile("" : "+v"(w[0]), "+v"(w[1]), "+v"(w[2]), "+v"(w[3]), "+v"(w[4]), "+v"(w[5]),
                 "+v"(w[6]), "+v"(w[7])::"memory");
  else
    asm volatile("" : "+v"(w[0]), "+v"(w[1]), "+v"(w[2]), "+v"(w[3])::"memory");
}

// slots in the block starting at K of the range ending at K1 (8, or the 4 of a tail)
template <int K, int K1>
constexpr int block_len() {
  return K1 - K >= 8 ? 8 : K1 - K;
}

// Software pipeline: the LDS reads of block K+8 are issued before the sums of block K, so
// every wave keeps 8 reads in flight while it adds instead of waiting for all of a part's
// reads at once.  The empty asm statements pin the order: each block's sum is complete
// before the next block's reads are issued (memory clobber), and cannot sink below them.
// PEND: the FP32 sum of the previous block is waiting in `pend` for its pair.
template <int K, int K1, int P, int B, bool REB, bool PEND, bool HOLD>
__device__ __forceinline__ void gather_pipe(const char* lds, const uint32_t (&a)[P],
                                            float (&v)[2][8], double (&acc)[4], uint32_t rb,
                                            float pend) {
  if constexpr (K < K1) {
    constexpr int N = block_len<K, K1>();
    static_assert(N == 8 || N == 4, "parts hold a multiple of 4 slots");
    if constexpr (K + N < K1) load_block<K + N, block_len<K + N, K1>(), P, REB>(lds, a, v[B ^ 1], rb);
    hold_block<N, HOLD>(v[B]);
    const float* w = v[B];
    float s8;
    if constexpr (N == 8)
      s8 = ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
    else
      s8 = (w[0] + w[1]) + (w[2] + w[3]);
    asm volatile("" : "+v"(s8) : : "memory");
    if constexpr (PEND) {
      acc[(K >> 4) & 3] += (double)(pend + s8);
      gather_pipe<K + N, K1, P, B ^ 1, REB, false, HOLD>(lds, a, v, acc, rb, 0.f);
    } else if constexpr (K + N >= K1) {
      acc[(K >> 4) & 3] += (double)s8;  // last block of the part, unpaired
    } else {
      gather_pipe<K + N, K1, P, B ^ 1, REB, true, HOLD>(lds, a, v, acc, rb, s8);
    }
  }
}

template <int K0, int K1, int P, bool REB, bool HOLD>
__device__ __forceinline__ void gather_range(const char* lds, const uint32_t (&a)[P],
                                             double (&acc)[4]) {
  if constexpr (K0 < K1) {
    // an opaque copy of kRebase made at every step: the adds that use it cannot be hoisted
    // out of the hypothesis loop (which would hold one address register per slot)
    uint32_t rb = 0;
    if constexpr (REB) asm volatile("" : "=s"(rb) : "0"(kRebase));
    float v[2][8];
    load_block<K0, block_len<K0, K1>(), P, REB>(lds, a, v[0], rb);
    gather_pipe<K0, K1, P, 0, REB, false, HOLD>(lds, a, v, acc, rb, 0.f);
  }
}

// Two frames per lane: each slot register packs the LDS byte addresses of frame 0 (low
// 16 bits) and frame 1 (high 16 bits).  One step walks the virtual slots k = 0 .. 2S-1 of
// part PP: frame k / S, register PP*S + k % S, with the same software pipeline as above
// (one extraction per read); blocks do not straddle the frames, a frame's last block is
// added unpaired, and frame f accumulates in acc[2f], acc[2f + 1].
template <int S, int K>
constexpr int block_len2() {
  return S - K % S >= 8 ? 8 : S - K % S;
}

// ext = {0xffff, 16} as opaque per-step values: the extractions cannot be hoisted out of
// the hypothesis loop (which would hold two address registers per slot register)
template <int S, int PP, int K, int N, int P, bool REB>
__device__ __forceinline__ void load_block2(const char* lds, const uint32_t (&a)[P],
                                            float (&v)[8], uint32_t rb, const uint32_t (&ext)[2]) {
  constexpr int f = K / S, r0 = PP * S + K % S;
#pragma unroll
  for (int u = 0; u < N; ++u) {
    if constexpr (REB) {
      // third buffer: the address with bit 16 set (v_and_or_b32 / v_alignbit_b32: one
      // operation, like the other buffers' extraction), and its base - 65536 as the
      // immediate offset
      const uint32_t x = f ? __builtin_amdgcn_alignbit(rb, a[r0 + u], 16u)
                           : ((a[r0 + u] & ext[0]) | (rb << 16));
      v[u] = *reinterpret_cast<const float*>(lds - 65536 + x);
    } else {
      const uint32_t x = f ? (a[r0 + u] >> ext[1]) : (a[r0 + u] & ext[0]);
      v[u] = *reinterpret_cast<const float*>(lds + x);
    }
  }
}

template <int S, int PP, int K, int P, int B, bool REB, bool PEND, bool HOLD>
__device__ __forceinline__ void gather_pipe2(const char* lds, const uint32_t (&a)[P],
                                             float (&v)[2][8], double (&acc)[4], uint32_t rb,
                                             float pend, const uint32_t (&ext)[2]) {
  if constexpr (K < 2 * S) {
    constexpr int N = block_len2<S, K>();
    static_assert(N == 8 || N == 4, "parts hold a multiple of 4 slots");
    constexpr int f = K / S;
    constexpr bool last = (K + N) % S == 0;  // the frame's last block
    if constexpr (K + N < 2 * S)
      load_block2<S, PP, K + N, block_len2<S, K + N>(), P, REB>(lds, a, v[B ^ 1], rb, ext);
    hold_block<N, HOLD>(v[B]);
    const float* w = v[B];
    float s8;
    if constexpr (N == 8)
      s8 = ((w[0] + w[1]) + (w[2] + w[3])) + ((w[4] + w[5]) + (w[6] + w[7]));
    else
      s8 = (w[0] + w[1]) + (w[2] + w[3]);
    asm volatile("" : "+v"(s8) : : "memory");
    constexpr int ai = 2 * f + (((K % S) >> 4) & 1);
    if constexpr (PEND) {
      acc[ai] += (double)(pend + s8);
      gather_pipe2<S, PP, K + N, P, B ^ 1, REB, false, HOLD>(lds, a, v, acc, rb, 0.f, ext);
    } else if constexpr (last) {
      acc[ai] += (double)s8;
      gather_pipe2<S, PP, K + N, P, B ^ 1, REB, false, HOLD>(lds, a, v, acc, rb, 0.f, ext);
    } else {
      gather_pipe2<S, PP, K + N, P, B ^ 1, REB, true, HOLD>(lds, a, v, acc, rb, s8, ext);
    }
  }
}

template <int S, int PP, int P, bool REB, bool HOLD>
__device__ __forceinline__ void gather_range2(const char* lds, const uint32_t (&a)[P],
                                              double (&acc)[4]) {
  // rb = 1 (opaque): the third buffer's addresses carry bit 16
  uint32_t rb = 0;
  if constexpr (REB) asm volatile("" : "=s"(rb) : "0"(1u));
  // the mask in a VGPR: v_and_or_b32 (VOP3) takes one scalar operand on gfx9
  uint32_t ext[2];
  asm volatile("" : "=v"(ext[0]) : "0"(0xffffu));
  asm volatile("" : "=s"(ext[1]) : "0"(16u));
  float v[2][8];
  load_block2<S, PP, 0, block_len2<S, 0>(), P, REB>(lds, a, v[0], rb, ext);
  gather_pipe2<S, PP, 0, P, 0, REB, false, HOLD>(lds, a, v, acc, rb, 0.f, ext);
}

template <int K0, int P>
__device__ __forceinline__ void load_slots(uint32_t (&a)[P], const uint32_t* __restrict__ src) {
  if constexpr (K0 < P) {
    a[K0] = src[(int64_t)K0 * 64];
    load_slots<K0 + 1, P>(a, src);
  }
}

// --------------------------------------------------------------------------------
// 3b: the gather.  Frames are processed in "sorted order" j (by scored-photon count);
// wave gw holds frames j = 64*gw + lane.  slots[(gw*P + k)*64 + lane] is the LDS byte
// address (within the table part resident in LDS) of the frame's k-th photon entry; slots
// p*S .. p*S+S-1 belong to part p (S a multiple of 4).
// --------------------------------------------------------------------------------
// The three LDS part buffers of the gather, as separate objects: the compiler then knows
// which buffer each direct-to-LDS load writes and waits only for the buffer it reads.
__shared__ float g_buf0[kPStride], g_buf1[kPStride], g_buf2[kPStride];

template <int B>
__device__ __forceinline__ float* part_buf() {
  if constexpr (B == 0) return g_buf0;
  else if constexpr (B == 1) return g_buf1;
  else return g_buf2;
}

// The gather walks steps q = (hypothesis, part) in groups of 3 hypotheses so that the
// part index and the LDS buffer (q % 3) of every step are compile-time constants.  Step
// q: wait until its part is in LDS (this wave's loads via vmcnt, every wave's via the
// barrier), start the loads of step q+2 into the buffer freed by step q-1, then gather the
// part's S slots.  Template: S slots per part and lane, NP parts, NW waves, L loads/wave.
template <int S, int NP, int NW, int L, int X, int F>
struct GatherStep {
  static constexpr int P = S * NP;
  static_assert(F == 1 || (F == 2 && X == 1), "two frames per lane: whole frames only");
  // VGPRs per lane at NW / 4 waves per SIMD (512 per SIMD lane, granules of 8), and whether
  // the slot registers leave room to hold a block's values past the next block's issue
  // (about 40 working registers with one frame per lane, 56 with two, held values
  // included; not for frames split over lanes, whose shuffles need more; measured on the
  // instantiations: beyond this the held values spill)
  static constexpr int kVgprBudget = (512 / (NW / 4)) & ~7;
  static constexpr bool kHold = X == 1 && P + (F == 2 ? 56 : 40) <= kVgprBudget;
  const float* tables;
  int TS, hb, h0, h1, nq, R, wid, lane;
  double scale[F];  // frame f of the lane (F = 2: low / high halves of the slot registers)
  // this lane's byte offset within each 1 KB chunk for parts 0 and 1 (the tile's table
  // layout: 16 * lane for the identity)
  unsigned off0, off1;

  // Loads of step q into dst: chunks 0 .. kChunks-2 as 16-byte-per-lane loads, L - 1 per
  // wave (chunk c from wave c % NW), and the last chunk as 4-byte-per-lane loads, one per
  // wave: wave w loads quarter w % 4 of it (waves 4.. repeat a quarter with the same bytes,
  // 256 B each instead of a whole 1 KB chunk).  Every wave issues exactly L loads on every
  // path -- steps past the end re-load the first part (in bounds) into a buffer that is
  // never read for a real hypothesis -- so the wait counts are the same everywhere and
  // the compiler adds no waits.
  // PP: the part of step q (q % NP) as a constant, which selects the lane's offset; a step
  // past the end loads part 0 with it, in bounds, into a buffer no real step reads.
  static_assert((L - 1) * NW == kChunks - 1, "chunks split as (L - 1) per wave + 1");
  template <int PP>
  __device__ __forceinline__ void issue(int q, float* dst) const {
    if (q >= nq) q = 0;
    const int h = h0 + q / NP, p = q - (q / NP) * NP;
    const float* src = tables + (int64_t)(h - hb) * TS + (int64_t)p * kPStride;
    // addresses as a wave-uniform base + one 32-bit lane offset (scalar-base loads: no
    // per-load address registers)
    // Buffer loads (MUBUF) rather than global ones: the compiler counts a global (FLAT)
    // load into LDS as an LDS access as well, and while one is pending it waits for all LDS
    // reads at every wait (lgkmcnt(0)); a buffer load is counted by vmcnt alone.  The
    // resource spans the part exactly (50,176 B: every load below stays inside it).
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(src), (short)0, kPStride * 4, kBufWord3);
    const unsigned l16 = NP > 2 ? (unsigned)lane * 16u : (PP == 0 ? off0 : off1),
                   l4 = (unsigned)lane * 4u;
#pragma unroll
    for (int u = 0; u < L - 1; ++u) {
      const int c = wid + u * NW;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rs, (__attribute__((address_space(3))) void*)(dst + c * 256), 16, l16, c * 1024, 0,
          0);
    }
    const int qo = (kChunks - 1) * 256 + (wid & 3) * 64;
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        rs, (__attribute__((address_space(3))) void*)(dst + qo), 4, l4, qo * 4, 0, 0);
  }

  // static step QS (0 .. 3*NP-1) of group g
  template <int QS>
  __device__ __forceinline__ void step(int g, const uint32_t (&a)[P], const double (&ini)[3],
                                       double (&acc)[4], Top3 (&top)[F]) const {
    constexpr int PP = QS % NP;
    const int q = g * 3 * NP + QS;
    const int h = h0 + g * 3 + QS / NP;
    // s_waitcnt through the builtin so that the compiler's own wait bookkeeping sees it
    // (gfx9 encoding: vmcnt[3:0], expcnt[6:4] = 7 (no wait), lgkmcnt[11:8])
    static_assert(L < 16, "vmcnt immediate");
    // vmcnt = L: this wave's loads of step q have landed (the L loads of step q+1, issued
    // after them, may still be in flight: vector memory completes in order);
    // lgkmcnt = 0: this wave's LDS reads of step q-1 have completed -- the loads issued
    // after the barrier overwrite the buffer step q-1 read, so no read of it may still be
    // in flight in any wave
    __builtin_amdgcn_s_waitcnt(0x070 | L);
    // a bare barrier: a fenced __syncthreads would also wait for the loads of step q+1
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    issue<(QS + 2) % NP>(q + 2, part_buf<(QS + 2) % 3>());
    if constexpr (F == 2) {
      if constexpr (PP == 0) {
        const double in = ini[QS / NP];
        acc[0] = in * scale[0];
        acc[2] = in * scale[1];
        acc[1] = acc[3] = 0.;
      }
      gather_range2<S, PP, P, QS % 3 == 2, kHold>(reinterpret_cast<const char*>(part_buf<QS % 3>()),
                                           a, acc);
      if constexpr (PP == NP - 1) {
        const double t0 = acc[0] + acc[1], t1 = acc[2] + acc[3];
        if (h < h1 && t0 >= top[0].v3) top3_insert(top[0], t0, h);
        if (h < h1 && t1 >= top[1].v3) top3_insert(top[1], t1, h);
      }
      return;
    }
    if constexpr (PP == 0) {
      acc[0] = ini[QS / NP] * scale[0];
      acc[1] = acc[2] = acc[3] = 0.;
    }
    gather_range<PP * S, PP * S + S, P, QS % 3 == 2, kHold>(
        reinterpret_cast<const char*>(part_buf<QS % 3>()), a, acc);
    if constexpr (PP == NP - 1) {
      double tot = (acc[0] + acc[1]) + (acc[2] + acc[3]);
      // a frame split over X lanes: sum the X partial sums (only its first lane carries
      // the init term; every lane of the group then holds the same total)
#pragma unroll
      for (int off = 1; off < X; off <<= 1) tot += __shfl_xor(tot, off);
      // below the third value nothing changes (v3 <= v2 <= v1): most hypotheses stop here,
      // and a wave with no lane above its v3 skips the insertion (not for 96 slots: the
      // branch's registers would push slot addresses to scratch)
      if constexpr (P < 96) {
        if (h < h1 && tot >= top[0].v3) top3_insert(top[0], tot, h);
      } else {
        if (h < h1) top3_insert(top[0], tot, h);
      }
    }
  }

  template <int QS>
  __device__ __forceinline__ void group_steps(int g, const uint32_t (&a)[P],
                                              const double (&ini)[3], double (&acc)[4],
                                              Top3 (&top)[F]) const {
    if constexpr (QS < 3 * NP) {
      step<QS>(g, a, ini, acc, top);
      group_steps<QS + 1>(g, a, ini, acc, top);
    }
  }

  // The group's three init terms (init[s] of its hypotheses, wave-uniform) are read at its
  // start and waited for at once (an empty asm reads them; they stay in SGPRs): a scalar
  // load shares lgkmcnt with the LDS reads and may complete out of order, so while one is
  // pending every wait for the walk's reads drains all of them.
  __device__ __forceinline__ void group(int g, const uint32_t (&a)[P],
                                        const double* __restrict__ init, double (&acc)[4],
                                        Top3 (&top)[F]) const {
    double ini[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      ini[t] = init[min(h0 + g * 3 + t, h1 - 1) / R];
      asm volatile("" : "+s"(ini[t]));
    }
    group_steps<0>(g, a, ini, acc, top);
  }
};

template <int S, int NP, int NW, int L, int X, int F>
__global__ __launch_bounds__(64 * NW) void k_gather(
    const float* __restrict__ tables, int TS, int hb, int nb, int nseg, int n_tiles, int gw0,
    const uint32_t* __restrict__ slots, const double* __restrict__ scale_sorted,
    const double* __restrict__ init, int R, Top3* __restrict__ partials, int Npad,
    const uint32_t* __restrict__ lay) {
  static_assert(L * NW >= kChunks, "L loads per wave must cover the part");
  constexpr int P = S * NP;
  const int b = blockIdx.x;
  int seg, tile;
  if ((nseg & 7) == 0) {  // keep one segment's workgroups on one XCD (round-robin dispatch)
    const int q = b >> 3;
    seg = (b & 7) + 8 * (q / n_tiles);
    tile = q % n_tiles;
  } else {
    seg = b / n_tiles;
    tile = b % n_tiles;
  }
  // the wave index as a scalar: the table loads' LDS destinations and chunk offsets are then
  // scalar too (no address registers per load)
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int lw = tile * NW + wid;  // wave within this slot class
  // frames (global sorted order) of this lane: F per lane, the waves gw0 + F*lw + f
  const int j = (gw0 + F * lw) * 64 + lane;
  const int h0 = hb + (int)((int64_t)seg * nb / nseg);
  const int h1 = hb + (int)((int64_t)(seg + 1) * nb / nseg);

  uint32_t a[P];
  load_slots<0, P>(a, slots + (int64_t)lw * P * 64 + lane);
  Top3 top[F];
#pragma unroll
  for (int f = 0; f < F; ++f) top[f] = top3_empty();
  // every wave issues exactly L direct-to-LDS loads per step; two steps are in flight
  // ahead of the gather
  // lay[p * Npad + j]: the lane's chunk offset for part p (parts 0 and 1; more parts keep
  // the identity), at the lane's first frame
  GatherStep<S, NP, NW, L, X, F> st{tables, TS, hb, h0, h1, (h1 - h0) * NP, R, wid, lane,
                                    {}, NP <= 2 ? lay[j] : 0u, NP == 2 ? lay[Npad + j] : 0u};
#pragma unroll
  for (int f = 0; f < F; ++f) st.scale[f] = scale_sorted[j + 64 * f];
  st.template issue<0>(0, g_buf0);
  st.template issue<1 % NP>(1, g_buf1);
  double acc[4];
  const int ngroups = (h1 - h0 + 2) / 3;
  for (int g = 0; g < ngroups; ++g) st.group(g, a, init, acc, top);
  // no direct-to-LDS write may still be in flight when the workgroup's LDS is released
  __builtin_amdgcn_s_waitcnt(0xF70);
#pragma unroll
  for (int f = 0; f < F; ++f) partials[(int64_t)seg * Npad + j + 64 * f] = top[f];
}

// per-class segment counts for the merge (waves [gw_end[c-1], gw_end[c]) use nseg[c])
constexpr int kMaxClass = 32;
struct SegMap {
  int n;
  int gw_end[kMaxClass];
  int nseg[kMaxClass];
};

// kMergeLanes lanes per frame: lane u of the group merges segments u, u + kMergeLanes, ...;
// the partial top-3 are then combined by shuffles.  The top two are selected by (value, lower index)
// and v3 is a maximum, so the result does not depend on the order of the merges.
__device__ __forceinline__ void top3_merge(Top3& t, const Top3& p) {
  top3_insert(t, p.v1, p.h1);
  top3_insert(t, p.v2, p.h2);
  if (p.v3 > t.v3) t.v3 = p.v3;
}

constexpr int kMergeLanes = 16;

__global__ __launch_bounds__(256) void k_merge_top3(const Top3* __restrict__ partials, SegMap sm,
                                                    int Npad, Top3* __restrict__ running,
                                                    int first) {
  const int tid = blockIdx.x * 256 + threadIdx.x;
  const int j = tid / kMergeLanes, u = tid % kMergeLanes;
  const bool live = j < Npad;   // no early return: the shuffles below need all lanes
  const int gw = j >> 6;  // the frame's wave of sorted frames
  int nseg = 0;
  for (int c = sm.n - 1; c >= 0; --c)
    if (gw < sm.gw_end[c]) nseg = sm.nseg[c];
  if (!live) nseg = 0;
  Top3 t = (first || u != 0 || !live) ? top3_empty() : running[j];
  // the partials of 8 of this lane's segments are loaded before any is merged
  constexpr int kU = 8;
  for (int g0 = u; g0 < nseg; g0 += kMergeLanes * kU) {
    Top3 p[kU];
#pragma unroll
    for (int v = 0; v < kU; ++v)
      if (g0 + kMergeLanes * v < nseg) p[v] = partials[(int64_t)(g0 + kMergeLanes * v) * Npad + j];
#pragma unroll
    for (int v = 0; v < kU; ++v)
      if (g0 + kMergeLanes * v < nseg) top3_merge(t, p[v]);
  }
#pragma unroll
  for (int off = 1; off < kMergeLanes; off <<= 1) {
    Top3 o;
    o.v1 = __shfl_xor(t.v1, off);
    o.v2 = __shfl_xor(t.v2, off);
    o.v3 = __shfl_xor(t.v3, off);
    o.h1 = __shfl_xor(t.h1, off);
    o.h2 = __shfl_xor(t.h2, off);
    top3_merge(t, o);
  }
  if (live && u == 0) running[j] = t;
}

// --------------------------------------------------------------------------------
// 4: exact resolution.  Per frame (frame id d), the scored photons in reference order:
// ex_off[d]..ex_off[d]+ex_n1[d] are 'ones', the rest up to ex_off[d+1] 'multi'.
// --------------------------------------------------------------------------------
struct ExactArgs {
  const double* I;
  const double* bg;
  const double* rcos;
  const double* rsin;
  const double* init;
  const double* scales;
  const int64_t* ex_off;
  const int32_t* ex_n1;
  const int32_t* ex_pix;
  const int32_t* ex_cnt;
  int size, R;
};

__device__ double exact_ll(const ExactArgs& A, int d, int h) {
  const int s = h / A.R, r = h - s * (A.R);
  const double* Is = A.I + (int64_t)s * A.size * A.size;
  const double ac = A.rcos[r], as = A.rsin[r];
  double acc = A.init[s] * A.scales[d];
  const int64_t e0 = A.ex_off[d], e1 = A.ex_off[d + 1], em = e0 + A.ex_n1[d];
  for (int64_t e = e0; e < em; ++e) acc += logview_exact(Is, A.bg, A.size, A.ex_pix[e], ac, as);
  for (int64_t e = em; e < e1; ++e)
    acc += logview_exact(Is, A.bg, A.size, A.ex_pix[e], ac, as) * A.ex_cnt[e];
  return acc;
}

// Per sorted frame j: decide, re-evaluate the top-2 in FP64 if needed, or queue a full
// FP64 search.  errmax holds per scored pixel the max |fl32(lv) - lv| over hypotheses.
__global__ __launch_bounds__(256) void k_resolve(ExactArgs A, const Top3* __restrict__ running,
                                                 const int32_t* __restrict__ perm, int Npad,
                                                 const int32_t* __restrict__ compact_of,
                                                 const unsigned* __restrict__ errmax,
                                                 const unsigned* __restrict__ absmax,
                                                 int exact_mode, int32_t* __restrict__ rmax,
                                                 double* __restrict__ maxll,
                                                 int32_t* __restrict__ full_list,
                                                 unsigned* __restrict__ counters) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= Npad) return;
  const int d = perm[j];
  if (d < 0) return;
  if (exact_mode) {
    full_list[atomicAdd(counters + 2, 1u)] = j;
    return;
  }
  const Top3 t = running[j];
  // bound on |LL_fp32tables - LL_fp64| for any hypothesis of this frame
  double E = 0.;
  for (int64_t e = A.ex_off[d]; e < A.ex_off[d + 1]; ++e) {
    const int c = (e < A.ex_off[d] + A.ex_n1[d]) ? 1 : A.ex_cnt[e];
    const int ci = compact_of[A.ex_pix[e]];
    // table rounding + the FP32 pairwise 16-sums of the gather (gamma_4 * |value|,
    // gamma_4 = 4u/(1-4u) = 2.384e-7 for u = 2^-24)
    E += (double)c * ((double)__uint_as_float(errmax[ci]) +
                      2.4e-7 * (double)__uint_as_float(absmax[ci]));
  }
  // plus the FP64 summation error of both paths: n*2^-53*sum|terms| with |log-view| <= 100
  // (values of -1000 mark out-of-range samples and never win); 1000x safety factor
  const double n = (double)(A.ex_off[d + 1] - A.ex_off[d]) + 1.;
  E = 2. * E + 1.e-13 * n * (fabs(t.v1) + 100. * n + 1.);
  int h = t.h1;
  double ll;
  if (E == 0. || t.v1 - t.v2 > 2. * E) {
    ll = exact_ll(A, d, h);
    atomicAdd(counters + 0, 1u);
  } else if (t.v1 - t.v3 > 2. * E) {
    const double l1 = exact_ll(A, d, t.h1), l2 = exact_ll(A, d, t.h2);
    if (better(l2, t.h2, l1, t.h1)) {
      h = t.h2;
      ll = l2;
    } else {
      ll = l1;
    }
    atomicAdd(counters + 1, 1u);
  } else {
    full_list[atomicAdd(counters + 2, 1u)] = j;
    return;
  }
  rmax[d] = h;
  maxll[d] = ll;
}

// Full FP64 search for the queued frames: block (bx, f) covers hypotheses
// [bx*256, bx*256+256) of frame full_list[f]; per-block first-occurrence argmax.
__global__ __launch_bounds__(256) void k_full_search(ExactArgs A, const int32_t* __restrict__ perm,
                                                     const int32_t* __restrict__ full_list,
                                                     int H, double* __restrict__ bval,
                                                     int32_t* __restrict__ bidx, int nbx) {
  __shared__ double sv[256];
  __shared__ int si[256];
  const int f = blockIdx.y;
  const int d = perm[full_list[f]];
  const int h = blockIdx.x * 256 + threadIdx.x;
  double v = -INFINITY;
  int hi = 0x7fffffff;
  if (h < H) {
    v = exact_ll(A, d, h);
    hi = h;
  }
  sv[threadIdx.x] = v;
  si[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      if (better(sv[threadIdx.x + w], si[threadIdx.x + w], sv[threadIdx.x], si[threadIdx.x])) {
        sv[threadIdx.x] = sv[threadIdx.x + w];
        si[threadIdx.x] = si[threadIdx.x + w];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bval[(int64_t)f * nbx + blockIdx.x] = sv[0];
    bidx[(int64_t)f * nbx + blockIdx.x] = si[0];
  }
}

__global__ __launch_bounds__(256) void k_full_reduce(const int32_t* __restrict__ perm,
                                                     const int32_t* __restrict__ full_list,
                                                     int nfull, const double* __restrict__ bval,
                                                     const int32_t* __restrict__ bidx, int nbx,
                                                     int32_t* __restrict__ rmax,
                                                     double* __restrict__ maxll) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  if (f >= nfull) return;
  double v = -INFINITY;
  int h = 0x7fffffff;
  for (int b = 0; b < nbx; ++b) {
    const double bv = bval[(int64_t)f * nbx + b];
    const int bi = bidx[(int64_t)f * nbx + b];
    if (better(bv, bi, v, h)) {
      v = bv;
      h = bi;
    }
  }
  const int d = perm[full_list[f]];
  rmax[d] = h;
  maxll[d] = v;
}

__global__ void k_scatter_scale(const double* __restrict__ scales, const int32_t* __restrict__ perm,
                                int Npad, double* __restrict__ scale_sorted) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= Npad) return;
  const int d = perm[j];
  scale_sorted[j] = d < 0 ? 0. : scales[d];
}

}  // namespace rsp

// ====================================================================================
// host side
// ====================================================================================
using namespace rsp;

struct rsp_emc {
  int device = 0;
  int size = 0, npix = 0, R = 0, S = 0;
  int64_t H = 0;
  std::vector<double> rcos, rsin, ssx, ssy, sdia;
  std::vector<int32_t> probmask, scored_t, compact_of;
  int n_scored = 0, NP = 1, PSTRIDE = 0, TS = 0, NB = 32;
  // device geometry
  double *d_rcos = nullptr, *d_rsin = nullptr, *d_sx = nullptr, *d_sy = nullptr,
         *d_dia = nullptr, *d_bg = nullptr, *d_I = nullptr, *d_init = nullptr,
         *d_vscale = nullptr;
  int32_t *d_probmask = nullptr, *d_scored_t = nullptr, *d_compact_of = nullptr;
  float* d_tables = nullptr;
  unsigned* d_errmax = nullptr;
  unsigned* d_absmax = nullptr;
  int64_t batch = 0;
  // frames
  int64_t N = 0, Npad = 0;
  int P = 0, NW = 0, n_waves = 0, n_tiles = 0;
  double powder_sel_sum = 0.;
  uint32_t* d_slots = nullptr;
  uint32_t* d_lay = nullptr;  // [part 0, 1][Npad]: per-lane chunk offsets of the table loads
  int32_t *d_meta = nullptr, *d_perm = nullptr;
  double* d_scale_sorted = nullptr;
  Top3 *d_partials = nullptr, *d_running = nullptr;
  int64_t partial_cap = 0;
  int64_t* d_ex_off = nullptr;
  int32_t *d_ex_n1 = nullptr, *d_ex_pix = nullptr, *d_ex_cnt = nullptr;
  int32_t* d_full_list = nullptr;
  unsigned* d_counters = nullptr;
  double* d_bval = nullptr;
  int32_t* d_bidx = nullptr;
  int64_t bcap = 0;
  // options / stats
  int exact_mode = 0;
  int64_t opt_batch = 0, opt_nseg = 0;
  int regroup_iters = 4000;  // local-search swaps per window of 8 lane groups (0: off)
  int layout = 1;            // per-tile LDS layout of the table parts (0: identity)
  int bcast = 1;             // broadcast-aware slot schedule (0: distinct banks per slot)
  int fpl = 2;               // frames per lane of whole-frame waves (1 or 2)
  int class_streams = 2;     // side streams the slot-class launches of a batch share (0: one)
  hipStream_t side[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  struct Class {
    int cfg, n_tiles, gw0;
    int64_t w_src, n_src, slot_off;
  };
  std::vector<Class> classes;
  // optional per-kernel timing with HIP events on the plan's stream
  int timing = 0;
  std::vector<hipEvent_t> ev;  // 4 per batch: tables begin/end, gather begin/end
  double t_tables_ms = 0., t_gather_ms = 0.;
  int64_t n_batches = 0, gather_bytes_alg = 0;
};

namespace {

template <typename T>
int dev_alloc(T** p, int64_t n) {
  if (*p) {
    hipFree(*p);
    *p = nullptr;
  }
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  return kOk;
}

template <typename T>
int dev_upload(T** p, const std::vector<T>& v) {
  int rc = dev_alloc(p, (int64_t)v.size());
  if (rc) return rc;
  if (!v.empty()) RSP_HIP(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return kOk;
}

void free_all(rsp_emc* e) {
  void* ptrs[] = {e->d_rcos, e->d_rsin, e->d_sx, e->d_sy, e->d_dia, e->d_bg, e->d_I,
                  e->d_init, e->d_vscale, e->d_probmask, e->d_scored_t, e->d_compact_of,
                  e->d_tables, e->d_errmax, e->d_absmax, e->d_slots, e->d_lay, e->d_meta,
                  e->d_perm,
                  e->d_scale_sorted, e->d_partials, e->d_running, e->d_ex_off, e->d_ex_n1,
                  e->d_ex_pix, e->d_ex_cnt, e->d_full_list, e->d_counters, e->d_bval,
                  e->d_bidx};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (hipEvent_t ev : e->ev) hipEventDestroy(ev);
  e->ev.clear();
  for (int k = 0; k < 4; ++k) {
    if (e->side[k]) hipStreamDestroy(e->side[k]);
    if (e->ev_join[k]) hipEventDestroy(e->ev_join[k]);
    e->side[k] = nullptr;
    e->ev_join[k] = nullptr;
  }
  if (e->ev_fork) hipEventDestroy(e->ev_fork);
  e->ev_fork = nullptr;
}

// Bank-aware slot assignment for one group of lanes of one wave and one table part
// (schedule_group).  entries[l] lists LDS entry indices (within the part) for lane l; the
// result places every entry of lane l in a distinct slot so that, within the group, no two
// lanes read the same bank in the same slot.  Returns the number of slots used: the larger
// of the longest lane and the busiest bank, which no assignment can beat.
// Regrouping of frames into lane groups before the waves are built.  The slot count of a
// group of G lanes (a half-wave for 32 banks) on one part is
// delta = max(largest per-part entry count of its frames, busiest bank of its entries),
// which schedule_group reaches.  With frames in descending order of their largest per-part
// entry count, the lane term is flat over a window of similar frames, but the bank term is a sum of G frames' bank histograms and sits well
// above its mean.  Within a window of ng groups this randomised local search swaps frames
// between groups -- always out of the group with the largest delta, the frame that loads
// its busiest (part, bank) cell the most -- and keeps a swap when the larger of the two
// groups' deltas falls, or stays and their sum falls.  The groups are then emitted in
// descending order of delta, so that the two halves of a wave are alike.  The seed is
// fixed: the same frames give the same plan.
struct FrameHist {
  int NP = 0, NB = 0;
  std::vector<uint16_t> h;    // [frame][part][bank] entry counts
  std::vector<int32_t> pc;    // [frame][part] entry counts
  const uint16_t* hist(int32_t d) const { return h.data() + (size_t)d * NP * NB; }
  const int32_t* parts(int32_t d) const { return pc.data() + (size_t)d * NP; }
};

void regroup_window(std::vector<int32_t>& order, int64_t f0, int64_t f1, int G,
                    const FrameHist& fh, int iters, uint64_t seed) {
  const int ng = (int)((f1 - f0) / G);
  if (ng < 2) return;
  const int K = fh.NP * fh.NB;
  std::vector<int32_t> gf(order.begin() + f0, order.begin() + f0 + (int64_t)ng * G);
  std::vector<int> load((size_t)ng * K, 0);
  for (int g = 0; g < ng; ++g)
    for (int i = 0; i < G; ++i) {
      const uint16_t* h = fh.hist(gf[(size_t)g * G + i]);
      for (int k = 0; k < K; ++k) load[(size_t)g * K + k] += h[k];
    }
  // delta of group g with frame slot i holding frame `x` instead (i < 0: as it is)
  auto lane_max = [&](int g, int i, int32_t x) {
    int m = 0;
    for (int j = 0; j < G; ++j) {
      const int32_t* p = fh.parts(j == i ? x : gf[(size_t)g * G + j]);
      for (int q = 0; q < fh.NP; ++q) m = std::max(m, p[q]);
    }
    return m;
  };
  std::vector<int> delta(ng);
  for (int g = 0; g < ng; ++g)
    delta[g] = std::max(*std::max_element(load.begin() + (size_t)g * K,
                                          load.begin() + (size_t)(g + 1) * K),
                        lane_max(g, -1, 0));
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  auto rnd = [&](uint64_t n) {  // splitmix64
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (z ^ (z >> 31)) % n;
  };
  std::vector<int> la(K), lb(K);
  for (int it = 0; it < iters; ++it) {
    const int a = (int)(std::max_element(delta.begin(), delta.end()) - delta.begin());
    const int* La = load.data() + (size_t)a * K;
    const int cell = (int)(std::max_element(La, La + K) - La);
    int ia = 0;
    for (int i = 1; i < G; ++i)
      if (fh.hist(gf[(size_t)a * G + i])[cell] > fh.hist(gf[(size_t)a * G + ia])[cell]) ia = i;
    int b = (int)rnd(ng - 1);
    b += b >= a;
    const int ib = (int)rnd(G);
    const int32_t x = gf[(size_t)a * G + ia], y = gf[(size_t)b * G + ib];
    const uint16_t *hx = fh.hist(x), *hy = fh.hist(y);
    const int* Lb = load.data() + (size_t)b * K;
    int ma = 0, mb = 0;
    for (int k = 0; k < K; ++k) {
      la[k] = La[k] - hx[k] + hy[k];
      lb[k] = Lb[k] - hy[k] + hx[k];
      ma = std::max(ma, la[k]);
      mb = std::max(mb, lb[k]);
    }
    const int na = std::max(ma, lane_max(a, ia, y)), nb = std::max(mb, lane_max(b, ib, x));
    const int hi = std::max(na, nb);
    if (hi < delta[a] || (hi == delta[a] && na + nb < delta[a] + delta[b])) {
      std::copy(la.begin(), la.end(), load.begin() + (size_t)a * K);
      std::copy(lb.begin(), lb.end(), load.begin() + (size_t)b * K);
      gf[(size_t)a * G + ia] = y;
      gf[(size_t)b * G + ib] = x;
      delta[a] = na;
      delta[b] = nb;
    }
  }
  std::vector<int> gord(ng);
  std::iota(gord.begin(), gord.end(), 0);
  std::stable_sort(gord.begin(), gord.end(), [&](int p, int q) { return delta[p] > delta[q]; });
  for (int r = 0; r < ng; ++r)
    std::copy(gf.begin() + (size_t)gord[r] * G, gf.begin() + (size_t)(gord[r] + 1) * G,
              order.begin() + f0 + (int64_t)r * G);
}

int schedule_group(const std::vector<std::vector<int>>& entries, int nbanks,
                   std::vector<std::vector<int>>& slot_of) {
  const int G = (int)entries.size();
  slot_of.assign(G, {});
  int max_len = 0;
  for (int l = 0; l < G; ++l) {
    slot_of[l].assign(entries[l].size(), -1);
    max_len = std::max(max_len, (int)entries[l].size());
  }
  if (nbanks == 0) {  // plain packing: entry r of every lane in slot r
    for (int l = 0; l < G; ++l)
      for (size_t r = 0; r < entries[l].size(); ++r) slot_of[l][r] = (int)r;
    return max_len;
  }
  // An edge colouring of the bipartite multigraph lanes x banks (one edge per entry) with
  // D = max degree colours (Koenig's theorem), colours being slots: entry (l, r) goes to
  // the lowest slot a free at lane l when a is also free at its bank b; else to the lowest
  // slot c free at b when c is free at l; else the path from b along edges of slots a, c
  // alternately has its two slots exchanged -- it cannot reach l, which has no a-edge, and
  // leaves a free at b -- and the entry takes a.
  std::vector<int> bdeg(nbanks, 0);
  for (int l = 0; l < G; ++l)
    for (int v : entries[l]) ++bdeg[v % nbanks];
  const int D = std::max(max_len, *std::max_element(bdeg.begin(), bdeg.end()));
  if (D == 0) return 0;
  std::vector<int> lane_at((size_t)G * D, -1);       // entry index r at (lane, slot)
  std::vector<int> bank_at((size_t)nbanks * D, -1);  // (l << 16 | r) at (bank, slot)
  std::vector<std::pair<int, int>> path;
  for (int l = 0; l < G; ++l)
    for (int r = 0; r < (int)entries[l].size(); ++r) {
      const int b = entries[l][r] % nbanks;
      int a = 0, c = 0;
      while (lane_at[(size_t)l * D + a] >= 0) ++a;
      while (bank_at[(size_t)b * D + c] >= 0) ++c;
      int k;
      if (bank_at[(size_t)b * D + a] < 0) {
        k = a;
      } else if (lane_at[(size_t)l * D + c] < 0) {
        k = c;
      } else {
        path.clear();
        int node = b, col = a;
        bool at_bank = true;
        for (;;) {
          if (at_bank) {
            const int id = bank_at[(size_t)node * D + col];
            if (id < 0) break;
            path.emplace_back(id >> 16, id & 0xffff);
            node = id >> 16;
          } else {
            const int r2 = lane_at[(size_t)node * D + col];
            if (r2 < 0) break;
            path.emplace_back(node, r2);
            node = entries[node][r2] % nbanks;
          }
          at_bank = !at_bank;
          col = col == a ? c : a;
        }
        for (const auto& e : path) {
          const int k0 = slot_of[e.first][e.second];
          lane_at[(size_t)e.first * D + k0] = -1;
          bank_at[(size_t)(entries[e.first][e.second] % nbanks) * D + k0] = -1;
        }
        for (const auto& e : path) {
          const int k1 = slot_of[e.first][e.second] == a ? c : a;
          slot_of[e.first][e.second] = k1;
          lane_at[(size_t)e.first * D + k1] = e.second;
          bank_at[(size_t)(entries[e.first][e.second] % nbanks) * D + k1] =
              (e.first << 16) | e.second;
        }
        k = a;
      }
      slot_of[l][r] = k;
      lane_at[(size_t)l * D + k] = r;
      bank_at[(size_t)b * D + k] = (l << 16) | r;
    }
  return D;
}

// Broadcast-aware slot schedule.  Lanes of a half-wave that read the same LDS entry in the
// same slot take one bank access between them (gfx950 serves such a ds_read_b32 at the
// conflict-free rate; two entries of one bank halve it), so a (bank, slot) may hold one
// entry for any number of lanes.  The bound becomes max(longest lane, per bank the sum over
// its entries of the most times one lane holds that entry).  Occurrence r of entry v in a
// lane belongs to item (v, r), whose lanes should share a slot.  Items of two or more
// lanes are placed first, largest first, each in the lowest slot free at its bank and all
// its lanes (else in the free bank slot that takes the most of its lanes, the rest
// following as a new piece); single-lane items, busiest banks first, then take the lowest
// slot free at both their lane and bank, or the alternating-path exchange of
// schedule_group restricted to single-lane items (a path that meets a shared slot is
// abandoned and the next pair of slots tried).  Tries D = bound, bound + 1, .. < limit
// slots; returns the slots used, or -1 (then slot_of is meaningless).
using SlotSet = std::bitset<256>;

int schedule_bcast_try(const std::vector<std::vector<int>>& entries, int nbanks, int D,
                       const std::vector<std::pair<int, int>>& vl,
                       std::vector<std::vector<int>>& slot_of) {
  const int G = (int)entries.size();
  SlotSet mask;
  for (int k = 0; k < D; ++k) mask.set(k);
  std::vector<SlotSet> lu(G), bu(nbanks);
  // lane_at: -1 free, -2 shared item, else the bank of the single-lane item there;
  // bank_at: -1 free, -2 shared item, else the lane of the single-lane item there
  std::vector<int> lane_at((size_t)G * D, -1), bank_at((size_t)nbanks * D, -1);
  std::vector<int64_t> sval((size_t)G * D, -1);  // item key of the single at (lane, slot)
  std::vector<int> bload(nbanks, 0);
  struct Item {
    int v, n0, n;  // entry, its lanes ilane[n0 .. n0+n)
    int r;         // occurrence number within each of those lanes
  };
  std::vector<Item> multi, single;
  std::vector<int> ilane;
  // vl is sorted by (entry, lane): a run of equal pairs is one lane's repeats of the entry
  for (size_t i = 0; i < vl.size();) {
    size_t j = i;
    int rmax = 0;
    while (j < vl.size() && vl[j].first == vl[i].first) {
      size_t b = j;
      while (b < vl.size() && vl[b] == vl[j]) ++b;
      rmax = std::max(rmax, (int)(b - j));
      j = b;
    }
    bload[vl[i].first % nbanks] += rmax;
    for (int r = 0; r < rmax; ++r) {
      Item it{vl[i].first, (int)ilane.size(), 0, r};
      for (size_t a = i; a < j;) {
        size_t b = a;
        while (b < j && vl[b] == vl[a]) ++b;
        if ((int)(b - a) > r) {
          ilane.push_back(vl[a].second);
          ++it.n;
        }
        a = b;
      }
      (it.n > 1 ? multi : single).push_back(it);
    }
    i = j;
  }
  std::stable_sort(multi.begin(), multi.end(), [&](const Item& a, const Item& b) {
    return a.n != b.n ? a.n > b.n : bload[a.v % nbanks] > bload[b.v % nbanks];
  });
  std::stable_sort(single.begin(), single.end(), [&](const Item& a, const Item& b) {
    return bload[a.v % nbanks] > bload[b.v % nbanks];
  });
  auto key = [](int v, int r) { return ((int64_t)v << 32) | (uint32_t)r; };
  std::vector<std::vector<std::pair<int64_t, int>>> got(G);  // per lane: (item key, slot)
  std::vector<int> L, rest;
  for (const Item& it : multi) {
    const int b = it.v % nbanks;
    L.assign(ilane.begin() + it.n0, ilane.begin() + it.n0 + it.n);
    while (!L.empty()) {
      SlotSet c = ~bu[b] & mask;
      for (int l : L) c &= ~lu[l];
      int k = -1;
      if (c.any()) {
        k = (int)c._Find_first();
      } else {
        int best = 0;
        for (int s = 0; s < D; ++s)
          if (!bu[b][s]) {
            int n = 0;
            for (int l : L) n += !lu[l][s];
            if (n > best) {
              best = n;
              k = s;
            }
          }
        if (k < 0) return -1;
      }
      rest.clear();
      for (int l : L) {
        if (lu[l][k]) {
          rest.push_back(l);
          continue;
        }
        lu[l].set(k);
        lane_at[(size_t)l * D + k] = -2;
        got[l].push_back({key(it.v, it.r), k});
      }
      bu[b].set(k);
      bank_at[(size_t)b * D + k] = -2;
      L.swap(rest);
    }
  }
  std::vector<std::pair<int, int>> path;  // (lane, bank) of each edge, colours a, c, a, ..
  std::vector<int64_t> pv;
  for (const Item& it : single) {
    const int l = ilane[it.n0], b = it.v % nbanks;
    const SlotSet c0 = ~bu[b] & ~lu[l] & mask;
    int slot = -1;
    if (c0.any()) {
      slot = (int)c0._Find_first();
    } else {
      const SlotSet fa = ~lu[l] & mask, fc = ~bu[b] & mask;
      for (int a = (int)fa._Find_first(); a < D && slot < 0; a = (int)fa._Find_next(a))
        for (int c = (int)fc._Find_first(); c < D && slot < 0; c = (int)fc._Find_next(c)) {
          // from bank b along its a-edge, then that lane's c-edge, ..: it cannot reach l
          // (free at a) nor return to b (free at c)
          path.clear();
          int node = b, col = a;
          bool at_bank = true, ok = true;
          for (;;) {
            const int nx = at_bank ? bank_at[(size_t)node * D + col]
                                   : lane_at[(size_t)node * D + col];
            if (nx == -1) break;
            if (nx == -2) {
              ok = false;
              break;
            }
            path.push_back(at_bank ? std::make_pair(nx, node) : std::make_pair(node, nx));
            node = nx;
            at_bank = !at_bank;
            col = col == a ? c : a;
          }
          if (!ok) continue;
          pv.resize(path.size());
          for (size_t i = 0; i < path.size(); ++i) {
            const int k0 = i % 2 ? c : a, pl = path[i].first, pb = path[i].second;
            pv[i] = sval[(size_t)pl * D + k0];
            lane_at[(size_t)pl * D + k0] = bank_at[(size_t)pb * D + k0] = -1;
            sval[(size_t)pl * D + k0] = -1;
            lu[pl].reset(k0);
            bu[pb].reset(k0);
          }
          for (size_t i = 0; i < path.size(); ++i) {
            const int k1 = i % 2 ? a : c, pl = path[i].first, pb = path[i].second;
            lane_at[(size_t)pl * D + k1] = pb;
            bank_at[(size_t)pb * D + k1] = pl;
            sval[(size_t)pl * D + k1] = pv[i];
            lu[pl].set(k1);
            bu[pb].set(k1);
          }
          slot = a;
        }
      if (slot < 0) return -1;
    }
    lu[l].set(slot);
    bu[b].set(slot);
    lane_at[(size_t)l * D + slot] = b;
    bank_at[(size_t)b * D + slot] = l;
    sval[(size_t)l * D + slot] = key(it.v, it.r);
  }
  int used = 0;
  slot_of.assign(G, {});
  std::vector<std::pair<int, int>> ord;  // (entry, position in the lane)
  for (int l = 0; l < G; ++l) {
    for (int k = 0; k < D; ++k)
      if (sval[(size_t)l * D + k] >= 0) got[l].push_back({sval[(size_t)l * D + k], k});
    std::sort(got[l].begin(), got[l].end());
    ord.clear();
    for (int r = 0; r < (int)entries[l].size(); ++r) ord.push_back({entries[l][r], r});
    std::sort(ord.begin(), ord.end());
    slot_of[l].assign(entries[l].size(), -1);
    int occ = 0;
    for (size_t i = 0; i < ord.size(); ++i) {
      occ = (i > 0 && ord[i - 1].first == ord[i].first) ? occ + 1 : 0;
      const auto hit = std::lower_bound(got[l].begin(), got[l].end(),
                                        std::make_pair(key(ord[i].first, occ), -1));
      if (hit == got[l].end() || hit->first != key(ord[i].first, occ)) return -1;
      slot_of[l][ord[i].second] = hit->second;
      used = std::max(used, hit->second + 1);
    }
  }
  return used;
}

int schedule_bcast(const std::vector<std::vector<int>>& entries, int nbanks, int limit,
                   std::vector<std::vector<int>>& slot_of) {
  const int G = (int)entries.size();
  std::vector<std::pair<int, int>> vl;
  int lane = 0;
  for (int l = 0; l < G; ++l) {
    lane = std::max(lane, (int)entries[l].size());
    for (int v : entries[l]) vl.push_back({v, l});
  }
  if (vl.empty()) return -1;
  std::sort(vl.begin(), vl.end());
  std::vector<int> bb(nbanks, 0);
  for (size_t i = 0; i < vl.size();) {
    size_t j = i;
    int rmax = 0;
    while (j < vl.size() && vl[j].first == vl[i].first) {
      size_t b = j;
      while (b < vl.size() && vl[b] == vl[j]) ++b;
      rmax = std::max(rmax, (int)(b - j));
      j = b;
    }
    bb[vl[i].first % nbanks] += rmax;
    i = j;
  }
  const int bound = std::max(lane, *std::max_element(bb.begin(), bb.end()));
  for (int D = bound; D < limit && D <= (int)SlotSet().size(); ++D) {
    const int u = schedule_bcast_try(entries, nbanks, D, vl, slot_of);
    if (u >= 0) return u;
  }
  return -1;
}

// The slot schedule of one lane group: schedule_group's exact colouring, replaced by the
// broadcast-aware one when that lowers the slot count past a multiple of 4 (a wave's slots
// per part are rounded up to 4; a smaller saving changes no instantiation).
int schedule_slots(const std::vector<std::vector<int>>& entries, int nbanks, bool bcast,
                   std::vector<std::vector<int>>& slot_of) {
  const int ns = schedule_group(entries, nbanks, slot_of);
  if (!bcast || nbanks == 0 || ns <= 4) return ns;
  std::vector<std::vector<int>> so;
  const int nb = schedule_bcast(entries, nbanks, (int)ceil_div(ns, 4) * 4 - 3, so);
  if (nb < 0) return ns;
  slot_of.swap(so);
  return nb;
}

// Per-tile LDS layout of the table parts (32-bank model, whole-frame waves, <= 2 parts).
// A 16-byte direct-to-LDS load writes, for lane l, entry group l % 8 (4 entries) of LDS row
// 8c + l / 8 (32 entries) of chunk c, from any 16-byte group of the same 128-byte table row
// the lane addresses.  The waves of a tile load the chunks c == wid (mod NW) of every part
// (c < kChunks - 1; the last chunk, 4-byte loads, keeps the identity), so one permutation
// of the 8 groups per (part, wave, row in chunk) costs each lane one fixed source offset
// per part: sigma[p][w][rr][k] is the source group of LDS group k, inv its inverse.  A bank
// is an LDS entry index mod 32: the permutation moves quads of entries between the 8 bank
// quads.  Chosen greedily per tile to flatten the busiest bank of each half-wave -- a bound
// of the slot schedule besides the lane's own entry count -- it lowers the tile's slots.
using Placed = std::vector<std::vector<std::vector<int>>>;  // [part][lane][slot], -1 free

struct TileLayout {
  int NW = 0;
  std::vector<uint8_t> sigma, inv;  // [p][w][rr][k]
  // LDS entry of table entry i (within part p)
  int map(int p, int i) const {
    const int r = i >> 5;
    if (r >= (kChunks - 1) * 8) return i;
    const int c = r >> 3, rr = r & 7, w = c % NW, g = (i & 31) >> 2;
    const int k = inv[(((size_t)p * NW + w) * 8 + rr) * 8 + g];
    return (r << 5) + (k << 2) + (i & 3);
  }
};

// Greedy fit of a tile's layout: per part, the (wave, row-in-chunk) classes in descending
// order of entry count; per class its 8 groups in descending order, each to the free
// position that minimises (busiest of the 4 banks it lands on over the half-waves, sum of
// squared per-half-wave maxima, sum of squared loads).  ws: the tile's waves' schedules,
// entries as identity LDS entries.
void tile_layout(int NP, int NW, const std::vector<const Placed*>& ws, TileLayout& T) {
  T.NW = NW;
  T.sigma.assign((size_t)NP * NW * 64, 0);
  T.inv.assign((size_t)NP * NW * 64, 0);
  const int NH = 2 * (int)ws.size();         // half-waves
  const int nrl = (kChunks - 1) * 8;         // rows with a layout
  std::vector<int> A((size_t)NW * 8 * NH * 32), L((size_t)NH * 32), rmax(NH);
  for (int p = 0; p < NP; ++p) {
    std::fill(A.begin(), A.end(), 0);
    std::fill(L.begin(), L.end(), 0);
    for (int wi = 0; wi < (int)ws.size(); ++wi)
      for (int lane = 0; lane < 64; ++lane) {
        const int hw = 2 * wi + lane / 32;
        for (int v : (*ws[wi])[p][lane]) {
          if (v < 0) continue;
          const int r = v >> 5;
          if (r >= nrl) {
            ++L[(size_t)hw * 32 + (v & 31)];
          } else {
            const int cls = ((r >> 3) % NW) * 8 + (r & 7);
            ++A[((size_t)cls * NH + hw) * 32 + (v & 31)];
          }
        }
      }
    for (int hw = 0; hw < NH; ++hw)
      rmax[hw] = *std::max_element(L.begin() + (size_t)hw * 32, L.begin() + (size_t)hw * 32 + 32);
    std::vector<int64_t> tot(NW * 8, 0);
    for (int cls = 0; cls < NW * 8; ++cls)
      for (size_t q = 0; q < (size_t)NH * 32; ++q) tot[cls] += A[(size_t)cls * NH * 32 + q];
    std::vector<int> cord(NW * 8);
    std::iota(cord.begin(), cord.end(), 0);
    std::stable_sort(cord.begin(), cord.end(), [&](int a, int b) { return tot[a] > tot[b]; });
    for (int cls : cord) {
      const int* a = A.data() + (size_t)cls * NH * 32;
      uint8_t* sg = T.sigma.data() + ((size_t)p * NW * 8 + cls) * 8;
      uint8_t* iv = T.inv.data() + ((size_t)p * NW * 8 + cls) * 8;
      int64_t gt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int hw = 0; hw < NH; ++hw)
        for (int col = 0; col < 32; ++col) gt[col >> 2] += a[hw * 32 + col];
      int gord[8];
      std::iota(gord, gord + 8, 0);
      std::stable_sort(gord, gord + 8, [&](int x, int y) { return gt[x] > gt[y]; });
      bool used[8] = {false, false, false, false, false, false, false, false};
      for (int gi : gord) {
        int best = -1;
        int64_t b1 = 0, b2 = 0, b3 = 0;
        for (int pos = 0; pos < 8; ++pos) {
          if (used[pos]) continue;
          int64_t m1 = 0, m2 = 0, m3 = 0;
          for (int hw = 0; hw < NH; ++hw) {
            int mx = rmax[hw];
            for (int j = 0; j < 4; ++j) {
              const int l0 = L[(size_t)hw * 32 + pos * 4 + j], g = a[hw * 32 + gi * 4 + j];
              const int v = l0 + g;
              m1 = std::max<int64_t>(m1, v);
              mx = std::max(mx, v);
              m3 += (int64_t)2 * l0 * g + (int64_t)g * g;
            }
            m2 += (int64_t)mx * mx;
          }
          if (best < 0 || m1 < b1 || (m1 == b1 && (m2 < b2 || (m2 == b2 && m3 < b3)))) {
            best = pos;
            b1 = m1;
            b2 = m2;
            b3 = m3;
          }
        }
        used[best] = true;
        sg[best] = (uint8_t)gi;
        iv[gi] = (uint8_t)best;
        for (int hw = 0; hw < NH; ++hw)
          for (int j = 0; j < 4; ++j) {
            int& l0 = L[(size_t)hw * 32 + best * 4 + j];
            l0 += a[hw * 32 + gi * 4 + j];
            rmax[hw] = std::max(rmax[hw], l0);
          }
      }
    }
  }
}

using GatherFn = void (*)(const float*, int, int, int, int, int, int, const uint32_t*,
                          const double*, const double*, int, Top3*, int, const uint32_t*);

// Gather instantiations: S slot registers per table part and lane, NP parts, NW waves per
// workgroup so that the S*NP slots + ~30 working VGPRs fit NW/4 waves per SIMD (512
// registers per SIMD lane: 128 per wave at 16 waves, 170 at 12, 256 at 8), and
// L = ceil(kChunks / NW) direct-to-LDS loads per wave per part.
constexpr int nw_of(int P, int NP, int X) {
  // 16 waves need <= 128 VGPRs: up to 96 slots for whole frames (one 8-byte value of the
  // epilogue spills, outside the hypothesis loop), up to 80 for split frames (X > 1: the
  // lane-group reduction's registers would push slot addresses to scratch inside the loop)
  const int p16 = X == 1 ? 96 : 80;
  return P <= p16 ? 16 : (NP <= 2 ? (P <= 120 ? 12 : 8) : (P <= 96 ? 12 : 8));
}
// two frames per lane: P packed slot registers + ~52 working VGPRs within 512 / (NW / 4)
constexpr int nw2_of(int P) { return P <= 72 ? 16 : (P <= 112 ? 12 : 8); }
constexpr int l_of(int P, int NP, int X) {
  return (kChunks + nw_of(P, NP, X) - 1) / nw_of(P, NP, X);
}

struct GatherCfg {
  int S, NP, NW, X;
  int F;  // frames per lane (2: slot registers hold two 16-bit addresses)
  GatherFn fn;
};

#define RSP_CFG(S, NP) \
  {S, NP, nw_of(S * NP, NP, 1), 1, 1,  \
   &k_gather<S, NP, nw_of(S * NP, NP, 1), l_of(S * NP, NP, 1), 1, 1>}
#define RSP_CFGX(S, NP, X) \
  {S, NP, nw_of(S * NP, NP, X), X, 1,  \
   &k_gather<S, NP, nw_of(S * NP, NP, X), l_of(S * NP, NP, X), X, 1>}
#define RSP_CFG2(S, NP) \
  {S, NP, nw2_of(S * NP), 1, 2,  \
   &k_gather<S, NP, nw2_of(S * NP), (kChunks + nw2_of(S * NP) - 1) / nw2_of(S * NP), 1, 2>}
const GatherCfg kGatherCfgs[] = {
    RSP_CFG(16, 1), RSP_CFG(32, 1), RSP_CFG(48, 1), RSP_CFG(64, 1), RSP_CFG(80, 1),
    RSP_CFG(96, 1), RSP_CFG(112, 1), RSP_CFG(128, 1), RSP_CFG(160, 1),
    RSP_CFG(8, 2),  RSP_CFG(16, 2), RSP_CFG(24, 2), RSP_CFG(32, 2), RSP_CFG(36, 2),
    RSP_CFG(40, 2), RSP_CFG(44, 2), RSP_CFG(48, 2), RSP_CFG(52, 2), RSP_CFG(56, 2),
    RSP_CFG(60, 2), RSP_CFG(64, 2),
    RSP_CFG(8, 3),  RSP_CFG(16, 3), RSP_CFG(20, 3), RSP_CFG(24, 3), RSP_CFG(28, 3),
    RSP_CFG(32, 3),
    RSP_CFG(8, 4),  RSP_CFG(12, 4), RSP_CFG(16, 4), RSP_CFG(20, 4), RSP_CFG(24, 4),
    // frames split over X lanes
    RSP_CFGX(96, 1, 2), RSP_CFGX(96, 1, 4), RSP_CFGX(96, 1, 8),
    RSP_CFGX(96, 1, 16), RSP_CFGX(96, 1, 32), RSP_CFGX(96, 1, 64),
    RSP_CFGX(40, 2, 2), RSP_CFGX(40, 2, 4), RSP_CFGX(40, 2, 8),
    RSP_CFGX(40, 2, 16), RSP_CFGX(40, 2, 32), RSP_CFGX(40, 2, 64),
    RSP_CFGX(24, 3, 2), RSP_CFGX(24, 3, 4), RSP_CFGX(24, 3, 8),
    RSP_CFGX(24, 3, 16), RSP_CFGX(24, 3, 32), RSP_CFGX(24, 3, 64),
    RSP_CFGX(16, 4, 2), RSP_CFGX(16, 4, 4), RSP_CFGX(16, 4, 8),
    RSP_CFGX(16, 4, 16), RSP_CFGX(16, 4, 32), RSP_CFGX(16, 4, 64),
    // two frames per lane (whole frames)
    RSP_CFG2(16, 1), RSP_CFG2(32, 1), RSP_CFG2(48, 1), RSP_CFG2(64, 1),
    RSP_CFG2(8, 2), RSP_CFG2(16, 2), RSP_CFG2(24, 2), RSP_CFG2(28, 2), RSP_CFG2(32, 2),
    RSP_CFG2(36, 2), RSP_CFG2(40, 2), RSP_CFG2(44, 2), RSP_CFG2(48, 2),
};
#undef RSP_CFG
#undef RSP_CFGX
#undef RSP_CFG2
constexpr int kNumCfg = sizeof(kGatherCfgs) / sizeof(kGatherCfgs[0]);

// smallest instantiation with NP parts, lane split X and F frames per lane whose part
// capacity holds `need`
int choose_cfg(int NP, int X, int need, int F = 1) {
  for (int i = 0; i < kNumCfg; ++i)
    if (kGatherCfgs[i].NP == NP && kGatherCfgs[i].X == X && kGatherCfgs[i].F == F &&
        need <= kGatherCfgs[i].S)
      return i;
  return -1;
}

int max_slots(int NP, int X) {
  int m = 0;
  for (int i = 0; i < kNumCfg; ++i)
    if (kGatherCfgs[i].NP == NP && kGatherCfgs[i].X == X && kGatherCfgs[i].F == 1)
      m = std::max(m, kGatherCfgs[i].S);
  return m;
}

}  // namespace

extern "C" int rsp_emc_create(int device, int64_t size, int64_t num_rot, const double* rot_cos,
                              const double* rot_sin, const double* state_sx,
                              const double* state_sy, const double* state_dia,
                              int64_t num_states, rsp_emc** out) {
  if (!out || !rot_cos || !rot_sin || !state_sx || !state_sy || !state_dia)
    return fail(kErrNull, "rsp_emc_create: null pointer");
  *out = nullptr;
  if (size < 8 || size > 4096) return fail(kErrSize, "rsp_emc_create: size out of range");
  if (num_rot < 1 || num_states < 1) return fail(kErrSize, "rsp_emc_create: empty grid");
  if (num_rot * num_states > 0x7fffffffLL)
    return fail(kErrSize, "rsp_emc_create: more than 2^31-1 hypotheses");
  RSP_HIP(hipSetDevice(device));
  rsp_emc* e = new rsp_emc();
  e->device = device;
  e->size = (int)size;
  e->npix = (int)(size * size);
  e->R = (int)num_rot;
  e->S = (int)num_states;
  e->H = num_rot * num_states;
  e->rcos.assign(rot_cos, rot_cos + num_rot);
  e->rsin.assign(rot_sin, rot_sin + num_rot);
  e->ssx.assign(state_sx, state_sx + num_states);
  e->ssy.assign(state_sy, state_sy + num_states);
  e->sdia.assign(state_dia, state_dia + num_states);
  // emc.py:89-98 probmask and the scored annulus (probmask == 0)
  e->probmask.assign(e->npix, 0);
  e->compact_of.assign(e->npix, -1);
  const int c = e->size / 2;
  for (int x = 0; x < e->size; ++x)
    for (int y = 0; y < e->size; ++y) {
      const double rad = std::sqrt((double)(x - c) * (x - c) + (double)(y - c) * (y - c));
      int v = 0;
      if (rad >= e->size / 2) v = 2;
      if (rad < e->size / 8) v = 1;
      if (rad < 4) v = 2;
      e->probmask[x * e->size + y] = v;
      if (v == 0) {
        e->compact_of[x * e->size + y] = (int)e->scored_t.size();
        e->scored_t.push_back(x * e->size + y);
      }
    }
  e->n_scored = (int)e->scored_t.size();
  e->NP = std::max(1, (int)ceil_div(e->n_scored, kPartMax));
  int rc = 0;
  if (e->NP > kNPMax) {
    delete e;
    return fail(kErrUnsupported, "rsp_emc_create: scored annulus too large for LDS parts");
  }
  e->PSTRIDE = kPStride;
  e->TS = e->PSTRIDE * e->NP;
  if ((rc = dev_upload(&e->d_rcos, e->rcos)) || (rc = dev_upload(&e->d_rsin, e->rsin)) ||
      (rc = dev_upload(&e->d_sx, e->ssx)) || (rc = dev_upload(&e->d_sy, e->ssy)) ||
      (rc = dev_upload(&e->d_dia, e->sdia)) || (rc = dev_upload(&e->d_probmask, e->probmask)) ||
      (rc = dev_upload(&e->d_scored_t, e->scored_t)) ||
      (rc = dev_upload(&e->d_compact_of, e->compact_of)) ||
      (rc = dev_alloc(&e->d_I, (int64_t)e->S * e->npix)) ||
      (rc = dev_alloc(&e->d_init, e->S)) || (rc = dev_alloc(&e->d_vscale, 1)) ||
      (rc = dev_alloc(&e->d_errmax, e->n_scored)) || (rc = dev_alloc(&e->d_absmax, e->n_scored)) || (rc = dev_alloc(&e->d_counters, 4))) {
    free_all(e);
    delete e;
    return rc;
  }
  *out = e;
  return kOk;
}

extern "C" int rsp_emc_set_option(rsp_emc* e, int key, int64_t value) {
  if (!e) return fail(kErrNull, "rsp_emc_set_option: null plan");
  switch (key) {
    case 1: e->exact_mode = value != 0; return kOk;
    case 2:
      if (value < 0) return fail(kErrSize, "batch < 0");
      e->opt_batch = value;
      return kOk;
    case 3:
      if (value != 0 && value != 32 && value != 64)
        return fail(kErrSize, "bank count must be 0 (no scheduling), 32 or 64");
      e->NB = (int)value;
      return kOk;
    case 4:
      if (value < 0) return fail(kErrSize, "segments < 0");
      e->opt_nseg = value;
      return kOk;
    case 5: e->timing = value != 0; return kOk;
    case 6:
      if (value < 0 || value > (1 << 24)) return fail(kErrSize, "regroup iterations out of range");
      e->regroup_iters = (int)value;
      return kOk;
    case 7: e->layout = value != 0; return kOk;
    case 8: e->bcast = value != 0; return kOk;
    case 9:
      if (value != 1 && value != 2) return fail(kErrSize, "frames per lane must be 1 or 2");
      e->fpl = (int)value;
      return kOk;
    case 10:
      if (value < 0 || value > 4) return fail(kErrSize, "class streams must be 0 .. 4");
      e->class_streams = (int)value;
      return kOk;
    default: return fail(kErrSize, "rsp_emc_set_option: unknown key");
  }
}

extern "C" int rsp_emc_set_frames(rsp_emc* e, int64_t num_data, const int32_t* ones,
                                  const int32_t* multi, const int32_t* place_ones,
                                  const int32_t* place_multi, const int32_t* count_multi,
                                  const double* bg) {
  if (!e) return fail(kErrNull, "rsp_emc_set_frames: null plan");
  if (num_data < 1 || num_data > (1LL << 30)) return fail(kErrSize, "rsp_emc_set_frames: num_data");
  if (!ones || !multi) return fail(kErrNull, "rsp_emc_set_frames: null counts");
  RSP_HIP(hipSetDevice(e->device));
  const int64_t N = num_data;
  int64_t no = 0, nm = 0;
  for (int64_t d = 0; d < N; ++d) {
    if (ones[d] < 0 || multi[d] < 0) return fail(kErrSize, "negative photon count");
    no += ones[d];
    nm += multi[d];
  }
  if ((no && !place_ones) || (nm && (!place_multi || !count_multi)))
    return fail(kErrNull, "rsp_emc_set_frames: null photon list");
  for (int64_t i = 0; i < no; ++i)
    if (place_ones[i] < 0 || place_ones[i] >= e->npix) return fail(kErrSize, "place_ones out of range");
  for (int64_t i = 0; i < nm; ++i) {
    if (place_multi[i] < 0 || place_multi[i] >= e->npix)
      return fail(kErrSize, "place_multi out of range");
    // a multi photon's count is at least 2 in the reference's files; below 1 would shrink
    // the exactness bound of the FP32 pass, and the slot lists replicate each count
    if (count_multi[i] < 1 || count_multi[i] >= (1 << 22))
      return fail(kErrSize, "count_multi out of range (1 <= count < 2^22)");
  }
  e->N = N;
  // background (dset.py:56-60) and powder (dset.py:74-83)
  std::vector<double> hbg(e->npix, 0.);
  if (bg) hbg.assign(bg, bg + e->npix);
  std::vector<double> powder(e->npix, 0.);
  for (int64_t i = 0; i < no; ++i) powder[place_ones[i]] += 1.;
  for (int64_t i = 0; i < nm; ++i) powder[place_multi[i]] += count_multi[i];
  double ps = 0.;
  for (int t = 0; t < e->npix; ++t)
    if (e->probmask[t] < 1) ps += powder[t] / (double)N;
  e->powder_sel_sum = ps;
  // exact-path CSR (reference order: scored ones, then scored multi)
  std::vector<int64_t> ex_off(N + 1, 0);
  std::vector<int32_t> ex_n1(N), ex_pix, ex_cnt;
  std::vector<int> nentries(N, 0);  // entries with multiplicity for the FP32 path
  {
    int64_t po = 0, pm = 0;
    for (int64_t d = 0; d < N; ++d) {
      int n1 = 0, ne = 0;
      for (int k = 0; k < ones[d]; ++k) {
        const int t = place_ones[po + k];
        if (e->probmask[t] < 1) {
          ex_pix.push_back(t);
          ex_cnt.push_back(1);
          ++n1;
          ++ne;
        }
      }
      for (int k = 0; k < multi[d]; ++k) {
        const int t = place_multi[pm + k];
        if (e->probmask[t] < 1) {
          ex_pix.push_back(t);
          ex_cnt.push_back(count_multi[pm + k]);
          ne += count_multi[pm + k];
        }
      }
      po += ones[d];
      pm += multi[d];
      ex_n1[d] = n1;
      ex_off[d + 1] = (int64_t)ex_pix.size();
      nentries[d] = ne;
    }
  }
  // sort frames into waves of 64 by their largest per-part entry count, then by entry
  // count (both descending): the slots of a wave's part are at least the largest count
  // of any of its lanes on that part, and frames alike in both parts share waves
  std::vector<int> maxpart(N, 0);
  {
    std::vector<int> pc(e->NP);
    for (int64_t d = 0; d < N; ++d) {
      std::fill(pc.begin(), pc.end(), 0);
      for (int64_t q = ex_off[d]; q < ex_off[d + 1]; ++q) {
        const int c = (q < ex_off[d] + ex_n1[d]) ? 1 : ex_cnt[q];
        pc[e->compact_of[ex_pix[q]] / kPartMax] += c;
      }
      maxpart[d] = *std::max_element(pc.begin(), pc.end());
    }
  }
  std::vector<int32_t> order(N);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return maxpart[a] != maxpart[b] ? maxpart[a] > maxpart[b] : nentries[a] > nentries[b];
  });
  const int NB = e->NB;
  // lanes per conflict group: 32 banks -> half waves, 64 -> full waves, 0 -> no scheduling
  const int G = NB == 0 ? 64 : NB;
  // Waves of 64 lanes over the sorted frames.  A frame whose entries do not fit the largest
  // instantiation is split over X lanes (X = 2 .. 64; entry r goes to lane r % X of its
  // group), each wave then holding 64/X frames.  Per wave and part: slot -> LDS entry
  // (within the part) per lane, -1 = free slot.
  struct Wave {
    int X;
    int64_t f0, nf;  // frames order[f0 .. f0+nf)
    int need;        // slots per part (max over parts, multiple of 4)
    Placed placed;   // [part][lane][slot]: LDS entry within the part, -1 free
  };
  std::vector<Wave> waves;
  auto build_wave = [&](int64_t f0, int X, Wave& wv) {
    wv.X = X;
    wv.f0 = f0;
    wv.nf = std::min<int64_t>(64 / X, N - f0);
    wv.need = 0;
    wv.placed.assign(e->NP, std::vector<std::vector<int>>(64));
    for (int p = 0; p < e->NP; ++p) {
      int ns_max = 0;
      for (int g0 = 0; g0 < 64; g0 += G) {
        std::vector<std::vector<int>> ent(G);
        for (int l = 0; l < G; ++l) {
          const int lane = g0 + l, k = lane / X;
          if (k >= wv.nf) continue;
          const int d = order[f0 + k];
          int r = 0;
          for (int64_t q = ex_off[d]; q < ex_off[d + 1]; ++q) {
            const int ci = e->compact_of[ex_pix[q]];
            const int c = (q < ex_off[d] + ex_n1[d]) ? 1 : ex_cnt[q];
            for (int m = 0; m < c; ++m, ++r)
              if (ci / kPartMax == p && r % X == lane % X) ent[l].push_back(ci - p * kPartMax);
          }
        }
        std::vector<std::vector<int>> slot_of;
        const int ns = schedule_slots(ent, NB, e->bcast, slot_of);
        ns_max = std::max(ns_max, ns);
        for (int l = 0; l < G; ++l) {
          auto& pl = wv.placed[p][g0 + l];
          pl.assign(ns, -1);
          for (size_t r = 0; r < ent[l].size(); ++r) pl[slot_of[l][r]] = ent[l][r];
        }
      }
      wv.need = std::max(wv.need, (int)ceil_div(ns_max, 4) * 4);
    }
  };
  // per-frame (part, bank) histograms for the regrouping of whole-frame waves
  FrameHist fh;
  const bool regroup = e->regroup_iters > 0 && NB > 0;
  if (regroup) {
    fh.NP = e->NP;
    fh.NB = NB;
    fh.h.assign((size_t)N * e->NP * NB, 0);
    fh.pc.assign((size_t)N * e->NP, 0);
    for (int64_t d = 0; d < N; ++d)
      for (int64_t q = ex_off[d]; q < ex_off[d + 1]; ++q) {
        const int ci = e->compact_of[ex_pix[q]];
        const int c = (q < ex_off[d] + ex_n1[d]) ? 1 : ex_cnt[q];
        const int p = ci / kPartMax, bank = (ci - p * kPartMax) % NB;
        uint16_t& hv = fh.h[((size_t)d * e->NP + p) * NB + bank];
        hv = (uint16_t)std::min(65535, (int)hv + c);
        fh.pc[(size_t)d * e->NP + p] += c;
      }
  }
  int64_t next_window = 0;
  for (int64_t f0 = 0; f0 < N;) {
    Wave wv;
    if (regroup && f0 >= next_window) {
      // regroup the next 8 lane groups once the waves hold whole frames (X = 1)
      build_wave(f0, 1, wv);
      if (wv.need <= max_slots(e->NP, 1)) {
        next_window = std::min<int64_t>(N, f0 + 8 * (int64_t)G);
        regroup_window(order, f0, next_window, G, fh, e->regroup_iters, (uint64_t)f0);
      }
    }
    int X = 1;
    for (;; X *= 2) {
      build_wave(f0, X, wv);
      if (wv.need <= max_slots(e->NP, X)) break;
      if (X == 64)
        return fail(kErrUnsupported, "rsp_emc_set_frames: a frame has too many scored photons");
    }
    f0 += wv.nf;
    waves.push_back(std::move(wv));
  }
  const int64_t nw0 = (int64_t)waves.size();
  // waves in descending order of (lane split, slot need); classes = runs of equal
  // instantiation
  std::vector<int64_t> worder(nw0);
  std::iota(worder.begin(), worder.end(), 0);
  std::stable_sort(worder.begin(), worder.end(), [&](int64_t a, int64_t b) {
    return waves[a].X != waves[b].X ? waves[a].X > waves[b].X : waves[a].need > waves[b].need;
  });
  // whole-frame waves go two to a hardware wave (option 9) where an instantiation holds
  // their slots; split frames and heavier waves one per hardware wave
  auto cfg_of = [&](int X, int need) {
    const int c2 = X == 1 && e->fpl == 2 ? choose_cfg(e->NP, 1, need, 2) : -1;
    return c2 >= 0 ? c2 : choose_cfg(e->NP, X, need);
  };
  std::vector<int> wcfg(nw0);
  for (int64_t i = 0; i < nw0; ++i) {
    const Wave& wv = waves[worder[i]];
    const int c = cfg_of(wv.X, std::max(wv.need, 8));
    if (c < 0) return fail(kErrUnsupported, "rsp_emc_set_frames: a wave needs too many slots");
    wcfg[i] = c;
  }
  // Table layouts (TileLayout): runs of whole-frame waves get a layout fitted to their
  // frames and are rescheduled on it; a run whose slot need then fits a cheaper
  // instantiation with its workgroup size keeps it and becomes one workgroup tile of that
  // class (tiles first in every class, so that they stay aligned to the workgroups); the
  // other waves keep the identity.
  std::vector<TileLayout> layouts;
  std::vector<int> wlay(nw0, -1), wwid(nw0, 0);  // per wave: layout id, wave within tile
  {
    const bool use_layout = e->layout && NB == 32 && e->NP <= 2;
    auto resched = [&](const Wave& wv, const TileLayout& T, Wave& out) {
      out.X = wv.X;
      out.f0 = wv.f0;
      out.nf = wv.nf;
      out.need = 0;
      out.placed.assign(e->NP, std::vector<std::vector<int>>(64));
      for (int p = 0; p < e->NP; ++p) {
        int ns_max = 0;
        for (int g0 = 0; g0 < 64; g0 += G) {
          std::vector<std::vector<int>> ent(G);
          for (int l = 0; l < G; ++l)
            for (int v : wv.placed[p][g0 + l])
              if (v >= 0) ent[l].push_back(T.map(p, v));
          std::vector<std::vector<int>> slot_of;
          const int ns = schedule_slots(ent, NB, e->bcast, slot_of);
          ns_max = std::max(ns_max, ns);
          for (int l = 0; l < G; ++l) {
            auto& pl = out.placed[p][g0 + l];
            pl.assign(ns, -1);
            for (size_t r = 0; r < ent[l].size(); ++r) pl[slot_of[l][r]] = ent[l][r];
          }
        }
        out.need = std::max(out.need, (int)ceil_div(ns_max, 4) * 4);
      }
    };
    struct Unit {
      int X, cfg;
      bool tile;
      std::vector<int64_t> w;
    };
    std::vector<Unit> units;
    std::vector<char> in_tile(nw0, 0);  // by position in worder
    if (use_layout) {
      // Candidate tiles: whole-frame waves by their lane floor (the longest lane on any
      // part, rounded up to 4: no layout takes a wave below it), then by need, in runs of
      // the workgroup size of the floor's instantiation.  A run keeps its layout when the
      // tile's need then fits an instantiation of that workgroup size whose slots, over the
      // tile, are fewer than its waves' own instantiations'.
      std::vector<int> lfl(nw0, 0);
      std::vector<int64_t> cand;
      for (int64_t i = 0; i < nw0; ++i) {
        const Wave& wv = waves[worder[i]];
        if (wv.X != 1) continue;
        int m = 0;
        for (const auto& pp : wv.placed)
          for (const auto& pl : pp) {
            int n = 0;
            for (int v : pl) n += v >= 0;
            m = std::max(m, n);
          }
        lfl[i] = (int)ceil_div(m, 4) * 4;
        cand.push_back(i);
      }
      std::stable_sort(cand.begin(), cand.end(), [&](int64_t x, int64_t y) {
        return lfl[x] != lfl[y] ? lfl[x] > lfl[y] : waves[worder[x]].need > waves[worder[y]].need;
      });
      for (size_t t0 = 0; t0 < cand.size();) {
        // NWt: frames' waves per workgroup (hardware waves x frames per lane)
        const int c0 = cfg_of(1, std::max(lfl[cand[t0]], 8));
        const int NWh = kGatherCfgs[c0].NW, Ft = kGatherCfgs[c0].F, NWt = NWh * Ft;
        if (t0 + NWt > cand.size()) break;
        std::vector<const Placed*> ps;
        int64_t old_cost = 0;
        for (int k = 0; k < NWt; ++k) {
          ps.push_back(&waves[worder[cand[t0 + k]]].placed);
          old_cost += kGatherCfgs[wcfg[cand[t0 + k]]].S;
        }
        TileLayout T;
        tile_layout(e->NP, NWh, ps, T);
        std::vector<Wave> nws(NWt);
        int need = 0;
        for (int k = 0; k < NWt; ++k) {
          resched(waves[worder[cand[t0 + k]]], T, nws[k]);
          need = std::max(need, nws[k].need);
        }
        const int c2 = cfg_of(1, std::max(need, 8));
        if (c2 >= 0 && kGatherCfgs[c2].NW == NWh && kGatherCfgs[c2].F == Ft &&
            (int64_t)NWt * kGatherCfgs[c2].S < old_cost) {
          Unit u{1, c2, true, {}};
          for (int k = 0; k < NWt; ++k) {
            const int64_t wi = worder[cand[t0 + k]];
            waves[wi] = std::move(nws[k]);
            wlay[wi] = (int)layouts.size();
            wwid[wi] = k / Ft;  // the hardware wave that loads its chunks
            in_tile[cand[t0 + k]] = 1;
            u.w.push_back(wi);
          }
          layouts.push_back(std::move(T));
          units.push_back(std::move(u));
        }
        t0 += NWt;
      }
    }
    for (int64_t i = 0; i < nw0; ++i)
      if (!in_tile[i]) units.push_back(Unit{waves[worder[i]].X, wcfg[i], false, {worder[i]}});
    // lane split descending, frames per lane descending, then capacity descending (the
    // order above), tiles first
    std::stable_sort(units.begin(), units.end(), [](const Unit& a, const Unit& b) {
      if (a.X != b.X) return a.X > b.X;
      if (kGatherCfgs[a.cfg].F != kGatherCfgs[b.cfg].F)
        return kGatherCfgs[a.cfg].F > kGatherCfgs[b.cfg].F;
      if (a.cfg != b.cfg) return a.cfg > b.cfg;
      return a.tile && !b.tile;
    });
    // Runs of one instantiation (the classes): tiles, then single waves.  The last
    // workgroup of a class would be completed with empty waves, each reading its slots as
    // zero pads; it takes the neediest single waves of the following classes with the
    // same lane split and frames per lane instead (their capacity is smaller, so they
    // fit).
    struct Run {
      int X, cfg;
      std::vector<int64_t> tiles, singles;
    };
    std::vector<Run> runs;
    for (const Unit& u : units) {
      if (runs.empty() || runs.back().cfg != u.cfg) runs.push_back(Run{u.X, u.cfg, {}, {}});
      auto& dst = u.tile ? runs.back().tiles : runs.back().singles;
      dst.insert(dst.end(), u.w.begin(), u.w.end());
    }
    for (size_t a = 0; a < runs.size(); ++a) {
      const int Fa = kGatherCfgs[runs[a].cfg].F, NWa = kGatherCfgs[runs[a].cfg].NW * Fa;
      const int64_t n = (int64_t)(runs[a].tiles.size() + runs[a].singles.size());
      int64_t short_by = (NWa - n % NWa) % NWa;
      for (size_t b = a + 1; short_by > 0 && b < runs.size() && runs[b].X == runs[a].X &&
                             kGatherCfgs[runs[b].cfg].F == Fa;
           ++b) {
        auto& src = runs[b].singles;
        const int64_t take = std::min<int64_t>(short_by, (int64_t)src.size());
        runs[a].singles.insert(runs[a].singles.end(), src.begin(), src.begin() + take);
        src.erase(src.begin(), src.begin() + take);
        short_by -= take;
      }
    }
    int64_t k = 0;
    for (const Run& r : runs)
      for (const auto* part : {&r.tiles, &r.singles})
        for (int64_t wi : *part) {
          worder[k] = wi;
          wcfg[k] = r.cfg;
          ++k;
        }
  }
  e->classes.clear();
  int64_t gw_total = 0, slot_total = 0;
  for (int64_t i = 0; i < nw0;) {
    int64_t i1 = i;
    while (i1 < nw0 && wcfg[i1] == wcfg[i]) ++i1;
    rsp_emc::Class c;
    c.cfg = wcfg[i];
    const GatherCfg& g = kGatherCfgs[c.cfg];
    c.w_src = i;
    c.n_src = i1 - i;
    c.n_tiles = (int)ceil_div(c.n_src, g.NW * g.F);
    c.gw0 = (int)gw_total;
    c.slot_off = slot_total;
    gw_total += (int64_t)c.n_tiles * g.NW * g.F;  // waves of frames
    slot_total += (int64_t)c.n_tiles * g.NW * g.S * g.NP * 64;
    e->classes.push_back(c);
    i = i1;
  }
  const int64_t n_waves = gw_total;
  const int64_t Npad = n_waves * 64;
  std::vector<uint32_t> slots((size_t)slot_total);
  std::vector<int32_t> perm(Npad, -1);
  std::vector<uint32_t> lay((size_t)2 * Npad);
  for (int64_t j = 0; j < Npad; ++j) lay[j] = lay[Npad + j] = (uint32_t)(j % 64) * 16u;
  for (const auto& c : e->classes) {
    const GatherCfg& g = kGatherCfgs[c.cfg];
    const int P = g.S * g.NP;
    // waves of frames lwf; hardware wave lw = lwf / F holds them in the low (f = 0) and
    // high (f = 1) halves of its slot registers when F = 2
    for (int64_t lwf = 0; lwf < (int64_t)c.n_tiles * g.NW * g.F; ++lwf) {
      const int64_t lw = lwf / g.F, gw = c.gw0 + lwf;
      const int f = (int)(lwf % g.F);
      const bool real = lwf < c.n_src;
      const Wave* wv = real ? &waves[worder[c.w_src + lwf]] : nullptr;
      // the first lane of each frame's group stands for the frame; the others are -1
      for (int l = 0; l < 64; ++l)
        if (real && l % wv->X == 0 && l / wv->X < wv->nf)
          perm[gw * 64 + l] = order[wv->f0 + l / wv->X];
      // the chunk offsets of the hardware wave's loads, read at its first frame
      if (real && f == 0 && wlay[worder[c.w_src + lwf]] >= 0) {
        const TileLayout& T = layouts[wlay[worder[c.w_src + lwf]]];
        const int wid = wwid[worder[c.w_src + lwf]];
        if (wid != (int)(lw % g.NW))
          return fail(kErrUnsupported, "rsp_emc_set_frames: tile out of its workgroup");
        for (int p = 0; p < g.NP; ++p)
          for (int l = 0; l < 64; ++l)
            lay[(size_t)p * Npad + gw * 64 + l] =
                (uint32_t)(l >> 3) * 128u +
                16u * T.sigma[(((size_t)p * T.NW + wid) * 8 + (l >> 3)) * 8 + (l & 7)];
      }
      for (int p = 0; p < g.NP; ++p) {
        const int part_len = std::min(kPartMax, e->n_scored - p * kPartMax);
        // zero pads read LDS entries >= part_len (zero in every table): for 32 banks the
        // first row with no table entry, which every layout leaves zero; for 64 banks the
        // tail [part_len, kPStride), where a bank with no zero entry there takes another
        const int zrow = (int)ceil_div(part_len, 32) * 32;
        for (int g0 = 0; g0 < 64; g0 += G) {
          for (int k = 0; k < g.S; ++k) {
            // banks used by real entries in this slot, then conflict-free zero pads
            uint64_t used = 0;
            for (int l = 0; l < G && NB > 0 && real; ++l) {
              const auto& pl = wv->placed[p][g0 + l];
              if (k < (int)pl.size() && pl[k] >= 0) used |= 1ull << (pl[k] % NB);
            }
            for (int l = 0; l < G; ++l) {
              int entry = -1;
              if (real) {
                const auto& pl = wv->placed[p][g0 + l];
                if (k < (int)pl.size()) entry = pl[k];
              }
              if (entry < 0) {
                if (NB == 0) {
                  entry = part_len + (l % 32);
                } else {
                  int bank = 0;
                  while (bank < NB && ((used >> bank) & 1ull)) ++bank;
                  if (bank == NB) bank = 0;
                  if (NB == 32) {
                    entry = zrow + bank;
                  } else {
                    entry = part_len + ((bank - part_len) % NB + NB) % NB;
                    if (entry >= kPStride) entry = part_len + (entry - part_len) % (kPStride - part_len);
                    bank = entry % NB;
                  }
                  used |= 1ull << bank;
                }
              }
              uint32_t& sl =
                  slots[(size_t)(c.slot_off + ((int64_t)lw * P + p * g.S + k) * 64 + g0 + l)];
              if (g.F == 1) {
                sl = (uint32_t)entry * 4u;
              } else {
                // two 16-bit LDS byte addresses (a part buffer is 50,176 bytes)
                if ((uint32_t)entry * 4u > 0xffffu)
                  return fail(kErrUnsupported, "rsp_emc_set_frames: slot address above 16 bits");
                sl = f == 0 ? (uint32_t)entry * 4u : sl | ((uint32_t)entry * 4u << 16);
              }
            }
          }
        }
      }
    }
  }
  int rc = 0;
  std::vector<double> hb(hbg);
  if ((rc = dev_upload(&e->d_bg, hb)) || (rc = dev_upload(&e->d_slots, slots)) ||
      (rc = dev_upload(&e->d_lay, lay)) ||
      (rc = dev_upload(&e->d_perm, perm)) ||
      (rc = dev_upload(&e->d_ex_off, ex_off)) || (rc = dev_upload(&e->d_ex_n1, ex_n1)) ||
      (rc = dev_upload(&e->d_ex_pix, ex_pix)) || (rc = dev_upload(&e->d_ex_cnt, ex_cnt)) ||
      (rc = dev_alloc(&e->d_scale_sorted, Npad)) || (rc = dev_alloc(&e->d_running, Npad)) ||
      (rc = dev_alloc(&e->d_full_list, Npad)))
    return rc;
  e->P = kGatherCfgs[e->classes.front().cfg].S * kGatherCfgs[e->classes.front().cfg].NP;
  e->NW = kGatherCfgs[e->classes.front().cfg].NW;
  e->n_waves = (int)n_waves;
  e->Npad = Npad;
  return kOk;
}
extern "C" int rsp_emc_estep(rsp_emc* e, const double* model_c, const double* scales,
                             int32_t* rmax, double* maxll, double* msums, double* vscale,
                             rsp_stream_t stream_) {
  if (!e) return fail(kErrNull, "rsp_emc_estep: null plan");
  if (!model_c || !scales || !rmax || !maxll || !msums || !vscale)
    return fail(kErrNull, "rsp_emc_estep: null pointer");
  if (e->N == 0 || !e->d_slots) return fail(kErrState, "rsp_emc_estep: frames not set");
  RSP_HIP(hipSetDevice(e->device));
  hipStream_t st = (hipStream_t)stream_;
  int rc;
  // batch size: tables of `batch` hypotheses (default: <= 2 GiB and <= 16384)
  int64_t batch = e->opt_batch > 0 ? e->opt_batch
                                   : std::min<int64_t>(16384, (2LL << 30) / (4LL * e->TS));
  batch = std::max<int64_t>(1, std::min(batch, e->H));
  if (batch != e->batch) {
    if ((rc = dev_alloc(&e->d_tables, batch * e->TS))) return rc;
    RSP_HIP(hipMemsetAsync(e->d_tables, 0, sizeof(float) * batch * e->TS, st));
    e->batch = batch;
  }
  // segments per batch and slot class: at most 128 hypotheses per workgroup, and at least
  // about 4 workgroups per CU -- in every class launch when the classes run one after
  // another (a class with few tiles gets more hypothesis segments, so that every launch
  // fills the 256 CUs), over the batch's classes together when they share side streams
  // (class_streams >= 2: concurrent classes fill the CUs between them, and extra segments
  // only add partials and workgroup starts).  Measured at 185^2 with 100k frames (8
  // classes of 2 to 40 tiles), one stream: 4 workgroups per CU alone 1.86 s per E-step, a
  // fixed 64 / 96 / 128 / 160 / 192 / 256 / 512 segments 1.87 / 1.87 / 1.81 / 1.81 / 1.83 /
  // 1.83 / 1.92 s, the per-class rule 1.80 s; two class streams: the per-class rule 1.563
  // s, a fixed 64 / 96 / 128 / 192 / 256 segments 1.596 / 1.537 / 1.527 / 1.572 / 1.589 s.
  if ((int)e->classes.size() > kMaxClass) return fail(kErrUnsupported, "too many slot classes");
  int64_t tiles_all = 0;
  for (const auto& c : e->classes) tiles_all += c.n_tiles;
  auto nseg_of = [&](const rsp_emc::Class& c, int64_t nb) {
    const int64_t tiles = e->class_streams >= 2 ? std::max<int64_t>(1, tiles_all) : c.n_tiles;
    int64_t ns = e->opt_nseg > 0 ? e->opt_nseg
                                 : std::max(ceil_div(4 * 256, tiles), ceil_div(nb, 128));
    ns = std::max<int64_t>(1, std::min<int64_t>(ns, nb));
    if (ns >= 8 && e->opt_nseg == 0) ns = ns / 8 * 8;
    return (int)ns;
  };
  int nseg_max = 1;
  for (const auto& c : e->classes) nseg_max = std::max(nseg_max, nseg_of(c, batch));
  if ((int64_t)nseg_max * e->Npad > e->partial_cap) {
    if ((rc = dev_alloc(&e->d_partials, (int64_t)nseg_max * e->Npad))) return rc;
    e->partial_cap = (int64_t)nseg_max * e->Npad;
  }
  // 1-2
  hipLaunchKernelGGL(k_holo_states, dim3(e->S), dim3(256), 0, st, model_c, e->d_sx, e->d_sy,
                     e->d_dia, e->d_bg, e->d_probmask, e->size, e->d_I, msums);
  RSP_LAUNCH_CHECK("k_holo_states");
  hipLaunchKernelGGL(k_vscale_init, dim3(1), dim3(256), 0, st, msums, e->S, e->powder_sel_sum,
                     e->d_init, e->d_vscale);
  RSP_LAUNCH_CHECK("k_vscale_init");
  hipLaunchKernelGGL(k_scatter_scale, dim3((unsigned)ceil_div(e->Npad, 256)), dim3(256), 0, st,
                     scales, e->d_perm, (int)e->Npad, e->d_scale_sorted);
  RSP_LAUNCH_CHECK("k_scatter_scale");
  RSP_HIP(hipMemsetAsync(e->d_errmax, 0, sizeof(unsigned) * e->n_scored, st));
  RSP_HIP(hipMemsetAsync(e->d_absmax, 0, sizeof(unsigned) * e->n_scored, st));
  RSP_HIP(hipMemsetAsync(e->d_counters, 0, sizeof(unsigned) * 4, st));
  const size_t lds_bytes = 0;  // the gather's three part buffers are static LDS
  // The slot classes of a batch read the same tables and write disjoint partials: with
  // class_streams side streams (default priority) they run concurrently, so that the
  // last, partly filled wave of one class's workgroups shares the CUs with the next
  // class's instead of leaving them idle.  Fork after k_tables, join before the merge.
  const int ncs = e->class_streams;
  for (int k = 0; k < ncs; ++k)
    if (!e->side[k]) RSP_HIP(hipStreamCreateWithFlags(&e->side[k], hipStreamNonBlocking));
  for (int k = 0; k < ncs; ++k)
    if (!e->ev_join[k]) RSP_HIP(hipEventCreateWithFlags(&e->ev_join[k], hipEventDisableTiming));
  if (ncs > 0 && !e->ev_fork) RSP_HIP(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
  if (!e->exact_mode) {
    const int HC = 32;  // hypotheses per k_tables thread (same-process A/B: 8, 16, 32, 64,
                        // 128 -> 180, 168, 155-162, 163, 176 ms per iteration)
    int first = 1;
    const int64_t nbatch = ceil_div(e->H, batch);
    if (e->timing) {
      while ((int64_t)e->ev.size() < 4 * nbatch) {
        hipEvent_t ev;
        RSP_HIP(hipEventCreate(&ev));
        e->ev.push_back(ev);
      }
    }
    int64_t ib = 0;
    for (int64_t hb = 0; hb < e->H; hb += batch, ++ib) {
      const int nb = (int)std::min<int64_t>(batch, e->H - hb);
      SegMap sm;
      sm.n = (int)e->classes.size();
      dim3 tg((unsigned)ceil_div(e->n_scored, 256), (unsigned)ceil_div(nb, HC));
      if (e->timing) RSP_HIP(hipEventRecord(e->ev[4 * ib + 0], st));
      hipLaunchKernelGGL(k_tables, tg, dim3(256), 0, st, e->d_I, e->d_bg, e->d_scored_t,
                         e->n_scored, e->size, e->R, e->d_rcos, e->d_rsin, (int)hb, nb, HC,
                         e->PSTRIDE, e->TS, e->d_tables, e->d_errmax, e->d_absmax);
      RSP_LAUNCH_CHECK("k_tables");
      if (e->timing) RSP_HIP(hipEventRecord(e->ev[4 * ib + 1], st));
      if (e->timing) RSP_HIP(hipEventRecord(e->ev[4 * ib + 2], st));
      if (ncs > 0) {
        RSP_HIP(hipEventRecord(e->ev_fork, st));
        for (int k = 0; k < ncs; ++k) RSP_HIP(hipStreamWaitEvent(e->side[k], e->ev_fork, 0));
      }
      for (int ci = 0; ci < sm.n; ++ci) {
        const auto& c = e->classes[ci];
        const GatherCfg& gc = kGatherCfgs[c.cfg];
        const int ns = nseg_of(c, nb);
        sm.gw_end[ci] = c.gw0 + c.n_tiles * gc.NW * gc.F;
        sm.nseg[ci] = ns;
        hipStream_t cs = ncs > 0 ? e->side[ci % ncs] : st;
        hipLaunchKernelGGL(gc.fn, dim3((unsigned)(ns * c.n_tiles)), dim3(64 * gc.NW), lds_bytes,
                           cs, e->d_tables, e->TS, (int)hb, nb, ns, c.n_tiles, c.gw0,
                           e->d_slots + c.slot_off, e->d_scale_sorted, e->d_init, e->R,
                           e->d_partials, (int)e->Npad, (const uint32_t*)e->d_lay);
      }
      RSP_LAUNCH_CHECK("k_gather");
      for (int k = 0; k < ncs; ++k) {
        RSP_HIP(hipEventRecord(e->ev_join[k], e->side[k]));
        RSP_HIP(hipStreamWaitEvent(st, e->ev_join[k], 0));
      }
      if (e->timing) RSP_HIP(hipEventRecord(e->ev[4 * ib + 3], st));
      hipLaunchKernelGGL(k_merge_top3, dim3((unsigned)ceil_div(kMergeLanes * e->Npad, 256)),
                         dim3(256), 0, st,
                         e->d_partials, sm, (int)e->Npad, e->d_running, first);
      RSP_LAUNCH_CHECK("k_merge_top3");
      first = 0;
    }
  }
  ExactArgs A{e->d_I, e->d_bg, e->d_rcos, e->d_rsin, e->d_init, scales, e->d_ex_off,
              e->d_ex_n1, e->d_ex_pix, e->d_ex_cnt, e->size, e->R};
  hipLaunchKernelGGL(k_resolve, dim3((unsigned)ceil_div(e->Npad, 256)), dim3(256), 0, st, A,
                     e->d_running, e->d_perm, (int)e->Npad, e->d_compact_of, e->d_errmax,
                     e->d_absmax, e->exact_mode, rmax, maxll, e->d_full_list, e->d_counters);
  RSP_LAUNCH_CHECK("k_resolve");
  unsigned cnt[4] = {0, 0, 0, 0};
  RSP_HIP(hipMemcpyAsync(cnt, e->d_counters, sizeof(cnt), hipMemcpyDeviceToHost, st));
  RSP_HIP(hipMemcpyAsync(vscale, e->d_vscale, sizeof(double), hipMemcpyDeviceToHost, st));
  RSP_HIP(hipStreamSynchronize(st));
  const int nfull = (int)cnt[2];
  if (nfull > 0) {
    const int nbx = (int)ceil_div(e->H, 256);
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(nfull, (64LL << 20) / nbx));
    if ((int64_t)chunk * nbx > e->bcap) {
      if ((rc = dev_alloc(&e->d_bval, (int64_t)chunk * nbx)) ||
          (rc = dev_alloc(&e->d_bidx, (int64_t)chunk * nbx)))
        return rc;
      e->bcap = (int64_t)chunk * nbx;
    }
    for (int f0 = 0; f0 < nfull; f0 += chunk) {
      const int nf = std::min(chunk, nfull - f0);
      hipLaunchKernelGGL(k_full_search, dim3((unsigned)nbx, (unsigned)nf), dim3(256), 0, st, A,
                         e->d_perm, e->d_full_list + f0, (int)e->H, e->d_bval, e->d_bidx, nbx);
      RSP_LAUNCH_CHECK("k_full_search");
      hipLaunchKernelGGL(k_full_reduce, dim3((unsigned)ceil_div(nf, 256)), dim3(256), 0, st,
                         e->d_perm, e->d_full_list + f0, nf, e->d_bval, e->d_bidx, nbx, rmax,
                         maxll);
      RSP_LAUNCH_CHECK("k_full_reduce");
    }
    RSP_HIP(hipStreamSynchronize(st));
  }
  if (e->timing && !e->exact_mode) {
    const int64_t nbatch = ceil_div(e->H, batch);
    e->t_tables_ms = e->t_gather_ms = 0.;
    for (int64_t ib = 0; ib < nbatch; ++ib) {
      float a = 0.f, b = 0.f;
      RSP_HIP(hipEventElapsedTime(&a, e->ev[4 * ib + 0], e->ev[4 * ib + 1]));
      RSP_HIP(hipEventElapsedTime(&b, e->ev[4 * ib + 2], e->ev[4 * ib + 3]));
      e->t_tables_ms += a;
      e->t_gather_ms += b;
    }
    e->n_batches = nbatch;
  }
  e->stats[0] = e->N;
  e->stats[1] = e->H;
  e->stats[2] = cnt[0];
  e->stats[3] = cnt[1];
  e->stats[4] = cnt[2];
  e->stats[5] = e->P;
  e->stats[6] = e->NW;
  e->stats[7] = e->NP;
  return kOk;
}

// Host-side check of the photon slot scheduler and the table layouts (no device calls):
// `trials` random lane x bank groups (32 and 64 banks, uniform and clustered entries)
// through schedule_group, and random tiles (1-2 parts, 8/12/16 waves) through tile_layout.
// Each group is scheduled twice: by schedule_group and by schedule_slots with broadcast
// (one entry per bank and slot, shared by any number of lanes).
// out[0] groups, out[1] conflicts (a lane twice in one slot, a bank twice in one slot --
// with broadcast: two different entries of a bank in one slot --, an entry without a valid
// slot), out[2] groups whose exact slot count is not the maximum degree or whose broadcast
// schedule takes more slots than the exact one, out[3] layouts,
// out[4] layout faults (an LDS row that is not a permutation of its table row, sigma and
// inv not inverse, a direct-to-LDS source offset outside its chunk or not landing where
// map() puts the entry), out[5] slots the broadcast schedules saved in total, out[6] the
// broadcast slot count of a fixed half-wave whose lanes share five entries of one bank
// (6 is optimal).
extern "C" int rsp_emc_schedule_check(int64_t trials, uint64_t seed, int64_t* out) {
  if (!out) return fail(kErrNull, "rsp_emc_schedule_check: null pointer");
  if (trials < 0) return fail(kErrSize, "rsp_emc_schedule_check: trials < 0");
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
  auto rnd = [&](uint64_t n) {  // splitmix64
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (z ^ (z >> 31)) % n;
  };
  for (int i = 0; i < 7; ++i) out[i] = 0;
  for (int64_t it = 0; it < trials; ++it) {
    const int NB = it % 3 == 0 ? 64 : 32, G = NB;
    const int mean = 1 + (int)rnd(40);
    // entries: uniform, clustered in one bank, or drawn half of the time from 48 shared
    // "bright" entries (lanes then hold common entries, and repeats within a lane)
    const int kind = it % 5 == 0 ? 1 : it % 5 == 1 ? 2 : 0;
    std::vector<std::vector<int>> ent(G);
    for (int l = 0; l < G; ++l) {
      const int n = (int)rnd(2 * mean);
      for (int k = 0; k < n; ++k)
        ent[l].push_back(kind == 1 ? (int)rnd(7) * 32
                         : kind == 2 && rnd(2) ? (int)rnd(48) * 37 % kPartMax
                                               : (int)rnd(kPartMax));
    }
    std::vector<std::vector<int>> so, sb;
    const int ns = schedule_group(ent, NB, so);
    int D = 0;
    std::vector<int> bd(NB, 0);
    for (int l = 0; l < G; ++l) {
      D = std::max(D, (int)ent[l].size());
      for (int v : ent[l]) ++bd[v % NB];
    }
    for (int x : bd) D = std::max(D, x);
    out[2] += ns != D;
    std::vector<char> lu((size_t)G * (ns + 1), 0), bu((size_t)NB * (ns + 1), 0);
    for (int l = 0; l < G; ++l)
      for (size_t r = 0; r < ent[l].size(); ++r) {
        const int k = (l < (int)so.size() && r < so[l].size()) ? so[l][r] : -1;
        if (k < 0 || k >= ns) {
          ++out[1];
          continue;
        }
        out[1] += lu[(size_t)l * (ns + 1) + k]++ > 0;
        out[1] += bu[(size_t)(ent[l][r] % NB) * (ns + 1) + k]++ > 0;
      }
    // the broadcast-aware schedule: one slot per lane entry, and in each (bank, slot) one
    // entry, whatever the number of lanes reading it; never more slots than the above
    const int nb = schedule_slots(ent, NB, true, sb);
    out[2] += nb > ns;
    std::vector<int> lv((size_t)G * (nb + 1), -1), bv((size_t)NB * (nb + 1), -1);
    for (int l = 0; l < G; ++l)
      for (size_t r = 0; r < ent[l].size(); ++r) {
        const int k = (l < (int)sb.size() && r < sb[l].size()) ? sb[l][r] : -1;
        if (k < 0 || k >= nb) {
          ++out[1];
          continue;
        }
        out[1] += lv[(size_t)l * (nb + 1) + k] >= 0;
        lv[(size_t)l * (nb + 1) + k] = ent[l][r];
        int& b = bv[(size_t)(ent[l][r] % NB) * (nb + 1) + k];
        out[1] += b >= 0 && b != ent[l][r];
        b = ent[l][r];
      }
    out[5] += ns - nb;
    ++out[0];
  }
  {
    // every lane of a half-wave holds the same five entries of bank 0 and one of its own:
    // the exact schedule needs 161 slots (bank 0's degree), the broadcast one 6
    std::vector<std::vector<int>> ent(32), sb;
    for (int l = 0; l < 32; ++l) ent[l] = {0, 32, 64, 96, 128, 200 + l};
    const int nb = schedule_slots(ent, 32, true, sb);
    std::vector<int> bv((size_t)32 * 256, -1);
    for (int l = 0; l < 32 && nb > 0; ++l) {
      std::vector<char> lu(nb, 0);
      for (size_t r = 0; r < ent[l].size(); ++r) {
        const int k = sb[l][r];
        if (k < 0 || k >= nb || lu[k]++) {
          ++out[1];
          continue;
        }
        int& b = bv[(size_t)(ent[l][r] % 32) * 256 + k];
        out[1] += b >= 0 && b != ent[l][r];
        b = ent[l][r];
      }
    }
    out[6] = nb;
  }
  const int64_t ntiles = std::max<int64_t>(1, trials / 500);
  for (int64_t it = 0; it < ntiles; ++it)
    for (int NP = 1; NP <= 2; ++NP)
      for (int NW : {16, 12, 8}) {
        std::vector<Placed> pl(NW);
        for (auto& P : pl) {
          P.assign(NP, std::vector<std::vector<int>>(64));
          for (int p = 0; p < NP; ++p)
            for (int l = 0; l < 64; ++l)
              for (int k = 0; k < 30; ++k)
                P[p][l].push_back(rnd(3) == 0 ? -1 : (int)rnd(kPartMax));
        }
        std::vector<const Placed*> ps;
        for (auto& P : pl) ps.push_back(&P);
        TileLayout T;
        tile_layout(NP, NW, ps, T);
        int64_t bad = 0;
        for (int p = 0; p < NP; ++p) {
          for (int r = 0; r < kChunks * 8; ++r) {
            int seen[32] = {0};
            for (int i = r * 32; i < r * 32 + 32; ++i) {
              const int m = T.map(p, i);
              if (m / 32 != r) ++bad;
              else ++seen[m % 32];
            }
            for (int x : seen) bad += x != 1;
          }
          for (int q = 0; q < NW * 64; ++q) {
            const uint8_t* sg = &T.sigma[((size_t)p * NW * 8 + q / 8) * 8];
            const uint8_t* iv = &T.inv[((size_t)p * NW * 8 + q / 8) * 8];
            bad += sg[iv[q % 8]] != q % 8;
          }
          for (int wid = 0; wid < NW; ++wid)
            for (int c = wid; c < kChunks - 1; c += NW)
              for (int l = 0; l < 64; ++l) {
                const unsigned off = (unsigned)(l >> 3) * 128u +
                    16u * T.sigma[(((size_t)p * NW + wid) * 8 + (l >> 3)) * 8 + (l & 7)];
                if (off + 16 > 1024) ++bad;
                for (int j = 0; j < 4; ++j)
                  bad += T.map(p, c * 256 + (int)off / 4 + j) != c * 256 + 4 * l + j;
              }
        }
        ++out[3];
        out[4] += bad;
      }
  return kOk;
}

extern "C" int rsp_emc_stats(const rsp_emc* e, int64_t* out) {
  if (!e || !out) return fail(kErrNull, "rsp_emc_stats: null pointer");
  for (int i = 0; i < 8; ++i) out[i] = e->stats[i];
  return kOk;
}

extern "C" int rsp_emc_timing(const rsp_emc* e, double* out) {
  if (!e || !out) return fail(kErrNull, "rsp_emc_timing: null pointer");
  out[0] = e->t_tables_ms;
  out[1] = e->t_gather_ms;
  out[2] = (double)e->n_batches;
  out[3] = (double)e->batch;
  // LDS traffic of the gather per hypothesis, from the plan's slot classes: every slot
  // register is one 4-byte read per lane (zero pads included); every tile loads NW * L
  // 1 KB chunks per part (surplus loads included)
  double rd = 0., wr = 0.;
  for (const auto& c : e->classes) {
    const GatherCfg& g = kGatherCfgs[c.cfg];
    const int L = (kChunks + g.NW - 1) / g.NW;
    rd += (double)c.n_tiles * g.NW * g.F * 64. * g.S * g.NP * 4.;
    wr += (double)c.n_tiles * g.NP * g.NW * L * 1024.;
  }
  out[4] = rd;
  out[5] = wr;
  return kOk;
}

extern "C" int rsp_emc_destroy(rsp_emc* e) {
  if (!e) return kOk;
  hipSetDevice(e->device);
  free_all(e);
  delete e;
  return kOk;
}
