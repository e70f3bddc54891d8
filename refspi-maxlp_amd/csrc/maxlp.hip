// MaxLP object update for gfx950: the replacement of maxLP.py:304-340 (_run_phaser).
//
// Reference path: for every model pixel t (rmask annulus first, then the full mask) a
// sequential Python pattern search (maxLP.py:154-179) evaluates 33x33 complex candidates
// f_c = f + pattern_c*step, each with LL_c = mean_d[K_d log w_dc] - mean_d[w_dc],
// w_dc = rescale*|f_c + R_d(t)|^2 over all N frames (maxLP.py:181-197), K_d(t) being the
// photon count of frame d at the detector pixel rint(R(-a_d) q_t size + c)
// (maxLP.py:245-253).
//
// Device form (all pixels of a phase in parallel, one workgroup per pixel):
//   * model-space photon lists: every detector photon (d, p, K) is pulled back to the
//     model pixels t whose detector index is p (checked with the exact forward formula
//     in a 5x5 window), giving per pixel the frames with K_d(t) > 0.  Counting, scan and
//     a radix sort by (t, d) make the lists deterministic.
//   * mean_d w_dc = rescale*(|f_c + m1|^2 + v) from the per-pixel moments
//     m1 = mean_d R_d(t), v = mean_d |R_d(t) - m1|^2 (two non-negative terms).
//   * sum_d K_d log|f_c + R_d|^2 = log prod_d |f_c + R_d|^(2K_d): running FP64 products
//     with frexp renormalisation, one log per candidate at the end.
//   * 4 entry groups x 256 threads x 5 candidates per workgroup (each group every fourth
//     of the pixel's entries, partial products combined in LDS); first-occurrence argmax
//     (NaN first, as NumPy/CuPy argmax) by wave shuffles and LDS.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "refspi_amd.h"
#include "rsp_common.hpp"

namespace rsp {

constexpr int kCand = 1089;       // 33 x 33 pattern (maxLP.py:45-48, nmax = 8)
constexpr int kCentre = 544;      // pattern index of 0 + 0j
// 256 threads x 5 candidates (1280 >= 1089) per team: measured on the 185^2 workload
// (100k frames), 0.20 s per MaxLP pass against 0.27 s for 128 x 9 and 0.42 s for 64 x 18 --
// more waves per pixel hide the FP64 latency better than more candidates per thread save
// on per-entry work
constexpr int kTeam = 256;        // threads holding the candidates of one pixel
constexpr int kPerThread = 5;     // candidates per thread (7 x 5 rows x 33 columns >= 33 x 33)
// Entry groups: a pixel's workgroup is kGroups teams, each taking every kGroups-th of the
// pixel's entries for all candidates; their partial products are combined in LDS every
// round.  Four waves per SIMD on one pixel keep the FP64 pipelines issuing where a single
// wave per SIMD waits on its dependency chains: the heaviest pixels (tens of thousands of
// entries x 10 rounds) otherwise run alone at the end of the launch at a fraction of the
// issue rate (185^2, 100k frames: CUs busy 40% of the launch, VALU 87% of the busy time).
constexpr int kGroups = 4;
constexpr int kThreads = kTeam * kGroups;
constexpr int kWaves = kThreads / 64;
constexpr int kTeamWaves = kTeam / 64;
constexpr int kMaxPiece = 8;      // largest photon count of one model-space entry

struct Entry {
  double rr, ri, k;
};

struct Geo {
  int size, npix;
  const double* qx;
  const double* qy;
  const double* qrad;
};

// maxLP.py:250-252: x, y = rint(rots @ q * size + size//2), flat index x*size + y with no
// clamp.  At even sizes x or y reaches `size` for pixels near |q| = 0.5: y = size is the
// first pixel of row x+1; x = size (index >= npix) wraps modulo npix, as the photons
// matrix's fancy indexing reduces column indices (cupyx _asindices: x % length).
__device__ __forceinline__ int detector_index(const Geo& g, int t, double c, double s) {
  const double qx = g.qx[t], qy = g.qy[t];
  const double px = (c * qx + (-s) * qy) * g.size + (double)(g.size / 2);
  const double py = (s * qx + c * qy) * g.size + (double)(g.size / 2);
  const int x = (int)rint(px), y = (int)rint(py);
  int idx = (x * g.size + y) % g.npix;
  if (idx < 0) idx += g.npix;
  return idx;
}

// maxLP.py:238-243 get_fref_d for one (pixel, frame)
__device__ __forceinline__ void fref(const Geo& g, int t, double dia, double sx, double sy,
                                     double& rr, double& ri) {
  const double pi = M_PI;
  const double sval = pi * g.qrad[t] * dia;
  // s**3 as (s*s)*s, as the E-step's sphere_ff: within two ulps of pow(s, 3), without
  // pow's log and exp
  double ss, cc;
  sincos(sval, &ss, &cc);
  const double f = (ss - sval * cc) / ((sval * sval) * sval);
  const double theta = (2. * pi) * (g.qx[t] * sx + g.qy[t] * sy);
  double sn, cs;
  sincos(theta, &sn, &cs);
  rr = f * cs;
  ri = f * sn;
}

// Entry keys: model pixel t (bits 47-63), photon count K of the piece (43-46), frame d
// (18-42).  Sorted ascending, a pixel's entries run by count, then by frame: the pattern
// search then meets long runs of K = 1 entries (its straight-line path) instead of a
// K > 1 entry in a third of its blocks of four.
constexpr int kKeyD = 18, kKeyK = 43, kKeyT = 47;
__host__ __device__ inline uint64_t entry_key(int t, int K, int d) {
  return ((uint64_t)t << kKeyT) | ((uint64_t)K << kKeyK) | ((uint64_t)d << kKeyD);
}
__device__ __forceinline__ int key_t(uint64_t k) { return (int)(k >> kKeyT); }
__device__ __forceinline__ int key_k(uint64_t k) { return (int)((k >> kKeyK) & 15); }
__device__ __forceinline__ int key_d(uint64_t k) { return (int)((k >> kKeyD) & ((1u << 25) - 1)); }

// Pull-back of detector photons to model pixels.  Photon e belongs to frame fr[e] at
// detector pixel pix[e] with count cnt[e].  mode 0: count matches; mode 1: write keys.
__global__ __launch_bounds__(256) void k_pull(Geo g, const int32_t* __restrict__ mask,
                                              int64_t nph, const int32_t* __restrict__ fr,
                                              const int32_t* __restrict__ pix,
                                              const int32_t* __restrict__ cnt,
                                              const double* __restrict__ rc,
                                              const double* __restrict__ rs, int mode,
                                              int32_t* __restrict__ nmatch,
                                              const int64_t* __restrict__ pos,
                                              uint64_t* __restrict__ keys) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= nph) return;
  const int d = fr[e], p = pix[e];
  const double c = rc[d], s = rs[d];
  const int c0 = g.size / 2;
  const int px = p / g.size, py = p % g.size;
  int n = 0;
  int64_t w = mode ? pos[e] : 0;
  // Model pixels whose detector index is p lie within 0.71 px of the inverse rotation of
  // an unreduced index (X, Y) with X*size + Y == p (mod npix): (px, py) itself, and at the
  // wrap-around of detector_index (even sizes) (px-1, size) for py == 0 and (size, py) for
  // px == 0.  Each is searched in a 5x5 window, skipping the cells of earlier windows (so
  // no model pixel is found twice), and every candidate is checked with the forward index.
  int u0[3], v0[3], nw = 0;
  for (int k = 0; k < 3; ++k) {
    int X, Y;
    if (k == 0) {
      X = px; Y = py;
    } else if (k == 1) {
      if (py != 0 || px < 1) continue;
      X = px - 1; Y = g.size;
    } else {
      if (px != 0) continue;
      X = g.size; Y = py;
    }
    const double Xc = (double)(X - c0), Yc = (double)(Y - c0);
    // inverse of [[c, -s], [s, c]]
    const double u = c * Xc + s * Yc + c0, v = -s * Xc + c * Yc + c0;
    u0[nw] = (int)rint(u);
    v0[nw] = (int)rint(v);
    ++nw;
  }
  for (int k = 0; k < nw; ++k)
    for (int du = -2; du <= 2; ++du)
      for (int dv = -2; dv <= 2; ++dv) {
        const int uu = u0[k] + du, vv = v0[k] + dv;
        if (uu < 0 || uu >= g.size || vv < 0 || vv >= g.size) continue;
        bool earlier = false;
        for (int j = 0; j < k; ++j) earlier |= (abs(uu - u0[j]) <= 2 && abs(vv - v0[j]) <= 2);
        if (earlier) continue;
        const int t = uu * g.size + vv;
        if (!mask[t]) continue;
        if (detector_index(g, t, c, s) != p) continue;
        // a count above 8 becomes pieces of at most 8 (k_pattern raises each factor to its
        // piece's count between renormalisations; the pieces multiply to the same power)
        int left = cnt[e];
        while (left > 0) {
          const int piece = left < kMaxPiece ? left : kMaxPiece;
          if (mode) keys[w++] = entry_key(t, piece, d);
          ++n;
          left -= piece;
        }
      }
  if (!mode) nmatch[e] = n;
}

__global__ __launch_bounds__(256) void k_bounds(const uint64_t* __restrict__ keys, int64_t n,
                                                int64_t* __restrict__ seg_beg,
                                                int64_t* __restrict__ seg_end) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int t = key_t(keys[i]);
  if (i == 0 || key_t(keys[i - 1]) != t) seg_beg[t] = i;
  if (i == n - 1 || key_t(keys[i + 1]) != t) seg_end[t] = i + 1;
}

__global__ __launch_bounds__(256) void k_entries(Geo g, const uint64_t* __restrict__ keys,
                                                 int64_t n, const double* __restrict__ sx,
                                                 const double* __restrict__ sy,
                                                 const double* __restrict__ dia,
                                                 Entry* __restrict__ ent) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  const int t = key_t(k);
  const int d = key_d(k);
  const int K = key_k(k);
  Entry e;
  fref(g, t, dia[d], sx[d], sy[d], e.rr, e.ri);
  e.k = (double)K;
  ent[i] = e;
}

// Per-pixel moments m1 = mean_d R_d(t) and v = mean_d |R_d(t) - m1|^2, so that
// mean_d |f + R_d|^2 = |f + m1|^2 + v is a sum of two non-negative terms with no
// cancellation, whatever f.  One pass over the frames with sums shifted by c = R_0(t) (the
// first frame's term, the same for every thread): S1 = sum_d (R_d - c), S2 = sum_d
// |R_d - c|^2, m1 = c + S1/N, v = S2/N - |S1/N|^2 (clamped at 0).  Relative to the two-pass
// form the subtraction costs a factor 1 + |m1 - c|^2 / v in v's rounding error, a few ulps
// for a frame of the set as the shift.  Fixed reduction order (deterministic).
__global__ __launch_bounds__(256) void k_moments(Geo g, const int32_t* __restrict__ pixels,
                                                 int N, const double* __restrict__ sx,
                                                 const double* __restrict__ sy,
                                                 const double* __restrict__ dia,
                                                 double* __restrict__ mom) {
  __shared__ double red[3][256];
  const int t = pixels[blockIdx.x];
  double cr, ci;
  fref(g, t, dia[0], sx[0], sy[0], cr, ci);
  double a = 0., b = 0., q = 0.;
  for (int d = threadIdx.x; d < N; d += 256) {
    double rr, ri;
    fref(g, t, dia[d], sx[d], sy[d], rr, ri);
    const double er = rr - cr, ei = ri - ci;
    a += er;
    b += ei;
    q += er * er + ei * ei;
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  red[2][threadIdx.x] = q;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double ar = red[0][0] / N, ai = red[1][0] / N;
    mom[4 * t + 0] = cr + ar;
    mom[4 * t + 1] = ci + ai;
    mom[4 * t + 2] = fmax(red[2][0] / N - (ar * ar + ai * ai), 0.);
  }
}

// Dispatch order of a k_pattern launch: key = (2^31 - 1 - entries of pixel t) << 32 | i for
// list position i, so that an ascending sort puts the pixels with the most entries first
// (ties in list order).  A pixel's work is proportional to its entries, which range over
// two orders of magnitude (bright pixels near the centre); started last, the heaviest
// would set the launch's tail.
__global__ __launch_bounds__(256) void k_order_keys(const int32_t* __restrict__ pixels, int n,
                                                    const int64_t* __restrict__ seg_beg,
                                                    const int64_t* __restrict__ seg_end,
                                                    uint64_t* __restrict__ keys) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int t = pixels[i];
  const int64_t len = min(seg_end[t] - seg_beg[t], (int64_t)0x7fffffff);
  keys[i] = ((uint64_t)(0x7fffffff - len) << 32) | (uint32_t)i;
}

__device__ __forceinline__ bool cand_better(double a, int ia, double b, int ib) {
  const bool na = isnan(a), nb = isnan(b);
  if (na || nb) return na && (!nb || ia < ib);
  return a > b || (a == b && ia < ib);
}

// mean_d |c + R_d|^2 from the moments (no cancellation: see k_moments)
__device__ __forceinline__ double mean_w(double cr, double ci, double m1r, double m1i, double v) {
  const double ar = cr + m1r, ai = ci + m1i;
  return (ar * ar + ai * ai) + v;
}

// One workgroup per pixel: up to 10 pattern-search steps (maxLP.py:154-179).
// rescale < 0 selects the per-candidate rescale of phase 1 (maxLP.py:187).  Work-group b
// takes list position i = low 32 bits of order[b] (k_order_keys, sorted); the pixel's
// model value and resc_out[i] are its only outputs, so the order changes no result.
// kProbe: one evaluation of the caller's candidates probe_cand[0..probe_n) at pixel probe_t
// (get_logq_pixel, maxLP.py:181-197) written to probe_out; same arithmetic as the search.
template <bool kProbe>
__global__ __launch_bounds__(kThreads) void k_pattern(
    const uint64_t* __restrict__ order, const int32_t* __restrict__ pixels,
    const int64_t* __restrict__ seg_beg,
    const int64_t* __restrict__ seg_end, const Entry* __restrict__ ent,
    const double* __restrict__ mom, int N, double rescale, double* __restrict__ model,
    double* __restrict__ resc_out, int probe_t, const double* __restrict__ probe_cand,
    int probe_n, double* __restrict__ probe_out, unsigned long long* __restrict__ counters) {
  // partial products of entry groups 1 .. kGroups-1 (normalised mantissa, exponent)
  __shared__ double cpm[kGroups - 1][kPerThread][kTeam];
  __shared__ int cpe[kGroups - 1][kPerThread][kTeam];
  __shared__ double wv[kWaves];
  __shared__ int wi[kWaves];
  __shared__ double skt;
  const int ct = (int)threadIdx.x % kTeam, grp = (int)threadIdx.x / kTeam;
  const int li = kProbe ? 0 : (int)(uint32_t)order[blockIdx.x];
  const int t = kProbe ? probe_t : pixels[li];
  const int64_t e0 = seg_beg[t], e1 = seg_end[t];
  // issue priority by work: the heaviest pixels (dispatched first) would otherwise share
  // their SIMDs evenly with the light ones that fill the CU around them and finish last;
  // with priority the arbiter issues their waves first and the light waves fill the gaps
  if (!kProbe) {
    const int64_t n = e1 - e0;
    if (n >= 16384) __builtin_amdgcn_s_setprio(3);
    else if (n >= 4096) __builtin_amdgcn_s_setprio(2);
    else if (n >= 1024) __builtin_amdgcn_s_setprio(1);
  }
  const double m1r = mom[4 * t + 0], m1i = mom[4 * t + 1], vr = mom[4 * t + 2];
  // K-bar = mean_d K_d (exact integer sum)
  {
    double ks = 0.;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kThreads) ks += ent[e].k;
    for (int off = 32; off > 0; off >>= 1) ks += __shfl_xor(ks, off);
    if ((threadIdx.x & 63) == 0) wv[threadIdx.x >> 6] = ks;
    __syncthreads();
    if (threadIdx.x == 0) {
      double st = 0.;
      for (int w = 0; w < kWaves; ++w) st += wv[w];
      skt = st / N;
    }
    __syncthreads();
  }
  const double kbar = skt;
  double fr = kProbe ? 0. : model[2 * t], fi = kProbe ? 0. : model[2 * t + 1];
  double step = hypot(fr, fi) / 4.;
  int rounds = 0;  // candidate evaluation rounds (search steps evaluated)
  // Candidates of this thread (the same in every entry group).  Search: pattern column
  // iy = ct % 33 and rows ix = 5 * (ct / 33) + u (231 threads of the team cover the 33 x 33
  // pattern; ix >= 33 and threads >= 231 hold none), so the imaginary part -- and im^2
  // below -- is shared by the thread's candidates: 17 FP64 operations per entry (2 shared,
  // then an add, a fused multiply-add and the product's multiply per candidate).  Probe: the
  // caller's candidates ct + 256 u.
  int cidx[kPerThread];
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    if (kProbe) {
      const int c = ct + kTeam * u;
      cidx[u] = c < probe_n ? c : -1;
    } else {
      const int ix = 5 * (ct / 33) + u;
      cidx[u] = ix < 33 ? ix * 33 + ct % 33 : -1;
    }
  }
  // this group's entries: every kGroups-th of the pixel's, from its grp-th (the same mix
  // of counts in every group)
  const int64_t s0 = e0 + grp;
  const int64_t gn = e1 > s0 ? (e1 - s0 + kGroups - 1) / kGroups : 0;
  for (int it = 0; it < (kProbe ? 1 : 10); ++it) {
    ++rounds;
    double cr[kPerThread], ci[kProbe ? kPerThread : 1], prod[kPerThread];
    int ex[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int cc = cidx[u] >= 0 ? cidx[u] : 0;
      if (kProbe) {
        cr[u] = probe_cand[2 * cc];
        ci[u] = probe_cand[2 * cc + 1];
      } else {
        const double px = -8. + 0.5 * (double)(cc / 33);
        cr[u] = fr + px * step;
        if (u == 0) {
          const double py = -8. + 0.5 * (double)(ct % 33);
          ci[0] = fi + py * step;
        }
      }
      prod[u] = 1.;
      ex[u] = 0;
    }
    // One entry: prod[u] *= |c_u + R|^2 (K times for a count K > 1), in entry order.
    auto factor = [&](double rr, double ri, int K) {
      // im^2 of the shared imaginary part (search), per candidate (probe)
      double b0 = 0.;
      if (!kProbe) {
        const double im = ci[0] + ri;
        b0 = im * im;
      }
#pragma unroll
      for (int u = 0; u < kPerThread; ++u) {
        const double re = cr[u] + rr;
        double b = b0;
        if (kProbe) {
          const double im = ci[kProbe ? u : 0] + ri;
          b = im * im;
        }
        // |c + R|^2 as one fused multiply-add (the re^2 product unrounded; the VALU issue
        // bounds this loop, and the fusion takes a quarter of its instructions)
        const double a = fma(re, re, b);
        double pr = prod[u] * a;
        for (int k = 1; k < K; ++k) pr = pr * a;  // 2 <= K <= 8: k_pull splits larger counts
        prod[u] = pr;
      }
    };
    auto renorm = [&]() {
#pragma unroll
      for (int u = 0; u < kPerThread; ++u) {
        int e;
        prod[u] = frexp(prod[u], &e);
        ex[u] += e;
      }
    };
    // The entries are the same for every thread of a group: they are read by scalar loads
    // (uniform addresses, into SGPRs), four at a time with the next four in flight, and used
    // as scalar operands -- no LDS staging, no barriers.  Renormalisation before an entry
    // whose K factors would bring more than 8 since the last one (nf counts them, a
    // wave-uniform scalar): at most 8 factors between renormalisations, so the products stay
    // within FP64's normal range whenever |f_c + R_d|^2 is in [2^-120, 2^120] (scaling by
    // powers of two is exact, so where the renormalisations fall does not change the
    // product's rounding; an exact zero stays zero, giving -inf as log 0 does in the
    // reference).
    int nf = 0;
    const int64_t n = gn;
    const int64_t n4 = n & ~(int64_t)3;
    Entry nx[4];
    if (n4 > 0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) nx[j] = ent[s0 + (int64_t)kGroups * j];
    }
    for (int64_t j0 = 0; j0 < n4; j0 += 4) {
      Entry cu[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) cu[j] = nx[j];
      if (j0 + 4 < n4) {
#pragma unroll
        for (int j = 0; j < 4; ++j) nx[j] = ent[s0 + (int64_t)kGroups * (j0 + 4 + j)];
      }
      if (cu[0].k == 1. && cu[1].k == 1. && cu[2].k == 1. && cu[3].k == 1.) {
        // four single-photon entries, straight-line (the common case)
        if (nf > 4) {
          renorm();
          nf = 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) factor(cu[j].rr, cu[j].ri, 1);
        nf += 4;
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int K = (int)cu[j].k;
          if (nf + K > 8) {
            renorm();
            nf = 0;
          }
          factor(cu[j].rr, cu[j].ri, K);
          nf += K;
        }
      }
    }
    for (int64_t j = n4; j < n; ++j) {
      const Entry E = ent[s0 + (int64_t)kGroups * j];
      const int K = (int)E.k;
      if (nf + K > 8) {
        renorm();
        nf = 0;
      }
      factor(E.rr, E.ri, K);
      nf += K;
    }
    // the groups' products: each normalised (at most 7 factors since the last
    // renormalisation), then group 0 multiplies the mantissas in group order (each in
    // [0.5, 1): no overflow or underflow) and adds the exponents
    renorm();
    if (grp > 0) {
#pragma unroll
      for (int u = 0; u < kPerThread; ++u) {
        cpm[grp - 1][u][ct] = prod[u];
        cpe[grp - 1][u][ct] = ex[u];
      }
    }
    __syncthreads();
    if (grp == 0) {
#pragma unroll
      for (int g = 0; g < kGroups - 1; ++g)
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
          prod[u] = prod[u] * cpm[g][u][ct];
          ex[u] += cpe[g][u][ct];
        }
    }
    // candidate values and the local first-occurrence argmax (group 0)
    double bv = -INFINITY;
    int bi = 0x7fffffff;
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
      const int c = cidx[u];
      if (c < 0 || grp > 0) continue;
      const double mc = mean_w(cr[u], ci[kProbe ? u : 0], m1r, m1i, vr);
      const double slog = (log(prod[u]) + (double)ex[u] * M_LN2) / N;
      const double r = rescale < 0. ? kbar / mc : rescale;
      const double v = kbar * log(r) + slog - r * mc;
      if (kProbe) probe_out[c] = v;
      if (cand_better(v, c, bv, bi)) {
        bv = v;
        bi = c;
      }
    }
    if (kProbe) return;
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off);
      const int oi = __shfl_xor(bi, off);
      if (cand_better(ov, oi, bv, bi)) {
        bv = ov;
        bi = oi;
      }
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0 && grp == 0) {
      wv[threadIdx.x >> 6] = bv;
      wi[threadIdx.x >> 6] = bi;
    }
    __syncthreads();
    double v0 = wv[0];
    int imax = wi[0];
    for (int w = 1; w < kTeamWaves; ++w)
      if (cand_better(wv[w], wi[w], v0, imax)) {
        v0 = wv[w];
        imax = wi[w];
      }
    const double px = -8. + 0.5 * (double)(imax / 33), py = -8. + 0.5 * (double)(imax % 33);
    const double nr = fr + px * step, ni = fi + py * step;
    const double ns = imax == kCentre ? step / 6.6 : step;
    if (ns / hypot(nr, ni) < 1.e-3) break;
    fr = nr;
    fi = ni;
    step = ns;
  }
  if (threadIdx.x == 0) {
    // work counters: entry x candidate products and search steps evaluated
    atomicAdd(&counters[0], (unsigned long long)rounds * (unsigned long long)(e1 - e0) *
                                (unsigned long long)kCand);
    atomicAdd(&counters[1], (unsigned long long)rounds);
    model[2 * t] = fr;
    model[2 * t + 1] = fi;
    if (resc_out) {
      // maxLP.py:216-228 get_rescale_pixel at the final value
      resc_out[li] = kbar / mean_w(fr, fi, m1r, m1i, vr);
    }
  }
}

// get_photons_pixel (maxLP.py:245-253) from the pulled-back entries of pixel t: K_d(t) for
// every frame (duplicate detector photons of a frame add up, as the photons CSR sums them).
__global__ __launch_bounds__(256) void k_photons_pixel(int t, const int64_t* __restrict__ seg_beg,
                                                       const int64_t* __restrict__ seg_end,
                                                       const uint64_t* __restrict__ keys,
                                                       double* __restrict__ k_out) {
  const int64_t e0 = seg_beg[t], e1 = seg_end[t];
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
    const uint64_t k = keys[e];
    atomicAdd(&k_out[key_d(k)], (double)key_k(k));
  }
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ x, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) x[i] = v;
}

}  // namespace rsp

using namespace rsp;

struct rsp_maxlp {
  int device = 0, size = 0, npix = 0;
  int64_t N = 0, nph = 0;
  std::vector<int32_t> rmask_pix, mask_pix;
  double *d_qx = nullptr, *d_qy = nullptr, *d_qrad = nullptr, *d_mom = nullptr,
         *d_resc = nullptr;
  int32_t *d_mask = nullptr, *d_rmask_pix = nullptr, *d_mask_pix = nullptr, *d_fr = nullptr,
          *d_pix = nullptr, *d_cnt = nullptr, *d_nmatch = nullptr;
  int64_t *d_pos = nullptr, *d_seg_beg = nullptr, *d_seg_end = nullptr;
  uint64_t *d_keys = nullptr, *d_keys2 = nullptr, *d_ord = nullptr, *d_ord2 = nullptr;
  Entry* d_ent = nullptr;
  int64_t key_cap = 0;
  void* d_tmp = nullptr;
  size_t tmp_cap = 0;
  double last_rescale = 0.;
  int64_t n_entries = 0;
  bool ran = false;
  // work counters of the last run (entry x candidate products, search steps) and the
  // k_pattern time of both phases (HIP events on the run's stream)
  unsigned long long* d_counters = nullptr;
  unsigned long long counters[2] = {0, 0};
  hipEvent_t ev[4] = {};
  // k_moments on a side stream beside the photon lists' pull-back, sort and entries
  hipStream_t side = nullptr;
  hipEvent_t sev[2] = {};
  double pattern_ms = 0.;
};

namespace {

template <typename T>
int alloc(T** p, int64_t n) {
  if (*p) hipFree(*p);
  *p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  return kOk;
}

template <typename T>
int upload(T** p, const std::vector<T>& v) {
  int rc = alloc(p, (int64_t)v.size());
  if (rc) return rc;
  if (!v.empty()) RSP_HIP(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return kOk;
}

int ensure_tmp(rsp_maxlp* m, size_t bytes) {
  if (bytes <= m->tmp_cap) return kOk;
  if (m->d_tmp) hipFree(m->d_tmp);
  m->d_tmp = nullptr;
  RSP_HIP(hipMalloc(&m->d_tmp, bytes));
  m->tmp_cap = bytes;
  return kOk;
}

}  // namespace

extern "C" int rsp_maxlp_create(int device, int64_t size, rsp_maxlp** out) {
  if (!out) return fail(kErrNull, "rsp_maxlp_create: null out");
  *out = nullptr;
  if (size < 8 || size > 362) return fail(kErrSize, "rsp_maxlp_create: size out of range");
  RSP_HIP(hipSetDevice(device));
  rsp_maxlp* m = new rsp_maxlp();
  m->device = device;
  m->size = (int)size;
  m->npix = (int)(size * size);
  // maxLP.py:79-95
  std::vector<double> qx(m->npix), qy(m->npix), qrad(m->npix);
  std::vector<int32_t> mask(m->npix);
  const int c = m->size / 2;
  for (int x = 0; x < m->size; ++x)
    for (int y = 0; y < m->size; ++y) {
      const int t = x * m->size + y;
      qx[t] = (double)(x - c) / (double)m->size;
      qy[t] = (double)(y - c) / (double)m->size;
      double q = std::sqrt(qx[t] * qx[t] + qy[t] * qy[t]);
      if (q == 0.) q = 1.e-6;
      qrad[t] = q;
      const bool mk = (q > 4. / m->size) && (q < 0.5);
      const bool rk = (q > 0.39) && (q < 0.41);
      mask[t] = mk;
      if (mk) m->mask_pix.push_back(t);
      if (rk) m->rmask_pix.push_back(t);
    }
  int rc;
  if ((rc = upload(&m->d_qx, qx)) || (rc = upload(&m->d_qy, qy)) ||
      (rc = upload(&m->d_qrad, qrad)) || (rc = upload(&m->d_mask, mask)) ||
      (rc = upload(&m->d_rmask_pix, m->rmask_pix)) || (rc = upload(&m->d_mask_pix, m->mask_pix)) ||
      (rc = alloc(&m->d_mom, 4LL * m->npix)) || (rc = alloc(&m->d_seg_beg, m->npix)) ||
      (rc = alloc(&m->d_seg_end, m->npix)) ||
      (rc = alloc(&m->d_resc, std::max<int64_t>(1, (int64_t)m->rmask_pix.size()))) ||
      (rc = alloc(&m->d_ord, (int64_t)std::max(m->rmask_pix.size(), m->mask_pix.size()))) ||
      (rc = alloc(&m->d_ord2, (int64_t)std::max(m->rmask_pix.size(), m->mask_pix.size()))) ||
      (rc = alloc(&m->d_counters, 2))) {
    rsp_maxlp_destroy(m);
    return rc;
  }
  for (hipEvent_t& e : m->ev) {
    const hipError_t he = hipEventCreate(&e);
    if (he != hipSuccess) {
      e = nullptr;
      rsp_maxlp_destroy(m);
      return hip_fail(he, "hipEventCreate");
    }
  }
  *out = m;
  return kOk;
}

extern "C" int rsp_maxlp_set_frames(rsp_maxlp* m, int64_t num_data, const int32_t* ones,
                                    const int32_t* multi, const int32_t* place_ones,
                                    const int32_t* place_multi, const int32_t* count_multi) {
  if (!m || !ones || !multi) return fail(kErrNull, "rsp_maxlp_set_frames: null pointer");
  if (num_data < 1 || num_data >= (1LL << 25))
    return fail(kErrSize, "rsp_maxlp_set_frames: num_data out of range");
  RSP_HIP(hipSetDevice(m->device));
  std::vector<int32_t> fr, pix, cnt;
  int64_t po = 0, pm = 0;
  for (int64_t d = 0; d < num_data; ++d) {
    if (ones[d] < 0 || multi[d] < 0) return fail(kErrSize, "negative photon count");
    for (int k = 0; k < ones[d]; ++k) {
      const int p = place_ones[po + k];
      if (p < 0 || p >= m->npix) return fail(kErrSize, "place_ones out of range");
      fr.push_back((int32_t)d);
      pix.push_back(p);
      cnt.push_back(1);
    }
    for (int k = 0; k < multi[d]; ++k) {
      const int p = place_multi[pm + k];
      const int cm = count_multi[pm + k];
      if (p < 0 || p >= m->npix) return fail(kErrSize, "place_multi out of range");
      if (cm < 1 || cm >= (1 << 22)) return fail(kErrSize, "count_multi out of range");
      fr.push_back((int32_t)d);
      pix.push_back(p);
      cnt.push_back(cm);
    }
    po += ones[d];
    pm += multi[d];
  }
  int rc;
  if ((rc = upload(&m->d_fr, fr)) || (rc = upload(&m->d_pix, pix)) || (rc = upload(&m->d_cnt, cnt)) ||
      (rc = alloc(&m->d_nmatch, (int64_t)fr.size())) || (rc = alloc(&m->d_pos, (int64_t)fr.size() + 1)))
    return rc;
  m->N = num_data;
  m->nph = (int64_t)fr.size();
  m->ran = false;
  return kOk;
}

extern "C" int rsp_maxlp_run(rsp_maxlp* m, double* model_c, const double* sx, const double* sy,
                             const double* dia, const double* rot_cos, const double* rot_sin,
                             double* rescale_out, rsp_stream_t stream_) {
  if (!m || !model_c || !sx || !sy || !dia || !rot_cos || !rot_sin)
    return fail(kErrNull, "rsp_maxlp_run: null pointer");
  if (m->N == 0) return fail(kErrState, "rsp_maxlp_run: frames not set");
  RSP_HIP(hipSetDevice(m->device));
  hipStream_t st = (hipStream_t)stream_;
  Geo g{m->size, m->npix, m->d_qx, m->d_qy, m->d_qrad};
  int rc;
  const int nmask = (int)m->mask_pix.size(), nr = (int)m->rmask_pix.size();
  // The per-pixel moments need only the frames' parameters: on a side stream they run beside
  // the photon lists' pull-back, sort and entries, after everything queued on st so far (the
  // previous run's pattern search read them; the parameters come from st), and the pattern
  // search waits for them.
  if (!m->side) RSP_HIP(hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking));
  for (hipEvent_t& e : m->sev)
    if (!e) RSP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  RSP_HIP(hipEventRecord(m->sev[0], st));
  RSP_HIP(hipStreamWaitEvent(m->side, m->sev[0], 0));
  if (nmask > 0) {
    hipLaunchKernelGGL(k_moments, dim3(nmask), dim3(256), 0, m->side, g, m->d_mask_pix,
                       (int)m->N, sx, sy, dia, m->d_mom);
    RSP_LAUNCH_CHECK("k_moments");
  }
  RSP_HIP(hipEventRecord(m->sev[1], m->side));
  // model-space photon lists
  int64_t total = 0;
  if (m->nph > 0) {
    hipLaunchKernelGGL(k_pull, dim3((unsigned)ceil_div(m->nph, 256)), dim3(256), 0, st, g,
                       m->d_mask, m->nph, m->d_fr, m->d_pix, m->d_cnt, rot_cos, rot_sin, 0,
                       m->d_nmatch, (const int64_t*)nullptr, (uint64_t*)nullptr);
    RSP_LAUNCH_CHECK("k_pull(count)");
    size_t tb = 0;
    RSP_HIP(rocprim::exclusive_scan(nullptr, tb, m->d_nmatch, m->d_pos, (int64_t)0,
                                    (size_t)m->nph, rocprim::plus<int64_t>(), st));
    if ((rc = ensure_tmp(m, tb))) return rc;
    RSP_HIP(rocprim::exclusive_scan(m->d_tmp, tb, m->d_nmatch, m->d_pos, (int64_t)0,
                                    (size_t)m->nph, rocprim::plus<int64_t>(), st));
    int64_t last_pos = 0;
    int32_t last_n = 0;
    RSP_HIP(hipMemcpyAsync(&last_pos, m->d_pos + (m->nph - 1), sizeof(int64_t),
                           hipMemcpyDeviceToHost, st));
    RSP_HIP(hipMemcpyAsync(&last_n, m->d_nmatch + (m->nph - 1), sizeof(int32_t),
                           hipMemcpyDeviceToHost, st));
    RSP_HIP(hipStreamSynchronize(st));
    total = last_pos + last_n;
  }
  m->n_entries = total;
  if (total > m->key_cap) {
    if ((rc = alloc(&m->d_keys, total)) || (rc = alloc(&m->d_keys2, total)) ||
        (rc = alloc(&m->d_ent, total)))
      return rc;
    m->key_cap = total;
  }
  RSP_HIP(hipMemsetAsync(m->d_counters, 0, 2 * sizeof(unsigned long long), st));
  RSP_HIP(hipMemsetAsync(m->d_seg_beg, 0, sizeof(int64_t) * m->npix, st));
  RSP_HIP(hipMemsetAsync(m->d_seg_end, 0, sizeof(int64_t) * m->npix, st));
  if (total > 0) {
    hipLaunchKernelGGL(k_pull, dim3((unsigned)ceil_div(m->nph, 256)), dim3(256), 0, st, g,
                       m->d_mask, m->nph, m->d_fr, m->d_pix, m->d_cnt, rot_cos, rot_sin, 1,
                       m->d_nmatch, (const int64_t*)m->d_pos, m->d_keys);
    RSP_LAUNCH_CHECK("k_pull(fill)");
    size_t tb = 0;
    // all 64 bits (bits 0-17 are zero; restricting the sort to bits 18-63 gave a wrong order
    // of the entries here: the MaxLP oracle test failed)
    RSP_HIP(rocprim::radix_sort_keys(nullptr, tb, m->d_keys, m->d_keys2, (size_t)total, 0, 64, st));
    if ((rc = ensure_tmp(m, tb))) return rc;
    RSP_HIP(rocprim::radix_sort_keys(m->d_tmp, tb, m->d_keys, m->d_keys2, (size_t)total, 0, 64, st));
    hipLaunchKernelGGL(k_bounds, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st,
                       m->d_keys2, total, m->d_seg_beg, m->d_seg_end);
    RSP_LAUNCH_CHECK("k_bounds");
    hipLaunchKernelGGL(k_entries, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st, g,
                       m->d_keys2, total, sx, sy, dia, m->d_ent);
    RSP_LAUNCH_CHECK("k_entries");
  }
  RSP_HIP(hipStreamWaitEvent(st, m->sev[1], 0));
  // the pixel lists of a phase, heaviest first: d_ord2 (sorted keys of k_order_keys)
  auto order = [&](const int32_t* pixels, int n) -> int {
    hipLaunchKernelGGL(k_order_keys, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, pixels,
                       n, m->d_seg_beg, m->d_seg_end, m->d_ord);
    RSP_LAUNCH_CHECK("k_order_keys");
    size_t tb = 0;
    RSP_HIP(rocprim::radix_sort_keys(nullptr, tb, m->d_ord, m->d_ord2, (size_t)n, 0, 64, st));
    int rc2 = ensure_tmp(m, tb);
    if (rc2) return rc2;
    RSP_HIP(rocprim::radix_sort_keys(m->d_tmp, tb, m->d_ord, m->d_ord2, (size_t)n, 0, 64, st));
    return kOk;
  };
  // phase 1: rmask annulus with per-candidate rescale (maxLP.py:317-320)
  RSP_HIP(hipEventRecord(m->ev[0], st));
  if (nr > 0) {
    if ((rc = order(m->d_rmask_pix, nr))) return rc;
    hipLaunchKernelGGL(k_pattern<false>, dim3(nr), dim3(kThreads), 0, st,
                       (const uint64_t*)m->d_ord2, m->d_rmask_pix, m->d_seg_beg, m->d_seg_end,
                       m->d_ent, m->d_mom, (int)m->N, -1., model_c, m->d_resc, 0,
                       (const double*)nullptr, 0, (double*)nullptr, m->d_counters);
    RSP_LAUNCH_CHECK("k_pattern(phase1)");
  }
  RSP_HIP(hipEventRecord(m->ev[1], st));
  // phase 2: mean rescale over the annulus (maxLP.py:324-329)
  double rescale = 0.;
  if (nr > 0) {
    std::vector<double> r(nr);
    RSP_HIP(hipMemcpyAsync(r.data(), m->d_resc, sizeof(double) * nr, hipMemcpyDeviceToHost, st));
    RSP_HIP(hipStreamSynchronize(st));
    // pairwise summation, as numpy.mean does for a contiguous float64 array
    std::vector<double> buf(r);
    double s = 0.;
    // numpy's pairwise sum: blocks of 8 accumulators up to 128 elements, recursive halves
    struct PW {
      static double sum(const double* a, int64_t n) {
        if (n < 8) {
          double res = 0.;
          for (int64_t i = 0; i < n; ++i) res += a[i];
          return res;
        }
        if (n <= 128) {
          double r8[8];
          for (int k = 0; k < 8; ++k) r8[k] = a[k];
          int64_t i;
          for (i = 8; i < n - (n % 8); i += 8)
            for (int k = 0; k < 8; ++k) r8[k] += a[i + k];
          double res = ((r8[0] + r8[1]) + (r8[2] + r8[3])) + ((r8[4] + r8[5]) + (r8[6] + r8[7]));
          for (; i < n; ++i) res += a[i];
          return res;
        }
        int64_t n2 = n / 2;
        n2 -= n2 % 8;
        return sum(a, n2) + sum(a + n2, n - n2);
      }
    };
    s = PW::sum(buf.data(), nr);
    rescale = s / (double)nr;
  }
  m->last_rescale = rescale;
  if (rescale_out) *rescale_out = rescale;
  // phase 3: full mask with the global rescale (maxLP.py:332-336)
  RSP_HIP(hipEventRecord(m->ev[2], st));
  if (nmask > 0) {
    if ((rc = order(m->d_mask_pix, nmask))) return rc;
    hipLaunchKernelGGL(k_pattern<false>, dim3(nmask), dim3(kThreads), 0, st,
                       (const uint64_t*)m->d_ord2, m->d_mask_pix, m->d_seg_beg, m->d_seg_end,
                       m->d_ent, m->d_mom, (int)m->N, rescale, model_c, (double*)nullptr, 0,
                       (const double*)nullptr, 0, (double*)nullptr, m->d_counters);
    RSP_LAUNCH_CHECK("k_pattern(phase3)");
  }
  RSP_HIP(hipEventRecord(m->ev[3], st));
  RSP_HIP(hipMemcpyAsync(m->counters, m->d_counters, 2 * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, st));
  RSP_HIP(hipStreamSynchronize(st));
  // both phases' k_pattern launches (the dispatch-order sort between them is excluded)
  float t1 = 0.f, t3 = 0.f;
  RSP_HIP(hipEventElapsedTime(&t1, m->ev[0], m->ev[1]));
  RSP_HIP(hipEventElapsedTime(&t3, m->ev[2], m->ev[3]));
  m->pattern_ms = (double)t1 + (double)t3;
  m->ran = true;
  return kOk;
}

extern "C" int rsp_maxlp_photons_pixel(const rsp_maxlp* m, int64_t t, double* k_out,
                                       rsp_stream_t stream_) {
  if (!m || !k_out) return fail(kErrNull, "rsp_maxlp_photons_pixel: null pointer");
  if (t < 0 || t >= m->npix) return fail(kErrSize, "rsp_maxlp_photons_pixel: pixel out of range");
  if (!m->ran) return fail(kErrState, "rsp_maxlp_photons_pixel: no rsp_maxlp_run yet");
  RSP_HIP(hipSetDevice(m->device));
  hipStream_t st = (hipStream_t)stream_;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)ceil_div(m->N, 256)), dim3(256), 0, st, k_out, m->N,
                     0.);
  RSP_LAUNCH_CHECK("k_fill");
  if (m->n_entries > 0) {
    hipLaunchKernelGGL(k_photons_pixel, dim3(1), dim3(256), 0, st, (int)t, m->d_seg_beg,
                       m->d_seg_end, m->d_keys2, k_out);
    RSP_LAUNCH_CHECK("k_photons_pixel");
  }
  return kOk;
}

extern "C" int rsp_maxlp_logq_pixel(const rsp_maxlp* m, int64_t t, const double* cand,
                                    int64_t ncand, double rescale, double* out,
                                    rsp_stream_t stream_) {
  if (!m || !cand || !out) return fail(kErrNull, "rsp_maxlp_logq_pixel: null pointer");
  if (ncand < 1 || ncand > kTeam * kPerThread)
    return fail(kErrSize, "rsp_maxlp_logq_pixel: 1 <= ncand <= 1280");
  if (t < 0 || t >= m->npix) return fail(kErrSize, "rsp_maxlp_logq_pixel: pixel out of range");
  if (!std::binary_search(m->mask_pix.begin(), m->mask_pix.end(), (int32_t)t))
    return fail(kErrSize, "rsp_maxlp_logq_pixel: pixel outside the MaxLP mask");
  if (!m->ran) return fail(kErrState, "rsp_maxlp_logq_pixel: no rsp_maxlp_run yet");
  RSP_HIP(hipSetDevice(m->device));
  hipStream_t st = (hipStream_t)stream_;
  hipLaunchKernelGGL(k_pattern<true>, dim3(1), dim3(kThreads), 0, st, (const uint64_t*)nullptr,
                     (const int32_t*)nullptr, m->d_seg_beg, m->d_seg_end, m->d_ent, m->d_mom,
                     (int)m->N, rescale, (double*)nullptr, (double*)nullptr, (int)t, cand,
                     (int)ncand, out, (unsigned long long*)nullptr);
  RSP_LAUNCH_CHECK("k_pattern(probe)");
  return kOk;
}

extern "C" int rsp_maxlp_info(const rsp_maxlp* m, int64_t* out) {
  if (!m || !out) return fail(kErrNull, "rsp_maxlp_info: null pointer");
  out[0] = m->N;
  out[1] = m->nph;
  out[2] = m->n_entries;
  out[3] = (int64_t)m->rmask_pix.size();
  out[4] = (int64_t)m->mask_pix.size();
  out[5] = (int64_t)m->counters[0];
  out[6] = (int64_t)m->counters[1];
  out[7] = (int64_t)(m->pattern_ms * 1000. + 0.5);
  return kOk;
}

extern "C" int rsp_maxlp_destroy(rsp_maxlp* m) {
  if (!m) return kOk;
  hipSetDevice(m->device);
  void* ptrs[] = {m->d_qx, m->d_qy, m->d_qrad, m->d_mom, m->d_resc, m->d_mask, m->d_rmask_pix,
                  m->d_mask_pix, m->d_fr, m->d_pix, m->d_cnt, m->d_nmatch, m->d_pos,
                  m->d_seg_beg, m->d_seg_end, m->d_keys, m->d_keys2, m->d_ord, m->d_ord2,
                  m->d_ent, m->d_tmp};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (m->side) hipStreamDestroy(m->side);
  for (hipEvent_t e : m->sev)
    if (e) hipEventDestroy(e);
  if (m->d_counters) hipFree(m->d_counters);
  for (hipEvent_t e : m->ev)
    if (e) hipEventDestroy(e);
  delete m;
  return kOk;
}
