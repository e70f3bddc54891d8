/* FP64 C restatement of the RefSPI-MaxLP reconstruction path (the oracle's fast form).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke(), tests/golden/ fixture
 * scripts and the cpu_baseline leg of bench.py load this library, and only as the checker
 * or the timed CPU baseline.  The product (refspi-maxlp_amd/) never loads it.
 *
 * It restates the same reference lines as oracle/emc_oracle.py (kernels.cu, emc.py,
 * maxLP.py), in the same operation order, so that the two agree to the last bits of the
 * transcendental functions (C library here, NumPy there); tests/test_oracle_c.py pins this
 * file to emc_oracle.py on small inputs, function by function.  It exists because the
 * NumPy form needs hours for the full 6069 x 360 hypothesis grid at 185^2 and for MaxLP
 * over thousands of pixels; this form runs them in minutes on a few cores (OpenMP).
 *
 * Built with -O2 -ffp-contract=off (no fused multiply-add: every product and sum is
 * rounded as the NumPy oracle rounds it), no fast-math.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_PI 3.141592653589793238462643383279502884

/* ------------------------------------------------------------------------------------
 * kernels.cu restatements (one pixel at a time, as the CUDA threads do)
 * ---------------------------------------------------------------------------------- */

/* kernels.cu:44-82 slice_gen_holo for all pixels; model_c interleaved (re, im). */
void ora_slice_gen_holo(const double* model_c, double shiftx, double shifty, double diameter,
                        double rel_scale, double scale, int64_t size, const double* bg,
                        int64_t log_flag, double* view) {
  const double cen = floor(size / 2.);
  for (int64_t x = 0; x < size; ++x)
    for (int64_t y = 0; y < size; ++y) {
      const int64_t t = x * size + y;
      const double phase = 2. * ORA_PI * ((x - cen) * shiftx + (y - cen) * shifty) / size;
      const double ramp_r = cos(phase), ramp_i = sin(phase);
      double s = sqrt((x - cen) * (x - cen) + (y - cen) * (y - cen)) * ORA_PI * diameter / size;
      if (s == 0.) s = 1.e-5;
      const double sphere = rel_scale * (sin(s) - s * cos(s)) / (s * s * s);
      /* ramp * complex(sphere, 0) + model[t] */
      const double cr = ramp_r * sphere + model_c[2 * t];
      const double ci = ramp_i * sphere + model_c[2 * t + 1];
      const double a = hypot(cr, ci);
      double v = a * a;
      v *= scale;
      v += bg[t];
      if (log_flag) v = v < 1.e-20 ? -1000. : log(v);
      view[t] = v;
    }
}

/* kernels.cu:19-33 for one detector pixel: bilinear sample (scale and bg not applied).
 * Returns 0 when the rotated position is out of bounds (kernels.cu:24-25). */
static inline int rot_sample(const double* model, int64_t size, int64_t x, int64_t y, double ac,
                             double as, double* out) {
  const int64_t cen = size / 2;
  const double tx = (double)(x - cen) * ac - (double)(y - cen) * as + (double)cen;
  const double ty = (double)(x - cen) * as + (double)(y - cen) * ac + (double)cen;
  const double fxl = floor(tx), fyl = floor(ty);
  if (fxl < 0. || fxl > (double)(size - 2) || fyl < 0. || fyl > (double)(size - 2)) return 0;
  const int64_t ix = (int64_t)fxl, iy = (int64_t)fyl;
  const double fx = tx - ix, fy = ty - iy;
  const double cx = 1. - fx, cy = 1. - fy;
  *out = model[ix * size + iy] * cx * cy + model[(ix + 1) * size + iy] * fx * cy +
         model[ix * size + (iy + 1)] * cx * fy + model[(ix + 1) * size + (iy + 1)] * fx * fy;
  return 1;
}

/* kernels.cu:6-42 slice_gen for all pixels; ac, as = cos, sin of the angle. */
void ora_slice_gen(const double* model, double ac, double as, double scale, int64_t size,
                   const double* bg, int64_t log_flag, double* view) {
  for (int64_t x = 0; x < size; ++x)
    for (int64_t y = 0; y < size; ++y) {
      const int64_t t = x * size + y;
      double v;
      if (!rot_sample(model, size, x, y, ac, as, &v)) {
        view[t] = log_flag ? -1000. : 0.;
        continue;
      }
      v *= scale;
      v += bg[t];
      if (log_flag) v = v < 1.e-20 ? -1000. : log(v);
      view[t] = v;
    }
}

/* kernels.cu:84-114 slice_merge: scatter of view through the bilinear weights
 * (accumulates into model and mweights; sequential over t, in place of atomicAdd). */
void ora_slice_merge(const double* view, double ac, double as, int64_t size, double* model,
                     double* mweights) {
  const int64_t cen = size / 2;
  for (int64_t x = 0; x < size; ++x)
    for (int64_t y = 0; y < size; ++y) {
      const int64_t t = x * size + y;
      const double tx = (double)(x - cen) * ac - (double)(y - cen) * as + (double)cen;
      const double ty = (double)(x - cen) * as + (double)(y - cen) * ac + (double)cen;
      const double fxl = floor(tx), fyl = floor(ty);
      if (fxl < 0. || fxl > (double)(size - 2) || fyl < 0. || fyl > (double)(size - 2)) continue;
      const int64_t ix = (int64_t)fxl, iy = (int64_t)fyl;
      const double fx = tx - ix, fy = ty - iy;
      const double cx = 1. - fx, cy = 1. - fy;
      model[ix * size + iy] += view[t] * cx * cy;
      mweights[ix * size + iy] += cx * cy;
      model[(ix + 1) * size + iy] += view[t] * fx * cy;
      mweights[(ix + 1) * size + iy] += fx * cy;
      model[ix * size + (iy + 1)] += view[t] * cx * fy;
      mweights[ix * size + (iy + 1)] += cx * fy;
      model[(ix + 1) * size + (iy + 1)] += view[t] * fx * fy;
      mweights[(ix + 1) * size + (iy + 1)] += fx * fy;
    }
}

/* kernels.cu:116-137 calc_prob_all (sequential per frame: init*scale, ones, multi). */
void ora_calc_prob_all(const double* lview, const int32_t* mask, int64_t ndata,
                       const int32_t* ones, const int32_t* multi, const int64_t* o_acc,
                       const int64_t* m_acc, const int32_t* p_o, const int32_t* p_m,
                       const int32_t* c_m, double init, const double* scales, double* prob_r) {
  for (int64_t d = 0; d < ndata; ++d) {
    double acc = init * scales[d];
    for (int64_t t = o_acc[d]; t < o_acc[d] + ones[d]; ++t)
      if (mask[p_o[t]] < 1) acc += lview[p_o[t]];
    for (int64_t t = m_acc[d]; t < m_acc[d] + multi[d]; ++t)
      if (mask[p_m[t]] < 1) acc += lview[p_m[t]] * c_m[t];
    prob_r[d] = acc;
  }
}

/* kernels.cu:139-152 merge_all: view[p] += prob_r[d] * count (sequential over frames). */
void ora_merge_all(const double* prob_r, int64_t ndata, const int32_t* ones,
                   const int32_t* multi, const int64_t* o_acc, const int64_t* m_acc,
                   const int32_t* p_o, const int32_t* p_m, const int32_t* c_m, double* view) {
  for (int64_t d = 0; d < ndata; ++d) {
    for (int64_t t = o_acc[d]; t < o_acc[d] + ones[d]; ++t) view[p_o[t]] += prob_r[d];
    for (int64_t t = m_acc[d]; t < m_acc[d] + multi[d]; ++t) view[p_m[t]] += prob_r[d] * c_m[t];
  }
}

/* ------------------------------------------------------------------------------------
 * emc.py:183-216: the joint (state x rotation) search of _calculate_prob pass 2 and
 * _get_rmax, with the per-frame best, first-occurrence argmax h = s*R + r and the
 * second-best value (v1 - v2 = the top-2 margin; 0 when the best value is tied).
 * ---------------------------------------------------------------------------------- */

typedef struct {
  double v1, v2;
  int64_t h;
} Top2;

static inline void top2_add(Top2* t, double v, int64_t h) {
  if (v > t->v1 || (v == t->v1 && h < t->h)) {
    if (t->h >= 0) t->v2 = t->v1 > t->v2 ? t->v1 : t->v2;
    t->v1 = v;
    t->h = h;
  } else if (v > t->v2) {
    t->v2 = v;
  }
}

/* Merge b into a (order-independent: ties in value go to the smaller h). */
static inline void top2_merge(Top2* a, const Top2* b) {
  if (b->h < 0) return;
  if (a->h < 0) {
    *a = *b;
    return;
  }
  if (b->v1 > a->v1 || (b->v1 == a->v1 && b->h < a->h)) {
    const double r = a->v1 > b->v2 ? a->v1 : b->v2;
    a->v1 = b->v1;
    a->h = b->h;
    a->v2 = r;
  } else {
    a->v2 = a->v2 > b->v1 ? a->v2 : b->v1;
  }
}

/* Full E-step search.  The holograms are slice_gen_holo(model, sx, sy, dia, 1, 1, size,
 * bg, 0) (emc.py:172-176), the log-views slice_gen(view, j*pi/R, 1, size, bg, 1)
 * (emc.py:190-193; rcos/rsin = cos/sin of j*pi/R from the caller), and the per-frame LL
 * calc_prob_all with init[s] = -msums[s]*vscale (emc.py:194-199, computed by the caller
 * with NumPy's summation order).  Only the pixels with probmask == 0 enter calc_prob_all,
 * so the log-views are evaluated there only; the order of the remaining terms is the
 * reference's.  Outputs per frame: rmax (int32), maxll, second-best LL.  Returns 0. */
int ora_estep_top2(int64_t size, const double* model_c, const double* bg, int64_t S,
                   const double* sx, const double* sy, const double* dia, const double* init,
                   int64_t R, const double* rcos, const double* rsin, int64_t N,
                   const int32_t* ones, const int32_t* multi, const int64_t* o_acc,
                   const int64_t* m_acc, const int32_t* p_o, const int32_t* p_m,
                   const int32_t* c_m, const int32_t* probmask, const double* scales,
                   int nthreads, int32_t* rmax, double* maxll, double* second) {
  const int64_t npix = size * size;
  /* scored pixels and their compact index */
  int32_t* compact = (int32_t*)malloc(sizeof(int32_t) * npix);
  int64_t K = 0;
  for (int64_t t = 0; t < npix; ++t) compact[t] = probmask[t] < 1 ? (int32_t)K++ : -1;
  int32_t* sxy = (int32_t*)malloc(sizeof(int32_t) * 2 * (K ? K : 1));
  for (int64_t t = 0; t < npix; ++t)
    if (compact[t] >= 0) {
      sxy[2 * compact[t]] = (int32_t)(t / size);
      sxy[2 * compact[t] + 1] = (int32_t)(t % size);
    }
  /* frames restricted to scored photons, order kept: ones then multi */
  int64_t* fo1 = (int64_t*)malloc(sizeof(int64_t) * (N + 1));
  int64_t* fom = (int64_t*)malloc(sizeof(int64_t) * (N + 1));
  fo1[0] = fom[0] = 0;
  for (int64_t d = 0; d < N; ++d) {
    int64_t n1 = 0, nm = 0;
    for (int64_t t = o_acc[d]; t < o_acc[d] + ones[d]; ++t) n1 += compact[p_o[t]] >= 0;
    for (int64_t t = m_acc[d]; t < m_acc[d] + multi[d]; ++t) nm += compact[p_m[t]] >= 0;
    fo1[d + 1] = fo1[d] + n1;
    fom[d + 1] = fom[d] + nm;
  }
  int32_t* j1 = (int32_t*)malloc(sizeof(int32_t) * (fo1[N] ? fo1[N] : 1));
  int32_t* jm = (int32_t*)malloc(sizeof(int32_t) * (fom[N] ? fom[N] : 1));
  double* cm = (double*)malloc(sizeof(double) * (fom[N] ? fom[N] : 1));
  for (int64_t d = 0; d < N; ++d) {
    int64_t w1 = fo1[d], wm = fom[d];
    for (int64_t t = o_acc[d]; t < o_acc[d] + ones[d]; ++t)
      if (compact[p_o[t]] >= 0) j1[w1++] = compact[p_o[t]];
    for (int64_t t = m_acc[d]; t < m_acc[d] + multi[d]; ++t)
      if (compact[p_m[t]] >= 0) {
        jm[wm] = compact[p_m[t]];
        cm[wm++] = (double)c_m[t];
      }
  }
  int nt = nthreads > 0 ? nthreads : 1;
#ifndef _OPENMP
  nt = 1;
#endif
  Top2* all = (Top2*)malloc(sizeof(Top2) * (size_t)nt * (N ? N : 1));
  for (int64_t i = 0; i < (int64_t)nt * N; ++i) {
    all[i].v1 = -INFINITY;
    all[i].v2 = -INFINITY;
    all[i].h = -1;
  }
#pragma omp parallel num_threads(nt)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    Top2* mine = all + (int64_t)tid * N;
    double* view = (double*)malloc(sizeof(double) * npix);
    double* lv = (double*)malloc(sizeof(double) * (K ? K : 1));
#pragma omp for schedule(dynamic, 1)
    for (int64_t s = 0; s < S; ++s) {
      ora_slice_gen_holo(model_c, sx[s], sy[s], dia[s], 1., 1., size, bg, 0, view);
      for (int64_t r = 0; r < R; ++r) {
        for (int64_t j = 0; j < K; ++j) {
          const int64_t x = sxy[2 * j], y = sxy[2 * j + 1];
          double v;
          if (!rot_sample(view, size, x, y, rcos[r], rsin[r], &v)) {
            lv[j] = -1000.;
            continue;
          }
          v *= 1.;
          v += bg[x * size + y];
          lv[j] = v < 1.e-20 ? -1000. : log(v);
        }
        const int64_t h = s * R + r;
        for (int64_t d = 0; d < N; ++d) {
          double acc = init[s] * scales[d];
          for (int64_t e = fo1[d]; e < fo1[d + 1]; ++e) acc += lv[j1[e]];
          for (int64_t e = fom[d]; e < fom[d + 1]; ++e) acc += lv[jm[e]] * cm[e];
          top2_add(&mine[d], acc, h);
        }
      }
    }
    free(view);
    free(lv);
  }
  for (int64_t d = 0; d < N; ++d) {
    Top2 t = all[d];
    for (int k = 1; k < nt; ++k) top2_merge(&t, &all[(int64_t)k * N + d]);
    rmax[d] = (int32_t)t.h;
    maxll[d] = t.v1;
    second[d] = t.v2;
  }
  free(all);
  free(j1);
  free(jm);
  free(cm);
  free(fo1);
  free(fom);
  free(sxy);
  free(compact);
  return 0;
}

/* ------------------------------------------------------------------------------------
 * MaxLP (maxLP.py)
 * ---------------------------------------------------------------------------------- */

/* maxLP.py:250-252 get_photons_pixel's detector index of model pixel t in a frame whose
 * rotation is rots = [[c, -s], [s, c]] (c, s = cos, sin of -angle, maxLP.py:311):
 * x, y = rint(rots @ [qx, qy] * size + size//2) (half to even), t_vals = x*size + y with
 * no clamp.  x or y can reach `size` at even sizes: y = size lands on the first pixel of
 * row x+1 (plain arithmetic); x*size + y >= npix (x = size) is reduced modulo the npix
 * columns of the photons matrix, as cupyx's sparse fancy indexing reduces array indices
 * (cupyx/scipy/sparse/_index.py, _asindices: `x % length`).  The products are rounded
 * separately (no FMA), as in emc_oracle.photon_index. */
static inline int32_t detector_index(int64_t size, double qx, double qy, double c, double s) {
  const double px = (c * qx + (-s) * qy) * (double)size + (double)(size / 2);
  const double py = (s * qx + c * qy) * (double)size + (double)(size / 2);
  const int64_t x = (int64_t)nearbyint(px), y = (int64_t)nearbyint(py);
  int64_t t = x * size + y;
  const int64_t npix = size * size;
  t %= npix;
  if (t < 0) t += npix;
  return (int32_t)t;
}

void ora_photon_index(int64_t size, int64_t n, const int32_t* t, const double* rc,
                      const double* rs, const double* qx, const double* qy, int32_t* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = detector_index(size, qx[t[i]], qy[t[i]], rc[i], rs[i]);
}

/* Frame photon lookup: per frame the distinct detector pixels (ascending) and their summed
 * counts (the cupyx CSR of maxLP.py:60-71 sums duplicates). */
static double frame_count(const int64_t* fo, const int32_t* fpix, const double* fcnt, int64_t d,
                          int32_t p) {
  int64_t lo = fo[d], hi = fo[d + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (fpix[mid] < p)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < fo[d + 1] && fpix[lo] == p) ? fcnt[lo] : 0.;
}

typedef struct {
  int64_t N;
  double* k;   /* K_d(t) */
  double* rr;  /* Re, Im R_d(t) (maxLP.py:238-243) */
  double* ri;
  double kmean;
} PixConst;

/* maxLP.py:107-114 get_pixel_constants (all frames) + 238-243 get_fref_d. */
static void pixel_constants(int64_t size, int64_t N, const int64_t* fo, const int32_t* fpix,
                            const double* fcnt, const double* sx, const double* sy,
                            const double* dia, const double* rc, const double* rs, double qx,
                            double qy, double qrad, PixConst* pc) {
  pc->N = N;
  for (int64_t d = 0; d < N; ++d) {
    const int32_t p = detector_index(size, qx, qy, rc[d], rs[d]);
    pc->k[d] = frame_count(fo, fpix, fcnt, d, p);
    const double sval = ORA_PI * qrad * dia[d];
    const double fref = (sin(sval) - sval * cos(sval)) / pow(sval, 3.);
    const double theta = (2. * ORA_PI) * (qx * sx[d] + qy * sy[d]);
    pc->rr[d] = fref * cos(theta);
    pc->ri[d] = fref * sin(theta);
  }
  /* k_d.mean(): the counts are integers, so every summation order gives this sum */
  double ks = 0.;
  for (int64_t d = 0; d < N; ++d) ks += pc->k[d];
  pc->kmean = ks / (double)N;
}

/* maxLP.py:216-228 get_rescale_pixel for one candidate: mean K / mean |f + R_d|^2 (the
 * mean over frames summed in frame order, NumPy's order for an axis-0 mean). */
static double rescale_pixel(const PixConst* pc, double fr, double fi) {
  double s = 0.;
  for (int64_t d = 0; d < pc->N; ++d) {
    const double a = hypot(fr + pc->rr[d], fi + pc->ri[d]);
    s += a * a;
  }
  return pc->kmean / (s / (double)pc->N);
}

/* maxLP.py:181-197 get_logq_pixel for one candidate: mean_d(K log w) - mean_d(w),
 * w = rescale * |f + R_d|^2, both means summed in frame order. */
static double logq_pixel(const PixConst* pc, double fr, double fi, double rescale) {
  double s1 = 0., s2 = 0.;
  for (int64_t d = 0; d < pc->N; ++d) {
    const double a = hypot(fr + pc->rr[d], fi + pc->ri[d]);
    const double w = rescale * (a * a);
    /* K = 0 with a finite positive w adds a zero (0 * log w = +-0, and s1 + -0 == s1):
     * the log is skipped; otherwise 0 * log(0) = NaN etc. propagate as in the reference */
    if (pc->k[d] != 0. || !(w > 0. && w < INFINITY)) s1 += pc->k[d] * log(w);
    s2 += w;
  }
  return s1 / (double)pc->N - s2 / (double)pc->N;
}

/* maxLP.py:154-179 run_pixel_pattern + iterate_pixel_pattern for one pixel, 33 x 33
 * pattern of maxLP.py:45-48 (offsets arange(-8, 8.5, 0.5), x + iy, 'ij' order).
 * rescale < 0 means per-candidate rescale (phase 1 of maxLP.py:315-319).
 * Outputs the new value, the number of accepted steps, and the smallest relative top-2
 * margin (LL best - LL second) / |LL best| over the steps evaluated. */
static void pixel_pattern(const PixConst* pc, double fr, double fi, double rescale, int num_iter,
                          double* out_r, double* out_i, int32_t* steps, double* min_margin) {
  double pat[33];
  for (int i = 0; i < 33; ++i) pat[i] = -8. + 0.5 * i;
  double step = hypot(fr, fi) / 4.;
  double mm = INFINITY;
  int nacc = 0;
  for (int it = 0; it < num_iter; ++it) {
    double best = 0., second = -INFINITY;
    int imax = -1, has_nan = 0;
    for (int c = 0; c < 1089; ++c) {
      const double px = pat[c / 33], py = pat[c % 33];
      /* fobj_t + pattern*step: complex pattern times a real step */
      const double cr = fr + (px * step - py * 0.), ci = fi + (px * 0. + py * step);
      const double rs = rescale < 0. ? rescale_pixel(pc, cr, ci) : rescale;
      const double v = logq_pixel(pc, cr, ci, rs);
      if (has_nan) continue;
      if (isnan(v)) {  /* NumPy/CuPy argmax: the first NaN wins */
        has_nan = 1;
        imax = c;
        continue;
      }
      if (imax < 0 || v > best) {
        if (imax >= 0) second = best;
        best = v;
        imax = c;
      } else if (v > second) {
        second = v;
      }
    }
    if (!has_nan) {
      const double m = (best - second) / fabs(best);
      if (m < mm) mm = m;
    }
    const double px = pat[imax / 33], py = pat[imax % 33];
    const double nr = fr + (px * step - py * 0.), ni = fi + (px * 0. + py * step);
    const double nstep = imax == 544 ? step / (sqrt(1089.) / 5.) : step;
    if (nstep / hypot(nr, ni) < 1.e-3) break;  /* maxLP.py:162-163: keep the pre-step value */
    fr = nr;
    fi = ni;
    step = nstep;
    ++nacc;
  }
  *out_r = fr;
  *out_i = fi;
  *steps = nacc;
  *min_margin = mm;
}

/* Pattern search over a list of model pixels (independent of each other within a phase of
 * maxLP.py:304-340).  Frames: fo (N+1), fpix/fcnt (distinct detector pixels, ascending,
 * summed counts).  Per frame sx, sy, dia and rc, rs = cos, sin of -angle.  f_in/f_out are
 * interleaved complex per listed pixel.  rescale < 0: per-candidate (phase 1).  resc_out
 * (may be NULL): get_rescale_pixel of the result (maxLP.py:321-327).  Returns 0. */
int ora_maxlp_pixels(int64_t size, int64_t N, const int64_t* fo, const int32_t* fpix,
                     const double* fcnt, const double* sx, const double* sy, const double* dia,
                     const double* rc, const double* rs, const double* qx, const double* qy,
                     const double* qrad, int64_t npx, const int32_t* pixels,
                     const double* f_in, double rescale, int num_iter, int nthreads,
                     double* f_out, double* resc_out, int32_t* steps, double* margin) {
  int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel num_threads(nt)
  {
    PixConst pc;
    pc.k = (double*)malloc(sizeof(double) * (N ? N : 1));
    pc.rr = (double*)malloc(sizeof(double) * (N ? N : 1));
    pc.ri = (double*)malloc(sizeof(double) * (N ? N : 1));
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < npx; ++i) {
      const int32_t t = pixels[i];
      pixel_constants(size, N, fo, fpix, fcnt, sx, sy, dia, rc, rs, qx[t], qy[t], qrad[t], &pc);
      double orr, oi, mm;
      int32_t ns;
      pixel_pattern(&pc, f_in[2 * i], f_in[2 * i + 1], rescale, num_iter, &orr, &oi, &ns, &mm);
      f_out[2 * i] = orr;
      f_out[2 * i + 1] = oi;
      if (steps) steps[i] = ns;
      if (margin) margin[i] = mm;
      if (resc_out) resc_out[i] = rescale_pixel(&pc, orr, oi);
    }
    free(pc.k);
    free(pc.rr);
    free(pc.ri);
  }
  return 0;
}

/* get_logq_pixel (maxLP.py:181-197) of given candidates at one pixel; rescale < 0 means
 * the per-candidate rescale of get_rescale_pixel.  Also returns K_d(t) and R_d(t). */
int ora_maxlp_logq(int64_t size, int64_t N, const int64_t* fo, const int32_t* fpix,
                   const double* fcnt, const double* sx, const double* sy, const double* dia,
                   const double* rc, const double* rs, const double* qx, const double* qy,
                   const double* qrad, int32_t t, int64_t ncand, const double* cand,
                   double rescale, int nthreads, double* out, double* k_out, double* rr_out,
                   double* ri_out) {
  PixConst pc;
  pc.k = k_out;
  pc.rr = rr_out;
  pc.ri = ri_out;
  pixel_constants(size, N, fo, fpix, fcnt, sx, sy, dia, rc, rs, qx[t], qy[t], qrad[t], &pc);
  int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (int64_t c = 0; c < ncand; ++c) {
    const double rsc = rescale < 0. ? rescale_pixel(&pc, cand[2 * c], cand[2 * c + 1]) : rescale;
    out[c] = logq_pixel(&pc, cand[2 * c], cand[2 * c + 1], rsc);
  }
  return 0;
}
