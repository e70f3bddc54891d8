/* C-ABI of the RefSPI-MaxLP reconstruction hot path on MI355X (gfx950).
 *
 * Conventions
 *  - Every entry point returns an int status: 0 = OK, < 0 = argument/state error,
 *    > 0 = a hipError_t from the runtime.  rsp_last_error() returns the message of the
 *    most recent failure on the calling thread.
 *  - Device buffers are plain pointers in device memory owned by the caller; complex
 *    arrays are interleaved (re, im) doubles, as numpy/CuPy 'c16'.  Host buffers are
 *    marked "host".  Sizes are int64_t, scalars double.
 *  - Calls that take an rsp_stream_t (a hipStream_t; NULL = default stream) are
 *    asynchronous on that stream unless stated otherwise.
 *  - Plans (rsp_emc, rsp_maxlp, rsp_rot) own their scratch device memory and, for
 *    rsp_rot, their rocBLAS handle; one plan per host thread at a time (thread-compatible,
 *    not thread-safe); different plans may be used from different threads.
 */
#ifndef REFSPI_AMD_H
#define REFSPI_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* rsp_stream_t;
typedef struct rsp_emc rsp_emc;
typedef struct rsp_maxlp rsp_maxlp;

/* Message of the last failing call on this thread ("" if none). */
const char* rsp_last_error(void);
/* Library ABI version (major*100 + minor). */
int rsp_version(void);

/* ---------------------------------------------------------------------------------
 * Single-kernel mirrors of the reference's CuPy RawKernels (FP64, device pointers).
 * ------------------------------------------------------------------------------- */

/* kernels.cu:6-42 slice_gen: bilinear in-plane rotation of a real size x size image,
 * times scale plus bg, optional log (<1e-20 or out of bounds -> -1000).  cos/sin of
 * `angle` are taken on the host with the C library. */
int rsp_slice_gen(const double* model, double angle, double scale, int64_t size,
                  const double* bg, int64_t log_flag, double* view, rsp_stream_t stream);

/* kernels.cu:44-82 slice_gen_holo: |model + rel*S(s)*e^{i phase}|^2 * scale + bg. */
int rsp_slice_gen_holo(const double* model_c, double shiftx, double shifty,
                       double diameter, double rel_scale, double scale, int64_t size,
                       const double* bg, int64_t log_flag, double* view,
                       rsp_stream_t stream);

/* kernels.cu:116-137 calc_prob_all: per-frame sparse Poisson log-likelihood, FP64,
 * sequential per frame in the reference order (init*scale, ones, multi). */
int rsp_calc_prob_all(const double* lview, const int32_t* mask, int64_t ndata,
                      const int32_t* ones, const int32_t* multi, const int64_t* o_acc,
                      const int64_t* m_acc, const int32_t* p_o, const int32_t* p_m,
                      const int32_t* c_m, double init, const double* scales,
                      double* prob_r, rsp_stream_t stream);

/* kernels.cu:172-183 get_f_dt: f_dt[d*npix + t] = fobj_t[t] + fref_d[d] (complex). */
int rsp_get_f_dt(const double* fobj_t, const double* fref_d, int64_t ndata, int64_t npix,
                 double* f_dt, rsp_stream_t stream);

/* kernels.cu:84-114 slice_merge: model[m] += sum view[t]*wx*wy and mweights[m] += sum
 * wx*wy over the view pixels t whose bilinear sample under `angle` (cos/sin on the host)
 * touches m.  Gather form (one thread per model pixel, no atomics): each pixel's terms are
 * summed in view-pixel order and then added, so the result differs from the reference's
 * atomic scatter only by the association of the additions. */
int rsp_slice_merge(const double* view, double angle, int64_t size, double* model,
                    double* mweights, rsp_stream_t stream);

/* kernels.cu:139-152 merge_all: view[p] += prob_r[d] for every single photon of frame d at
 * p and prob_r[d]*count for every multi photon.  npix (not in the reference's list) bounds
 * the photon pixels: any pixel outside [0, npix) fails the call with -2 before view is
 * touched.  No atomics: the photons are sorted by (pixel, frame-major position) and each
 * pixel's terms are added to view[p] one after another in frame order, ones before multi --
 * the order of a sequential loop over the reference's threads.  Allocates its sort buffers
 * on `stream`; synchronises `stream`. */
int rsp_merge_all(const double* prob_r, int64_t ndata, const int32_t* ones,
                  const int32_t* multi, const int64_t* o_acc, const int64_t* m_acc,
                  const int32_t* p_o, const int32_t* p_m, const int32_t* c_m, double* view,
                  int64_t npix, rsp_stream_t stream);

/* ---------------------------------------------------------------------------------
 * EMC E-step plan: replaces emc.py:159-216 (_calculate_prob + _get_rmax) for one GPU.
 * Hypotheses are h = s*num_rot + r (s = state index of the raveled 'ij' meshgrid of
 * emc.py:71-77, r = rotation j*pi/num_rot); the result is the first-occurrence argmax
 * over h of the FP64 log-likelihood of kernels.cu:116-137.
 * ------------------------------------------------------------------------------- */

/* Create a plan on `device`.  rot_cos/rot_sin (host, num_rot) are cos/sin of
 * j*pi/num_rot; state_sx/sy/dia (host, num_states) the raveled state grid. */
int rsp_emc_create(int device, int64_t size, int64_t num_rot, const double* rot_cos,
                   const double* rot_sin, const double* state_sx, const double* state_sy,
                   const double* state_dia, int64_t num_states, rsp_emc** out);

/* Load the sparse frames (host arrays, dset.py:23-60 layout; bg may be NULL = zeros).
 * The plan copies what it needs; the caller may free the arrays afterwards.
 * Synchronous. */
int rsp_emc_set_frames(rsp_emc* plan, int64_t num_data, const int32_t* ones,
                       const int32_t* multi, const int32_t* place_ones,
                       const int32_t* place_multi, const int32_t* count_multi,
                       const double* bg);

/* Tuning / mode switches.  Keys:
 *   1 = exact mode (1: every frame searched in FP64 over all hypotheses; default 0)
 *   2 = hypotheses per table batch (default chosen from memory)
 *   3 = LDS bank model of the photon scheduler (32 = conflict-free half-waves, the
 *       default and gfx950's layout; 64; 0 = no scheduling; set before frames)
 *   4 = hypothesis segments per batch (default automatic)
 *   5 = per-kernel timing with HIP events (1 on; read with rsp_emc_timing)
 *   6 = swaps of the local search that regroups frames into lane groups with flat LDS
 *       bank loads, per window of 8 groups (default 4000; 0 = off; set before frames).
 *       Only the slot layout changes, never the result.
 *   7 = per-workgroup LDS layout of the table parts, fitted to the workgroup's frames
 *       (1 = on, the default; 0 = identity; 32-bank model and <= 2 parts; set before
 *       frames).  Only the slot layout and the table loads' addresses change, never the
 *       result.
 *   8 = broadcast-aware photon slots (1 = on, the default; 0 = off; set before frames):
 *       lanes that read the same table entry may share a bank in one slot (one LDS access
 *       serves them).  Only the slot layout changes, never the result.
 *   9 = frames per lane of whole-frame waves (2, the default: each slot register holds
 *       the 16-bit LDS addresses of two frames, so a workgroup covers twice the frames per
 *       table load; 1 = one frame per lane; set before frames).  Only which lanes and
 *       registers hold a frame changes, never the result.
 *  10 = side streams shared by a batch's slot-class gather launches (2, the default;
 *       1 .. 4; 0 = all on the caller's stream): the classes run concurrently between the
 *       batch's table build and its merge, on streams of default priority created by the
 *       plan.  Never changes the result. */
int rsp_emc_set_option(rsp_emc* plan, int key, int64_t value);

/* Run the E-step for `model_c` (device, npix complex) with per-frame `scales` (device,
 * N).  Writes rmax (device int32 N: global hypothesis index), maxll (device f64 N: the
 * FP64 log-likelihood of that hypothesis), msums (device f64 num_states: emc.py:177) and
 * *vscale (host, emc.py:181).  Synchronises `stream` before returning. */
int rsp_emc_estep(rsp_emc* plan, const double* model_c, const double* scales,
                  int32_t* rmax, double* maxll, double* msums, double* vscale,
                  rsp_stream_t stream);

/* Counters of the last E-step (host int64[8]): [0] frames, [1] hypotheses,
 * [2] frames decided by the FP32-table pass, [3] frames resolved by top-2 FP64
 * re-evaluation, [4] frames re-searched in full in FP64, [5] slots per lane (P),
 * [6] waves per workgroup, [7] table parts. */
int rsp_emc_stats(const rsp_emc* plan, int64_t* out);

/* Timing of the last E-step with option 5 on (host double[6]): [0] total ms of the
 * table kernels, [1] total ms of the gather kernels, [2] batches (= launches of each),
 * [3] hypotheses per batch; from the plan: [4] LDS bytes the gather reads per hypothesis
 * (every slot, zero pads included), [5] LDS bytes its table loads write per hypothesis. */
int rsp_emc_timing(const rsp_emc* plan, double* out);

int rsp_emc_destroy(rsp_emc* plan);

/* Host-side check of the photon slot scheduler and the per-workgroup table layouts (no
 * device calls; no counterpart in the reference: the reference has no slot schedule).
 * `trials` random lane x bank groups and trials/500 (at least 1) rounds of random tiles,
 * from `seed`.  Each group is scheduled exactly (Koenig) and broadcast-aware (option 8).
 * out (host int64[7]): [0] groups, [1] slot conflicts (a lane twice in a slot; a bank
 * twice in a slot -- broadcast: two different entries of a bank in a slot), [2] groups
 * whose exact slot count differs from the maximum degree or whose broadcast schedule is
 * longer than the exact one, [3] layouts, [4] layout faults, [5] slots saved by the
 * broadcast schedules, [6] broadcast slot count of a fixed half-wave whose 32 lanes share
 * five entries of one bank plus one own entry (optimum 6, exact 161); [1], [2], [4] are 0
 * for a correct scheduler. */
int rsp_emc_schedule_check(int64_t trials, uint64_t seed, int64_t* out);

/* ---------------------------------------------------------------------------------
 * MaxLP plan: replaces maxLP.py:304-340 (MaxLPhaser._run_phaser) for one GPU.
 * ------------------------------------------------------------------------------- */

/* Create a plan for a size x size model (maxLP.py:14-35 geometry, masks and pattern). */
int rsp_maxlp_create(int device, int64_t size, rsp_maxlp** out);

/* Load the sparse frames (host arrays, as rsp_emc_set_frames; maxLP.py:50-71).
 * Synchronous. */
int rsp_maxlp_set_frames(rsp_maxlp* plan, int64_t num_data, const int32_t* ones,
                         const int32_t* multi, const int32_t* place_ones,
                         const int32_t* place_multi, const int32_t* count_multi);

/* Pattern-search update of model_c (device, npix complex, in place) given the per-frame
 * shift (sx, sy), diameter and cos/sin of the negated angle (device f64 N each:
 * rot_cos = cos(-ang), rot_sin = sin(-ang), maxLP.py:311).  Writes the ring rescale of
 * maxLP.py:324-329 to *rescale_out (host, may be NULL).  Synchronises `stream`. */
int rsp_maxlp_run(rsp_maxlp* plan, double* model_c, const double* sx, const double* sy,
                  const double* dia, const double* rot_cos, const double* rot_sin,
                  double* rescale_out, rsp_stream_t stream);

/* maxLP.py:245-253 get_photons_pixel for all frames: K_d(t) (device f64 N) of model pixel t
 * in the frames and rotations of the last rsp_maxlp_run, read from the model-space photon
 * lists that run pulled back (zero where frame d has no photon at t's detector pixel). */
int rsp_maxlp_photons_pixel(const rsp_maxlp* plan, int64_t t, double* k_out,
                            rsp_stream_t stream);

/* maxLP.py:181-197 get_logq_pixel: LL_c = mean_d K_d log w_dc - mean_d w_dc for ncand
 * (1..1280) candidates cand (device c16) at mask pixel t, with the frames and per-frame
 * parameters of the last rsp_maxlp_run; rescale < 0 takes each candidate's
 * get_rescale_pixel (maxLP.py:216-228).  out: device f64 ncand.  The same arithmetic as
 * the pattern search of rsp_maxlp_run (product form of the log term, moments for the mean). */
int rsp_maxlp_logq_pixel(const rsp_maxlp* plan, int64_t t, const double* cand, int64_t ncand,
                         double rescale, double* out, rsp_stream_t stream);

/* Sizes and work of the plan (host int64[8]): [0] frames, [1] photon entries,
 * [2] model-space photon entries of the last run, [3] ring pixels, [4] mask pixels;
 * of the last run: [5] entry x candidate products evaluated by the pattern search (each
 * one |c + R_d|^2 factor: 6 FP64 operations), [6] search steps evaluated (over all pixels
 * of both phases), [7] microseconds of the two pattern-search launches (HIP events). */
int rsp_maxlp_info(const rsp_maxlp* plan, int64_t* out);

int rsp_maxlp_destroy(rsp_maxlp* plan);

/* ---------------------------------------------------------------------------------
 * Rotation form (per-frame shift/diameter, rotations as the latent variable): the
 * dense cross term sum_p O_r(p) conj(R_d(p)) as a GEMM.  No reference counterpart: the
 * reference's nearest code is the per-state unrotated msums of emc.py:172-181.
 * ------------------------------------------------------------------------------- */

/* C[d*M + r] = sum_k A[r*K + k] * B[d*K + k] (row-major, device pointers).
 *   mode 0: FP32 in/out through rocBLAS sgemm (the plain FP32 form)
 *   mode 1: FP32 in/out, split-BF16 hi/lo products with FP32 accumulation on MFMA
 *   mode 2: FP64 in/out through rocBLAS dgemm
 * Mode 1 allocates its padded bf16 planes on `stream` for the call; modes 0 and 2 use a
 * rocBLAS handle created for the call (so concurrent callers never share one). */
int rsp_xterm_gemm(int mode, int64_t M, int64_t N, int64_t K, const void* A, const void* B,
                   void* C, rsp_stream_t stream);

/* The split-BF16 form of rsp_xterm_gemm (mode 1) with no allocation: its bf16 planes and
 * K-slice partial products live in a static device array of 32 MiB per GPU (held from the
 * library's load until the process ends) instead of memory allocated per call.  No
 * counterpart in the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Limit:
 * (2*pad256(M) + 2*pad256(N)) * pad32(K) * 2 + ceil(pad32(K)/12288) * N * pad256(M) * 4
 * bytes <= 33554432, else kErrSize (-2) before any device call.  Calls on one GPU share
 * the array: calls that overlap in time on the same GPU (two streams, two threads) race;
 * calls queued on one stream do not.  Results equal mode 1 of rsp_xterm_gemm bit for bit. */
int rsp_xterm_gemm_static(int64_t M, int64_t N, int64_t K, const float* A, const float* B,
                          float* C, rsp_stream_t stream);

/* Rotation-form plan over the scored pixels (probmask == 0 of emc.py:89-98), with
 * hypotheses h = c*R + r (class c < C, set by option 3; rotation r < R = num_rot):
 *   O_h = bilinear rotation of class c's complex model by the r-th angle (kernels.cu:6-42
 *         geometry), R_d = rel_scale * sphere(dia_d) * shift ramp(sx_d, sy_d) (the
 *         reference term of kernels.cu:44-82), W = scale_d * |O_h + R_d|^2,
 *   LL[d*H + h] = sum_photons K log W - sum_pixels W (log clamp of kernels.cu:37-40),
 *   prob[d*H + h] = softmax_h LL (FP64), logz[d] = log sum_h exp LL, H = C*R.
 * Models are C consecutive c16 npix arrays.  rot_cos/rot_sin: host, num_rot
 * (C * num_rot <= 1536). */
typedef struct rsp_rot rsp_rot;
int rsp_rot_create(int device, int64_t size, int64_t num_rot, const double* rot_cos,
                   const double* rot_sin, double rel_scale, rsp_rot** out);
/* Sparse frames (host arrays, rsp_emc_set_frames layout); photons outside the scored
 * pixels are dropped.  Synchronous. */
int rsp_rot_set_frames(rsp_rot* plan, int64_t num_data, const int32_t* ones,
                       const int32_t* multi, const int32_t* place_ones,
                       const int32_t* place_multi, const int32_t* count_multi);
/* Keys: 1 = cross-term precision (0 FP32 rocBLAS, 1 split BF16 MFMA = default, 2 FP64
 * with an FP64 log term), 2 = frames per GEMM chunk (0 = about 8 GB of operand rows;
 * 4096 for the adjoint), 3 = number of object classes C (default 1), 5 = per-kernel
 * timing (read with rsp_rot_timing), 6 = precision of the frame-gradient GEMM
 * Obar = P O of the adjoint (0 FP64 rocBLAS = default, 1 split BF16 MFMA: about 1e-5
 * relative in Obar, for gradient refinement), 7 = precision of the model-gradient GEMM
 * GT = Q^T B of the adjoint (Q = prob * scale, B = the reference rows; 0 FP64 rocBLAS =
 * default, 1 split BF16 MFMA with FP32 accumulation over each sub-chunk of at most 12288
 * frames, sub-chunks added in FP64), 8 = reference rows of the split cross term (1 =
 * generated inside the GEMM kernel from sx, sy, dia and the pixel coordinates, never
 * stored: the default for size <= 513; 0 = written to HBM as bf16 hi/lo rows and read by
 * the GEMM; sizes above 513 always store them).  Refused while a pass is open. */
int rsp_rot_set_option(rsp_rot* plan, int key, int64_t value);
/* model_c: device c16 npix; sx, sy, dia, scales: device f64 N; ll, prob: device f64 N*R
 * (either may be NULL); logz: device f64 N (may be NULL).  Synchronises `stream`. */
int rsp_rot_estep(rsp_rot* plan, const double* model_c, const double* sx, const double* sy,
                  const double* dia, const double* scales, double* ll, double* prob,
                  double* logz, rsp_stream_t stream);
/* As rsp_rot_estep, plus best (device i32 N, may be NULL): the first hypothesis index
 * attaining max_h LL[d, h] (the first-occurrence argmax of emc.py:207-209). */
int rsp_rot_estep_best(rsp_rot* plan, const double* model_c, const double* sx,
                       const double* sy, const double* dia, const double* scales, double* ll,
                       double* prob, double* logz, int32_t* best, rsp_stream_t stream);
/* Merged intensity models (device f64 C x npix) from prob (device f64 N*H): for every
 * hypothesis merge_all (kernels.cu:139-152) then slice_merge (kernels.cu:84-114), summed
 * over the rotations of each class and divided by the slice_merge weights times
 * sum_d prob*scale; 0 where the weight is 0.  acc_out and weights_out (device f64
 * C x npix) may be NULL.  Synchronises `stream`. */
int rsp_rot_merge(rsp_rot* plan, const double* prob, const double* scales, double* model_out,
                  double* acc_out, double* weights_out, rsp_stream_t stream);
/* Gradient of L = -sum_d log((1/H) sum_h exp LL[d, h]) (uniform prior over classes and
 * rotations) at the given inputs, in FP64, recomputing O_h and R_d (nothing of the forward
 * pass is kept but prob).  prob: device f64 N*H, the softmax of LL of rsp_rot_estep at the
 * same inputs.  Outputs (device): g_model c16 C x npix as dL/dRe + i dL/dIm; g_sx, g_sy,
 * g_dia f64 N.
 * dL/dW = -P (K/W - 1) (only -P where W < 1e-20, the log clamp of kernels.cu:37-40),
 * scattered through |O_r + R_d|^2, the sphere/ramp term of kernels.cu:44-82 and the
 * transpose of the bilinear rotation of kernels.cu:6-42 (slice_merge's scatter,
 * kernels.cu:84-114, in gather form).  No reference counterpart: maxLP.py:199-214
 * (get_grad_pixel, never called) is the nearest code.  Synchronises `stream`. */
int rsp_rot_adjoint(rsp_rot* plan, const double* model_c, const double* sx, const double* sy,
                    const double* dia, const double* scales, const double* prob,
                    double* g_model, double* g_sx, double* g_sy, double* g_dia,
                    rsp_stream_t stream);
/* Streamed frames.  rsp_rot_pack (host, synchronous) converts num_data frames of the
 * dset.py layout into the plan's scored CSR in both orders: fo (num_data+1 i64 offsets),
 * fj (i32 compact pixel), fk (f32 count) by frame; po (K+1 i64), pd (i32 frame index within
 * the chunk), pk (f32) by pixel.  With all six outputs NULL it only returns the entry count
 * in *nnz (sizes: fj/fk/pd/pk nnz, po K+1 with K = rsp_rot_info()[1]).
 * A pass evaluates one model over chunks of frames whose packed arrays the caller has
 * copied to the device (for example from pinned host memory on a copy stream):
 *   rsp_rot_pass_begin(plan, model, flags, max_chunk, stream)   flags: 1 merge,
 *       2 frame gradients, 4 model gradient (the loss of rsp_rot_adjoint summed over all
 *       chunks' frames)
 *   rsp_rot_pass_chunk(plan, n, fo, fj, fk, po, pd, pk, sx, sy, dia, scales, ll, prob,
 *       logz, best, g_sx, g_sy, g_dia, stream)   all device, chunk-local (n frames);
 *       po/pd/pk only needed for flags 1 and 4; ll, prob, logz, best may be NULL;
 *       g_* needed for flag 2.  Work is queued on `stream` (no synchronisation).
 *   rsp_rot_pass_end(plan, model_out, acc_out, weights_out, g_model, stream)   merged
 *       intensities (C x npix f64, flag 1) and model gradient (C x npix c16, flag 4);
 *       synchronises `stream`.
 * The chunks of a pass must be queued on one stream in order. */
int rsp_rot_pack(const rsp_rot* plan, int64_t num_data, const int32_t* ones,
                 const int32_t* multi, const int32_t* place_ones, const int32_t* place_multi,
                 const int32_t* count_multi, int64_t* fo, int32_t* fj, float* fk, int64_t* po,
                 int32_t* pd, float* pk, int64_t* nnz);
int rsp_rot_pass_begin(rsp_rot* plan, const double* model_c, int flags, int64_t max_chunk,
                       rsp_stream_t stream);
int rsp_rot_pass_chunk(rsp_rot* plan, int64_t n, const int64_t* fo, const int32_t* fj,
                       const float* fk, const int64_t* po, const int32_t* pd, const float* pk,
                       const double* sx, const double* sy, const double* dia,
                       const double* scales, double* ll, double* prob, double* logz,
                       int32_t* best, double* g_sx, double* g_sy, double* g_dia,
                       rsp_stream_t stream);
int rsp_rot_pass_end(rsp_rot* plan, double* model_out, double* acc_out, double* weights_out,
                     double* g_model, rsp_stream_t stream);
/* host double[5]: ms of slices, reference rows, GEMM, log term + LSE; chunks. */
int rsp_rot_timing(const rsp_rot* plan, double* out);
/* host int64[10]: frames, scored pixels K, padded GEMM K, padded hypotheses (GEMM rows),
 * chunk, photon entries in the scored pixels, distinct radii, classes C, hypotheses H,
 * hypothesis stride of the pixel-major arrays. */
int rsp_rot_info(const rsp_rot* plan, int64_t* out);
int rsp_rot_destroy(rsp_rot* plan);

#ifdef __cplusplus
}
#endif

#endif /* REFSPI_AMD_H */
