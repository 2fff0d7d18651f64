/*
 * wavehip.h -- C ABI of libwavehip.so, the MI355X (gfx950) wavelet engine.
 *
 * Two layers:
 *
 * 1. The reference package's `.C` entry points, same names, argument
 *    lists and meaning (reference src/rfunctions.cuh:17-35, bound by
 *    NAMESPACE:26-30 and called from R/rfunctions.R:394-403, :455-464,
 *    :596-609).  They take host pointers, copy to the GPU, run the HIP
 *    kernels and copy back; they never call exit() and report errors on
 *    stderr, leaving the outputs untouched.
 *
 * 2. `wh_*` functions on device pointers with a status return, an
 *    explicit element type, 64-bit lengths and a caller stream (a
 *    hipStream_t passed as void*; NULL = the null stream).  Work is
 *    enqueued asynchronously unless stated otherwise.  Device buffers of
 *    the transforms and thresholds must be aligned to two elements (16
 *    bytes for FP64, 8 for FP32; hipMalloc gives 256); WH_ERR_ARG otherwise.
 *
 *    Scratch: the library keeps per-device work buffers, grown on demand
 *    and held until exit (the MODWT inverse's level hand-off, DWT staging,
 *    the selection and cross-validation state, smoothing workspaces).
 *    Calls on one device that use them must not overlap in time: issue
 *    them on one stream, or synchronize between streams.  Batch plans own
 *    their workspace.  Error messages are per thread.
 *
 * Codes are the reference's: sense FWD=1/BWD=0, ttype DWT=0/MODWT=1
 * (time ordered), filter Haar=1/D4=2/C6=3/LA8=4, hardness HARD=0/SOFT=1,
 * threshtype MANUAL=0/UNIV=1/CV=2, levels 0-based for thresholding.
 *
 * Layouts (reference src/haar.cpp:9-30, src/daub4.cpp:525-528):
 *   DWT   in-place interleaved: y[m 2^J] = s_J[m], y[2^{j-1} + m 2^j] = d_j[m]
 *   MODWT time ordered: xmod[2n(j-1) + i] = S_j[i], xmod[2n(j-1) + n + i] = D_j[i]
 */
#ifndef WAVEHIP_H
#define WAVEHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define WH_OK 0
#define WH_ERR_ARG 1         /* invalid length / levels / code / pointer */
#define WH_ERR_HIP 2         /* HIP runtime error (message in wh_last_error) */
#define WH_ERR_NOMEM 3       /* device allocation failed */
#define WH_ERR_UNSUPPORTED 4 /* combination not provided */
#define WH_ERR_NODEVICE 5    /* no usable gfx950 device */

/* element types */
#define WH_F64 0
#define WH_F32 1

/* Thresholding specification applied to detail coefficients of the
   0-based levels minlevel..maxlevel (rule of reference
   src/thresh.cpp:142-157: |c| < t -> 0, else hard keeps c, soft gives
   sign(c)(|c|-t)).  `levels`, when not NULL, is a DEVICE array of one
   threshold per transform level (index = 0-based level) and overrides
   `value`. */
typedef struct wh_thresh {
  int enabled;
  int hardness;
  int minlevel;
  int maxlevel;
  double value;
  const double* levels;
} wh_thresh;

/* ---- library --------------------------------------------------------- */

/* Message of the last error on the calling thread (never NULL). */
const char* wh_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int wh_version(void);
/* Optional: select and initialise a device (the current one when < 0);
   also places wh_modwt_inv_noalloc's static scratch on it. */
int wh_init(int device);
/* Resolved level count: reference src/wavutils.cpp:48-65, 0 on error. */
int wh_resolve_levels(int64_t n, int nlevels, int filter);

/* ---- reference .C entry points (host pointers, synchronous) ----------- */

/* The five symbols the reference package binds (NAMESPACE:26-30,
   src/rfunctions.cuh:17-35), same argument lists and codes.  All five run
   on the GPU: the "cpu" names exist so that the package's CPU functions
   (CPUTransform, CPUReconstruct, CPUThreshold, CPUSmooth) find their
   symbols; they compute what the reference's CPU path computes (the
   parity target of every kernel here).  Host arrays move through a pinned
   staging ring and device buffers that persist across calls.  One call at
   a time per device (the calls serialise on a per-device lock).

   RcpuTransform / RgpuTransform: reference src/rfunctions.cu:3-57.  DWT:
   x transformed in place.  MODWT (ttype 1): FWD x -> xmod (2 len nlevels
   values), BWD xmod -> x.  ttype 2 is the packet-ordered MODWT of the
   reference's CPU dispatch (src/transform.cpp:39-55), here for all four
   filters.  A BWD MODWT leaves xmod as it was (the reference's CPU
   routine also rewrites its S_j slots, j < nlevels, with intermediate
   reconstructions that the R functions discard: they return x). */
void RcpuTransform(double* x, double* xmod, int* len, int* sense, int* nlevels, int* ttype,
                   int* filter, int* filterlen);
void RgpuTransform(double* x, double* xmod, int* len, int* sense, int* nlevels, int* ttype,
                   int* filter, int* filterlen);
/* RcpuThreshold: reference src/rfunctions.cu:59-69 (src/thresh.cpp:113-140).
   The details of the 0-based levels minlevel..maxlevel of x (DWT) or xmod
   (MODWT, both orders) thresholded in place at *thresh. */
void RcpuThreshold(double* x, double* xmod, int* len, int* nlevels, int* ttype, int* filter,
                   int* filterlen, double* thresh, int* hardness, int* minlevel, int* maxlevel);
/* RcpuSmooth / RgpuSmooth: reference src/rfunctions.cu:71-136.  x smoothed
   in place; threshtype MANUAL (*thresh is the value), UNIV or CV (*thresh
   receives the threshold chosen); all four filters, both transforms (the
   reference's RgpuSmooth offers MANUAL and CV only). */
void RcpuSmooth(double* x, int* len, int* nlevels, int* ttype, int* filter, int* filterlen,
                int* threshtype, double* thresh, int* hardness, int* minlevel, int* maxlevel,
                double* tol);
void RgpuSmooth(double* x, int* len, int* nlevels, int* ttype, int* filter, int* filterlen,
                int* threshtype, double* thresh, int* hardness, int* minlevel, int* maxlevel,
                double* tol);
/* RgpuSmoothList: the `.C` form of the batched smoother (wh_batch_* below),
   in place of the reference's list stub RgpuTransformList
   (src/rfunctions.cu:139-154; R/rfunctions.R:614-643 commented).  x holds
   *count signals back to back (signal i: lens[i] samples); each is smoothed
   in place with its own filter, transform (ttypes: 0 DWT, 1 MODWT) and
   level count (nlevels, 0 = maximum; MODWT depth <= 8) by universal
   thresholding of every level; thresh (may be NULL) receives the *count
   thresholds.  The plan of the last signal table is kept and reused. */
void RgpuSmoothList(double* x, int* count, int* lens, int* filters, int* ttypes, int* nlevels,
                    int* hardness, double* thresh);

/* ---- device transforms ------------------------------------------------ */

/* MODWT forward: x (n) -> xmod (2 n J).  reference src/daub4.cpp:510-533
   (time-ordered pattern; C6/LA8 extend it). */
int wh_modwt_fwd(const void* x, void* xmod, int64_t n, int nlevels, int filter, int dtype,
                 void* stream);
/* MODWT inverse: xmod (2 n J) -> x (n), with optional fused thresholding
   of the detail blocks as they are read.  xmod is not modified.
   reference src/daub4.cpp:535-589. */
int wh_modwt_inv(const void* xmod, void* x, int64_t n, int nlevels, int filter, int dtype,
                 const wh_thresh* thr, void* stream);
/* DWT forward: x (n) -> y (n), interleaved layout.  y may equal x.
   reference src/la8.cpp:10-76 (and the Haar/D4/C6 equivalents). */
int wh_dwt_fwd(const void* x, void* y, int64_t n, int nlevels, int filter, int dtype,
               void* stream);
/* DWT inverse: y (n) -> x (n) with optional fused thresholding.  x may
   equal y.  reference src/la8.cpp:78-115. */
int wh_dwt_inv(const void* y, void* x, int64_t n, int nlevels, int filter, int dtype,
               const wh_thresh* thr, void* stream);

/* In-place thresholding of a transformed vector (DWT: n values, MODWT:
   2 n J values).  reference src/thresh.cpp:15-83. */
int wh_threshold(void* coeffs, int64_t n, int nlevels, int ttype, int filter, int dtype,
                 const wh_thresh* thr, void* stream);

/* Every level of a MODWT coefficient vector (2 n J values) between time
   order (ttype 1) and packet order (reference src/wavutils.cpp:532-600,
   convert_modwt): to_po = 1 reads time order and writes packet order,
   to_po = 0 the reverse.  Level l (0-based) of the packet order holds
   2^{l+1} packets of n / 2^l values, (s, d) pairs, packet `shift` carrying
   the time positions rev(shift) + 2^{l+1} k (rev: the l + 1 low bits
   reversed).  Out of place (in != out); a permutation, so the result is
   exact.  The packet-ordered transforms are the time-ordered ones plus
   this conversion (RgpuTransform with ttype 2). */
int wh_modwt_convert(const void* in, void* out, int64_t n, int nlevels, int filter, int to_po,
                     int dtype, void* stream);

/* The coefficients of one level gathered into the compact device array
   `out`, in translate order: the values at the positions the reference's
   getCoeffLevel lists (R/rplotting.R:10-46).  DWT: details of the 1-based
   `level` (n / 2^level values from positions 2^{level-1} + m 2^level) or,
   with coeff_type 1 and level == nlevels, the scaling coefficients
   (n / 2^nlevels values from positions m 2^nlevels).  MODWT (time ordered):
   D_level (coeff_type 0) or S_level (coeff_type 1), n values.  `coeffs`
   holds a transform of length n with `nlevels` levels (0 = the maximum);
   *count, when not NULL, receives the number of values written. */
int wh_coeff_level(const void* coeffs, void* out, int64_t n, int nlevels, int ttype, int filter,
                   int level, int coeff_type, int dtype, int64_t* count, void* stream);

/* The label columns the reference's WSTtoDT attaches to every coefficient
   (R/rfunctions.R:925-1004; types as label.detail.scaling, :902-912), in
   storage order: n entries for a DWT, 2 n J for a MODWT.  level: 1-based
   level; translate: DWT 2^{l-1} (m + 1/2) for coefficient m of level l
   (scaling coefficients share level J's translates), MODWT i + 1 within each
   block; type: 0 = "d", 1 = "s".  Any of the three device arrays may be
   NULL. */
int wh_coeff_labels(int64_t n, int nlevels, int ttype, int filter, int* level, double* translate,
                    unsigned char* type, void* stream);

/* The data the reference's WSTtoWavethresh hands to the wavethresh R
   package (R/rfunctions.R:864-893): it builds a zero wd object of n values
   (wavethresh level count log2 n; type "wavelet" for a DWT, "station" for
   a MODWT; periodic boundary), then for l = 1..nlevels calls putD at
   wavethresh level log2(n) - l with the level-l details in translate order,
   then putC at level log2(n) - nlevels (R's loop variable keeps its last
   value) with the level-nlevels scaling coefficients.

   wh_wavethresh_layout (host only, no device needed) fills table[0..J] in
   that call order and returns J + 1 (negative status on error):
   wt_level = the wavethresh level, coeff_type 0 = putD / 1 = putC, count
   values at offset within the packed buffer.  *filter_number and family
   receive the wd arguments (Haar 1 and D4 2 "DaubExPhase", C6 3 and LA8 4
   "DaubLeAsymm"; the reference's switch keys LA8 as "LA", so its own call
   leaves filter.number unset for an "LA8" object); *warnings the reference's
   warning flags, 1: C6 (wavethresh has no coiflets, so it cannot
   reconstruct), 2: MODWT (wavethresh orders those coefficients
   differently).  wh_wavethresh_pack gathers every vector into `out`
   (DEVICE, sum of the counts: n for a DWT, n (J + 1) for a MODWT) in one
   stream-ordered pass per vector.  The wd object itself belongs to
   wavethresh (an R package absent here): these give its inputs. */
typedef struct wh_wt_entry {
  int wt_level;
  int coeff_type;
  int64_t offset;
  int64_t count;
} wh_wt_entry;
int wh_wavethresh_layout(int64_t n, int nlevels, int ttype, int filter, wh_wt_entry* table,
                         int cap, int* filter_number, char* family, int family_cap,
                         int* warnings);
int wh_wavethresh_pack(const void* coeffs, void* out, int64_t n, int nlevels, int ttype,
                       int filter, int dtype, void* stream);

/* Plan tuning for the fused MODWT: times the planner's alternatives (depth
   of the first fused level group, segments per CU, a 256-thread first
   inverse group) for this length, level count, filter and element type on
   scratch buffers of its own (allocated and freed here; about (2J + 2) n
   elements) and keeps, for later calls in this process, the default plan
   unless another is more than 2% faster.  Results do not change: every
   plan computes each coefficient with the same operations.
   Synchronous.  No counterpart in the reference, whose GPU MODWT runs one
   kernel per level (src/haarcuda.cu:1649-1914). */
int wh_modwt_tune(int64_t n, int nlevels, int filter, int dtype, void* stream);

/* Diagnostics of the MODWT planner (no counterpart in the reference).
   wh_debug_modwt_plan / _plan_many write the level groups a forward
   (inverse = 0) or inverse launch of 1 / `count` signals runs, as text
   "{j0= g= R= NT= M= nseg= P= warm= chunks= lds=} ...", into buf.
   wh_debug_modwt_set_plan sets the planner's choices for one shape as
   wh_modwt_tune would (deepest first group g1max, segments per CU of the
   first and deeper groups, 512 or 256 threads in the first inverse group);
   g1max <= 0 restores the default rules.  Results never depend on the
   plan; these let a caller run and check every plan the tuner can pick. */
int wh_debug_modwt_plan(int inverse, int64_t n, int nlevels, int filter, int dtype, char* buf,
                        int cap);
int wh_debug_modwt_plan_many(int inverse, int64_t n, int64_t count, int nlevels, int filter,
                             int dtype, char* buf, int cap);
int wh_debug_modwt_set_plan(int inverse, int64_t n, int nlevels, int filter, int dtype,
                            int g1max, int first_per_cu, int deep_per_cu, int first_nt);

/* ---- many signals, adjoints ------------------------------------------ */

/* The MODWT of `count` signals of n samples at once (1 <= count <= 65535):
   x and the inverse's output hold the signals n apart, xmod holds their
   coefficient blocks 2 n J apart.  Same arithmetic per signal as
   wh_modwt_fwd / wh_modwt_inv (reference src/daub4.cpp:510-589). */
int wh_modwt_fwd_many(const void* x, void* xmod, int64_t n, int64_t count, int nlevels,
                      int filter, int dtype, void* stream);
int wh_modwt_inv_many(const void* xmod, void* x, int64_t n, int64_t count, int nlevels,
                      int filter, int dtype, const wh_thresh* thr, void* stream);
/* wh_modwt_inv_many without any device allocation.  No counterpart in the
   reference; it is for callers that cannot allocate (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  The
   inverse's level groups hand S_j between them through a fixed 16 MiB
   device array that each GPU holds from wh_init (or this function's first
   call there) until the process ends, instead of the grown library
   scratch: n * count may be at most 2^20 (FP64) or 2^21 (FP32), larger
   inputs return WH_ERR_ARG.  Every call on a GPU shares that GPU's array
   (it needs no clearing: each group writes its part before the next reads
   it), so calls that overlap in time on one GPU race: order them on one
   stream.  Same results as wh_modwt_inv_many, bit for bit. */
int wh_modwt_inv_noalloc(const void* xmod, void* x, int64_t n, int64_t count, int nlevels,
                         int filter, int dtype, const wh_thresh* thr, void* stream);
/* Adjoints (vector-Jacobian products) for differentiation; no counterpart
   in the reference, derived from the transforms above.
   wh_modwt_fwd_adj: cotangent of xmod (every S_j and D_j) -> cotangent of x.
   wh_modwt_inv_adj: cotangent of x -> cotangent of xmod (zero in the S_j
   slots, j < J, which the inverse does not read).
   wh_dwt_fwd_adj / wh_dwt_inv_adj: the DWT is orthogonal; these run the
   inverse / forward kernels on the cotangent. */
int wh_modwt_fwd_adj(const void* gxmod, void* gx, int64_t n, int64_t count, int nlevels,
                     int filter, int dtype, void* stream);
int wh_modwt_inv_adj(const void* gx, void* gxmod, int64_t n, int64_t count, int nlevels,
                     int filter, int dtype, void* stream);
int wh_dwt_fwd_adj(const void* gy, void* gx, int64_t n, int nlevels, int filter, int dtype,
                   void* stream);
int wh_dwt_inv_adj(const void* gx, void* gy, int64_t n, int nlevels, int filter, int dtype,
                   void* stream);
/* wh_threshold over `count` transformed signals (DWT n apart, MODWT 2nJ
   apart).  reference src/thresh.cpp:15-83, :142-157. */
int wh_threshold_many(void* coeffs, int64_t n, int64_t count, int nlevels, int ttype, int filter,
                      int dtype, const wh_thresh* thr, void* stream);
/* Gradient of the thresholding y = threshold(c): given the coefficients c
   before thresholding and the cotangent gy of y, writes gc (the cotangent
   of c; gc may equal gy, or be NULL when only glam is wanted) and, when
   glam is not NULL, glam[l] (DEVICE, one
   double per level, J entries) = sum over all signals of dL/dt_l: for the
   soft rule -sign(c) gy over |c| > t_l, for the hard rule 0 (almost
   everywhere); levels outside [minlevel, maxlevel] get 0.  At |c| = t the
   soft derivative is taken as 0 (the rule maps it to 0). */
int wh_thresh_grad(const void* c, const void* gy, void* gc, double* glam, int64_t n,
                   int64_t count, int nlevels, int ttype, int filter, int dtype,
                   const wh_thresh* thr, void* stream);

/* ---- statistics ------------------------------------------------------- */

/* Exact median / MAD (1.4826 median|x - median x|) of `count` values,
   result written as one double to the DEVICE pointer `out`.
   reference src/utils.cpp:166-203. */
int wh_median(const void* x, int64_t count, int dtype, double* out, void* stream);
int wh_mad(const void* x, int64_t count, int dtype, double* out, void* stream);
/* Universal threshold of a transformed vector over 0-based levels
   [minlevel, maxlevel], one double to DEVICE `out`.
   reference src/thresh.cpp:170-201. */
int wh_univ_thresh(const void* coeffs, int64_t n, int nlevels, int ttype, int filter,
                   int dtype, int minlevel, int maxlevel, double* out, void* stream);
/* Interpolation error of the CV scheme, one double to DEVICE `out`.
   reference src/thresh.cpp:344-384. */
int wh_interp_mse(const void* xn, const void* ye, const void* yo, int64_t n, int dtype,
                  double* out, void* stream);

/* ---- smoothing -------------------------------------------------------- */

/* Forward transform -> threshold (MANUAL / UNIV / CV) -> inverse, x (device,
   n values) smoothed in place.  *thresh (host) is the manual threshold on
   input and receives the chosen one for UNIV / CV.  Synchronous.
   reference src/rfunctions.cu:71-136, src/thresh.cpp:203-342. */
int wh_smooth(void* x, int64_t n, int nlevels, int ttype, int filter, int dtype,
              int threshtype, double* thresh, int hardness, int minlevel, int maxlevel,
              double tol, void* stream);

/* ---- batched smoothing ------------------------------------------------ */

/* Universal-threshold smoothing of many independent signals in one
   persistent launch.  Replaces the reference's planned list API
   RgpuTransformList (src/rfunctions.cu:139-154, a stub returning its first
   argument; R/rfunctions.R:614-643 commented GPUTransformList), and per
   signal computes exactly what the smoothing entry points do with threshtype
   UNIV over the full level range (src/rfunctions.cu:82-87: univ_thresh(w,
   0, 0), then threshold levels 0..J-1, then the inverse; the reference's
   GPU RgpuSmooth offers MANUAL and CV only, src/rfunctions.cu:101-136).

   Signals are packed back to back in the order given: signal i occupies
   samples [off_i, off_i + lens[i]) with off_i = sum of the previous
   lengths.  Per signal: power-of-two length, filter 1..4, ttype 0 (DWT) or
   1 (MODWT), levels as RgpuSmooth's nlevels (0 = the reference's maximum;
   the MODWT is fused in one group, so its depth must be <= 8).  hardness is
   shared by the batch.  A plan owns its device tables, workspace and
   captured graph; calls on one plan must be ordered (one stream). */
typedef struct wh_batch wh_batch;

int wh_batch_create(const int* lens, const int* filters, const int* ttypes, const int* levels,
                    int count, int dtype, int hardness, wh_batch** out);
void wh_batch_destroy(wh_batch* plan);
/* x, y: DEVICE packed signals (y may equal x); thresh_out: optional DEVICE
   array, one threshold per signal. Asynchronous on `stream`. */
int wh_batch_smooth(wh_batch* plan, const void* x, void* y, double* thresh_out, void* stream);
/* Time the launch shapes (threads per signal, resident workgroups) on these
   buffers and keep the fastest in the plan and in a process-wide table keyed
   by the batch's composition (later plans of the same composition start
   tuned).  Synchronous; y receives the smoothed signals. */
int wh_batch_tune(wh_batch* plan, const void* x, void* y, double* thresh_out, void* stream);
/* Capture wh_batch_smooth on these buffers into a HIP graph held by the
   plan (replacing an earlier one), and replay it on `stream`. */
int wh_batch_capture(wh_batch* plan, const void* x, void* y, double* thresh_out);
int wh_batch_replay(wh_batch* plan, void* stream);
/* Plan facts: total samples, threads per signal, workgroups, dynamic LDS
   bytes per workgroup, workspace bytes, whether the shape came from a
   tuning run (1) or the default rule (0). Any pointer may be NULL. */
int wh_batch_info(const wh_batch* plan, int64_t* total, int* threads, int* workgroups,
                  int64_t* lds_bytes, int64_t* workspace_bytes, int* tuned);
/* Diagnostics (no counterpart in the reference): set the plan's launch
   shape as wh_batch_tune would -- threads per signal 128, 256 or 512 and
   1..the resident maximum of workgroups -- growing the workspace if needed
   and dropping a captured graph.  Results never depend on the shape; this
   lets a caller run and check every shape the tuner can pick. */
int wh_debug_batch_set_shape(wh_batch* plan, int threads, int workgroups);

#ifdef __cplusplus
}
#endif
#endif
