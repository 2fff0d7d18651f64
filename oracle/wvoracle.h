/*
 * wvoracle.h -- CPU oracle for the wavelet hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference package's serial CPU path (src/haar.cpp, src/daub4.cpp,
 * src/c6.cpp, src/la8.cpp, src/thresh.cpp, src/utils.cpp,
 * src/wavutils.cpp, src/transform.cpp, glue semantics of
 * src/rfunctions.cu:3-100).  It exists so that tests/, the smoke check
 * and the CPU-baseline leg of bench.py can check and time the HIP
 * library.  Nothing in the product (wavecuda_amd/) may link or call it.
 *
 * Conventions follow the reference exactly:
 *   sense      FWD=1, BWD=0
 *   ttype      DWT=0, MODWT (time ordered)=1
 *   filter     Haar=1, D4=2, C6=3, LA8=4
 *   hardness   HARD=0, SOFT=1
 *   threshtype MANUAL=0, UNIV=1, CV=2
 *   levels     0-based for thresholding (minlevel..maxlevel)
 *
 * Every function returns 0 on success and a non-zero code on error,
 * except the statistics, which return their value.
 */
#ifndef WVORACLE_H
#define WVORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/wavutils.cpp:48-65 : power-of-two check, maxlevels=floor(log2(n/L))+1,
   nlevels=0 -> max.  Returns the resolved level count, 0 on error. */
unsigned wvo_check_len_levels(unsigned len, unsigned nlevels, unsigned filtlen);
/* src/wavutils.cpp:67-77 */
unsigned wvo_check_len_levels_mm(unsigned len, unsigned nlevels, unsigned minlevel,
                                 unsigned maxlevel, unsigned filtlen);

int wvo_filter_length(int filt);
/* Scaling filter C_0..C_{L-1} of the given filter code (reference constants). */
int wvo_filter_coeffs(int filt, double* c_out);

/* In-place periodic DWT, convolution alignment (src/haar.cpp:32-93,
   src/daub4.cpp:5-88, src/c6.cpp:15-111, src/la8.cpp:10-115). */
int wvo_dwt(double* x, unsigned len, int sense, unsigned nlevels, int filt);

/* Time-ordered MODWT.  FWD reads x and writes xmod = [S1|D1|...|SJ|DJ];
   BWD reads xmod, writes x and overwrites the S_{j-1} slots of xmod
   (src/haar.cpp:590-685, src/daub4.cpp:495-589; C6/LA8 follow the same
   pattern, which the reference does not provide). */
int wvo_modwt(double* x, double* xmod, unsigned len, int sense, unsigned nlevels, int filt);

/* Same transforms, OpenMP over positions within a level (CPU baseline). */
int wvo_modwt_omp(double* x, double* xmod, unsigned len, int sense, unsigned nlevels, int filt);
int wvo_dwt_omp(double* x, unsigned len, int sense, unsigned nlevels, int filt);
/* timing helper for the CPU baseline: reps x (copy, forward, inverse) DWT,
   wall seconds */
double wvo_time_dwt_fwd_inv(const double* x, double* work, unsigned len, unsigned nlevels,
                            int filt, int reps);

/* src/thresh.cpp:142-157 */
double wvo_thresh_coef(double coef, double thresh, int hardness);

/* Threshold details of levels [minlevel,maxlevel] (0-based).  in==out
   allowed (in place, RcpuThreshold semantics, src/rfunctions.cu:59-69 ->
   src/thresh.cpp:113-140).  For MODWT `in`/`out` are xmod arrays of
   2*len*nlevels.  Out-of-place DWT copies every untouched entry
   (see wvo_cvt for the difference to the reference). */
int wvo_threshold(const double* in, double* out, unsigned len, unsigned nlevels,
                  int ttype, int filt, double thresh, int hardness,
                  unsigned minlevel, unsigned maxlevel);

/* src/utils.cpp:166-203 */
double wvo_median(const double* x, unsigned len);
double wvo_mad(const double* x, unsigned len);

/* src/thresh.cpp:170-201 on a transformed vector: x (DWT) or xmod (MODWT). */
double wvo_univ_thresh(const double* x, const double* xmod, unsigned len, unsigned nlevels,
                       int ttype, unsigned minlevel, unsigned maxlevel);

/* src/thresh.cpp:344-384 */
double wvo_interp_mse(const double* xn, const double* ye, const double* yo, unsigned len);

/* src/thresh.cpp:203-342.  x is overwritten with the smoothed signal.
   Returns the chosen threshold; *status gets 0 or an error code.  When
   trace != NULL, up to trace_cap rows of (t0,t1,t2,t3,m1,m2) are written,
   one per golden-section state (initial state first), and *ntrace gets
   the number of rows. */
double wvo_cvt(double* x, unsigned len, unsigned nlevels, int ttype, int filt,
               int hardness, double tol, unsigned minlevel, unsigned maxlevel,
               double* trace, int trace_cap, int* ntrace, int* status);

/* RcpuSmooth semantics (src/rfunctions.cu:71-100): fwd -> threshold
   (manual / universal / CV) -> inverse, x in place; *thresh in/out. */
int wvo_smooth(double* x, unsigned len, unsigned nlevels, int ttype, int filt,
               int threshtype, double* thresh, int hardness,
               unsigned minlevel, unsigned maxlevel, double tol);

/* getCoeffLevel (R/rplotting.R:10-46): the 1-based positions of level
   `level`'s detail (coeff_s = 0) or scaling (coeff_s = 1) coefficients in a
   transformed vector; returns their count (0 on an argument the R code
   stops on).  DWT: seq(2^(level-1) * d + 1, n, by = 2^level); MODWT:
   seq(1, n) whatever the level and type, as the R code has it. */
long wvo_coeff_positions(long len, unsigned nlevels, int ttype, unsigned level, int coeff_s,
                         long* pos);

/* The label columns of WSTtoDT (R/rfunctions.R:925-1004) without its
   plotting rows: CoeffType from label.detail.scaling (:902-912) as 1 = "s",
   0 = "d"; Level; Translate.  DWT: len entries, MODWT: 2 len nlevels. */
int wvo_coeff_labels(long len, unsigned nlevels, int ttype, int* level, double* translate,
                     unsigned char* type);

/* convert_modwt / convert_modwt_level (src/wavutils.cpp:532-600): every
   level's 2 len block of `in` from intype (1 = time ordered, 2 = packet
   ordered) to the other order, into `out`.  Packet `shift` of level l
   (0-based; 2^{l+1} packets of len / 2^l values, (s, d) pairs) holds the
   time positions shift_to + 2^{l+1} k, shift_to = shift with its l + 1 low
   bits reversed (:579-583). */
int wvo_modwt_convert(const double* in, double* out, unsigned len, unsigned nlevels, int intype);

/* The packet-ordered MODWT of src/transform.cpp:39-55 (the reference has
   Haar and D4; the same construction serves C6 and LA8): the time-ordered
   transform and the conversion (forward), or the conversion and the
   time-ordered inverse (BWD; xmod is left in packet order). */
int wvo_modwt_po(double* x, double* xmod, unsigned len, int sense, unsigned nlevels, int filt);

#ifdef __cplusplus
}
#endif
#endif
