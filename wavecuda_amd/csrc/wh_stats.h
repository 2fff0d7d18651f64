// wh_stats.h -- thresholding / statistics back ends shared by stats.hip and
// smooth.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wavehip.h"

namespace wh {

// per-level (or single) threshold handed to the in-place kernels
struct Thr1 {
  int hardness, minlevel, maxlevel;
  double value;
  const double* levels;  // device, one per level, or nullptr
};

// Values taking part in a median:
//   mode 0: base[off .. off+count)
//   mode 1: MODWT detail blocks 0..(count/n - 1): base[2 n b + n + i]
//   mode 2: DWT details of 0-based levels [lmin, lmax] (positions with
//           ctz in range and < J) of an n-long interleaved vector
struct SelSrc {
  const void* base;
  int mode;
  int64_t n;
  int64_t count;
  int64_t off;
  int lmin, lmax, J;
};

// device-resident selection state
constexpr int kSelCandCap = 16384;  // keys the finishing workgroup keeps in LDS

struct SelState {
  uint64_t prefix[2];
  long long k[2];
  double med;
  double mad;
  unsigned long long ncand;  // keys appended to cand by the compaction pass
  uint64_t cmin[2], cmax[2];  // smallest / largest key under each prefix
  unsigned long long hist[2 * 2048];
  uint64_t cand[kSelCandCap];
};

constexpr int kMaxPartials = 1024;

// Golden-section search of the cross-validated threshold, kept on the device
// (reference src/thresh.cpp:296-330): the bracket t0 <= t1 <= t2 <= t3, the
// errors m1, m2 at t1, t2, the loop count, which point the pending
// evaluation is for (1: t1, 2: t2), evaluations consumed so far, and stop.
struct CvState {
  double t0, t1, t2, t3, m1, m2;
  int iter, pending, nevals, stop;
};

// in-place threshold of `count` transformed signals (DWT: n apart, MODWT:
// 2nJ apart)
int threshold_inplace(void* c, int64_t n, int64_t count, int J, int ttype, int dtype,
                      const wh_thresh* thr, hipStream_t st);
int select_median(const SelSrc& s, int dtype, bool deviation, SelState* st, double factor,
                  double* out, hipStream_t stream);
int select_mad(const SelSrc& s, int dtype, SelState* st, double factor, double* out,
               hipStream_t stream);
// search state the CV error reduction advances (cs == nullptr: none)
struct CvArgs {
  CvState* cs = nullptr;
  double* levels = nullptr;
  int J = 0;
  double tol = 0.0;
};
int interp_mse(const void* xn, const void* ye, const void* yo, int64_t len, int dtype,
               double* partial_buf, double* out, hipStream_t st, const CvArgs& cv = CvArgs{});
int split_halves(const void* x, void* e, void* o, int64_t len, int dtype, hipStream_t st);
int broadcast_threshold(double* dst, const double* src, double scale, int J, hipStream_t st);
// bracket [0, *tc] and the first point t1 into levels[0..J)
int cv_init(CvState* cs, const double* tc, double* levels, int J, hipStream_t st);


}  // namespace wh
