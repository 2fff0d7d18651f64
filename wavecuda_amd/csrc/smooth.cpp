// smooth.cpp -- forward transform -> threshold -> inverse on the GPU
// (reference src/rfunctions.cu:71-136 and src/thresh.cpp:170-342).
//
// The coefficient arrays stay on the device for the whole pipeline.  The
// threshold is applied inside the inverse transform as the details are
// read, so smoothing costs one forward and one inverse pass.  Cross
// validation follows the reference's CPU algorithm (src/thresh.cpp:203-342)
// step for step: the universal threshold of the full transform brackets a
// golden-section search over the threshold of the even/odd half-length
// transforms, each candidate costing one fused threshold+inverse launch over
// both halves and one interpolation-error reduction.  The search's scalar
// steps run on the device with the reference's arithmetic (cv_advance), so
// candidates follow one another without the host; the host reads the
// search state back only to learn how many more candidates the loop is
// certain to evaluate.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "wh_common.h"
#include "wh_internal.h"
#include "wh_stats.h"

namespace wh {

namespace {

struct StatsBuf {
  SelState sel;
  double partial[kMaxPartials];
  double vals[16];
  double levels[64];
  double cvlev[64];  // the candidate threshold of every level
  CvState cv;
};

int stats_buf(StatsBuf** out) {
  void* p = nullptr;
  int rc = stats_get(sizeof(StatsBuf), &p);
  if (rc) return rc;
  *out = static_cast<StatsBuf*>(p);
  return WH_OK;
}

int64_t coeff_len(int ttype, int64_t n, int J) { return ttype == kDWT ? n : 2 * n * J; }

int fwd(int ttype, const void* x, void* w, int64_t n, int J, int filt, int dtype,
        hipStream_t st) {
  return ttype == kDWT ? dwt_forward(x, w, n, J, filt, dtype, st)
                       : modwt_forward(x, w, n, 1, J, filt, dtype, st);
}

int inv(int ttype, const void* w, void* x, int64_t n, int J, int filt, int dtype,
        const wh_thresh* thr, hipStream_t st) {
  return ttype == kDWT ? dwt_inverse(w, x, n, J, filt, dtype, thr, st)
                       : modwt_inverse(w, x, n, 1, J, filt, dtype, thr, st);
}

// Candidates the search will certainly evaluate from state `cs` on (not
// stopped): the pending one, then one more for every loop test that must
// pass -- the bracket narrows by R per step and t1 + t2 <= 2 t3, so
// |t3 - t0| R^j > 2 tol t3 (with a margin for rounding) keeps the loop going,
// as long as the iteration cap allows.
int cv_sure_evals(const CvState& cs, double tol) {
  const double R = 0.61803399, margin = 1. - 1e-6;
  double w = std::fabs(cs.t3 - cs.t0) * margin;
  const double lim = tol * 2. * cs.t3;
  int it = cs.iter + (cs.nevals >= 2 ? 1 : 0);
  int k = 1;
  while (k < 64 && it <= 50 && w > lim) {
    ++k;
    ++it;
    w *= R * margin;
  }
  return k;
}

int d2h(double* host, const double* dev, hipStream_t st) {
  WH_HIP_CHECK(hipMemcpyAsync(host, dev, sizeof(double), hipMemcpyDeviceToHost, st));
  WH_HIP_CHECK(hipStreamSynchronize(st));
  return WH_OK;
}

}  // namespace

// src/thresh.cpp:170-201 (with the time-ordered MODWT gathering detail
// blocks from level 0 whatever minlevel is, src/wavutils.cpp:711-717)
int univ_thresh_dev(const void* c, int64_t n, int J, int ttype, int dtype, int minlevel,
                    int maxlevel, double* out, hipStream_t st) {
  StatsBuf* sb = nullptr;
  int rc = stats_buf(&sb);
  if (rc) return rc;
  SelSrc s{};
  s.base = c;
  s.n = n;
  double factor;
  if (ttype == kDWT) {
    if (minlevel == 0 && maxlevel == J - 1) {
      s.mode = 0;
      s.off = 1;
      s.count = n - 1;
      factor = std::sqrt(2. * std::log(double(n - 1)));
    } else {
      s.mode = 2;
      s.lmin = minlevel;
      s.lmax = maxlevel;
      s.J = J;
      s.count = 0;
      for (int l = minlevel; l <= maxlevel; ++l) s.count += n >> (l + 1);
      factor = std::sqrt(2. * std::log(double(n)));
    }
  } else {
    s.mode = 1;
    s.count = int64_t(maxlevel - minlevel + 1) * n;
    factor = std::sqrt(2. * std::log(double(n)));
  }
  return select_mad(s, dtype, &sb->sel, factor, out, st);
}

int smooth(void* x, int64_t n, int J, int ttype, int filt, int dtype, int threshtype,
           double* thresh, int hardness, int minlevel, int maxlevel, double tol,
           hipStream_t st) {
  const size_t esz = dtype == WH_F32 ? 4 : 8;
  StatsBuf* sb = nullptr;
  int rc = stats_buf(&sb);
  if (rc) return rc;
  if (minlevel < 0 || minlevel > maxlevel || maxlevel >= J)
    return set_error(WH_ERR_ARG, "levels must satisfy 0 <= minlevel <= maxlevel < nlevels");
  if (hardness != kHARD && hardness != kSOFT) return set_error(WH_ERR_ARG, "hardness must be 0 or 1");

  if (threshtype == 0 || threshtype == 1) {  // MANUAL / UNIV
    void* w = nullptr;
    if ((rc = ws_get(size_t(coeff_len(ttype, n, J)) * esz, &w))) return rc;
    if ((rc = fwd(ttype, x, w, n, J, filt, dtype, st))) return rc;
    wh_thresh t{};
    t.enabled = 1;
    t.hardness = hardness;
    t.minlevel = minlevel;
    t.maxlevel = maxlevel;
    if (threshtype == 0) {
      if (!thresh) return set_error(WH_ERR_ARG, "manual threshold needs a value");
      t.value = *thresh;
    } else {
      // univ_thresh(w, 0, 0) (src/rfunctions.cu:82-87)
      if ((rc = univ_thresh_dev(w, n, J, ttype, dtype, 0, 0, &sb->vals[0], st))) return rc;
      if ((rc = broadcast_threshold(sb->levels, &sb->vals[0], 1.0, J, st))) return rc;
      t.levels = sb->levels;
    }
    if ((rc = inv(ttype, w, x, n, J, filt, dtype, &t, st))) return rc;
    if (threshtype == 1 && thresh) {
      if ((rc = d2h(thresh, &sb->vals[0], st))) return rc;
    } else {
      WH_HIP_CHECK(hipStreamSynchronize(st));
    }
    return WH_OK;
  }
  if (threshtype != 2) return set_error(WH_ERR_ARG, "threshtype must be 0 (manual), 1 (univ) or 2 (cv)");

  // ---- cross validation (src/thresh.cpp:203-342)
  if (J <= 1) return set_error(WH_ERR_ARG, "cross validation needs at least two levels");
  if (minlevel == maxlevel)
    return set_error(WH_ERR_ARG, "cross validation needs minlevel < maxlevel");
  const int64_t lenh = n / 2;
  const int Jh = J - 1;
  const int64_t cw = coeff_len(ttype, n, J), ch = coeff_len(ttype, lenh, Jh);
  void* ws = nullptr;
  const size_t bytes = size_t(cw + 2 * ch + 4 * lenh) * esz + 256;
  if ((rc = ws_get(bytes, &ws))) return rc;
  char* p = static_cast<char*>(ws);
  void* W = p;
  p += cw * esz;
  void* We = p;
  p += ch * esz;
  void* Wo = p;
  p += ch * esz;
  void* xe = p;
  p += lenh * esz;
  void* xo = p;
  p += lenh * esz;
  void* ye = p;
  p += lenh * esz;
  void* yo = p;

  if ((rc = fwd(ttype, x, W, n, J, filt, dtype, st))) return rc;
  if ((rc = univ_thresh_dev(W, n, J, ttype, dtype, minlevel, maxlevel, &sb->vals[0], st))) return rc;
  if ((rc = split_halves(x, xe, xo, n, dtype, st))) return rc;
  // the two half-length MODWTs run as one two-signal launch (xe, xo and
  // We, Wo and ye, yo are adjacent): twice the workgroups, half the launches
  if (ttype == kDWT) {
    if ((rc = fwd(ttype, xe, We, lenh, Jh, filt, dtype, st))) return rc;
    if ((rc = fwd(ttype, xo, Wo, lenh, Jh, filt, dtype, st))) return rc;
  } else if ((rc = modwt_forward(xe, We, lenh, 2, Jh, filt, dtype, st))) {
    return rc;
  }

  // one candidate: the halves thresholded at the point cv_advance last wrote
  wh_thresh th{};
  th.enabled = 1;
  th.hardness = hardness;
  th.minlevel = minlevel;
  th.maxlevel = maxlevel - 1;
  th.levels = sb->cvlev;
  CvArgs cva;
  cva.cs = &sb->cv;
  cva.levels = sb->cvlev;
  cva.J = Jh;
  cva.tol = tol;
  auto eval = [&]() -> int {
    int r;
    if (ttype == kDWT) {
      if ((r = inv(ttype, We, ye, lenh, Jh, filt, dtype, &th, st))) return r;
      if ((r = inv(ttype, Wo, yo, lenh, Jh, filt, dtype, &th, st))) return r;
    } else if ((r = modwt_inverse(We, ye, lenh, 2, Jh, filt, dtype, &th, st))) {
      return r;
    }
    // the error reduction's last block also advances the search
    return interp_mse(x, ye, yo, n, dtype, sb->partial, &sb->vals[1], st, cva);
  };
  if ((rc = cv_init(&sb->cv, &sb->vals[0], sb->cvlev, Jh, st))) return rc;
  CvState cs{};
  for (int batch = 2;;) {  // t1 and t2 first
    for (int i = 0; i < batch; ++i)
      if ((rc = eval())) return rc;
    WH_HIP_CHECK(hipMemcpyAsync(&cs, &sb->cv, sizeof(CvState), hipMemcpyDeviceToHost, st));
    WH_HIP_CHECK(hipStreamSynchronize(st));
    if (cs.stop) break;
    batch = cv_sure_evals(cs, tol);
  }
  const double t1 = cs.t1, t2 = cs.t2, m1 = cs.m1, m2 = cs.m2;
  double t = m1 < m2 ? t1 : t2;
  t = t / (std::sqrt(1. - std::log(2.) / std::log(double(n))));
  wh_thresh tf{};
  tf.enabled = 1;
  tf.hardness = hardness;
  tf.minlevel = minlevel;
  tf.maxlevel = maxlevel;
  tf.value = t;
  if ((rc = inv(ttype, W, x, n, J, filt, dtype, &tf, st))) return rc;
  WH_HIP_CHECK(hipStreamSynchronize(st));
  if (thresh) *thresh = t;
  return WH_OK;
}

}  // namespace wh
