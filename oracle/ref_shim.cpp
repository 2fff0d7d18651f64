// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points over a build of the reference package's own
// serial CPU sources (compiled in place from the reference tree by
// oracle/Makefile into oracle/_ref/libwvref.so).  They reproduce what the
// reference's R glue does with the same arguments
// (src/rfunctions.cu:3-100) so that tests/golden/make_golden.py can record
// golden vectors and tests can compare the oracle against the reference
// itself.  Never linked into the product.
#include "thresh.h"
#include "transform.h"
#include "utils.h"
#include "wavutils.h"

extern "C" {

// RcpuTransform (src/rfunctions.cu:3-19)
int ref_transform(double* x, double* xmod, int len, int sense, int nlevels, int ttype,
                  int filter, int filterlen) {
  wst* w = create_wvtstruct(x, xmod, (short)ttype, (short)filter, (short)filterlen,
                            (uint)nlevels, (uint)len);
  if (w == NULL) return -1;
  if (sense == BWD) w->transformed = 1;
  int res = transform(w, (short)sense);
  remove_wvtstruct(w);
  return res;
}

// RcpuThreshold (src/rfunctions.cu:59-69)
int ref_threshold(double* x, double* xmod, int len, int nlevels, int ttype, int filter,
                  int filterlen, double thresh, int hardness, int minlevel, int maxlevel) {
  wst* w = create_wvtstruct(x, xmod, (short)ttype, (short)filter, (short)filterlen,
                            (uint)nlevels, (uint)len);
  if (w == NULL) return -1;
  threshold(w, NULL, thresh, (short)hardness, (uint)minlevel, (uint)maxlevel);
  remove_wvtstruct(w);
  return 0;
}

// RcpuSmooth (src/rfunctions.cu:71-100)
int ref_smooth(double* x, int len, int nlevels, int ttype, int filter, int filterlen,
               int threshtype, double* thresh, int hardness, int minlevel, int maxlevel,
               double tol) {
  wst* w = create_wvtstruct(x, (short)ttype, (short)filter, (short)filterlen, (uint)nlevels,
                            (uint)len);
  if (w == NULL) return -1;
  switch (threshtype) {
    case 2:
      *thresh = CVT(w, (short)hardness, tol, (uint)minlevel, (uint)maxlevel);
      break;
    case 1:
      transform(w, FWD);
      *thresh = univ_thresh(w, 0, 0);
      threshold(w, NULL, *thresh, (short)hardness, (uint)minlevel, (uint)maxlevel);
      transform(w, BWD);
      break;
    case 0:
      transform(w, FWD);
      threshold(w, NULL, *thresh, (short)hardness, (uint)minlevel, (uint)maxlevel);
      transform(w, BWD);
      break;
    default:
      remove_wvtstruct(w);
      return -2;
  }
  if (ttype != DWT) free(w->xmod);
  remove_wvtstruct(w);
  return 0;
}

double ref_median(double* x, int len) { return median(x, (uint)len); }
double ref_mad(double* x, int len) { return mad(x, (uint)len); }

// univ_thresh on an already transformed vector (src/thresh.cpp:170-201)
double ref_univ_thresh(double* x, double* xmod, int len, int nlevels, int ttype, int filter,
                       int filterlen, int minlevel, int maxlevel) {
  wst* w = create_wvtstruct(x, xmod, (short)ttype, (short)filter, (short)filterlen,
                            (uint)nlevels, (uint)len);
  if (w == NULL) return -1.0;
  w->transformed = 1;
  double t = univ_thresh(w, (uint)minlevel, (uint)maxlevel);
  remove_wvtstruct(w);
  return t;
}

// interp_mse (src/thresh.cpp:344-384)
double ref_interp_mse(double* xn, double* ye, double* yo, int len) {
  wst* wn = create_wvtstruct(xn, NULL, DWT, HAAR, 2, 1, (uint)len);
  wst* we = create_wvtstruct(ye, NULL, DWT, HAAR, 2, 1, (uint)(len / 2));
  wst* wo = create_wvtstruct(yo, NULL, DWT, HAAR, 2, 1, (uint)(len / 2));
  double m = interp_mse(wn, we, wo);
  remove_wvtstruct(wn);
  remove_wvtstruct(we);
  remove_wvtstruct(wo);
  return m;
}

unsigned ref_check_len_levels(unsigned len, unsigned nlevels, unsigned filterlength) {
  return check_len_levels(len, nlevels, filterlength);
}

// convert_modwt (src/wavutils.cpp:532-567) on caller buffers: every level's
// 2 len block of xin converted from intype (1 = time ordered, 2 = packet
// ordered) to the other order, into xout
int ref_convert_modwt(double* xin, double* xout, int len, int nlevels, int intype) {
  for (int l = 0; l < nlevels; ++l)
    if (convert_modwt_level(xin + 2 * (size_t)len * l, xout + 2 * (size_t)len * l, (uint)l,
                            (short)intype, (uint)len))
      return 1;
  return 0;
}

}  // extern "C"
