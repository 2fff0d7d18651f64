// capi.cpp -- the C ABI of libwavehip.so: errors, device state, scratch,
// the wh_* device entry points and the reference's .C entry points.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "wh_common.h"
#include "wh_internal.h"
#include "wh_stats.h"

namespace wh {

namespace {
thread_local std::string g_err = "no error";
std::mutex g_mu;

constexpr int kMaxDev = 64;
DeviceInfo g_info[kMaxDev];
std::set<std::pair<int, const void*>> g_lds_set;  // (device, kernel) pairs already set
struct Scratch {
  void* p = nullptr;
  size_t bytes = 0;
};
Scratch g_scratch[kMaxDev][6];

int current_device() {
  int d = -1;
  if (hipGetDevice(&d) != hipSuccess) return -1;
  return d;
}

int scratch_impl(int which, size_t bytes, void** out) {
  int d = current_device();
  if (d < 0 || d >= kMaxDev) return set_error(WH_ERR_NODEVICE, "no current HIP device");
  std::lock_guard<std::mutex> lk(g_mu);
  Scratch& s = g_scratch[d][which];
  if (s.bytes < bytes) {
    if (s.p) {
      // work already queued may still read the old buffer
      if (hipDeviceSynchronize() != hipSuccess) return set_error(WH_ERR_HIP, "sync before scratch grow");
      (void)hipFree(s.p);
      s.p = nullptr;
      s.bytes = 0;
    }
    size_t want = bytes + (bytes >> 3) + 4096;
    if (hipMalloc(&s.p, want) != hipSuccess) {
      s.p = nullptr;
      return set_error(WH_ERR_NOMEM, "scratch allocation failed");
    }
    s.bytes = want;
  }
  *out = s.p;
  return WH_OK;
}
}  // namespace

int set_error(int code, const char* msg) {
  g_err = msg ? msg : "error";
  return code;
}

int set_hip_error(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return WH_ERR_HIP;
}

const DeviceInfo& device_info() {
  static DeviceInfo none;
  int d = current_device();
  if (d < 0 || d >= kMaxDev) return none;
  std::lock_guard<std::mutex> lk(g_mu);
  DeviceInfo& di = g_info[d];
  if (di.device != d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess) {
      di.num_cus = prop.multiProcessorCount;
      di.max_lds_per_block = prop.sharedMemPerBlock;
      di.warp_size = prop.warpSize;
      di.arch = prop.gcnArchName;
    }
    if (di.num_cus <= 0) di.num_cus = 256;
    if (di.max_lds_per_block == 0) di.max_lds_per_block = 65536;
    di.device = d;
  }
  return di;
}

void allow_dynamic_lds(const void* kern) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_lds_set.insert({current_device(), kern}).second) return;
  }
  hipFuncAttributes at{};
  size_t stat = 0;
  if (hipFuncGetAttributes(&at, kern) == hipSuccess) stat = at.sharedSizeBytes;
  const size_t maxb = device_info().max_lds_per_block;
  if (maxb > stat)
    (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(maxb - stat));
  (void)hipGetLastError();
}

int scratch_get(size_t bytes, void** out) { return scratch_impl(0, bytes, out); }
int scratch2_get(size_t bytes, void** out) { return scratch_impl(1, bytes, out); }
int ws_get(size_t bytes, void** out) { return scratch_impl(2, bytes, out); }
int stats_get(size_t bytes, void** out) { return scratch_impl(3, bytes, out); }
int host_buf_get(int which, size_t bytes, void** out) {
  return scratch_impl(4 + (which & 1), bytes, out);
}

}  // namespace wh

using namespace wh;

namespace wh {

bool valid_filter_code(int f) { return f >= kHAAR && f <= kLA8; }
bool valid_dtype_code(int t) { return t == WH_F64 || t == WH_F32; }

int check_common(const void* a, const void* b, int64_t n, int nlevels, int filter, int dtype,
                 int* J) {
  if (!a || !b) return set_error(WH_ERR_ARG, "null pointer");
  if (!valid_dtype_code(dtype)) return set_error(WH_ERR_ARG, "dtype must be WH_F64 or WH_F32");
  // the kernels move element pairs with single vector accesses
  const uintptr_t pair = dtype == WH_F64 ? 16 : 8;
  if ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & (pair - 1))
    return set_error(WH_ERR_ARG, "device buffers must be aligned to two elements (16 bytes FP64, 8 bytes FP32)");
  if (!valid_filter_code(filter)) return set_error(WH_ERR_ARG, "filter must be 1 (Haar), 2 (D4), 3 (C6) or 4 (LA8)");
  *J = resolve_levels(n, nlevels, filter_length(filter));
  if (*J <= 0) return set_error(WH_ERR_ARG, "length must be a power of two and nlevels <= floor(log2(n/L))+1");
  return WH_OK;
}

int check_thresh(const wh_thresh* thr, int J) {
  if (!thr || !thr->enabled) return WH_OK;
  if (thr->hardness != kHARD && thr->hardness != kSOFT) return set_error(WH_ERR_ARG, "hardness must be 0 (hard) or 1 (soft)");
  if (thr->minlevel < 0 || thr->minlevel > thr->maxlevel || thr->maxlevel >= J)
    return set_error(WH_ERR_ARG, "threshold levels must satisfy 0 <= minlevel <= maxlevel < nlevels");
  return WH_OK;
}

int check_count(int64_t count) {
  if (count < 1 || count > 65535) return set_error(WH_ERR_ARG, "count must be 1..65535");
  return WH_OK;
}

}  // namespace wh

namespace {
bool valid_filter(int f) { return wh::valid_filter_code(f); }
bool valid_dtype(int t) { return wh::valid_dtype_code(t); }
}  // namespace

extern "C" {

const char* wh_last_error(void) { return g_err.c_str(); }

int wh_version(void) { return 100; }

int wh_init(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return set_error(WH_ERR_NODEVICE, "no HIP device");
  if (device >= 0) {
    if (device >= count) return set_error(WH_ERR_ARG, "device index out of range");
    WH_HIP_CHECK(hipSetDevice(device));
  }
  const DeviceInfo& di = device_info();
  if (di.arch.rfind("gfx950", 0) != 0) {
    std::string m = "device arch " + di.arch + " is not gfx950";
    return set_error(WH_ERR_NODEVICE, m.c_str());
  }
  return noalloc_setup();
}

int wh_resolve_levels(int64_t n, int nlevels, int filter) {
  if (!valid_filter(filter)) return 0;
  return resolve_levels(n, nlevels, filter_length(filter));
}

int wh_modwt_fwd(const void* x, void* xmod, int64_t n, int nlevels, int filter, int dtype,
                 void* stream) {
  int J, rc = check_common(x, xmod, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  return modwt_forward(x, xmod, n, 1, J, filter, dtype, static_cast<hipStream_t>(stream));
}

int wh_modwt_inv(const void* xmod, void* x, int64_t n, int nlevels, int filter, int dtype,
                 const wh_thresh* thr, void* stream) {
  int J, rc = check_common(xmod, x, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  if ((rc = check_thresh(thr, J))) return rc;
  return modwt_inverse(xmod, x, n, 1, J, filter, dtype, thr, static_cast<hipStream_t>(stream));
}

int wh_dwt_fwd(const void* x, void* y, int64_t n, int nlevels, int filter, int dtype,
               void* stream) {
  int J, rc = check_common(x, y, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  return dwt_forward(x, y, n, J, filter, dtype, static_cast<hipStream_t>(stream));
}

int wh_dwt_inv(const void* y, void* x, int64_t n, int nlevels, int filter, int dtype,
               const wh_thresh* thr, void* stream) {
  int J, rc = check_common(y, x, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  if ((rc = check_thresh(thr, J))) return rc;
  return dwt_inverse(y, x, n, J, filter, dtype, thr, static_cast<hipStream_t>(stream));
}

int wh_threshold(void* coeffs, int64_t n, int nlevels, int ttype, int filter, int dtype,
                 const wh_thresh* thr, void* stream) {
  int J, rc = check_common(coeffs, coeffs, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  if (ttype != kDWT && ttype != kMODWT) return set_error(WH_ERR_ARG, "ttype must be 0 or 1");
  if (!thr || !thr->enabled) return WH_OK;
  if ((rc = check_thresh(thr, J))) return rc;
  return threshold_inplace(coeffs, n, 1, J, ttype, dtype, thr, static_cast<hipStream_t>(stream));
}

int wh_modwt_fwd_many(const void* x, void* xmod, int64_t n, int64_t count, int nlevels,
                      int filter, int dtype, void* stream) {
  int J, rc = check_common(x, xmod, n, nlevels, filter, dtype, &J);
  if (rc || (rc = check_count(count))) return rc;
  return modwt_forward(x, xmod, n, count, J, filter, dtype, static_cast<hipStream_t>(stream));
}

int wh_modwt_inv_many(const void* xmod, void* x, int64_t n, int64_t count, int nlevels,
                      int filter, int dtype, const wh_thresh* thr, void* stream) {
  int J, rc = check_common(xmod, x, n, nlevels, filter, dtype, &J);
  if (rc || (rc = check_count(count))) return rc;
  if ((rc = check_thresh(thr, J))) return rc;
  return modwt_inverse(xmod, x, n, count, J, filter, dtype, thr, static_cast<hipStream_t>(stream));
}

int wh_modwt_fwd_adj(const void* gxmod, void* gx, int64_t n, int64_t count, int nlevels,
                     int filter, int dtype, void* stream) {
  int J, rc = check_common(gxmod, gx, n, nlevels, filter, dtype, &J);
  if (rc || (rc = check_count(count))) return rc;
  return modwt_forward_adjoint(gxmod, gx, n, count, J, filter, dtype,
                               static_cast<hipStream_t>(stream));
}

int wh_modwt_inv_adj(const void* gx, void* gxmod, int64_t n, int64_t count, int nlevels,
                     int filter, int dtype, void* stream) {
  int J, rc = check_common(gx, gxmod, n, nlevels, filter, dtype, &J);
  if (rc || (rc = check_count(count))) return rc;
  return modwt_inverse_adjoint(gx, gxmod, n, count, J, filter, dtype,
                               static_cast<hipStream_t>(stream));
}

int wh_dwt_fwd_adj(const void* gy, void* gx, int64_t n, int nlevels, int filter, int dtype,
                   void* stream) {
  int J, rc = check_common(gy, gx, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  return dwt_inverse(gy, gx, n, J, filter, dtype, nullptr, static_cast<hipStream_t>(stream));
}

int wh_dwt_inv_adj(const void* gx, void* gy, int64_t n, int nlevels, int filter, int dtype,
                   void* stream) {
  int J, rc = check_common(gx, gy, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  return dwt_forward(gx, gy, n, J, filter, dtype, static_cast<hipStream_t>(stream));
}

int wh_threshold_many(void* coeffs, int64_t n, int64_t count, int nlevels, int ttype, int filter,
                      int dtype, const wh_thresh* thr, void* stream) {
  int J, rc = check_common(coeffs, coeffs, n, nlevels, filter, dtype, &J);
  if (rc || (rc = check_count(count))) return rc;
  if (ttype != kDWT && ttype != kMODWT) return set_error(WH_ERR_ARG, "ttype must be 0 or 1");
  if (!thr || !thr->enabled) return WH_OK;
  if ((rc = check_thresh(thr, J))) return rc;
  return threshold_inplace(coeffs, n, count, J, ttype, dtype, thr,
                           static_cast<hipStream_t>(stream));
}

int wh_thresh_grad(const void* c, const void* gy, void* gc, double* glam, int64_t n,
                   int64_t count, int nlevels, int ttype, int filter, int dtype,
                   const wh_thresh* thr, void* stream) {
  int J, rc = check_common(c, gy, n, nlevels, filter, dtype, &J);
  if (rc || (rc = check_count(count))) return rc;
  if (!gc && !glam) return set_error(WH_ERR_ARG, "threshold gradient: gc and glam both NULL");
  if (gc && (reinterpret_cast<uintptr_t>(gc) & (dtype == WH_F64 ? 15 : 7)))
    return set_error(WH_ERR_ARG, "device buffers must be aligned to two elements (16 bytes FP64, 8 bytes FP32)");
  if (ttype != kDWT && ttype != kMODWT) return set_error(WH_ERR_ARG, "ttype must be 0 or 1");
  if (!thr || !thr->enabled) return set_error(WH_ERR_ARG, "threshold gradient needs a threshold");
  if ((rc = check_thresh(thr, J))) return rc;
  return thresh_grad(c, gy, gc, glam, n, count, J, ttype, dtype, thr,
                     static_cast<hipStream_t>(stream));
}

static int stats_state(SelState** st) {
  void* p = nullptr;
  int rc = stats_get(sizeof(SelState) + (kMaxPartials + 80) * sizeof(double), &p);
  if (rc) return rc;
  *st = static_cast<SelState*>(p);
  return WH_OK;
}

int wh_median(const void* x, int64_t count, int dtype, double* out, void* stream) {
  if (!x || !out || count <= 0) return set_error(WH_ERR_ARG, "null pointer or empty input");
  if (!valid_dtype(dtype)) return set_error(WH_ERR_ARG, "bad dtype");
  SelState* st;
  int rc = stats_state(&st);
  if (rc) return rc;
  SelSrc s{};
  s.base = x;
  s.mode = 0;
  s.count = count;
  return select_median(s, dtype, false, st, 0.0, out, static_cast<hipStream_t>(stream));
}

int wh_mad(const void* x, int64_t count, int dtype, double* out, void* stream) {
  if (!x || !out || count <= 0) return set_error(WH_ERR_ARG, "null pointer or empty input");
  if (!valid_dtype(dtype)) return set_error(WH_ERR_ARG, "bad dtype");
  SelState* st;
  int rc = stats_state(&st);
  if (rc) return rc;
  SelSrc s{};
  s.base = x;
  s.mode = 0;
  s.count = count;
  return select_mad(s, dtype, st, 0.0, out, static_cast<hipStream_t>(stream));
}

int wh_univ_thresh(const void* coeffs, int64_t n, int nlevels, int ttype, int filter, int dtype,
                   int minlevel, int maxlevel, double* out, void* stream) {
  // the coefficients are read as element pairs; out is one double
  int J, rc = check_common(coeffs, coeffs, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  if (!out || (reinterpret_cast<uintptr_t>(out) & 7))
    return set_error(WH_ERR_ARG, "out must be a non-null, 8-byte aligned device double");
  if (ttype != kDWT && ttype != kMODWT) return set_error(WH_ERR_ARG, "ttype must be 0 or 1");
  if (minlevel < 0 || minlevel > maxlevel || maxlevel >= J)
    return set_error(WH_ERR_ARG, "levels must satisfy 0 <= minlevel <= maxlevel < nlevels");
  return univ_thresh_dev(coeffs, n, J, ttype, dtype, minlevel, maxlevel, out,
                         static_cast<hipStream_t>(stream));
}

int wh_interp_mse(const void* xn, const void* ye, const void* yo, int64_t n, int dtype,
                  double* out, void* stream) {
  if (!xn || !ye || !yo || !out) return set_error(WH_ERR_ARG, "null pointer");
  if (n < 4 || (n & 1)) return set_error(WH_ERR_ARG, "length must be even and >= 4");
  if (!valid_dtype(dtype)) return set_error(WH_ERR_ARG, "bad dtype");
  void* p = nullptr;
  int rc = stats_get(sizeof(SelState) + (kMaxPartials + 80) * sizeof(double), &p);
  if (rc) return rc;
  double* partial = reinterpret_cast<double*>(static_cast<char*>(p) + sizeof(SelState));
  return interp_mse(xn, ye, yo, n, dtype, partial, out, static_cast<hipStream_t>(stream));
}

int wh_smooth(void* x, int64_t n, int nlevels, int ttype, int filter, int dtype, int threshtype,
              double* thresh, int hardness, int minlevel, int maxlevel, double tol,
              void* stream) {
  int J, rc = check_common(x, x, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  if (ttype != kDWT && ttype != kMODWT) return set_error(WH_ERR_ARG, "ttype must be 0 or 1");
  return smooth(x, n, J, ttype, filter, dtype, threshtype, thresh, hardness, minlevel, maxlevel,
                tol, static_cast<hipStream_t>(stream));
}

int wh_coeff_level(const void* coeffs, void* out, int64_t n, int nlevels, int ttype, int filter,
                   int level, int coeff_type, int dtype, int64_t* count, void* stream) {
  if (!coeffs || !out) return set_error(WH_ERR_ARG, "null pointer");
  if (!valid_filter(filter)) return set_error(WH_ERR_ARG, "filter must be 1 (Haar), 2 (D4), 3 (C6) or 4 (LA8)");
  if (!valid_dtype(dtype)) return set_error(WH_ERR_ARG, "dtype must be WH_F64 or WH_F32");
  if (ttype != kDWT && ttype != kMODWT) return set_error(WH_ERR_ARG, "ttype must be 0 (DWT) or 1 (MODWT)");
  const int J = resolve_levels(n, nlevels, filter_length(filter));
  if (J <= 0) return set_error(WH_ERR_ARG, "length must be a power of two and nlevels <= floor(log2(n/L))+1");
  // getCoeffLevel's rules (R/rplotting.R:17-35)
  if (level < 1 || level > J) return set_error(WH_ERR_ARG, "level must be an integer between 1 and nlevels");
  if (coeff_type != 0 && coeff_type != 1) return set_error(WH_ERR_ARG, "coeff_type must be 0 (d) or 1 (s)");
  if (ttype == kDWT && coeff_type == 1 && level != J)
    return set_error(WH_ERR_ARG, "DWT scaling coefficients exist at level nlevels only");
  return level_extract(coeffs, out, n, J, ttype, level, coeff_type, dtype,
                       static_cast<hipStream_t>(stream), count);
}

int wh_modwt_convert(const void* in, void* out, int64_t n, int nlevels, int filter, int to_po,
                     int dtype, void* stream) {
  int J, rc = check_common(in, out, n, nlevels, filter, dtype, &J);
  if (rc) return rc;
  if (in == out) return set_error(WH_ERR_ARG, "modwt convert: input and output must differ");
  if (to_po != 0 && to_po != 1) return set_error(WH_ERR_ARG, "to_po must be 0 or 1");
  return modwt_order(in, out, n, J, to_po == 1, dtype, static_cast<hipStream_t>(stream));
}

int wh_coeff_labels(int64_t n, int nlevels, int ttype, int filter, int* level, double* translate,
                    unsigned char* type, void* stream) {
  if (!valid_filter(filter)) return set_error(WH_ERR_ARG, "filter must be 1 (Haar), 2 (D4), 3 (C6) or 4 (LA8)");
  if (ttype != kDWT && ttype != kMODWT) return set_error(WH_ERR_ARG, "ttype must be 0 (DWT) or 1 (MODWT)");
  const int J = resolve_levels(n, nlevels, filter_length(filter));
  if (J <= 0) return set_error(WH_ERR_ARG, "length must be a power of two and nlevels <= floor(log2(n/L))+1");
  if (!level && !translate && !type) return WH_OK;
  return level_labels(n, J, ttype, level, translate, type, static_cast<hipStream_t>(stream));
}

int wh_wavethresh_layout(int64_t n, int nlevels, int ttype, int filter, wh_wt_entry* table,
                         int cap, int* filter_number, char* family, int family_cap,
                         int* warnings) {
  if (!valid_filter(filter)) return -set_error(WH_ERR_ARG, "filter must be 1 (Haar), 2 (D4), 3 (C6) or 4 (LA8)");
  if (ttype != kDWT && ttype != kMODWT) return -set_error(WH_ERR_ARG, "ttype must be 0 (DWT) or 1 (MODWT)");
  const int J = resolve_levels(n, nlevels, filter_length(filter));
  if (J <= 0) return -set_error(WH_ERR_ARG, "length must be a power of two and nlevels <= floor(log2(n/L))+1");
  if (!table || cap < J + 1) return -set_error(WH_ERR_ARG, "wavethresh layout: table needs nlevels + 1 entries");
  int lg = 0;
  while ((int64_t(1) << lg) < n) ++lg;  // wavethresh's nlevels for a wd of length n
  int64_t off = 0;
  for (int e = 0; e <= J; ++e) {
    const int l = e < J ? e + 1 : J;  // R's loop variable keeps its last value for putC
    wh_wt_entry& t = table[e];
    t.wt_level = lg - l;
    t.coeff_type = e < J ? 0 : 1;
    t.count = (ttype == kDWT) ? (n >> l) : n;
    t.offset = off;
    off += t.count;
  }
  // switch(XW$filt, "Haar"=1, "D4"=2, "C6"=3, "LA"=4) and the family
  // switch (R/rfunctions.R:866-867)
  static const char* fam[5] = {"", "DaubExPhase", "DaubExPhase", "DaubLeAsymm", "DaubLeAsymm"};
  if (filter_number) *filter_number = filter;
  if (family && family_cap > 0) snprintf(family, size_t(family_cap), "%s", fam[filter]);
  // the reference's warnings (:869-873): 1 = C6 (no coiflets in
  // wavethresh), 2 = MODWT (wavethresh orders it differently)
  if (warnings) *warnings = (filter == kC6 ? 1 : 0) | (ttype == kMODWT ? 2 : 0);
  return J + 1;
}

int wh_wavethresh_pack(const void* coeffs, void* out, int64_t n, int nlevels, int ttype,
                       int filter, int dtype, void* stream) {
  if (!coeffs || !out) return set_error(WH_ERR_ARG, "null pointer");
  if (!valid_filter(filter)) return set_error(WH_ERR_ARG, "filter must be 1 (Haar), 2 (D4), 3 (C6) or 4 (LA8)");
  if (!valid_dtype(dtype)) return set_error(WH_ERR_ARG, "dtype must be WH_F64 or WH_F32");
  if (ttype != kDWT && ttype != kMODWT) return set_error(WH_ERR_ARG, "ttype must be 0 (DWT) or 1 (MODWT)");
  const int J = resolve_levels(n, nlevels, filter_length(filter));
  if (J <= 0) return set_error(WH_ERR_ARG, "length must be a power of two and nlevels <= floor(log2(n/L))+1");
  return wavethresh_pack(coeffs, out, n, J, ttype, dtype, static_cast<hipStream_t>(stream));
}

int wh_modwt_tune(int64_t n, int nlevels, int filter, int dtype, void* stream) {
  if (!valid_filter(filter)) return set_error(WH_ERR_ARG, "filter must be 1..4");
  if (!valid_dtype(dtype)) return set_error(WH_ERR_ARG, "dtype must be WH_F64 or WH_F32");
  const int J = resolve_levels(n, nlevels, filter_length(filter));
  if (J <= 0) return set_error(WH_ERR_ARG, "invalid length / levels");
  return modwt_tune(n, J, filter, dtype, static_cast<hipStream_t>(stream));
}

int wh_debug_modwt_plan_many(int inverse, int64_t n, int64_t count, int nlevels, int filter,
                             int dtype, char* buf, int cap) {
  if (!valid_filter(filter) || !valid_dtype(dtype) || !buf || cap <= 0)
    return set_error(WH_ERR_ARG, "modwt plan: bad filter, dtype or buffer");
  if (check_count(count)) return WH_ERR_ARG;
  const int J = resolve_levels(n, nlevels, filter_length(filter));
  if (J <= 0) return set_error(WH_ERR_ARG, "invalid length / levels");
  std::string s = modwt_plan_string(inverse != 0, n, J, filter, dtype, count);
  snprintf(buf, size_t(cap), "%s", s.c_str());
  return WH_OK;
}

int wh_debug_modwt_plan(int inverse, int64_t n, int nlevels, int filter, int dtype, char* buf,
                        int cap) {
  return wh_debug_modwt_plan_many(inverse, n, 1, nlevels, filter, dtype, buf, cap);
}

int wh_debug_modwt_set_plan(int inverse, int64_t n, int nlevels, int filter, int dtype,
                            int g1max, int first_per_cu, int deep_per_cu, int first_nt) {
  if (!valid_filter(filter)) return set_error(WH_ERR_ARG, "filter must be 1..4");
  if (!valid_dtype(dtype)) return set_error(WH_ERR_ARG, "dtype must be WH_F64 or WH_F32");
  const int J = resolve_levels(n, nlevels, filter_length(filter));
  if (J <= 0) return set_error(WH_ERR_ARG, "invalid length / levels");
  return modwt_force_plan(inverse != 0, n, J, filter, dtype, g1max, first_per_cu, deep_per_cu,
                          first_nt);
}

}  // extern "C"
