// rglue.cpp -- the reference package's `.C` entry points (host arrays in and
// out, synchronous) over the HIP path, and a `.C` entry for the batched
// smoother.
//
// The reference binds five `.C` symbols (NAMESPACE:26-30, declared at
// src/rfunctions.cuh:17-35): RcpuTransform, RgpuTransform, RcpuThreshold,
// RcpuSmooth and RgpuSmooth.  Here all five run on the GPU, so an unmodified
// R package finds every symbol it binds in this library.  Their arguments,
// codes and in/out meaning are the reference's (src/rfunctions.cu:3-136);
// errors print a message and leave the outputs untouched instead of ending
// the process.
//
// Host <-> device traffic goes through a per-device ring of pinned staging
// slots: a chunk is copied between the caller's (pageable) array and a slot
// by a small pool of host threads while the DMA engine moves the previous
// chunk, and the device copies of the arguments live in two library
// regions kept across calls (no device allocation per call once they have
// grown).  One call at a time per device: the calls hold a per-device lock.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "wh_common.h"
#include "wh_internal.h"

namespace wh {
namespace {

// ------------------------------------------------------------- copy pool
// memcpy split over the calling thread and a few workers, each taking one
// fixed part of the range (no shared work counter, so a worker can never
// run a part of the next copy with the fields of the previous one).
class CopyPool {
 public:
  static CopyPool& get() {
    static CopyPool* p = new CopyPool();  // never destroyed: workers may be blocked at exit
    return *p;
  }
  void copy(void* dst, const void* src, size_t bytes) {
    const int nw = int(workers_.size());
    if (nw == 0 || bytes < (size_t(4) << 20)) {
      std::memcpy(dst, src, bytes);
      return;
    }
    const int parts = nw + 1;
    size_t part = (bytes + parts - 1) / parts;
    part = (part + 63) & ~size_t(63);
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = static_cast<char*>(dst);
      src_ = static_cast<const char*>(src);
      bytes_ = bytes;
      part_ = part;
      remaining_.store(nw);
      ++gen_;
    }
    cv_.notify_all();
    run_part(0, static_cast<char*>(dst), static_cast<const char*>(src), bytes, part);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return remaining_.load() == 0; });
  }

 private:
  CopyPool() {
    int nt = int(std::thread::hardware_concurrency());
    if (const char* e = std::getenv("WH_COPY_THREADS")) nt = std::atoi(e);
    nt = std::max(1, std::min(nt, 16));
    for (int i = 1; i < nt; ++i) workers_.emplace_back([this, i] { loop(i); });
    for (auto& t : workers_) t.detach();
  }
  static void run_part(int i, char* d, const char* s, size_t bytes, size_t part) {
    const size_t lo = size_t(i) * part;
    if (lo >= bytes) return;
    std::memcpy(d + lo, s + lo, std::min(part, bytes - lo));
  }
  void loop(int idx) {
    uint64_t seen = 0;
    for (;;) {
      char* d;
      const char* s;
      size_t bytes, part;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        d = dst_;
        s = src_;
        bytes = bytes_;
        part = part_;
      }
      run_part(idx, d, s, bytes, part);
      if (remaining_.fetch_sub(1) == 1) {
        std::lock_guard<std::mutex> lk(mu_);
        done_.notify_all();
      }
    }
  }
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t bytes_ = 0, part_ = 0;
  uint64_t gen_ = 0;
  std::atomic<int> remaining_{0};
};

// ------------------------------------------------------------ staging ring
constexpr size_t kSlotBytes = size_t(32) << 20;
constexpr int kSlots = 4;
constexpr int kMaxDev = 64;

struct Ring {
  bool ready = false;
  void* slot[kSlots] = {};
  hipEvent_t ev[kSlots] = {};
  bool inflight[kSlots] = {};  // a copy out of / into the slot was recorded
  std::mutex call_mu;  // one host-pointer call at a time on this device
};
Ring g_ring[kMaxDev];
std::mutex g_ring_mu;

int ring_of(Ring** out) {
  int d = -1;
  WH_HIP_CHECK(hipGetDevice(&d));
  if (d < 0 || d >= kMaxDev) return set_error(WH_ERR_NODEVICE, "no current HIP device");
  Ring& r = g_ring[d];
  std::lock_guard<std::mutex> lk(g_ring_mu);
  if (!r.ready) {
    for (int s = 0; s < kSlots; ++s) {
      if (hipHostMalloc(&r.slot[s], kSlotBytes, hipHostMallocDefault) != hipSuccess) {
        r.slot[s] = nullptr;
        return set_error(WH_ERR_NOMEM, "pinned staging allocation failed");
      }
      WH_HIP_CHECK(hipEventCreateWithFlags(&r.ev[s], hipEventDisableTiming));
    }
    r.ready = true;
  }
  *out = &r;
  return WH_OK;
}

// host -> device through the ring: chunk i is staged by the host threads
// while chunk i - 1 is in flight
int h2d(Ring& r, void* dev, const void* host, size_t bytes, hipStream_t st) {
  size_t off = 0;
  for (int i = 0; off < bytes; ++i, off += kSlotBytes) {
    const int s = i % kSlots;
    const size_t len = std::min(kSlotBytes, bytes - off);
    // the slot's previous copy (of this call or an earlier one) must be done
    if (r.inflight[s]) WH_HIP_CHECK(hipEventSynchronize(r.ev[s]));
    r.inflight[s] = false;
    CopyPool::get().copy(r.slot[s], static_cast<const char*>(host) + off, len);
    WH_HIP_CHECK(hipMemcpyAsync(static_cast<char*>(dev) + off, r.slot[s], len,
                                hipMemcpyHostToDevice, st));
    WH_HIP_CHECK(hipEventRecord(r.ev[s], st));
    r.inflight[s] = true;
  }
  return WH_OK;
}

// device -> host: up to kSlots chunks in flight, each unstaged by the host
// threads as it lands; returns when the data is in `host`
int d2h(Ring& r, void* host, const void* dev, size_t bytes, hipStream_t st) {
  const size_t nch = (bytes + kSlotBytes - 1) / kSlotBytes;
  auto issue = [&](size_t i) -> int {
    const int s = int(i % kSlots);
    const size_t off = i * kSlotBytes, len = std::min(kSlotBytes, bytes - off);
    WH_HIP_CHECK(hipMemcpyAsync(r.slot[s], static_cast<const char*>(dev) + off, len,
                                hipMemcpyDeviceToHost, st));
    WH_HIP_CHECK(hipEventRecord(r.ev[s], st));
    r.inflight[s] = true;
    return WH_OK;
  };
  int rc;
  for (size_t i = 0; i < std::min<size_t>(nch, kSlots); ++i)
    if ((rc = issue(i))) return rc;
  for (size_t i = 0; i < nch; ++i) {
    const int s = int(i % kSlots);
    const size_t off = i * kSlotBytes, len = std::min(kSlotBytes, bytes - off);
    WH_HIP_CHECK(hipEventSynchronize(r.ev[s]));
    r.inflight[s] = false;
    CopyPool::get().copy(static_cast<char*>(host) + off, r.slot[s], len);
    if (i + kSlots < nch && (rc = issue(i + kSlots))) return rc;
  }
  return WH_OK;
}

bool valid_filter(int f) { return f >= kHAAR && f <= kLA8; }

void report(const char* where, int rc) {
  std::fprintf(stderr, "\n Error in GPU %s: %s (code %d)\n", where, wh_last_error(), rc);
}

// device copies of a call's arguments: region 0 (x-sized, or 2nJ for the
// packet-order conversions) and region 1 (xmod-sized)
int bufs(size_t b0, size_t b1, void** p0, void** p1) {
  int rc;
  if (b0 && (rc = host_buf_get(0, b0, p0))) return rc;
  if (b1 && (rc = host_buf_get(1, b1, p1))) return rc;
  return WH_OK;
}

// RcpuTransform / RgpuTransform (src/rfunctions.cu:3-57): DWT in place on x;
// MODWT forward x -> xmod (2 n J values), backward xmod -> x; ttype 2 is the
// packet-ordered MODWT (src/transform.cpp:39-55), here the time-ordered
// transform plus the exact conversion of src/wavutils.cpp:532-600.
int transform_host(double* x, double* xmod, int n32, int sense, int nlevels, int ttype,
                   int filter) {
  if (!x || n32 <= 0) return set_error(WH_ERR_ARG, "null argument or empty input");
  const int64_t n = n32;
  const int J = valid_filter(filter) ? resolve_levels(n, nlevels, filter_length(filter)) : 0;
  if (J <= 0) return set_error(WH_ERR_ARG, "invalid length, levels or filter");
  if (sense != kFWD && sense != kBWD) return set_error(WH_ERR_ARG, "sense must be 1 (FWD) or 0 (BWD)");
  if (ttype != kDWT && ttype != kMODWT && ttype != kMODWT_PO)
    return set_error(WH_ERR_ARG, "ttype must be 0 (DWT), 1 (MODWT) or 2 (packet-ordered MODWT)");
  if (ttype != kDWT && !xmod) return set_error(WH_ERR_ARG, "null xmod");
  Ring* r;
  int rc = ring_of(&r);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(r->call_mu);
  hipStream_t st = nullptr;
  const size_t nb = size_t(n) * 8, mb = size_t(2) * size_t(n) * size_t(J) * 8;
  void *a = nullptr, *b = nullptr;
  if (ttype == kDWT) {
    if ((rc = bufs(nb, 0, &a, &b))) return rc;
    if ((rc = h2d(*r, a, x, nb, st))) return rc;
    rc = (sense == kFWD) ? wh_dwt_fwd(a, a, n, J, filter, WH_F64, st)
                         : wh_dwt_inv(a, a, n, J, filter, WH_F64, nullptr, st);
    if (rc) return rc;
    return d2h(*r, x, a, nb, st);
  }
  const bool po = ttype == kMODWT_PO;
  if ((rc = bufs(po ? mb : nb, mb, &a, &b))) return rc;
  if (sense == kFWD) {
    if ((rc = h2d(*r, a, x, nb, st))) return rc;
    if ((rc = wh_modwt_fwd(a, b, n, J, filter, WH_F64, st))) return rc;
    // packet order: the coefficients go from b to a (x is no longer needed)
    if (po && (rc = wh_modwt_convert(b, a, n, J, filter, 1, WH_F64, st))) return rc;
    return d2h(*r, xmod, po ? a : b, mb, st);
  }
  if ((rc = h2d(*r, b, xmod, mb, st))) return rc;
  const void* coef = b;
  if (po) {
    if ((rc = wh_modwt_convert(b, a, n, J, filter, 0, WH_F64, st))) return rc;
    coef = a;
  }
  void* out = po ? b : a;  // the packet-ordered input in b is consumed by then
  if ((rc = wh_modwt_inv(coef, out, n, J, filter, WH_F64, nullptr, st))) return rc;
  return d2h(*r, x, out, nb, st);
}

// RcpuThreshold (src/rfunctions.cu:59-69 -> src/thresh.cpp:113-140): the
// details of levels minlevel..maxlevel of x (DWT) or xmod (MODWT) in place
int threshold_host(double* x, double* xmod, int n32, int nlevels, int ttype, int filter,
                   double thresh, int hardness, int minlevel, int maxlevel) {
  if (n32 <= 0) return set_error(WH_ERR_ARG, "empty input");
  const int64_t n = n32;
  const int J = valid_filter(filter) ? resolve_levels(n, nlevels, filter_length(filter)) : 0;
  if (J <= 0) return set_error(WH_ERR_ARG, "invalid length, levels or filter");
  if (minlevel < 0 || minlevel > maxlevel || maxlevel >= J)
    return set_error(WH_ERR_ARG, "levels must satisfy 0 <= minlevel <= maxlevel < nlevels");
  if (hardness != kHARD && hardness != kSOFT) return set_error(WH_ERR_ARG, "hardness must be 0 (hard) or 1 (soft)");
  if (ttype != kDWT && ttype != kMODWT && ttype != kMODWT_PO)
    return set_error(WH_ERR_ARG, "ttype must be 0 (DWT), 1 (MODWT) or 2 (packet-ordered MODWT)");
  double* c = (ttype == kDWT) ? x : xmod;
  if (!c) return set_error(WH_ERR_ARG, "null coefficient vector");
  Ring* r;
  int rc = ring_of(&r);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(r->call_mu);
  hipStream_t st = nullptr;
  const size_t bytes = size_t(ttype == kDWT ? n : 2 * n * J) * 8;
  void *a = nullptr, *b = nullptr;
  if ((rc = bufs(bytes, ttype == kMODWT_PO ? bytes : 0, &a, &b))) return rc;
  wh_thresh t{};
  t.enabled = 1;
  t.hardness = hardness;
  t.minlevel = minlevel;
  t.maxlevel = maxlevel;
  t.value = thresh;
  if (ttype == kMODWT_PO) {
    // the details of a packet-ordered level sit at its odd positions
    // (src/thresh.cpp:56-63): threshold them in time order, then convert back
    if ((rc = h2d(*r, b, c, bytes, st))) return rc;
    if ((rc = wh_modwt_convert(b, a, n, J, filter, 0, WH_F64, st))) return rc;
    if ((rc = wh_threshold(a, n, J, kMODWT, filter, WH_F64, &t, st))) return rc;
    if ((rc = wh_modwt_convert(a, b, n, J, filter, 1, WH_F64, st))) return rc;
    return d2h(*r, c, b, bytes, st);
  }
  if ((rc = h2d(*r, a, c, bytes, st))) return rc;
  if ((rc = wh_threshold(a, n, J, ttype, filter, WH_F64, &t, st))) return rc;
  return d2h(*r, c, a, bytes, st);
}

// RcpuSmooth / RgpuSmooth (src/rfunctions.cu:71-136): x smoothed in place,
// *thresh receives the universal / cross-validated threshold
int smooth_host(double* x, int n32, int nlevels, int ttype, int filter, int threshtype,
                double* thresh, int hardness, int minlevel, int maxlevel, double tol) {
  if (!x || n32 <= 0) return set_error(WH_ERR_ARG, "null argument or empty input");
  const int64_t n = n32;
  const int J = valid_filter(filter) ? resolve_levels(n, nlevels, filter_length(filter)) : 0;
  if (J <= 0) return set_error(WH_ERR_ARG, "invalid length, levels or filter");
  if (threshtype == 0 && !thresh) return set_error(WH_ERR_ARG, "manual threshold needs a value");
  Ring* r;
  int rc = ring_of(&r);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(r->call_mu);
  hipStream_t st = nullptr;
  void *a = nullptr, *b = nullptr;
  if ((rc = bufs(size_t(n) * 8, 0, &a, &b))) return rc;
  if ((rc = h2d(*r, a, x, size_t(n) * 8, st))) return rc;
  double t = thresh ? *thresh : 0.0;
  if ((rc = wh_smooth(a, n, J, ttype, filter, WH_F64, threshtype, &t, hardness, minlevel,
                      maxlevel, tol, st)))
    return rc;
  if ((rc = d2h(*r, x, a, size_t(n) * 8, st))) return rc;
  if (thresh && threshtype != 0) *thresh = t;
  return WH_OK;
}

// the batch plan of the last RgpuSmoothList call per device, reused while
// the signal table stays the same
struct ListPlan {
  std::vector<int> key;
  wh_batch* plan = nullptr;
};
ListPlan g_list[kMaxDev];

int smooth_list_host(double* x, int count, const int* lens, const int* filters,
                     const int* ttypes, const int* nlevels, int hardness, double* thresh) {
  if (!x || !lens || !filters || !ttypes || !nlevels || count <= 0)
    return set_error(WH_ERR_ARG, "null argument or empty signal table");
  int64_t total = 0;
  for (int i = 0; i < count; ++i) {
    if (lens[i] < 2) return set_error(WH_ERR_ARG, "batch: lengths must be powers of two >= 2");
    total += lens[i];
  }
  Ring* r;
  int rc = ring_of(&r);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(r->call_mu);
  int dev = 0;
  WH_HIP_CHECK(hipGetDevice(&dev));
  ListPlan& lp = g_list[dev];
  std::vector<int> key;
  key.reserve(size_t(count) * 4 + 2);
  key.push_back(count);
  key.push_back(hardness);
  for (int i = 0; i < count; ++i) {
    key.push_back(lens[i]);
    key.push_back(filters[i]);
    key.push_back(ttypes[i]);
    key.push_back(nlevels[i]);
  }
  if (!lp.plan || lp.key != key) {
    if (lp.plan) wh_batch_destroy(lp.plan);
    lp.plan = nullptr;
    lp.key.clear();
    if ((rc = wh_batch_create(lens, filters, ttypes, nlevels, count, WH_F64, hardness,
                              &lp.plan)))
      return rc;
    lp.key = std::move(key);
  }
  hipStream_t st = nullptr;
  void *a = nullptr, *b = nullptr;
  if ((rc = bufs(size_t(total) * 8, size_t(count) * 8, &a, &b))) return rc;
  if ((rc = h2d(*r, a, x, size_t(total) * 8, st))) return rc;
  if ((rc = wh_batch_smooth(lp.plan, a, a, static_cast<double*>(b), st))) return rc;
  if ((rc = d2h(*r, x, a, size_t(total) * 8, st))) return rc;
  if (thresh) WH_HIP_CHECK(hipMemcpy(thresh, b, size_t(count) * 8, hipMemcpyDeviceToHost));
  return WH_OK;
}

}  // namespace
}  // namespace wh

using namespace wh;

extern "C" {

void RcpuTransform(double* x, double* xmod, int* len, int* sense, int* nlevels, int* ttype,
                   int* filter, int* filterlen) {
  (void)filterlen;
  if (!len || !sense || !nlevels || !ttype || !filter)
    return report("transform", set_error(WH_ERR_ARG, "null argument"));
  if (int rc = transform_host(x, xmod, *len, *sense, *nlevels, *ttype, *filter))
    report("transform", rc);
}

void RgpuTransform(double* x, double* xmod, int* len, int* sense, int* nlevels, int* ttype,
                   int* filter, int* filterlen) {
  RcpuTransform(x, xmod, len, sense, nlevels, ttype, filter, filterlen);
}

void RcpuThreshold(double* x, double* xmod, int* len, int* nlevels, int* ttype, int* filter,
                   int* filterlen, double* thresh, int* hardness, int* minlevel, int* maxlevel) {
  (void)filterlen;
  if (!len || !nlevels || !ttype || !filter || !thresh || !hardness || !minlevel || !maxlevel)
    return report("threshold", set_error(WH_ERR_ARG, "null argument"));
  if (int rc = threshold_host(x, xmod, *len, *nlevels, *ttype, *filter, *thresh, *hardness,
                              *minlevel, *maxlevel))
    report("threshold", rc);
}

void RcpuSmooth(double* x, int* len, int* nlevels, int* ttype, int* filter, int* filterlen,
                int* threshtype, double* thresh, int* hardness, int* minlevel, int* maxlevel,
                double* tol) {
  (void)filterlen;
  if (!len || !nlevels || !ttype || !filter || !threshtype || !hardness || !minlevel ||
      !maxlevel)
    return report("smooth", set_error(WH_ERR_ARG, "null argument"));
  if (int rc = smooth_host(x, *len, *nlevels, *ttype, *filter, *threshtype, thresh, *hardness,
                           *minlevel, *maxlevel, tol ? *tol : 0.01))
    report("smooth", rc);
}

void RgpuSmooth(double* x, int* len, int* nlevels, int* ttype, int* filter, int* filterlen,
                int* threshtype, double* thresh, int* hardness, int* minlevel, int* maxlevel,
                double* tol) {
  RcpuSmooth(x, len, nlevels, ttype, filter, filterlen, threshtype, thresh, hardness, minlevel,
             maxlevel, tol);
}

void RgpuSmoothList(double* x, int* count, int* lens, int* filters, int* ttypes, int* nlevels,
                    int* hardness, double* thresh) {
  if (!count || !hardness) return report("smooth list", set_error(WH_ERR_ARG, "null argument"));
  if (int rc = smooth_list_host(x, *count, lens, filters, ttypes, nlevels, *hardness, thresh))
    report("smooth list", rc);
}

}  // extern "C"
