// noalloc.hip -- wh_modwt_inv_noalloc: the MODWT inverse (many signals,
// optional fused thresholding) without any device allocation, for callers
// that cannot allocate (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// The inverse's multi-group plans hand S_{j0} from one level group to the
// next through 2 n count elements of device memory.  wh_modwt_inv_many takes
// them from a library arena grown by hipMalloc; here they come from
// `scratch`, a fixed array in the code object's device memory (each GPU the
// library is loaded on holds its own copy until the process ends; the
// runtime places it when it loads this file's code object, which the
// address lookup below triggers).  Host code never uses the array's name as
// an address: its device address is looked up once per device with
// hipGetSymbolAddress -- by wh_init, or else by the first call on that
// device -- and kept.
#include <hip/hip_runtime.h>

#include <mutex>

#include "wh_common.h"
#include "wh_internal.h"

namespace wh {
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;  // 16 MiB per GPU
constexpr int kMaxDev = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDev] = {};

// the device address of `scratch` on the current device (looked up on the
// first call there)
int scratch_address(void** out) {
  int d = -1;
  WH_HIP_CHECK(hipGetDevice(&d));
  if (d < 0 || d >= kMaxDev) return set_error(WH_ERR_NODEVICE, "no current HIP device");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[d]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return set_hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(WH_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[d] = p;
  }
  *out = g_scratch_ptr[d];
  return WH_OK;
}

}  // namespace

// set-up (from wh_init): look the array up on the current device, which
// makes the runtime load this file's code object and place the array there
int noalloc_setup() {
  void* p = nullptr;
  return scratch_address(&p);
}

}  // namespace wh

using namespace wh;

extern "C" int wh_modwt_inv_noalloc(const void* xmod, void* x, int64_t n, int64_t count,
                                    int nlevels, int filter, int dtype, const wh_thresh* thr,
                                    void* stream) {
  int J, rc = check_common(xmod, x, n, nlevels, filter, dtype, &J);
  if (rc || (rc = check_count(count))) return rc;
  if ((rc = check_thresh(thr, J))) return rc;
  const size_t esz = dtype == WH_F32 ? 4 : 8;
  if (size_t(n) * size_t(count) > kNoAllocScratchBytes / (2 * esz))
    return set_error(WH_ERR_ARG,
                     "wh_modwt_inv_noalloc: n * count exceeds its fixed 16 MiB scratch "
                     "(at most 2^20 samples in FP64, 2^21 in FP32)");
  void* ws = nullptr;
  if ((rc = scratch_address(&ws))) return rc;
  return modwt_inverse_ws(xmod, x, n, count, J, filter, dtype, thr, ws, kNoAllocScratchBytes,
                          static_cast<hipStream_t>(stream));
}
