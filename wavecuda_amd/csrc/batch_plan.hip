// batch_plan.hip -- host side of the batched smoothing pipeline: the plan
// (item table, workspace, launch shape), the autotuner and the captured
// graph.  The kernel is in batch_kern.h, instantiated in batch_<type>.hip.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "batch_kern.h"
#include "wh_common.h"
#include "wh_internal.h"

struct wh_batch {
  int device = -1;
  int dtype = WH_F64;
  int hardness = 1;
  int count = 0;
  int64_t total = 0;
  std::string key;            // composition signature (tuning table key)
  wh::BatchItem* items = nullptr;
  int* counter = nullptr;     // queue head, finished workgroups
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  int64_t stride = 0;         // workspace elements per workgroup
  size_t lds[3] = {0, 0, 0};  // dynamic LDS for 128 / 256 / 512 threads
  int NT = 256;
  unsigned grid = 0;
  int tuned = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
};

namespace wh {
namespace {

constexpr int kNTs[3] = {128, 256, 512};

std::mutex g_tune_mu;
std::map<std::string, std::pair<int, unsigned>> g_tune;  // key -> (NT, grid)

int nt_index(int NT) { return NT == 512 ? 2 : (NT == 256 ? 1 : 0); }

// dynamic LDS of one item for NT threads
size_t item_lds(int ttype, int filt, int64_t n, int J, int NT, size_t esz) {
  const int L = filter_length(filt);
  size_t b = (NT == 512) ? block_sel_lds<512>()
             : (NT == 256 ? block_sel_lds<256>() : block_sel_lds<128>());
  b = std::max(b, batch_pre_lds(NT, esz));  // the forward's first-digit histogram
  if (ttype == kDWT) {
    const size_t tile = (esz == 8) ? size_t(batch_lds_elems<double>()) * 8
                                   : size_t(batch_lds_elems<float>()) * 4;
    return std::max(b, 2 * tile);
  }
  const int M = int(std::min<int64_t>(2 * NT, n));
  const size_t fwd = size_t(fwd_off(L, M, J)) * esz;
  const size_t inv = size_t(2) * lin_off(L, M, J + 1) * esz;
  return std::max(b, std::max(fwd, inv));
}

int occupancy(int dtype, int NT, size_t lds, int* blocks) {
  return dtype == WH_F32 ? batch_occupancy<float>(NT, lds, blocks)
                         : batch_occupancy<double>(NT, lds, blocks);
}

// largest resident grid for NT threads
int max_grid(const wh_batch* p, int NT, unsigned* grid) {
  int occ = 0;
  int rc = occupancy(p->dtype, NT, p->lds[nt_index(NT)], &occ);
  if (rc) return rc;
  if (occ <= 0) return set_error(WH_ERR_UNSUPPORTED, "batch: kernel does not fit a CU");
  const int64_t g = int64_t(occ) * std::max(1, device_info().num_cus);
  *grid = unsigned(std::min<int64_t>(g, p->count));
  return WH_OK;
}

int ensure_scratch(wh_batch* p, unsigned grid) {
  const size_t esz = p->dtype == WH_F32 ? 4 : 8;
  const size_t need = size_t(grid) * size_t(p->stride) * esz;
  if (need <= p->scratch_bytes) return WH_OK;
  if (p->scratch) {
    WH_HIP_CHECK(hipDeviceSynchronize());
    WH_HIP_CHECK(hipFree(p->scratch));
    p->scratch = nullptr;
    p->scratch_bytes = 0;
  }
  if (hipMalloc(&p->scratch, need) != hipSuccess) {
    p->scratch = nullptr;
    return set_error(WH_ERR_NOMEM, "batch: workspace allocation failed");
  }
  p->scratch_bytes = need;
  return WH_OK;
}

int run(wh_batch* p, int NT, unsigned grid, const void* x, void* y, double* thresh_out,
        hipStream_t st) {
  BatchArgs a{};
  a.items = p->items;
  a.count = p->count;
  a.counter = p->counter;
  a.x = x;
  a.y = y;
  a.thresh_out = thresh_out;
  a.scratch = p->scratch;
  a.scratch_stride = p->stride;
  a.hardness = p->hardness;
  BatchLaunch bl{NT, grid, p->lds[nt_index(NT)]};
  return p->dtype == WH_F32 ? batch_launch<float>(bl, a, st) : batch_launch<double>(bl, a, st);
}

void drop_graph(wh_batch* p) {
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->gexec = nullptr;
  p->graph = nullptr;
}

}  // namespace
}  // namespace wh

using namespace wh;

extern "C" {

int wh_batch_create(const int* lens, const int* filters, const int* ttypes, const int* levels,
                    int count, int dtype, int hardness, wh_batch** out) {
  if (!out) return set_error(WH_ERR_ARG, "null plan pointer");
  *out = nullptr;
  if (!lens || !filters || !ttypes || !levels || count <= 0)
    return set_error(WH_ERR_ARG, "batch: empty or null signal table");
  if (dtype != WH_F64 && dtype != WH_F32) return set_error(WH_ERR_ARG, "bad dtype");
  if (hardness != kHARD && hardness != kSOFT) return set_error(WH_ERR_ARG, "hardness must be 0 or 1");
  int dev = 0;
  WH_HIP_CHECK(hipGetDevice(&dev));
  const size_t esz = dtype == WH_F32 ? 4 : 8;

  std::vector<BatchItem> items(static_cast<size_t>(count));
  std::map<int64_t, int> comp;  // (kind, log2 n, J) -> count
  int64_t off = 0, stride = 0;
  size_t lds[3] = {0, 0, 0};
  for (int i = 0; i < count; ++i) {
    const int64_t n = lens[i];
    const int filt = filters[i], tt = ttypes[i];
    const int L = filter_length(filt);
    if (L == 0) return set_error(WH_ERR_ARG, "batch: filter must be 1..4");
    if (tt != kDWT && tt != kMODWT) return set_error(WH_ERR_ARG, "batch: ttype must be 0 or 1");
    if (n < 2 || (n & (n - 1)) || n > (int64_t(1) << 30))
      return set_error(WH_ERR_ARG, "batch: lengths must be powers of two");
    const int J = resolve_levels(n, levels[i], L);
    if (J <= 0) return set_error(WH_ERR_ARG, "batch: invalid length/levels for the filter");
    if (tt == kMODWT && J > kBatchMaxModwtJ)
      return set_error(WH_ERR_UNSUPPORTED, "batch: MODWT depth above 8 (use wh_smooth)");
    BatchItem& I = items[size_t(i)];
    I.off = off;
    I.n = int32_t(n);
    I.J = J;
    I.kind = batch_kind(tt, filt);
    I.sig = i;
    // factor of univ_thresh(w, 0, 0) (src/thresh.cpp:170-201): a one-level
    // DWT takes the MAD of x[1..n-1] scaled by sqrt(2 ln(n-1))
    const double nn = (tt == kDWT && J == 1) ? double(n - 1) : double(n);
    I.factor = std::sqrt(2.0 * std::log(nn));
    off += n;
    stride = std::max<int64_t>(stride, tt == kDWT ? 2 * n : 2 * n * J);
    for (int k = 0; k < 3; ++k)
      lds[k] = std::max(lds[k], item_lds(tt, filt, n, J, kNTs[k], esz));
    int lg = 0;
    while ((int64_t(1) << lg) < n) ++lg;
    comp[(int64_t(I.kind) << 16) | (lg << 8) | J] += 1;
  }
  // longest work first: MODWT moves about 2J + 4 words per sample, the DWT
  // about 6 plus its level-serial latency
  std::stable_sort(items.begin(), items.end(), [](const BatchItem& a, const BatchItem& b) {
    auto cost = [](const BatchItem& I) {
      return kind_ttype(I.kind) == kMODWT ? int64_t(I.n) * (2 * I.J + 4) : int64_t(I.n) * 6 + 256 * I.J;
    };
    return cost(a) > cost(b);
  });

  wh_batch* p = new wh_batch();
  p->device = dev;
  p->dtype = dtype;
  p->hardness = hardness;
  p->count = count;
  p->total = off;
  p->stride = (stride + 63) / 64 * 64;
  for (int k = 0; k < 3; ++k) p->lds[k] = lds[k];
  p->key = std::to_string(dtype) + "/" + std::to_string(hardness) + ":";
  for (auto& kv : comp) p->key += std::to_string(kv.first) + "x" + std::to_string(kv.second) + ",";

  auto fail = [&](int rc) {
    wh_batch_destroy(p);
    return rc;
  };
  if (hipMalloc(reinterpret_cast<void**>(&p->items), sizeof(BatchItem) * size_t(count)) !=
      hipSuccess) {
    p->items = nullptr;
    return fail(set_error(WH_ERR_NOMEM, "batch: item table allocation failed"));
  }
  if (hipMalloc(reinterpret_cast<void**>(&p->counter), 2 * sizeof(int)) != hipSuccess) {
    p->counter = nullptr;
    return fail(set_error(WH_ERR_NOMEM, "batch: queue allocation failed"));
  }
  if (hipMemcpy(p->items, items.data(), sizeof(BatchItem) * size_t(count),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(p->counter, 0, 2 * sizeof(int)) != hipSuccess)
    return fail(set_hip_error(hipGetLastError(), "batch: table upload"));

  // launch shape: a tuned one for this composition, else 256 threads and
  // every workgroup that fits
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(p->key);
    if (it != g_tune.end()) {
      p->NT = it->second.first;
      p->grid = it->second.second;
      p->tuned = 1;
    }
  }
  if (!p->tuned) {
    // untuned: the shape the tuner picks for mixed batches on MI355X (one
    // 512-thread workgroup per CU in FP64; in FP32 128 threads and every
    // workgroup that fits); wh_batch_tune times the others
    const int NT = dtype == WH_F32 ? 128 : 512;
    int rc = max_grid(p, NT, &p->grid);
    if (rc) return fail(rc);
    p->NT = NT;
  }
  if (int rc = ensure_scratch(p, p->grid)) return fail(rc);
  *out = p;
  return WH_OK;
}

void wh_batch_destroy(wh_batch* p) {
  if (!p) return;
  (void)hipDeviceSynchronize();
  drop_graph(p);
  if (p->items) (void)hipFree(p->items);
  if (p->counter) (void)hipFree(p->counter);
  if (p->scratch) (void)hipFree(p->scratch);
  delete p;
}

// the packed signals start at even offsets; the kernels move element pairs
// with single vector accesses, so the buffers must be aligned to two elements
static int check_io(const wh_batch* p, const void* x, const void* y) {
  if (!p || !x || !y) return set_error(WH_ERR_ARG, "null pointer");
  const uintptr_t pair = p->dtype == WH_F64 ? 16 : 8;
  if ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & (pair - 1))
    return set_error(WH_ERR_ARG, "device buffers must be aligned to two elements (16 bytes FP64, 8 bytes FP32)");
  return WH_OK;
}

int wh_batch_smooth(wh_batch* p, const void* x, void* y, double* thresh_out, void* stream) {
  if (int rc = check_io(p, x, y)) return rc;
  return run(p, p->NT, p->grid, x, y, thresh_out, static_cast<hipStream_t>(stream));
}

int wh_batch_tune(wh_batch* p, const void* x, void* y, double* thresh_out, void* stream) {
  if (int rc = check_io(p, x, y)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int cus = std::max(1, device_info().num_cus);
  struct Cand {
    int NT;
    unsigned grid;
  };
  std::vector<Cand> cands;
  unsigned gmax_all = 0;
  for (int NT : kNTs) {
    unsigned gmax = 0;
    if (int rc = max_grid(p, NT, &gmax)) return rc;
    gmax_all = std::max(gmax_all, gmax);
    for (unsigned g = unsigned(cus); g < gmax; g += unsigned(cus)) cands.push_back({NT, g});
    cands.push_back({NT, gmax});
  }
  if (int rc = ensure_scratch(p, gmax_all)) return rc;
  hipEvent_t e0, e1;
  WH_HIP_CHECK(hipEventCreate(&e0));
  WH_HIP_CHECK(hipEventCreate(&e1));
  float best = 1e30f;
  Cand bc = {p->NT, p->grid};
  int rc = WH_OK;
  const bool log = std::getenv("WH_TUNE_LOG") != nullptr;  // candidate timings to stderr
  for (const Cand& c : cands) {
    if ((rc = run(p, c.NT, c.grid, x, y, thresh_out, st))) break;  // warm-up
    float tmin = 1e30f;
    for (int r = 0; r < 3 && !rc; ++r) {
      if (hipEventRecord(e0, st) != hipSuccess) rc = set_hip_error(hipGetLastError(), "event");
      if (!rc) rc = run(p, c.NT, c.grid, x, y, thresh_out, st);
      if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = set_hip_error(hipGetLastError(), "event");
      if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = set_hip_error(hipGetLastError(), "sync");
      float ms = 0.f;
      if (!rc && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) tmin = std::min(tmin, ms);
    }
    if (rc) break;
    if (log) std::fprintf(stderr, "wh_batch_tune: %d threads x %u workgroups: %.3f ms\n", c.NT, c.grid, tmin);
    if (tmin < best) {
      best = tmin;
      bc = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  p->NT = bc.NT;
  p->grid = bc.grid;
  p->tuned = 1;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[p->key] = {p->NT, p->grid};
  }
  drop_graph(p);
  // leave y holding the result of the chosen shape
  return run(p, p->NT, p->grid, x, y, thresh_out, st);
}

int wh_batch_capture(wh_batch* p, const void* x, void* y, double* thresh_out) {
  if (int rc = check_io(p, x, y)) return rc;
  drop_graph(p);
  hipStream_t cs;
  WH_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  int rc = WH_OK;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = set_hip_error(hipGetLastError(), "hipStreamBeginCapture");
  } else {
    rc = run(p, p->NT, p->grid, x, y, thresh_out, cs);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cs, &g);
    if (!rc && e != hipSuccess) rc = set_hip_error(e, "hipStreamEndCapture");
    if (!rc) {
      hipGraphExec_t ge = nullptr;
      e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        rc = set_hip_error(e, "hipGraphInstantiate");
        (void)hipGraphDestroy(g);
      } else {
        p->graph = g;
        p->gexec = ge;
      }
    } else if (g) {
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipStreamDestroy(cs);
  return rc;
}

int wh_batch_replay(wh_batch* p, void* stream) {
  if (!p) return set_error(WH_ERR_ARG, "null pointer");
  if (!p->gexec) return set_error(WH_ERR_ARG, "batch: no captured graph");
  WH_HIP_CHECK(hipGraphLaunch(p->gexec, static_cast<hipStream_t>(stream)));
  return WH_OK;
}

int wh_batch_info(const wh_batch* p, int64_t* total, int* threads, int* workgroups,
                  int64_t* lds_bytes, int64_t* workspace_bytes, int* tuned) {
  if (!p) return set_error(WH_ERR_ARG, "null pointer");
  if (total) *total = p->total;
  if (threads) *threads = p->NT;
  if (workgroups) *workgroups = int(p->grid);
  if (lds_bytes) *lds_bytes = int64_t(p->lds[nt_index(p->NT)]);
  if (workspace_bytes) *workspace_bytes = int64_t(p->scratch_bytes);
  if (tuned) *tuned = p->tuned;
  return WH_OK;
}

int wh_debug_batch_set_shape(wh_batch* p, int threads, int workgroups) {
  if (!p) return set_error(WH_ERR_ARG, "null plan");
  if (threads != 128 && threads != 256 && threads != 512)
    return set_error(WH_ERR_ARG, "batch: threads must be 128, 256 or 512");
  unsigned gmax = 0;
  if (int rc = max_grid(p, threads, &gmax)) return rc;
  if (workgroups < 1 || unsigned(workgroups) > gmax)
    return set_error(WH_ERR_ARG, "batch: workgroups must be 1..the resident maximum");
  if (int rc = ensure_scratch(p, unsigned(workgroups))) return rc;
  p->NT = threads;
  p->grid = unsigned(workgroups);
  drop_graph(p);
  return WH_OK;
}

}  // extern "C"
