// wh_internal.h -- host-side internals shared by the wavehip source files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/wavehip.h"

namespace wh {

struct DeviceInfo {
  int device = -1;
  int num_cus = 0;
  size_t max_lds_per_block = 0;
  int warp_size = 64;
  std::string arch;
};

// Info of the current device (queried once per device).
const DeviceInfo& device_info();

// Library-owned device scratch of at least `bytes`, per device, grown on
// demand and kept until process exit.  Calls that overlap on one device
// through different streams must not share it; the library's own entry
// points order their use on the caller's stream.
int scratch_get(size_t bytes, void** out);
// Independent regions for buffers that are live at the same time:
// scratch2 (in-place staging of the DWT), ws (pipeline workspace of the
// smoothing / batched paths) and stats (small selection/reduction state).
int scratch2_get(size_t bytes, void** out);
int ws_get(size_t bytes, void** out);
int stats_get(size_t bytes, void** out);
// Two more regions (which = 0, 1) for the device copies of the host-pointer
// entry points' arguments, kept across calls like the others.
int host_buf_get(int which, size_t bytes, void** out);

// Argument checks shared by the entry points (status WH_ERR_ARG with a
// message on failure): codes, two-element alignment of a and b, the
// resolved level count *J; a threshold specification against J levels; a
// signal count of 1..65535.
bool valid_filter_code(int f);
bool valid_dtype_code(int t);
int check_common(const void* a, const void* b, int64_t n, int nlevels, int filter, int dtype,
                 int* J);
int check_thresh(const wh_thresh* thr, int J);
int check_count(int64_t count);

// The static scratch of wh_modwt_inv_noalloc placed on the current device
// and its address looked up (wh_init calls it).
int noalloc_setup();

// Let kernel `kern` take all of the current device's LDS beyond its static
// LDS as dynamic LDS; done once per (device, kernel), so it is cheap to call
// before every launch.  A refused request is cleared from HIP's last error
// so it cannot surface in a later, unrelated call.
void allow_dynamic_lds(const void* kern);

// transform back ends (device pointers)
// MODWT of `count` signals of n samples (x, out: n apart; xmod: 2nJ apart)
int modwt_forward(const void* x, void* xmod, int64_t n, int64_t count, int J, int filt,
                  int dtype, hipStream_t st);
int modwt_inverse(const void* xmod, void* x, int64_t n, int64_t count, int J, int filt, int dtype,
                  const wh_thresh* thr, hipStream_t st);
// the same with the level hand-off in the caller's device workspace ws of
// ws_bytes (2 n count elements needed when the plan has several groups)
int modwt_inverse_ws(const void* xmod, void* x, int64_t n, int64_t count, int J, int filt,
                     int dtype, const wh_thresh* thr, void* ws, size_t ws_bytes, hipStream_t st);
// adjoints: of the inverse (x-cotangent -> xmod-cotangent, the fused forward
// with exact 2^-j scalings) and of the forward (xmod-cotangent -> x-cotangent)
int modwt_inverse_adjoint(const void* gx, void* gxmod, int64_t n, int64_t count, int J, int filt,
                          int dtype, hipStream_t st);
int modwt_forward_adjoint(const void* gxmod, void* gx, int64_t n, int64_t count, int J, int filt,
                          int dtype, hipStream_t st);
int thresh_grad(const void* c, const void* gy, void* gc, double* glam, int64_t n, int64_t count,
                int J, int ttype, int dtype, const wh_thresh* thr, hipStream_t st);
// the level groups a launch of `count` signals runs, as text
std::string modwt_plan_string(bool inverse, int64_t n, int J, int filt, int dtype,
                              int64_t count = 1);
// time the planner's alternatives for (n, J, filter, element type) and keep
// the fastest forward and inverse plans for the rest of the process
int modwt_tune(int64_t n, int J, int filt, int dtype, hipStream_t st);
// set the planner's knobs for (n, J, filter, element type) as the tuner
// would (g1max <= 0: back to the default rules)
int modwt_force_plan(bool inverse, int64_t n, int J, int filt, int dtype, int g1max,
                     int first_per_cu, int deep_per_cu, int first_nt);

// levels.hip: one level's coefficients gathered into out (count written to
// *count_out), and the per-coefficient label columns
int level_extract(const void* c, void* out, int64_t n, int J, int ttype, int level, int coeff_s,
                  int dtype, hipStream_t st, int64_t* count_out);
int level_labels(int64_t n, int J, int ttype, int* level, double* translate, unsigned char* type,
                 hipStream_t st);
// WSTtoWavethresh's putD / putC vectors packed in its loop order
int wavethresh_pack(const void* c, void* out, int64_t n, int J, int ttype, int dtype,
                    hipStream_t st);

// modwt_order.hip: every level of a 2nJ MODWT vector between time order and
// packet order (to_po: time ordered in), out of place
int modwt_order(const void* in, void* out, int64_t n, int J, bool to_po, int dtype,
                hipStream_t st);

int dwt_forward(const void* x, void* y, int64_t n, int J, int filt, int dtype, hipStream_t st);
int dwt_inverse(const void* y, void* x, int64_t n, int J, int filt, int dtype,
                const wh_thresh* thr, hipStream_t st);

int univ_thresh_dev(const void* c, int64_t n, int J, int ttype, int dtype, int minlevel,
                    int maxlevel, double* out, hipStream_t st);
int smooth(void* x, int64_t n, int J, int ttype, int filt, int dtype, int threshtype,
           double* thresh, int hardness, int minlevel, int maxlevel, double tol,
           hipStream_t st);

}  // namespace wh
