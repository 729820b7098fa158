// Host-side internals shared by the translation units of libfvs_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/fvs.h"

struct fvs_handle {
  fvs_method_desc desc;
  hipStream_t stream;
  int device;
  double* scratch_dt; // 4 device doubles for reductions of entry points
  unsigned long long* scratch_keys; // 8 device keys
  unsigned int* scratch_counter;    // 4 device counters
};

namespace fvs {

void SetError(const char* fmt, ...);

#define FVS_HIP_CHECK(expr)                                                       \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::fvs::SetError("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                      __FILE__, __LINE__);                                        \
      return FVS_HIP_ERROR;                                                       \
    }                                                                             \
  } while (0)

#define FVS_LAUNCH_CHECK()                                                        \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess) {                                                       \
      ::fvs::SetError("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),  \
                      __FILE__, __LINE__);                                        \
      return FVS_HIP_ERROR;                                                       \
    }                                                                             \
  } while (0)

// Parameters shared by the fused sweep kernels.
struct SweepArgs {
  const double* in[5];
  double* out[5];
  int64_t n[3];
  double dx;
  const double* dt_ptr;   // split dt on the device (nullptr: use dt_val)
  double dt_val;
  double gamma;
  double gm1inv;
  int limiter;
  int muscl;              // 1: MUSCL-Hancock, 0: first order
  int quirk3d;
  int check;              // compute the CFL check of the input in this sweep
  int seg;                // pencil segment length (Y/Z sweeps)
  unsigned long long* check_key; // atomic-min accumulator of the check
  unsigned int* done_counter;    // last-block detection (nullable)
  int* halt;              // nonzero: skip all work (nullable)
  // Finalisation of the check by the last block (level driver only).
  struct Ctl* ctl;
  int sweep_slot;
};

// Device control block of the single-patch level driver.
struct Ctl {
  double t;
  double final_time;
  double cfl;
  double dt;          // limited dt of the current step
  double split_dt;    // dt or dt / 2
  double failure_dt;  // estimate after a TimeStepTooLarge
  int have_failure;   // use failure_dt for the next step
  int halt;           // 0 running, 1 TimeStepTooLarge, 2 final time reached
  int64_t cycle;
  unsigned long long cfl_key[3];
  unsigned long long check_key;
  unsigned int cfl_counter;
  unsigned int check_counter;
  double dt_dir[3];   // step-start stable dt per direction (decoded)
  double last_check;
};

int LaunchSweep(const fvs_method_desc& desc, hipStream_t stream, int dir,
                const SweepArgs& a);
int LaunchPatchCfl(const fvs_method_desc& desc, hipStream_t stream,
                   const double* const* cons, const int64_t n[3],
                   const double dx[3], unsigned long long* keys,
                   unsigned int* counter, Ctl* ctl, int* halt, int first_dir,
                   int strang);
int LaunchStepEnd(hipStream_t s, Ctl* ctl, double* dt_log, int64_t dt_log_base,
                  int dt_log_len);
int LaunchKeysToDt(hipStream_t s, unsigned long long* keys, int n, double* out);
int LaunchResetKeys(hipStream_t s, unsigned long long* keys, int n,
                    unsigned int* counters, int nc);

} // namespace fvs
