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
  unsigned long long* partials;     // per-workgroup minima of reductions
  int64_t partials_cap;             // capacity of partials (keys)
  unsigned long long* newton_stats; // 5 device counters: iterations, max,
                                    // non-converged faces, faces, fallbacks
  struct fvs_mlp* mlp;              // learned initial guess (nullable)
};

namespace fvs {

void SetError(const char* fmt, ...);

// Single-patch level of the flux methods without a fused sweep, composed
// from the patch operators (fvs_level_composed.hip).
struct fvs_composed_level;

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
  int bound_check;        // the check decides pass / fail only (level driver):
                          // waves whose faces pass a cheap upper bound of
                          // their signal speeds skip the exact speeds
  int seg;                // pencil segment length (Y/Z sweeps)
  unsigned long long* partials;  // per-workgroup minimum of the check
  int* halt;              // nonzero: skip all work (nullable)
};

// RunOptions::smallest_time_step_size of a patch descriptor (1e-12, the
// reference's default, when the field is not positive).
inline double SmallestDt(const fvs_patch_desc& p) {
  return p.smallest_dt > 0.0 ? p.smallest_dt : 1e-12;
}

// Relative margin of the CFL pass's speed floor below the previous pass's
// largest speed: four times the largest relative drop of that speed over
// the last step, at least half the previous margin and kFloorMarginMin, at
// most kFloorMarginMax; a floor found unreached quadruples it.
constexpr double kFloorMarginMin = 0x1p-10;
constexpr double kFloorMarginFirst = 0x1p-6;
constexpr double kFloorMarginMax = 0.25;

// Device control block of the single-patch level driver.
struct Ctl {
  double t;
  double final_time;
  double eps;         // smallest time step size: stop once t + eps >= final_time
  double cfl;
  double dt;          // limited dt of the current step
  double split_dt;    // dt or dt / 2
  double failure_dt;  // estimate after a TimeStepTooLarge
  int have_failure;   // use failure_dt for the next step
  int halt;           // 0 running, 1 TimeStepTooLarge, 2 final time reached,
                      // 3 the speed floor of the step-start CFL pass was wrong
                      // (the host repeats the step without it)
  int64_t cycle;
  double dt_dir[3];   // step-start stable dt per direction (decoded)
  double last_check;
  // 3D step-start CFL pass: a speed floor per direction (0: none), the
  // previous pass's largest signal speed less floor_margin.  The pass
  // evaluates the exact speed only of faces whose cheap upper bound reaches
  // it; floor_hit[d] records that some evaluated face did reach it, which
  // proves that the skipped faces are below the maximum.
  double cfl_floor[3];
  int floor_hit[3];
  double floor_margin;    // relative margin of the last floor (0: none yet)
  double floor_speed[3];  // largest speeds of the last pass (0: none yet)
  double* dt_log;       // optional log of the steps' dt (set per advance call)
  int64_t dt_log_base;  // cycle of dt_log[0]
  int64_t dt_log_len;
};

// Number of workgroups (= partial minima) of a sweep / of the CFL pass.
int64_t SweepBlocks(int dim, int dir, const int64_t n[3], int seg);
int64_t CflBlocks(int dim, const int64_t n[3]);
// Cells per thread of a Y/Z pencil sweep: 32, fewer when the patch is too
// small to give the device about 1,024 workgroups (down to 1).
int PencilSeg(int dim, int dir, const int64_t n[3]);
int LaunchSweep(const fvs_method_desc& desc, hipStream_t stream, int dir,
                const SweepArgs& a);
// Fused sweeps and CFL pass of the tolerance-checked arithmetic modes
// (FVS_FP64_RCP, FVS_FP32_FLUX), compiled with multiply-add contraction.
int LaunchSweepFast(int dim, int precision, bool larr, hipStream_t stream, int dir,
                    const SweepArgs& a);
int LaunchPatchCflFast(int dim, unsigned gx, unsigned gy, hipStream_t stream,
                       const double* const* cons, const int64_t n[3], const double dx[3],
                       double gamma, double gm1inv, unsigned long long* partials, int* halt,
                       Ctl* floor_ctl);
// Reduces a sweep's partial minima: writes the value to out (nullable) and,
// with a control block, applies the split step's CFL check.
int LaunchFinalizeCheck(hipStream_t s, const unsigned long long* partials,
                        int64_t nblocks, double* out, Ctl* ctl);
// The CFL checks of a level step's sweeps (partial minima of check r at
// partials + r * stride, n[r] of them), reduced in order after the last
// sweep, then the step end.
constexpr int kMaxChecks = 5;
struct CheckList {
  int64_t n[kMaxChecks];
  int64_t stride;
  int count;
};
int LaunchFinalizeStep(hipStream_t s, const unsigned long long* partials, const CheckList& cl,
                       Ctl* ctl);
// Step-start CFL pass: per-direction partial minima, then a finalize kernel
// writing dt_dir (out, nullable) or, with a control block, the step's dt.
// In 3D with a control block the pass skips the faces below ctl->cfl_floor.
int LaunchPatchCfl(const fvs_method_desc& desc, hipStream_t stream,
                   const double* const* cons, const int64_t n[3],
                   const double dx[3], unsigned long long* partials,
                   double* out, Ctl* ctl, int* halt, int first_dir, int strang);
// Exact-Godunov flux method (fvs_godunov.hip).  Partials hold 5 keys per
// workgroup: min dt, iteration sum, iteration max, non-converged, faces.
int64_t GodunovSweepBlocks(int dim, int dir, const int64_t n[3], int seg);
int64_t GodunovCflBlocks(int dim, const int64_t n[3]);
int LaunchGodunovSweep(const fvs_method_desc& desc, hipStream_t s, int dir,
                       const double* const* in, double* const* out, const int64_t n[3],
                       double dx, const double* dt_ptr, double dt_val, const double* guess,
                       unsigned long long* partials, int* halt, int update,
                       const fvs_mlp* net = nullptr, int reuse = 0);
int LaunchGodunovFinalize(hipStream_t s, const unsigned long long* partials, int64_t nblocks,
                          double* out, unsigned long long* stats, Ctl* ctl = nullptr);
// Newton iteration cap from which the exact-Godunov CFL pass uses the speed
// floor (below it every face is solved, so a small cap is reported as the
// reference's would be).
constexpr int kGodunovFloorCap = 16;
// With floor_ctl (the level driver) and an iteration cap of at least 16, the
// faces of directions other than pm_dir that a certified speed bound keeps
// below ctl->cfl_floor skip their Newton solve.
int LaunchGodunovCfl(const fvs_method_desc& desc, hipStream_t s, const double* const* cons,
                     const int64_t n[3], const double dx[3], const double* const* guess,
                     unsigned long long* partials, int* halt, const fvs_mlp* net = nullptr,
                     double* pm_out = nullptr, int pm_dir = -1, Ctl* floor_ctl = nullptr);
int LaunchGodunovViewFlux(const fvs_method_desc& desc, hipStream_t s, const fvs_view* fluxes,
                          const fvs_view* states, double dt, double dx, int dir,
                          unsigned long long* nfail);
int64_t GodunovViewStableDtBlocks(const fvs_method_desc& desc, const fvs_view* states, int dir);
int LaunchGodunovViewStableDt(const fvs_method_desc& desc, hipStream_t s, const fvs_view* states,
                              double dx, int dir, unsigned long long* partials);
// Learned Newton guesses for the faces of direction dir of a ghost-free
// patch, launched on stream s (fvs_mlp.hip).
int MlpFaceGuess(fvs_mlp* m, hipStream_t s, int dim, double gamma, const double* const* cons,
                 const int64_t n[3], int dir, double* guess);
// The network's packed matrix-core operands, for kernels that infer inline.
void MlpOperands(const fvs_mlp* m, const short** tiles, const float** bias);
int LaunchMiddleState(int dim, hipStream_t s, const double* L, const double* R, int64_t n,
                      int dir, double gamma, double tol, int max_it, const double* guess,
                      double* pM, double* uM, int* iters, int* conv);
// Reduces `dim` consecutive groups of nblocks partial minima into the
// per-direction stable dts (out, nullable; ctl->dt_dir) and the step dt
// with the first split step's CFL check.
int LaunchFinalizeCfl(hipStream_t s, const unsigned long long* partials, int64_t nblocks,
                      int dim, double* out, Ctl* ctl, int first_dir, int strang);
int LaunchStepEnd(hipStream_t s, Ctl* ctl);
// Step dt from ctl->dt_dir; with use_floor (the exact-Godunov CFL pass ran
// with a speed floor) the floor is validated and renewed (dx: cell widths).
int LaunchStepDt(hipStream_t s, Ctl* ctl, int dim, int first_dir, int strang,
                 const double* dx = nullptr, int use_floor = 0);
int LaunchCheckValue(hipStream_t s, Ctl* ctl, const double* value);
int LaunchKeysToDt(hipStream_t s, unsigned long long* keys, int n, double* out);
int LaunchResetKeys(hipStream_t s, unsigned long long* keys, int n,
                    unsigned int* counters, int nc);

// Composed level (fvs_level_composed.hip): the methods it serves, and the
// level operations the fvs_level_* entry points forward to it.
bool ComposedLevelNeeded(const fvs_method_desc& d);
int ComposedCreate(fvs_handle* h, const fvs_patch_desc* patch, fvs_composed_level** out);
void ComposedDestroy(fvs_composed_level* C);
int ComposedSetState(fvs_composed_level* C, const double* const* comps, double t0);
int ComposedGetState(fvs_composed_level* C, double* const* comps);
void ComposedTime(const fvs_composed_level* C, double* t, int64_t* cycles);
int ComposedAdvance(fvs_composed_level* C, int max_steps, double final_time, double* dt_log,
                    int* steps_done, int* failures);

} // namespace fvs
