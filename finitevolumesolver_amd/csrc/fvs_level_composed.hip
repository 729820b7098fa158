// Single-patch level integrator for the flux methods that have no fused
// sweep: MUSCL-Hancock over the exact-Godunov base (the reference's default
// MusclHancockMethod, include/fub/flux_method/MusclHancockMethod.hpp:277-285),
// the HLL base, MusclHancockPrim and MusclHancockCharacteristic.
//
// The step is the reference's DimensionalSplitLevelIntegrator sequence
// (include/fub/solver/DimensionalSplitLevelIntegrator.hpp:221-276) composed
// from the library's patch operators on one device scratch of complete
// states with 2 x stencil-width ghost layers:
//   stable dt of every direction (transmissive fill of that direction's
//     ghosts, then FluxMethod::ComputeStableDt over the interior grown by
//     the ghosts in that direction), cfl times the minimum, capped at the
//     final time (RunSimulation's limited_dt, include/fub/RunSimulation.hpp:125-131);
//   per split direction (MakeSplitDirections order, Strang half steps): the
//     direction's CFL check (the same stable dt against the split dt; the
//     first split step's is the step-start value of its direction, computed
//     on the same state; TimeStepTooLarge repeats the step from a device
//     copy of its start state with the failing dt), ComputeNumericFluxes
//     over faces 0 .. n, UpdateConservatively of the interior and
//     CompleteFromCons.
// Each check reads one double back to the host, so this path synchronises
// a few times per step; the fused level kernels serve the HLLEM and
// first-order exact-Godunov methods.
#include "fvs_internal.hpp"

#include <algorithm>
#include <new>
#include <vector>

namespace fvs {

struct fvs_composed_level {
  fvs_handle* h = nullptr;
  fvs_patch_desc patch{};
  int D = 2;
  int gw = 2;               // ghost layers: twice the stencil width
  int64_t n[3] = {1, 1, 1};
  int64_t ng[3] = {1, 1, 1};
  int64_t ncell_g = 0;
  double* scratch = nullptr;  // (D+4) complete components over ng cells
  double* backup = nullptr;   // the scratch at the start of a step
  double* fluxes = nullptr;   // (D+2) components over (ng0+1)(ng1+1)(ng2+1) faces
  double* dt_dev = nullptr;
  double t = 0.0;
  int64_t cycle = 0;
};

namespace {

using CL = fvs_composed_level;

// View of the scratch over box [lo, hi] (ghosts allowed), ncomp components;
// data points at the box's lower corner.
fvs_view ScratchView(const CL* C, const int64_t lo[3], const int64_t hi[3], int ncomp) {
  fvs_view v{};
  v.ncomp = ncomp;
  v.stride[0] = 1;
  v.stride[1] = C->ng[0];
  v.stride[2] = C->ng[0] * C->ng[1];
  v.comp_stride = C->ncell_g;
  int64_t off = 0;
  for (int d = 0; d < 3; ++d) {
    v.lo[d] = lo[d];
    v.hi[d] = hi[d];
    const int64_t g = d < C->D ? C->gw : 0;
    off += (lo[d] + g) * v.stride[d];
  }
  v.data = C->scratch + off;
  return v;
}

void Interior(const CL* C, int64_t lo[3], int64_t hi[3]) {
  for (int d = 0; d < 3; ++d) {
    lo[d] = 0;
    hi[d] = d < C->D ? C->n[d] - 1 : 0;
  }
}

void Grown(const CL* C, int64_t lo[3], int64_t hi[3]) {
  for (int d = 0; d < 3; ++d) {
    lo[d] = d < C->D ? -C->gw : 0;
    hi[d] = d < C->D ? C->n[d] - 1 + C->gw : 0;
  }
}

// Transmissive fill of the ghost layers in direction dir: the stable dt and
// the fluxes of dir read the interior grown in dir only, so the other
// directions' ghosts (and the corners) are never read.
int FillGhosts(CL* C, int dir) {
  int64_t glo[3], ghi[3], ilo[3], ihi[3];
  Grown(C, glo, ghi);
  Interior(C, ilo, ihi);
  const fvs_view v = ScratchView(C, glo, ghi, C->D + 4);
  return fvs_transmissive_fill(C->h, &v, ilo, ihi, dir);
}

// Stable dt of direction dir on freshly filled ghosts (host value).
int DirStableDt(CL* C, int dir, double* dt) {
  if (int rc = FillGhosts(C, dir)) return rc;
  int64_t lo[3], hi[3];
  Interior(C, lo, hi);
  lo[dir] -= C->gw;
  hi[dir] += C->gw;
  const fvs_view s = ScratchView(C, lo, hi, C->D + 4);
  if (int rc = fvs_stable_dt(C->h, &s, C->patch.dx[dir], dir, C->dt_dev)) return rc;
  FVS_HIP_CHECK(hipMemcpyAsync(dt, C->dt_dev, sizeof(double), hipMemcpyDeviceToHost,
                               C->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(C->h->stream));
  return FVS_OK;
}

// Newton iterations that hit the cap so far (exact-Godunov base).
int NewtonCapHits(CL* C, uint64_t* hits) {
  uint64_t st[5];
  if (int rc = fvs_newton_stats(C->h, st, 0)) return rc;
  *hits = st[2];
  return FVS_OK;
}

// One split step in dir with the split dt; *level_dt receives the
// direction's stable dt.  With known_dt > 0 the ghosts of dir are filled and
// known_dt is that stable dt already (the step-start value of the first
// split direction: the same faces of the same state).
int SplitStep(CL* C, int dir, double split_dt, double known_dt, double* level_dt) {
  if (known_dt > 0.0) {
    if (int rc = FillGhosts(C, dir)) return rc;
    *level_dt = known_dt;
  } else if (int rc = DirStableDt(C, dir, level_dt)) {
    return rc;
  }
  if (*level_dt < split_dt) return FVS_TIME_STEP_TOO_LARGE;
  const int D = C->D;
  int64_t flo[3], fhi[3];
  Interior(C, flo, fhi);
  fhi[dir] += 1;  // faces 0 .. n
  fvs_view fv{};
  fv.ncomp = D + 2;
  fv.stride[0] = 1;
  fv.stride[1] = fhi[0] - flo[0] + 1;
  fv.stride[2] = fv.stride[1] * (fhi[1] - flo[1] + 1);
  fv.comp_stride = fv.stride[2] * (fhi[2] - flo[2] + 1);
  for (int d = 0; d < 3; ++d) {
    fv.lo[d] = flo[d];
    fv.hi[d] = fhi[d];
  }
  fv.data = C->fluxes;
  int64_t slo[3], shi[3];
  Interior(C, slo, shi);
  slo[dir] -= C->gw;
  shi[dir] += C->gw;
  const fvs_view sv = ScratchView(C, slo, shi, D + 4);
  const bool godunov = C->h->desc.base == FVS_BASE_GODUNOV_EXACT;
  uint64_t caps0 = 0, caps1 = 0;
  if (godunov) {
    if (int rc = NewtonCapHits(C, &caps0)) return rc;
  }
  if (int rc = fvs_numeric_fluxes(C->h, &fv, &sv, split_dt, C->patch.dx[dir], dir)) return rc;
  if (godunov) {
    if (int rc = NewtonCapHits(C, &caps1)) return rc;
    if (caps1 != caps0) {
      SetError("Newton iteration hit its cap in %llu faces",
               static_cast<unsigned long long>(caps1 - caps0));
      return FVS_NEWTON_NOT_CONVERGED;
    }
  }
  int64_t ilo[3], ihi[3];
  Interior(C, ilo, ihi);
  const fvs_view cons = ScratchView(C, ilo, ihi, D + 2);
  const fvs_view complete = ScratchView(C, ilo, ihi, D + 4);
  if (int rc = fvs_update_conservatively(C->h, &cons, &cons, &fv, split_dt, C->patch.dx[dir],
                                         dir))
    return rc;
  return fvs_complete_from_cons(C->h, &complete, &cons);
}

// Copies between ghost-free component arrays (n cells each) and the first
// D+2 components of the scratch interior.
__global__ void InteriorCopyKernel(double* interior, int64_t sy, int64_t sz, int64_t cs,
                                   double* dense, int64_t nx, int64_t ny, int64_t ncell, int ncomp,
                                   int to_scratch) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= ncell) return;
  const int64_t i = q % nx, j = (q / nx) % ny, k = q / (nx * ny);
  const int64_t at = i + j * sy + k * sz;
  for (int c = 0; c < ncomp; ++c) {
    if (to_scratch) interior[c * cs + at] = dense[c * ncell + q];
    else dense[c * ncell + q] = interior[c * cs + at];
  }
}

int CopyInterior(CL* C, const double* const* comps, double* const* out) {
  const int D = C->D;
  const int64_t ncell = C->n[0] * C->n[1] * C->n[2];
  int64_t ilo[3], ihi[3];
  Interior(C, ilo, ihi);
  const fvs_view iv = ScratchView(C, ilo, ihi, D + 2);
  // one component at a time: the caller's arrays need not be contiguous
  for (int c = 0; c < D + 2; ++c) {
    double* dense = comps ? const_cast<double*>(comps[c]) : out[c];
    hipLaunchKernelGGL(InteriorCopyKernel, dim3(static_cast<unsigned>((ncell + 255) / 256)),
                       dim3(256), 0, C->h->stream, iv.data + c * iv.comp_stride, iv.stride[1],
                       iv.stride[2], iv.comp_stride, dense, C->n[0], C->n[1], ncell, 1,
                       comps ? 1 : 0);
    FVS_LAUNCH_CHECK();
  }
  return FVS_OK;
}

} // namespace

bool ComposedLevelNeeded(const fvs_method_desc& d) {
  return (d.base == FVS_BASE_GODUNOV_EXACT && d.reconstruction != FVS_RECON_NONE) ||
         d.base == FVS_BASE_HLL_EINFELDT || d.reconstruction >= FVS_RECON_MUSCL_HANCOCK_PRIM;
}

int ComposedCreate(fvs_handle* h, const fvs_patch_desc* patch, fvs_composed_level** out) {
  *out = nullptr;
  CL* C = new (std::nothrow) CL();
  if (!C) {
    SetError("out of host memory");
    return FVS_BAD_ARGUMENT;
  }
  C->h = h;
  C->patch = *patch;
  C->D = h->desc.dim;
  C->gw = 2 * (h->desc.reconstruction != FVS_RECON_NONE ? 2 : 1);
  for (int d = 0; d < 3; ++d) {
    C->n[d] = d < C->D ? patch->n[d] : 1;
    C->ng[d] = d < C->D ? C->n[d] + 2 * C->gw : 1;
  }
  C->ncell_g = C->ng[0] * C->ng[1] * C->ng[2];
  const int64_t nfaces = (C->ng[0] + 1) * (C->ng[1] + 1) * (C->ng[2] + 1);
  auto fail = [&](int rc) {
    ComposedDestroy(C);
    return rc;
  };
  const size_t sbytes = static_cast<size_t>(C->ncell_g) * (C->D + 4) * sizeof(double);
  if (hipMalloc(&C->scratch, sbytes) != hipSuccess || hipMalloc(&C->backup, sbytes) != hipSuccess ||
      hipMalloc(&C->fluxes, static_cast<size_t>(nfaces) * (C->D + 2) * sizeof(double)) !=
          hipSuccess ||
      hipMalloc(&C->dt_dev, sizeof(double)) != hipSuccess) {
    SetError("device allocation failed");
    return fail(FVS_HIP_ERROR);
  }
  if (hipMemsetAsync(C->scratch, 0, sbytes, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  *out = C;
  return FVS_OK;
}

void ComposedDestroy(fvs_composed_level* C) {
  if (!C) return;
  (void)hipStreamSynchronize(C->h->stream);
  void* ps[] = {C->scratch, C->backup, C->fluxes, C->dt_dev};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete C;
}

int ComposedSetState(fvs_composed_level* C, const double* const* comps, double t0) {
  if (int rc = CopyInterior(C, comps, nullptr)) return rc;
  int64_t ilo[3], ihi[3];
  Interior(C, ilo, ihi);
  const fvs_view cons = ScratchView(C, ilo, ihi, C->D + 2);
  const fvs_view complete = ScratchView(C, ilo, ihi, C->D + 4);
  if (int rc = fvs_complete_from_cons(C->h, &complete, &cons)) return rc;
  FVS_HIP_CHECK(hipStreamSynchronize(C->h->stream));
  C->t = t0;
  C->cycle = 0;
  return FVS_OK;
}

int ComposedGetState(fvs_composed_level* C, double* const* comps) {
  return CopyInterior(C, nullptr, comps);
}

void ComposedTime(const fvs_composed_level* C, double* t, int64_t* cycles) {
  if (t) *t = C->t;
  if (cycles) *cycles = C->cycle;
}

// RunSimulation's loop (include/fub/RunSimulation.hpp:86-191) over at most
// max_steps steps while t + smallest_dt < final_time.
int ComposedAdvance(fvs_composed_level* C, int max_steps, double final_time, double* dt_log,
                    int* steps_done, int* failures) {
  const int D = C->D;
  const double eps = C->patch.smallest_dt > 0.0 ? C->patch.smallest_dt : 1e-12;
  const size_t sbytes = static_cast<size_t>(C->ncell_g) * (D + 4) * sizeof(double);
  int done = 0, nfail = 0, consecutive = 0;
  bool have_failure = false;
  double failure_dt = 0.0;
  int rc = FVS_OK;
  while (C->t + eps < final_time && done < max_steps) {
    FVS_HIP_CHECK(hipMemcpyAsync(C->backup, C->scratch, sbytes, hipMemcpyDeviceToDevice,
                                 C->h->stream));
    double stable = failure_dt;
    double dir_dt[3] = {0.0, 0.0, 0.0};
    if (!have_failure) {
      stable = 1.7976931348623157e308;
      for (int d = 0; d < D && rc == FVS_OK; ++d) {
        rc = DirStableDt(C, d, &dir_dt[d]);
        stable = std::min(stable, dir_dt[d]);
      }
      if (rc) break;
    }
    const double limited = std::min(final_time - C->t, C->patch.cfl * stable);
    // MakeSplitDirections (include/fub/split_method/SplittingMethod.hpp:115-126)
    const bool odd = (C->cycle % 2) == 1;
    int dirs[3];
    for (int i = 0; i < D; ++i) dirs[i] = odd ? D - 1 - i : i;
    std::vector<std::pair<int, double>> seq;
    if (C->patch.splitting == FVS_SPLIT_STRANG) {
      for (int i = 0; i < D; ++i) seq.push_back({dirs[i], 0.5 * limited});
      for (int i = D - 1; i >= 0; --i) seq.push_back({dirs[i], 0.5 * limited});
    } else {
      for (int i = 0; i < D; ++i) seq.push_back({dirs[i], limited});
    }
    bool failed = false;
    for (size_t i = 0; i < seq.size(); ++i) {
      const auto& s = seq[i];
      // The first split step starts from the state the step-start stable
      // dts were computed on: its check is that direction's value.
      const double known = (i == 0 && !have_failure) ? dir_dt[s.first] : 0.0;
      double level_dt;
      const int r = SplitStep(C, s.first, s.second, known, &level_dt);
      if (r == FVS_TIME_STEP_TOO_LARGE) {
        failed = true;
        failure_dt = level_dt;
        break;
      }
      if (r != FVS_OK) {
        rc = r;
        break;
      }
    }
    if (rc) break;
    if (failed) {
      FVS_HIP_CHECK(hipMemcpyAsync(C->scratch, C->backup, sbytes, hipMemcpyDeviceToDevice,
                                   C->h->stream));
      have_failure = true;
      ++nfail;
      if (++consecutive > 64) {
        SetError("CFL check failed 64 times in a row");
        rc = FVS_TIME_STEP_TOO_LARGE;
        break;
      }
      continue;
    }
    consecutive = 0;
    have_failure = false;
    if (dt_log) dt_log[done] = limited;
    C->t += limited;
    ++C->cycle;
    ++done;
  }
  FVS_HIP_CHECK(hipStreamSynchronize(C->h->stream));
  if (steps_done) *steps_done = done;
  if (failures) *failures = nfail;
  return rc;
}

} // namespace fvs
