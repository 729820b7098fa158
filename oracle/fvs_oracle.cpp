// TEST INFRASTRUCTURE ONLY: extern "C" entry points of the FP64 CPU oracle.
// Only tests, the smoke check and the benchmark's CPU-baseline leg load the
// shared library built from this file.  See fvs_oracle.hpp for the
// arithmetic conventions.
//
// Patch-level functions take the same fvs_view descriptions as the product's
// C ABI (include/fvs.h) but with host pointers, and follow the reference's
// drivers box for box:
//   FluxMethod::ComputeNumericFluxes(simd)  include/fub/flux_method/FluxMethod.hpp:246-293
//   FluxMethod::ComputeStableDt(simd)       include/fub/flux_method/FluxMethod.hpp:341-382
//   FluxMethod::*(seq)                      include/fub/flux_method/FluxMethod.hpp:200-215,295-314
//   HyperbolicPatchIntegrator               src/HyperbolicPatchIntegrator.cpp:35-104
//   CompleteFromConsFn                      include/fub/CompleteFromCons.hpp:141-204
//   TransmissiveBoundary::FillBoundary      src/AMReX/boundary_condition/TransmissiveBoundary.cpp:31-93
//   AdvanceLevelNonRecursively              include/fub/solver/DimensionalSplitLevelIntegrator.hpp:221-276
//   RunSimulation                           include/fub/RunSimulation.hpp:86-191

#include "fvs_oracle.hpp"
#include "fvs_oracle_views.hpp"

#include "../include/fvs.h"

#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace fvs_oracle;

namespace {

template <int D>
int NumericFluxes(const Method& m, const fvs_view* fluxes, const fvs_view* states,
                  double dt, double dx, int dir) {
  const int w = m.width();
  const int S = 2 * w;
  ViewAcc fa{fluxes}, sa{states};
  int64_t lo[3], hi[3];
  BoxOf(fluxes, lo, hi);
  int newton_fail = 0;
  const int64_t nk = hi[2] - lo[2] + 1, nj = hi[1] - lo[1] + 1;
#pragma omp parallel for collapse(2) schedule(static) reduction(| : newton_fail)
  for (int64_t k = lo[2]; k <= hi[2]; ++k) {
    for (int64_t j = lo[1]; j <= hi[1]; ++j) {
      State<D> st[4];
      Cons<D> f;
      for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        int64_t cell[3] = {i, j, k};
        Shift(cell, dir, -w);
        for (int s = 0; s < S; ++s) {
          LoadState<D>(st[s], sa, cell);
          Shift(cell, dir, 1);
        }
        FaceFlux<D>(m, f, st, dt, dx, dir, &newton_fail);
        for (int c = 0; c < D + 2; ++c) *fa.at(i, j, k, c) = C(f, c);
      }
    }
  }
  (void)nk;
  (void)nj;
  return newton_fail ? FVS_NEWTON_NOT_CONVERGED : FVS_OK;
}

template <int D>
int StableDt(const Method& m, const fvs_view* states, double dx, int dir, double* out) {
  const int S = 2 * m.width();
  ViewAcc sa{states};
  int64_t lo[3], hi[3];
  BoxOf(states, lo, hi);
  // Faces: the pairs of every stencil whose cells lie in the box
  // (Shrink(states, dir, {i, S-1-i}), FluxMethod.hpp:346-352).  The
  // sequential driver shrinks by the full stencil size instead
  // (Shrink(Box<0>(states), dir, {0, StencilSize}), FluxMethod.hpp:299) and
  // so skips the box's last face.
  hi[dir] -= m.seq ? S : (S - 1);
  double min_dt = std::numeric_limits<double>::infinity();
#pragma omp parallel for collapse(2) schedule(static) reduction(min : min_dt)
  for (int64_t k = lo[2]; k <= hi[2]; ++k) {
    for (int64_t j = lo[1]; j <= hi[1]; ++j) {
      State<D> st[4];
      for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        int64_t cell[3] = {i, j, k};
        for (int s = 0; s < S; ++s) {
          LoadState<D>(st[s], sa, cell);
          Shift(cell, dir, 1);
        }
        const double dt = FaceStableDt<D>(m, st, dx, dir);
        min_dt = std::min(min_dt, dt);
      }
    }
  }
  *out = min_dt;
  return FVS_OK;
}

int Update(const fvs_view* next, const fvs_view* prev, const fvs_view* fluxes,
           double dt, double dx, int dir, int ncomp) {
  ViewAcc na{next}, pa{prev}, fa{fluxes};
  int64_t lo[3], hi[3];
  BoxOf(next, lo, hi);
  const double dt_dx = dt / dx;
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t k = lo[2]; k <= hi[2]; ++k) {
    for (int64_t j = lo[1]; j <= hi[1]; ++j) {
      for (int c = 0; c < ncomp; ++c) {
        for (int64_t i = lo[0]; i <= hi[0]; ++i) {
          int64_t r[3] = {i, j, k};
          Shift(r, dir, 1);
          const double U_prev = *pa.at(i, j, k, c);
          const double fl = *fa.at(i, j, k, c);
          const double fr = *fa.at(r[0], r[1], r[2], c);
          *na.at(i, j, k, c) = U_prev + dt_dx * (fl - fr);
        }
      }
    }
  }
  return FVS_OK;
}

template <int D>
int CompleteView(const Gas& g, const fvs_view* complete, const fvs_view* cons) {
  ViewAcc ca{complete}, ua{cons};
  int64_t lo[3], hi[3];
  BoxOf(complete, lo, hi);
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t k = lo[2]; k <= hi[2]; ++k) {
    for (int64_t j = lo[1]; j <= hi[1]; ++j) {
      for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        State<D> q{};
        for (int c = 0; c < D + 2; ++c) C(q, c) = *ua.at(i, j, k, c);
        CompleteFromConsArray<D>(g, q);
        for (int c = 0; c < D + 4; ++c) *ca.at(i, j, k, c) = C(q, c);
      }
    }
  }
  return FVS_OK;
}

int CheckDesc(const fvs_method_desc* d) {
  if (!d || (d->dim != 2 && d->dim != 3)) return FVS_BAD_ARGUMENT;
  return FVS_OK;
}

// ---------------------------------------------------------------------------
// Single-patch level integrator following the reference's sequence.

template <int D> struct Level {
  Method m;
  int64_t n[3];
  double dx[3];
  int gw;
  int64_t ng[3];
  std::vector<double> scratch; // Complete state with gw ghosts, (D+4) comps
  std::vector<double> fluxes;  // (D+2) comps over faces, per direction reuse
  int64_t ncell_g;

  Level(const fvs_method_desc* d, const fvs_patch_desc* p) : m(d) {
    // Two stencil widths of ghosts: enough that every face of the patch and
    // its boundary (a cell and its transmissive copy) lies inside the CFL
    // box for both the simd and the sequential face ranges.
    gw = 2 * m.width();
    for (int i = 0; i < 3; ++i) {
      n[i] = p->n[i];
      dx[i] = p->dx[i];
      ng[i] = (i < D) ? n[i] + 2 * gw : 1;
    }
    ncell_g = ng[0] * ng[1] * ng[2];
    scratch.assign(static_cast<size_t>(ncell_g) * (D + 4), 0.0);
    fluxes.assign(static_cast<size_t>((ng[0] + 1) * (ng[1] + 1) * (ng[2] + 1)) * (D + 2), 0.0);
  }

  fvs_view ScratchView(const int64_t lo[3], const int64_t hi[3]) {
    fvs_view v{};
    v.ncomp = D + 4;
    v.stride[0] = 1;
    v.stride[1] = ng[0];
    v.stride[2] = ng[0] * ng[1];
    v.comp_stride = ncell_g;
    int64_t off = 0;
    for (int d = 0; d < 3; ++d) {
      v.lo[d] = lo[d];
      v.hi[d] = hi[d];
      const int64_t g = (d < D) ? gw : 0;
      off += (lo[d] + g) * v.stride[d];
    }
    v.data = scratch.data() + off;
    return v;
  }

  // The whole scratch box (interior plus ghosts).
  void GrownBox(int64_t lo[3], int64_t hi[3]) {
    for (int d = 0; d < 3; ++d) {
      lo[d] = (d < D) ? -gw : 0;
      hi[d] = (d < D) ? n[d] - 1 + gw : 0;
    }
  }
  void Interior(int64_t lo[3], int64_t hi[3]) {
    for (int d = 0; d < 3; ++d) {
      lo[d] = 0;
      hi[d] = (d < D) ? n[d] - 1 : 0;
    }
  }

  void FillAllGhosts() {
    int64_t glo[3], ghi[3], ilo[3], ihi[3];
    GrownBox(glo, ghi);
    Interior(ilo, ihi);
    fvs_view v = ScratchView(glo, ghi);
    for (int d = 0; d < D; ++d) TransmissiveFill(&v, ilo, ihi, d);
  }

  // Stable dt in dir with freshly filled transmissive ghosts: the states box
  // is the interior grown by the stencil width in dir.
  double DirStableDt(int dir) {
    FillAllGhosts();
    int64_t lo[3], hi[3];
    Interior(lo, hi);
    lo[dir] -= gw;
    hi[dir] += gw;
    fvs_view s = ScratchView(lo, hi);
    double dt;
    StableDt<D>(m, &s, dx[dir], dir, &dt);
    return dt;
  }

  double ComputeStableDt() {
    double min_dt = std::numeric_limits<double>::max();
    for (int d = 0; d < D; ++d) min_dt = std::min(min_dt, DirStableDt(d));
    return min_dt;
  }

  // One split step in dir (DimensionalSplitLevelIntegrator.hpp:227-264).
  int SplitStep(int dir, double split_dt, double* level_dt) {
    const double local_dt = DirStableDt(dir); // fills ghosts too
    *level_dt = local_dt;
    if (local_dt < split_dt) return FVS_TIME_STEP_TOO_LARGE;
    int64_t flo[3], fhi[3];
    Interior(flo, fhi);
    fhi[dir] += 1; // faces 0 .. n
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
    fv.data = fluxes.data();
    int64_t slo[3], shi[3];
    Interior(slo, shi);
    slo[dir] -= gw;
    shi[dir] += gw;
    fvs_view sv = ScratchView(slo, shi);
    const int rc = NumericFluxes<D>(m, &fv, &sv, split_dt, dx[dir], dir);
    if (rc != FVS_OK) return rc;
    int64_t ilo[3], ihi[3];
    Interior(ilo, ihi);
    fvs_view iv = ScratchView(ilo, ihi);
    Update(&iv, &iv, &fv, split_dt, dx[dir], dir, D + 2);
    CompleteView<D>(m.gas, &iv, &iv);
    return FVS_OK;
  }
};

template <int D>
int LevelRun(const fvs_method_desc* d, const fvs_patch_desc* p, double* cons,
             double t0, int64_t cycle0, int max_steps, double final_time, double* dt_log,
             int* steps_done, int* failures) {
  Level<D> L(d, p);
  const int64_t ncell = L.n[0] * L.n[1] * (D == 3 ? L.n[2] : 1);
  {
    int64_t ilo[3], ihi[3];
    L.Interior(ilo, ihi);
    fvs_view iv = L.ScratchView(ilo, ihi);
    fvs_view uv{};
    uv.data = cons;
    uv.ncomp = D + 2;
    uv.stride[0] = 1;
    uv.stride[1] = L.n[0];
    uv.stride[2] = L.n[0] * L.n[1];
    uv.comp_stride = ncell;
    for (int i = 0; i < 3; ++i) {
      uv.lo[i] = ilo[i];
      uv.hi[i] = ihi[i];
    }
    CompleteView<D>(L.m.gas, &iv, &uv);
  }
  double t = t0;
  int64_t cycle = 0;  // steps taken here; the split order follows cycle0 + cycle
  int nfail = 0;
  int consecutive_fail = 0;
  bool have_failure = false;
  double failure_dt = 0.0;
  // RunOptions::smallest_time_step_size (RunSimulation.hpp:61-63, 97, 109)
  const double eps = p->smallest_dt > 0.0 ? p->smallest_dt : 1e-12;
  std::vector<double> backup;
  int rc = FVS_OK;
  while (t + eps < final_time && cycle < max_steps) {
    backup = L.scratch; // RunSimulation's deep copy (do_backup)
    const double stable_dt = have_failure ? failure_dt : L.ComputeStableDt();
    const double limited_dt = std::min(final_time - t, p->cfl * stable_dt);
    // MakeSplitDirections (SplittingMethod.hpp:115-126) with subcycle {0, 1}.
    const int is_odd = static_cast<int>(((cycle0 + cycle) % 2 + 0) % 2);
    int dirs[3];
    for (int i = 0; i < D; ++i) dirs[i] = is_odd ? (D - 1 - i) : i;
    std::vector<std::pair<int, double>> seq;
    if (p->splitting == FVS_SPLIT_STRANG) {
      for (int i = 0; i < D; ++i) seq.push_back({dirs[i], 0.5 * limited_dt});
      for (int i = D - 1; i >= 0; --i) seq.push_back({dirs[i], 0.5 * limited_dt});
    } else {
      for (int i = 0; i < D; ++i) seq.push_back({dirs[i], limited_dt});
    }
    bool failed = false;
    for (auto& s : seq) {
      double level_dt;
      const int r = L.SplitStep(s.first, s.second, &level_dt);
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
    if (rc != FVS_OK) break;
    if (failed) {
      L.scratch = backup;
      have_failure = true;
      ++nfail;
      if (++consecutive_fail > 64) {
        rc = FVS_TIME_STEP_TOO_LARGE;
        break;
      }
      continue;
    }
    consecutive_fail = 0;
    have_failure = false;
    if (dt_log) dt_log[cycle] = limited_dt;
    t += limited_dt;
    ++cycle;
  }
  // Copy interior conservative variables back.
  {
    int64_t ilo[3], ihi[3];
    L.Interior(ilo, ihi);
    fvs_view iv = L.ScratchView(ilo, ihi);
    ViewAcc a{&iv};
    for (int c = 0; c < D + 2; ++c)
      for (int64_t k = ilo[2]; k <= ihi[2]; ++k)
        for (int64_t j = ilo[1]; j <= ihi[1]; ++j)
          for (int64_t i = ilo[0]; i <= ihi[0]; ++i)
            cons[c * ncell + (k * L.n[1] + j) * L.n[0] + i] = *a.at(i, j, k, c);
  }
  if (steps_done) *steps_done = static_cast<int>(cycle);
  if (failures) *failures = nfail;
  return rc;
}

template <int D>
void SoaLoad(State<D>& q, const double* a, int64_t n, int64_t i, int ncomp) {
  for (int c = 0; c < ncomp; ++c) C(q, c) = a[c * n + i];
}

} // namespace

extern "C" {

// Switches a planted mistake on (enum Plant in fvs_oracle.hpp) or off (0);
// returns the previous setting.  Only the oracle's own pinning tests use it.
int oracle_plant(int which) {
  const int prev = g_plant;
  g_plant = which;
  return prev;
}

int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

// ---- pointwise functions over SoA arrays of states -----------------------

int oracle_complete_from_cons(int dim, double gamma, int64_t n, const double* cons,
                              double* complete) {
  const Gas g(gamma);
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      State<2> q{};
      SoaLoad<2>(q, cons, n, i, 4);
      CompleteFromConsArray<2>(g, q);
      for (int c = 0; c < 6; ++c) complete[c * n + i] = C(q, c);
    } else if (dim == 3) {
      State<3> q{};
      SoaLoad<3>(q, cons, n, i, 5);
      CompleteFromConsArray<3>(g, q);
      for (int c = 0; c < 7; ++c) complete[c * n + i] = C(q, c);
    } else {
      return FVS_BAD_ARGUMENT;
    }
  }
  return FVS_OK;
}

int oracle_complete_from_prim(int dim, double gamma, int64_t n, const double* rho,
                              const double* vel /* dim x n */, const double* p,
                              double* complete) {
  const Gas g(gamma);
  for (int64_t i = 0; i < n; ++i) {
    double v[3];
    for (int d = 0; d < dim; ++d) v[d] = vel[d * n + i];
    if (dim == 2) {
      State<2> q = CompleteFromPrim<2>(g, rho[i], v, p[i]);
      for (int c = 0; c < 6; ++c) complete[c * n + i] = C(q, c);
    } else if (dim == 3) {
      State<3> q = CompleteFromPrim<3>(g, rho[i], v, p[i]);
      for (int c = 0; c < 7; ++c) complete[c * n + i] = C(q, c);
    } else {
      return FVS_BAD_ARGUMENT;
    }
  }
  return FVS_OK;
}

int oracle_physical_flux(int dim, int dir, int64_t n, const double* complete,
                         double* flux) {
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      State<2> q{};
      SoaLoad<2>(q, complete, n, i, 6);
      Cons<2> f;
      Flux<2>(f, q, dir);
      for (int c = 0; c < 4; ++c) flux[c * n + i] = C(f, c);
    } else {
      State<3> q{};
      SoaLoad<3>(q, complete, n, i, 7);
      Cons<3> f;
      Flux<3>(f, q, dir);
      for (int c = 0; c < 5; ++c) flux[c * n + i] = C(f, c);
    }
  }
  return FVS_OK;
}

int oracle_limited_slope(int limiter, int64_t n, const double* qL, const double* qM,
                         const double* qR, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = LimitedSlope(limiter, qL[i], qM[i], qR[i]);
  return FVS_OK;
}

// HLLEM flux between left[i] and right[i] (Complete states, SoA (D+4) x n).
int oracle_hllem_flux(int dim, double gamma, int dir, int larrouturou, int quirk3d,
                      int64_t n, const double* left, const double* right,
                      double* flux, double* pstar) {
  const Gas g(gamma);
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      State<2> L{}, R{};
      SoaLoad<2>(L, left, n, i, 6);
      SoaLoad<2>(R, right, n, i, 6);
      Cons<2> f;
      const double ps = HllemFlux<2>(g, f, L, R, dir, larrouturou != 0, quirk3d != 0);
      for (int c = 0; c < 4; ++c) flux[c * n + i] = C(f, c);
      if (pstar) pstar[i] = ps;
    } else if (dim == 3) {
      State<3> L{}, R{};
      SoaLoad<3>(L, left, n, i, 7);
      SoaLoad<3>(R, right, n, i, 7);
      Cons<3> f;
      const double ps = HllemFlux<3>(g, f, L, R, dir, larrouturou != 0, quirk3d != 0);
      for (int c = 0; c < 5; ++c) flux[c * n + i] = C(f, c);
      if (pstar) pstar[i] = ps;
    } else {
      return FVS_BAD_ARGUMENT;
    }
  }
  return FVS_OK;
}

int oracle_hllem_stable_dt(int dim, double gamma, int dir, double dx, int64_t n,
                           const double* left, const double* right, double* dt) {
  const Gas g(gamma);
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      State<2> L{}, R{};
      SoaLoad<2>(L, left, n, i, 6);
      SoaLoad<2>(R, right, n, i, 6);
      dt[i] = HllemStableDt<2>(g, L, R, dx, dir);
    } else {
      State<3> L{}, R{};
      SoaLoad<3>(L, left, n, i, 7);
      SoaLoad<3>(R, right, n, i, 7);
      dt[i] = HllemStableDt<3>(g, L, R, dx, dir);
    }
  }
  return FVS_OK;
}

// Face flux of the configured method on n stencils; stencil holds S = 2*width
// Complete states per face: stencil[(s * (D+4) + c) * n + i].
int oracle_face_flux(const fvs_method_desc* desc, int dir, double dt, double dx,
                     int64_t n, const double* stencil, double* flux) {
  if (CheckDesc(desc)) return FVS_BAD_ARGUMENT;
  Method m(desc);
  const int S = 2 * m.width();
  int fail = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (desc->dim == 2) {
      State<2> st[4];
      for (int s = 0; s < S; ++s) SoaLoad<2>(st[s], stencil + s * 6 * n, n, i, 6);
      Cons<2> f;
      FaceFlux<2>(m, f, st, dt, dx, dir, &fail);
      for (int c = 0; c < 4; ++c) flux[c * n + i] = C(f, c);
    } else {
      State<3> st[4];
      for (int s = 0; s < S; ++s) SoaLoad<3>(st[s], stencil + s * 7 * n, n, i, 7);
      Cons<3> f;
      FaceFlux<3>(m, f, st, dt, dx, dir, &fail);
      for (int c = 0; c < 5; ++c) flux[c * n + i] = C(f, c);
    }
  }
  return fail ? FVS_NEWTON_NOT_CONVERGED : FVS_OK;
}

// Exact Riemann solver pieces on n state pairs.  x0 may be NULL (reference
// guess 0.5 (pL + pR)).  Outputs may be NULL.
int oracle_exact_middle_state(int dim, double gamma, int dir, int64_t n,
                              const double* left, const double* right,
                              const double* x0, double tol, int max_it, double* pM,
                              double* uM, int* iterations, int* converged) {
  const Gas g(gamma);
  for (int64_t i = 0; i < n; ++i) {
    MiddleState ms;
    const double guess = x0 ? x0[i] : -1.0;
    if (dim == 2) {
      State<2> L{}, R{};
      SoaLoad<2>(L, left, n, i, 6);
      SoaLoad<2>(R, right, n, i, 6);
      ms = ComputeMiddleState<2>(g, L, R, dir, guess, tol, max_it);
    } else {
      State<3> L{}, R{};
      SoaLoad<3>(L, left, n, i, 7);
      SoaLoad<3>(R, right, n, i, 7);
      ms = ComputeMiddleState<3>(g, L, R, dir, guess, tol, max_it);
    }
    if (pM) pM[i] = ms.pM;
    if (uM) uM[i] = ms.uM;
    if (iterations) iterations[i] = ms.iterations;
    if (converged) converged[i] = ms.converged ? 1 : 0;
  }
  return FVS_OK;
}

int oracle_exact_riemann_state(int dim, double gamma, int dir, int64_t n,
                               const double* left, const double* right,
                               double* state) {
  const Gas g(gamma);
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      State<2> L{}, R{}, S{};
      SoaLoad<2>(L, left, n, i, 6);
      SoaLoad<2>(R, right, n, i, 6);
      SolveRiemannProblem<2>(g, S, L, R, dir);
      for (int c = 0; c < 6; ++c) state[c * n + i] = C(S, c);
    } else {
      State<3> L{}, R{}, S{};
      SoaLoad<3>(L, left, n, i, 7);
      SoaLoad<3>(R, right, n, i, 7);
      SolveRiemannProblem<3>(g, S, L, R, dir);
      for (int c = 0; c < 7; ++c) state[c * n + i] = C(S, c);
    }
  }
  return FVS_OK;
}

int oracle_exact_signals(int dim, double gamma, int dir, int64_t n,
                         const double* left, const double* right, double* sL,
                         double* sR) {
  const Gas g(gamma);
  for (int64_t i = 0; i < n; ++i) {
    double s[2];
    if (dim == 2) {
      State<2> L{}, R{};
      SoaLoad<2>(L, left, n, i, 6);
      SoaLoad<2>(R, right, n, i, 6);
      ComputeSignals<2>(g, L, R, dir, s);
    } else {
      State<3> L{}, R{};
      SoaLoad<3>(L, left, n, i, 7);
      SoaLoad<3>(R, right, n, i, 7);
      ComputeSignals<3>(g, L, R, dir, s);
    }
    sL[i] = s[0];
    sR[i] = s[1];
  }
  return FVS_OK;
}

// ---- patch-level operators on host views -----------------------------------

int oracle_stable_dt(const fvs_method_desc* desc, const fvs_view* states, double dx,
                     int dir, double* dt) {
  if (CheckDesc(desc)) return FVS_BAD_ARGUMENT;
  Method m(desc);
  return desc->dim == 2 ? StableDt<2>(m, states, dx, dir, dt)
                        : StableDt<3>(m, states, dx, dir, dt);
}

int oracle_numeric_fluxes(const fvs_method_desc* desc, const fvs_view* fluxes,
                          const fvs_view* states, double dt, double dx, int dir) {
  if (CheckDesc(desc)) return FVS_BAD_ARGUMENT;
  Method m(desc);
  return desc->dim == 2 ? NumericFluxes<2>(m, fluxes, states, dt, dx, dir)
                        : NumericFluxes<3>(m, fluxes, states, dt, dx, dir);
}

int oracle_update_conservatively(const fvs_view* next, const fvs_view* prev,
                                 const fvs_view* fluxes, double dt, double dx,
                                 int dir) {
  return Update(next, prev, fluxes, dt, dx, dir, next->ncomp);
}

int oracle_complete_view(int dim, double gamma, const fvs_view* complete,
                         const fvs_view* cons) {
  const Gas g(gamma);
  return dim == 2 ? CompleteView<2>(g, complete, cons) : CompleteView<3>(g, complete, cons);
}

int oracle_transmissive_fill(const fvs_view* states, const int64_t dlo[3],
                             const int64_t dhi[3], int dir) {
  return TransmissiveFill(states, dlo, dhi, dir);
}

// ---- single-patch level integrator ------------------------------------------

int oracle_level_run_from(const fvs_method_desc* desc, const fvs_patch_desc* patch,
                          double* cons, double t0, int64_t cycle0, int max_steps,
                          double final_time, double* dt_log, int* steps_done, int* failures) {
  if (CheckDesc(desc) || !patch) return FVS_BAD_ARGUMENT;
  return desc->dim == 2
             ? LevelRun<2>(desc, patch, cons, t0, cycle0, max_steps, final_time, dt_log,
                           steps_done, failures)
             : LevelRun<3>(desc, patch, cons, t0, cycle0, max_steps, final_time, dt_log,
                           steps_done, failures);
}

int oracle_level_run(const fvs_method_desc* desc, const fvs_patch_desc* patch,
                     double* cons, double t0, int max_steps, double final_time,
                     double* dt_log, int* steps_done, int* failures) {
  return oracle_level_run_from(desc, patch, cons, t0, 0, max_steps, final_time, dt_log,
                               steps_done, failures);
}

} // extern "C"
