// TEST INFRASTRUCTURE ONLY: FP64 CPU oracle of the KBN h-box cut-cell
// method (KbnCutCellMethod and the cut-cell time integrators).  Only tests
// and the benchmark's CPU-baseline leg load the library built from this
// file; the product never includes, links or calls it.
//
// Followed, box for box and operation for operation:
//   FillCutCellData__               src/AMReX/cutcell/FillCutCellData.cpp:28-118
//   IsCutCell, GetBoundaryNormal    src/CutCellData.cpp:90-105,225-256
//   MakeRotation (2D, 3D)           include/fub/ext/Eigen.hpp:104-130
//   Rotate / Reflect                src/equations/PerfectGas.cpp:230-300
//   Hllem::SolveRiemannProblem      src/equations/perfect_gas/HllemMethod.hpp:25-369
//   KbnCutCellMethod                include/fub/cutcell_method/KbnStabilisation.hpp:188-429
//   ComputeStableFluxComponents     src/cutcell_method/KbnStabilisation.cpp:55-184
//   cutcell::TimeIntegrator(2)      src/AMReX/cutcell/TimeIntegrator.cpp:65-508
//   masked MUSCL-Hancock            include/fub/flux_method/MusclHancockMethod.hpp:487-571
//   masked HLLEM flux / stable dt   src/equations/perfect_gas/HllemMethod.hpp:1074-1093,1218-1293
//   volume-masked CompleteFromCons  include/fub/AMReX/cutcell/Reconstruction.hpp:123-157
//
// The 2D rotation is exact in any evaluation order (two-term sums).  The 3D
// rotation's sums are evaluated left to right ((a0 + a1) + a2); Eigen's
// small fixed-size product order for them is not pinned by any reference
// test, so 3D reference/boundary states are "parity unpinned" beyond the
// restatement.

#include "fvs_oracle.hpp"
#include "fvs_oracle_views.hpp"

#include "../include/fvs.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <vector>

using namespace fvs_oracle;

namespace {

// ---- geometry access ----------------------------------------------------------

inline double At(const fvs_view& v, int64_t i, int64_t j, int64_t k, int c = 0) {
  return *(v.data + (i - v.lo[0]) * v.stride[0] + (j - v.lo[1]) * v.stride[1] +
           (k - v.lo[2]) * v.stride[2] + c * v.comp_stride);
}
inline double* Ptr(const fvs_view& v, int64_t i, int64_t j, int64_t k, int c = 0) {
  return v.data + (i - v.lo[0]) * v.stride[0] + (j - v.lo[1]) * v.stride[1] +
         (k - v.lo[2]) * v.stride[2] + c * v.comp_stride;
}

// IsCutCell_ (src/CutCellData.cpp:225-245).
template <int D> bool IsCutCell(const fvs_cutcell_data* g, const int64_t idx[3]) {
  const double alpha = At(g->volume_fractions, idx[0], idx[1], idx[2]);
  if (alpha <= 0.0) return false;
  if (alpha < 1.0) return true;
  for (int d = 0; d < D; ++d) {
    int64_t r[3] = {idx[0], idx[1], idx[2]};
    r[d] += 1;
    if (At(g->face_fractions[d], idx[0], idx[1], idx[2]) < 1.0 ||
        At(g->face_fractions[d], r[0], r[1], r[2]) < 1.0)
      return true;
  }
  return false;
}

// ---- rotations (include/fub/ext/Eigen.hpp:104-130) ----------------------------

template <int D> struct Rot {
  double m[D][D];
};

// MakeRotation(a, b): the rotation taking a onto b.
template <int D> Rot<D> MakeRotation(const double* a, const double* b) {
  Rot<D> r;
  bool opposite = true;
  for (int i = 0; i < D; ++i) opposite = opposite && (a[i] == -b[i]);
  if (opposite) {  // -Identity
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) r.m[i][j] = (i == j) ? -1.0 : -0.0;
    return r;
  }
  if constexpr (D == 2) {
    const double s = a[0] * b[1] - a[1] * b[0];
    const double c = a[0] * b[0] + a[1] * b[1];
    r.m[0][0] = c;
    r.m[0][1] = -s;
    r.m[1][0] = s;
    r.m[1][1] = c;
  } else {
    // I + [v]x + [v]x^2 / (1 + c), v = a x b, c = a . b
    const double v[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                         a[0] * b[1] - a[1] * b[0]};
    const double c = (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
    const double sk[3][3] = {{0.0, -v[2], +v[1]}, {+v[2], 0.0, -v[0]}, {-v[1], +v[0], 0.0}};
    const double den = 1.0 + c;
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) {
        const double sq = (sk[i][0] * sk[0][j] + sk[i][1] * sk[1][j]) + sk[i][2] * sk[2][j];
        r.m[i][j] = ((i == j ? 1.0 : 0.0) + sk[i][j]) + sq / den;
      }
    }
  }
  return r;
}

// Rotate(Complete): momentum <- R momentum, the rest copied.
template <int D> void RotateMomentum(State<D>& q, const Rot<D>& r) {
  double out[D];
  for (int i = 0; i < D; ++i) {
    double acc = r.m[i][0] * q.mom[0];
    for (int k = 1; k < D; ++k) acc = acc + r.m[i][k] * q.mom[k];
    out[i] = acc;
  }
  for (int i = 0; i < D; ++i) q.mom[i] = out[i];
}

// Reflect(Complete, normal): m - 2 (m . n) n.
template <int D> State<D> Reflect(const State<D>& q, const double n[D]) {
  State<D> r = q;
  double dot = q.mom[0] * n[0];
  for (int k = 1; k < D; ++k) dot = dot + q.mom[k] * n[k];
  for (int i = 0; i < D; ++i) r.mom[i] = q.mom[i] - 2.0 * (dot * n[i]);
  return r;
}

// GetBoundaryNormal: minus the stored (AMReX) normal.
template <int D> void BoundaryNormal(const fvs_cutcell_data* g, const int64_t idx[3], double n[D]) {
  for (int c = 0; c < D; ++c) n[c] = -At(g->boundary_normals, idx[0], idx[1], idx[2], c);
}

// ---- Hllem::SolveRiemannProblem, scalar (HllemMethod.hpp:25-369) -----------------

template <int D>
void HllemRiemannState(const Gas& g, State<D>& solution, const State<D>& left,
                       const State<D>& right, int d, bool larrouturou) {
  State<D> L = left, R = right;
  const double Msq = 1.0;
  const double Minv = 1.0 / std::sqrt(Msq);
  const double rhoL = L.rho, rhoR = R.rho;
  double vL_[D], vR_[D];
  for (int i = 0; i < D; ++i) {
    vL_[i] = L.mom[i] / rhoL;
    vR_[i] = R.mom[i] / rhoR;
  }
  double eKinRL = 0.0, eKinRR = 0.0;
  if (larrouturou) {
    for (int i = 0; i < D; ++i) {
      const int mask = (i == d);
      eKinRL += (1 - mask) * vL_[i] * vL_[i];
      eKinRR += (1 - mask) * vR_[i] * vR_[i];
      L.mom[i] *= mask;
      R.mom[i] *= mask;
    }
    eKinRL *= 0.5 * Msq;
    eKinRR *= 0.5 * Msq;
    L.E -= rhoL * eKinRL;
    R.E -= rhoR * eKinRR;
  }
  Cons<D> fluxL, fluxR;
  Flux<D>(fluxL, L, d);
  Flux<D>(fluxR, R, d);
  const double uL = L.mom[d] / rhoL;
  const double uR = R.mom[d] / rhoR;
  const double aL = L.a, aR = R.a;
  const double rhoEL = L.E, rhoER = R.E;
  const double pL = L.p, pR = R.p;
  const double hL = (rhoEL + pL) / rhoL;
  const double hR = (rhoER + pR) / rhoR;
  const double sqRhoL = std::sqrt(rhoL);
  const double sqRhoR = std::sqrt(rhoR);
  const double sqRhoSum = sqRhoL + sqRhoR;
  double vL[D], vR[D], roeU[D];
  for (int i = 0; i < D; ++i) {
    vL[i] = L.mom[i] / rhoL;
    vR[i] = R.mom[i] / rhoR;
    roeU[i] = (sqRhoL * vL[i] + sqRhoR * vR[i]) / sqRhoSum;
  }
  const double roeU0 = roeU[d];
  const double roeH = (sqRhoL * hL + sqRhoR * hR) / sqRhoSum;
  const double roeGamma = (sqRhoL * g.gamma + sqRhoR * g.gamma) / sqRhoSum;
  const double gm1 = roeGamma - 1.0;
  const double sqU = SquaredNormScalar<D>(roeU);
  const double roeA2 = gm1 * (roeH - 0.5 * Msq * sqU);
  const double roeA = std::sqrt(roeA2);
  const double maxA = std::max(aL, aR);
  const double sL1 = uL - maxA * Minv;
  const double sL2 = roeU0 - roeA * Minv;
  const double sR1 = roeU0 + roeA * Minv;
  const double sR2 = uR + maxA * Minv;
  const double sL = std::min(sL1, sL2);
  const double sR = std::max(sR1, sR2);
  const double bL = std::min(sL, 0.0);
  const double bR = std::max(sR, 0.0);
  const double db = bR - bL;
  const double db_positive = int(db <= 0) + int(db > 0) * db;
  Cons<D> w;
  for (int k = 0; k < D + 2; ++k) {
    C(w, k) = (bR * C(R, k) - bL * C(L, k) + C(fluxL, k) - C(fluxR, k)) / db_positive;
  }
  const double squaredNormRoeU = Msq * sqU;
  const double squaredNormRoeU_half = 0.5 * squaredNormRoeU;
  const double u_bar = 0.5 * (sR + sL);
  const double u_bar_abs = std::abs(u_bar);
  const double delta = roeA / (roeA + u_bar_abs);
  const double deltaRho = rhoR - rhoL;
  double deltaRhoU[D];
  for (int i = 0; i < D; ++i) deltaRhoU[i] = R.mom[i] - L.mom[i];
  const double deltaRhoE = rhoER - rhoEL;
  Cons<D> wm = w;
  const int upwind = (u_bar > 0.0);
  if constexpr (D == 2) {
    const int ix = d, iy = (ix + 1) % 2;
    const double l21 = gm1 / roeA2 * (roeH - Msq * sqU);
    const double l22 = gm1 / roeA2 * Msq * roeU[0];
    const double l23 = gm1 / roeA2 * roeU[1];
    const double l24 = gm1 / roeA2 * (-1.0);
    const double alpha_2 =
        l21 * deltaRho + l22 * deltaRhoU[0] + l23 * deltaRhoU[1] + l24 * deltaRhoE;
    const double l31 = -roeU[iy];
    const double l33 = 1;
    const double alpha_3 = l31 * deltaRho + l33 * deltaRhoU[iy];
    const double ubda2 = u_bar * delta * alpha_2;
    const double ubda3 = u_bar * delta * alpha_3;
    wm.rho = w.rho - ubda2 * 1.0;
    wm.mom[ix] = w.mom[ix] - ubda2 * roeU[ix];
    if (larrouturou) {
      wm.mom[iy] = wm.rho * (upwind * vL_[iy] + (1 - upwind) * vR_[iy]);
      wm.E = w.E - ubda2 * squaredNormRoeU_half +
             wm.rho * (upwind * eKinRL + (1 - upwind) * eKinRR);
    } else {
      wm.mom[iy] = w.mom[iy] - ubda2 * roeU[iy] - ubda3 * 1.0;
      wm.E = w.E - ubda2 * squaredNormRoeU_half - ubda3 * roeU[iy];
    }
  } else {
    const int ix = d, iy = (ix + 1) % 3, iz = (iy + 1) % 3;
    const double l21 = gm1 / roeA2 * (roeH - Msq * sqU);
    const double l22 = gm1 / roeA2 * Msq * roeU[0];
    const double l23 = gm1 / roeA2 * roeU[1];
    const double l24 = gm1 / roeA2 * roeU[2];
    const double l25 = gm1 / roeA2 * (-1.0);
    const double l31 = -roeU[iy];
    const double l33 = 1;
    const double l41 = -roeU[iz];
    const double l44 = 1;
    const double alpha_2 = l21 * deltaRho + l22 * deltaRhoU[0] + l23 * deltaRhoU[1] +
                           l24 * deltaRhoU[2] + l25 * deltaRhoE;
    const double alpha_3 = l31 * deltaRho + l33 * deltaRhoU[iy];
    const double alpha_4 = l41 * deltaRho + l44 * deltaRhoU[iz];
    const double ubd = u_bar * delta;
    wm.rho = w.rho - ubd * alpha_2;
    wm.mom[ix] = w.mom[ix] - ubd * alpha_2 * roeU[ix];
    if (larrouturou) {
      wm.mom[iy] = wm.rho * (upwind * vL_[iy] + (1 - upwind) * vR_[iy]);
      wm.mom[iz] = wm.rho * (upwind * vL_[iz] + (1 - upwind) * vR_[iz]);
      wm.E = w.E - ubd * alpha_2 * squaredNormRoeU_half +
             wm.rho * (upwind * eKinRL + (1 - upwind) * eKinRR);
    } else {
      wm.mom[iy] = w.mom[iy] - ubd * (alpha_2 * roeU[iy] + alpha_3 * 1.0);
      wm.mom[iz] = w.mom[iz] - ubd * (alpha_2 * roeU[iz] + alpha_4 * 1.0);
      wm.E = w.E - ubd * (alpha_2 * squaredNormRoeU_half + alpha_3 * roeU[iy] +
                          alpha_4 * roeU[iz]);
    }
  }
  if (0.0 < sL) {
    solution = L;
  } else if (sR < 0.0) {
    solution = R;
  } else {
    for (int k = 0; k < D + 2; ++k) C(solution, k) = C(wm, k);
    CompleteFromConsScalar<D>(g, solution);
  }
}

template <int D>
void BoundaryRiemann(const fvs_method_desc* desc, const Gas& g, State<D>& sol,
                     const State<D>& left, const State<D>& right) {
  if (desc->boundary_solver == FVS_BOUNDARY_HLLEM) {
    HllemRiemannState<D>(g, sol, left, right, 0, desc->base == FVS_BASE_HLLEM_LARROUTUROU);
  } else {
    const double tol = desc->newton_tolerance > 0 ? desc->newton_tolerance : 1e-7;
    const int max_it = desc->max_newton_iterations > 0 ? desc->max_newton_iterations : 5000;
    SolveRiemannProblem<D>(g, sol, left, right, 0, -1.0, tol, max_it);
  }
}

template <int D> void UnitX(double u[D]) {
  for (int i = 0; i < D; ++i) u[i] = (i == 0) ? 1.0 : 0.0;
}

// ---- KbnCutCellMethod ------------------------------------------------------------

template <int D>
int ReferenceStates(const fvs_method_desc* desc, const fvs_view* refs, const fvs_view* states,
                    const fvs_cutcell_data* geom) {
  const Gas g(desc->gamma);
  ViewAcc sa{states}, ra{refs};
  double unit[D];
  UnitX<D>(unit);
  for (int64_t k = states->lo[2]; k <= states->hi[2]; ++k)
    for (int64_t j = states->lo[1]; j <= states->hi[1]; ++j)
      for (int64_t i = states->lo[0]; i <= states->hi[0]; ++i) {
        const int64_t idx[3] = {i, j, k};
        if (!IsCutCell<D>(geom, idx)) continue;
        State<D> state;
        LoadState<D>(state, sa, idx);
        if (!(state.rho > 0.0)) continue;
        double normal[D];
        BoundaryNormal<D>(geom, idx, normal);
        RotateMomentum<D>(state, MakeRotation<D>(normal, unit));
        const State<D> reflected = Reflect<D>(state, unit);
        State<D> solution{};
        BoundaryRiemann<D>(desc, g, solution, reflected, state);
        RotateMomentum<D>(solution, MakeRotation<D>(unit, normal));
        for (int c = 0; c < D + 4; ++c) *ra.at(i, j, k, c) = C(solution, c);
      }
  return FVS_OK;
}

template <int D>
int BoundaryFluxes(const fvs_method_desc* desc, const fvs_view* bflux, const fvs_view* states,
                   const fvs_view* refs, const fvs_cutcell_data* geom, int dir) {
  const Gas g(desc->gamma);
  ViewAcc sa{states}, ra{refs}, fa{bflux};
  double unit[D];
  UnitX<D>(unit);
  for (int64_t k = refs->lo[2]; k <= refs->hi[2]; ++k)
    for (int64_t j = refs->lo[1]; j <= refs->hi[1]; ++j)
      for (int64_t i = refs->lo[0]; i <= refs->hi[0]; ++i) {
        const int64_t idx[3] = {i, j, k};
        if (!IsCutCell<D>(geom, idx)) continue;
        State<D> state, ref;
        LoadState<D>(state, sa, idx);
        LoadState<D>(ref, ra, idx);
        double normal[D];
        BoundaryNormal<D>(geom, idx, normal);
        Cons<D> flux;
        Flux<D>(flux, ref, dir);
        RotateMomentum<D>(state, MakeRotation<D>(normal, unit));
        const State<D> reflected = Reflect<D>(state, unit);
        State<D> solution{};
        BoundaryRiemann<D>(desc, g, solution, reflected, state);
        RotateMomentum<D>(solution, MakeRotation<D>(unit, normal));
        flux.mom[dir] += solution.p - ref.p;
        for (int c = 0; c < D + 2; ++c) *fa.at(i, j, k, c) = C(flux, c);
      }
  return FVS_OK;
}

// Masked MUSCL-Hancock + masked HLLEM for one face (MusclHancockMethod.hpp:487-571,
// HllemMethod.hpp:1074-1093), lane semantics of the SIMD path.
template <int D>
void MaskedFaceFlux(const Method& m, Cons<D>& flux, double beta, const State<D> st[4],
                    const double alpha[4], double dt, double dx, int dir) {
  const bool valid_face = beta > 0.0;
  if (!valid_face) {
    for (int c = 0; c < D + 2; ++c) C(flux, c) = 0.0;
    return;
  }
  if (m.recon != FVS_RECON_MUSCL_HANCOCK) {
    HllemFlux<D>(m.gas, flux, st[0], st[1], dir, m.base == FVS_BASE_HLLEM_LARROUTUROU, m.quirk3d);
    return;
  }
  const double lambda_half = 0.5 * dt / dx;
  State<D> rec[2];
  for (int side = 0; side < 2; ++side) {
    const State<D>* s3 = st + side;
    const double* a3 = alpha + side;
    const bool slope_mask = a3[0] > 0.0 && a3[1] > 0.0 && a3[2] > 0.0;
    State<D> qL{}, qR{};
    for (int kk = 0; kk < D + 2; ++kk) {
      const double slope =
          slope_mask ? LimitedSlope(m.limiter, C(s3[0], kk), C(s3[1], kk), C(s3[2], kk)) : 0.0;
      const double q0 = C(s3[1], kk);
      C(qL, kk) = q0 - 0.5 * slope;
      C(qR, kk) = q0 + 0.5 * slope;
    }
    CompleteFromConsMasked<D>(m.gas, qL, true);
    CompleteFromConsMasked<D>(m.gas, qR, true);
    Cons<D> fL, fR;
    FluxMasked<D>(fL, qL, true, dir);
    FluxMasked<D>(fR, qR, true, dir);
    State<D>& r = rec[side];
    const State<D>& base = side == 0 ? qR : qL;
    for (int kk = 0; kk < D + 2; ++kk) C(r, kk) = C(base, kk) + lambda_half * (C(fL, kk) - C(fR, kk));
    CompleteFromConsMasked<D>(m.gas, r, true);
  }
  HllemFlux<D>(m.gas, flux, rec[0], rec[1], dir, m.base == FVS_BASE_HLLEM_LARROUTUROU, m.quirk3d);
}

template <int D>
int RegularFluxes(const fvs_method_desc* desc, const fvs_view* fluxes, const fvs_view* states,
                  const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  Method m(desc);
  const int w = m.width();
  const int S = 2 * w;
  ViewAcc fa{fluxes}, sa{states};
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t k = fluxes->lo[2]; k <= fluxes->hi[2]; ++k)
    for (int64_t j = fluxes->lo[1]; j <= fluxes->hi[1]; ++j) {
      State<D> st[4];
      double alpha[4];
      Cons<D> f;
      for (int64_t i = fluxes->lo[0]; i <= fluxes->hi[0]; ++i) {
        int64_t cell[3] = {i, j, k};
        Shift(cell, dir, -w);
        for (int s = 0; s < S; ++s) {
          LoadState<D>(st[s], sa, cell);
          alpha[s] = At(geom->volume_fractions, cell[0], cell[1], cell[2]);
          Shift(cell, dir, 1);
        }
        const double beta = At(geom->face_fractions[dir], i, j, k);
        MaskedFaceFlux<D>(m, f, beta, st, alpha, dt, dx, dir);
        for (int c = 0; c < D + 2; ++c) *fa.at(i, j, k, c) = C(f, c);
      }
    }
  return FVS_OK;
}

template <int D>
int KbnStableDt(const fvs_method_desc* desc, const fvs_view* states,
                const fvs_cutcell_data* geom, double dx, int dir, double* out) {
  Method m(desc);
  const int w = m.width();
  ViewAcc sa{states};
  int64_t lo[3], hi[3];
  BoxOf(states, lo, hi);
  lo[dir] += w;        // fluxbox = Shrink(cellbox, dir, {w, w - 1})
  hi[dir] -= (w - 1);  // face f uses cells f - w .. f + w - 1
  double min_dt = std::numeric_limits<double>::infinity();
  for (int64_t k = lo[2]; k <= hi[2]; ++k)
    for (int64_t j = lo[1]; j <= hi[1]; ++j)
      for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        const double beta = At(geom->face_fractions[dir], i, j, k);
        double dt = std::numeric_limits<double>::max();
        if (beta > 0.0) {
          int64_t l[3] = {i, j, k}, r[3] = {i, j, k};
          Shift(l, dir, -1);
          State<D> L, R;
          LoadState<D>(L, sa, l);
          LoadState<D>(R, sa, r);
          dt = HllemStableDt<D>(m.gas, L, R, dx, dir);
        }
        min_dt = std::min(min_dt, dt);
      }
  *out = min_dt;
  return FVS_OK;
}

int CutCellFluxes(int dim, const fvs_view* stab, const fvs_view* fsl, const fvs_view* fsr,
                  const fvs_view* bflux, const fvs_view* regular, const fvs_cutcell_data* geom,
                  int dir) {
  const fvs_view& us = geom->unshielded[dir];
  const fvs_view& sl = geom->shielded_left[dir];
  const fvs_view& sr = geom->shielded_right[dir];
  const fvs_view& xb = geom->boundary_centroids;
  for (int c = 0; c < regular->ncomp; ++c)
    for (int64_t k = regular->lo[2]; k <= regular->hi[2]; ++k)
      for (int64_t j = regular->lo[1]; j <= regular->hi[1]; ++j)
        for (int64_t i = regular->lo[0]; i <= regular->hi[0]; ++i) {
          int64_t l[3] = {i, j, k};
          Shift(l, dir, -1);
          const double centerL = At(xb, l[0], l[1], l[2], dir);
          const double centerR = At(xb, i, j, k, dir);
          const double dL = std::clamp(0.5 - centerL, 0.0, 1.0);
          const double dR = std::clamp(0.5 + centerR, 0.0, 1.0);
          const double f = At(*regular, i, j, k, c);
          const double fbL = At(*bflux, l[0], l[1], l[2], c);
          const double fbR = At(*bflux, i, j, k, c);
          const double sL = dL * f + (1.0 - dL) * fbL;
          const double sR = dR * f + (1.0 - dR) * fbR;
          const double betaUS = At(us, i, j, k, 1);
          const double betaL = At(sl, i, j, k, 1);
          const double betaR = At(sr, i, j, k, 1);
          *Ptr(*fsl, i, j, k, c) = sL;
          *Ptr(*fsr, i, j, k, c) = sR;
          *Ptr(*stab, i, j, k, c) = betaUS * f + betaL * sL + betaR * sR;
        }
  (void)dim;
  return FVS_OK;
}

int Update(const fvs_view* scratch, const fvs_view* fus, const fvs_view* fst,
           const fvs_view* fsl, const fvs_view* fsr, const fvs_view* fb,
           const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  const double dt_over_dx = dt / dx;
  const fvs_view& beta = geom->face_fractions[dir];
  const fvs_view& us = geom->unshielded[dir];
  const fvs_view& sl = geom->shielded_left[dir];
  const fvs_view& sr = geom->shielded_right[dir];
  for (int c = 0; c < scratch->ncomp; ++c)
    for (int64_t k = scratch->lo[2]; k <= scratch->hi[2]; ++k)
      for (int64_t j = scratch->lo[1]; j <= scratch->hi[1]; ++j)
        for (int64_t i = scratch->lo[0]; i <= scratch->hi[0]; ++i) {
          int64_t r[3] = {i, j, k};
          Shift(r, dir, 1);
          double* next = Ptr(*scratch, i, j, k, c);
          const double alpha = At(geom->volume_fractions, i, j, k);
          if (alpha == 0.0) {
            *next = 0.0;
            continue;
          }
          const double prev = *next;
          const double betaL = At(beta, i, j, k);
          const double betaR = At(beta, r[0], r[1], r[2]);
          const double fLs = At(*fst, i, j, k, c);
          const double fRs = At(*fst, r[0], r[1], r[2], c);
          const double regular = prev + dt_over_dx * (fLs - fRs);
          const double fB = At(*fb, i, j, k, c);
          const double center = At(geom->boundary_centroids, i, j, k, dir);
          // beta_L > beta_R
          const double betaL_us_over_alpha = At(us, i, j, k, 0) / alpha;
          const double betaL_sL_over_alpha = At(sl, i, j, k, 0) / alpha;
          const double dLg = std::clamp((At(sr, i, j, k, 0) * (0.5 + center)) / alpha, 0.0, 1.0);
          const double fLus = At(*fus, i, j, k, c);
          const double fLsL = At(*fsl, i, j, k, c);
          const double left_greater = prev + dt_over_dx * betaL_us_over_alpha * (fLus - fRs) +
                                      dt_over_dx * betaL_sL_over_alpha * (fLsL - fRs) +
                                      dt_over_dx * dLg * (fLus - fB);
          // beta_L < beta_R
          const double betaR_us_over_alpha = At(us, r[0], r[1], r[2], 0) / alpha;
          const double betaR_sR_over_alpha = At(sr, r[0], r[1], r[2], 0) / alpha;
          const double dRg =
              std::clamp((At(sl, r[0], r[1], r[2], 0) *
                          (Planted(PLANT_HBOX_RIGHT_DISTANCE) ? 0.5 + center : 0.5 - center)) /
                             alpha,
                         0.0, 1.0);
          const double fRus = At(*fus, r[0], r[1], r[2], c);
          const double fRsR = At(*fsr, r[0], r[1], r[2], c);
          const double right_greater = prev + dt_over_dx * betaR_us_over_alpha * (fLs - fRus) +
                                       dt_over_dx * betaR_sR_over_alpha * (fLs - fRsR) +
                                       dt_over_dx * dRg *
                                           (Planted(PLANT_HBOX_RIGHT_DB_SIGN) ? fRus - fB
                                                                              : fB - fRus);
          *next = betaL == betaR ? regular : betaL > betaR ? left_greater : right_greater;
        }
  return FVS_OK;
}

int Update2(const fvs_view* dest, const fvs_view* src, const fvs_view* fst, const fvs_view* fb,
            const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  const double dt_over_dx = dt / dx;
  const fvs_view& beta = geom->face_fractions[dir];
  for (int c = 0; c < dest->ncomp; ++c)
    for (int64_t k = dest->lo[2]; k <= dest->hi[2]; ++k)
      for (int64_t j = dest->lo[1]; j <= dest->hi[1]; ++j)
        for (int64_t i = dest->lo[0]; i <= dest->hi[0]; ++i) {
          const double volfrac = At(geom->volume_fractions, i, j, k);
          if (!(volfrac > 0.0)) continue;
          int64_t r[3] = {i, j, k};
          Shift(r, dir, 1);
          const double bL = At(beta, i, j, k);
          const double bR = At(beta, r[0], r[1], r[2]);
          const double fL = At(*fst, i, j, k, c);
          const double fR = At(*fst, r[0], r[1], r[2], c);
          const double prev = At(*src, i, j, k, c);
          double next;
          if (bL == bR) {
            next = prev + dt_over_dx * (fL - fR);
          } else {
            const double dB = bL - bR;
            const double fB = At(*fb, i, j, k, c);
            next = prev + dt_over_dx / volfrac * (bL * fL - bR * fR - dB * fB);
          }
          *Ptr(*dest, i, j, k, c) = next;
        }
  return FVS_OK;
}

template <int D>
int CompleteMasked(double gamma, const fvs_view* complete, const fvs_view* cons,
                   const fvs_cutcell_data* geom) {
  const Gas g(gamma);
  for (int64_t k = complete->lo[2]; k <= complete->hi[2]; ++k)
    for (int64_t j = complete->lo[1]; j <= complete->hi[1]; ++j)
      for (int64_t i = complete->lo[0]; i <= complete->hi[0]; ++i) {
        State<D> q{};
        for (int c = 0; c < D + 2; ++c) C(q, c) = At(*cons, i, j, k, c);
        const bool mask = At(geom->volume_fractions, i, j, k) > 0.0;
        CompleteFromConsMasked<D>(g, q, mask);
        for (int c = 0; c < D + 4; ++c) *Ptr(*complete, i, j, k, c) = C(q, c);
      }
  return FVS_OK;
}

// FillCutCellData__ scalar tail (FillCutCellData.cpp:70-93); the Vc packet
// loop computes the same lane values.
void FillFace(double betaL, double betaM, double betaR, double out[8]) {
  const double dBetaL = std::max(0.0, betaM - betaL);
  const double dBetaR = std::max(0.0, betaM - betaR);
  const double unshielded = std::max(0.0, betaM - std::max(dBetaL, dBetaR));
  const double doubly = std::min(dBetaL, dBetaR);
  const double sl = dBetaL - doubly;
  const double sr = dBetaR - doubly;
  betaM = betaM > 0 ? betaM : 1.0;
  out[0] = unshielded;
  out[1] = unshielded / betaM;
  out[2] = sl;
  out[3] = sl / betaM;
  out[4] = sr;
  out[5] = sr / betaM;
  out[6] = doubly;
  out[7] = doubly / betaM;
}

int FillCutCellData(const fvs_cutcell_data* geom, int dir, const int64_t lo[3],
                    const int64_t hi[3]) {
  const fvs_view& beta = geom->face_fractions[dir];
  for (int64_t k = lo[2]; k <= hi[2]; ++k)
    for (int64_t j = lo[1]; j <= hi[1]; ++j)
      for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        int64_t l[3] = {i, j, k}, r[3] = {i, j, k};
        Shift(l, dir, -1);
        Shift(r, dir, 1);
        double out[8];
        FillFace(At(beta, l[0], l[1], l[2]), At(beta, i, j, k), At(beta, r[0], r[1], r[2]), out);
        *Ptr(geom->unshielded[dir], i, j, k, 0) = out[0];
        *Ptr(geom->unshielded[dir], i, j, k, 1) = out[1];
        *Ptr(geom->shielded_left[dir], i, j, k, 0) = out[2];
        *Ptr(geom->shielded_left[dir], i, j, k, 1) = out[3];
        *Ptr(geom->shielded_right[dir], i, j, k, 0) = out[4];
        *Ptr(geom->shielded_right[dir], i, j, k, 1) = out[5];
        *Ptr(geom->doubly_shielded[dir], i, j, k, 0) = out[6];
        *Ptr(geom->doubly_shielded[dir], i, j, k, 1) = out[7];
      }
  return FVS_OK;
}

// ---- single-patch cut-cell level (2D) -----------------------------------------------
//
// The sequence of DimensionalSplitLevelIntegrator::AdvanceLevelNonRecursively
// (include/fub/solver/DimensionalSplitLevelIntegrator.hpp:221-276) with the
// cut-cell IntegratorContext: reference states at every level advance
// (src/AMReX/cutcell/IntegratorContext.cpp:800-806), then per split step the
// boundary fill of the direction, the KBN CFL check, boundary, regular and
// stabilised fluxes (include/fub/AMReX/cutcell/FluxMethod.hpp:153-257), the
// cut-cell update and the masked reconstruction.  Two ghost layers,
// transmissive boundaries; the geometry covers the grown box.
struct KbnLevel {
  static constexpr int D = 2;
  static constexpr int G = 2;
  const fvs_method_desc* desc;
  fvs_cutcell_data geom;
  int integrator;
  int64_t n[2];
  int64_t ng[2];
  double dx[2];
  std::vector<double> scratch, refs, fb, flux[4];

  fvs_view Grid(std::vector<double>& buf, int ncomp, const int64_t lo[3], const int64_t hi[3]) {
    fvs_view v{};
    v.ncomp = ncomp;
    v.stride[0] = 1;
    v.stride[1] = ng[0];
    v.stride[2] = ng[0] * ng[1];
    v.comp_stride = ng[0] * ng[1];
    int64_t off = 0;
    for (int d = 0; d < 3; ++d) {
      v.lo[d] = lo[d];
      v.hi[d] = hi[d];
      off += (lo[d] + (d < D ? G : 0)) * v.stride[d];
    }
    v.data = buf.data() + off;
    return v;
  }
  void Interior(int64_t lo[3], int64_t hi[3]) const {
    lo[0] = lo[1] = lo[2] = 0;
    hi[0] = n[0] - 1;
    hi[1] = n[1] - 1;
    hi[2] = 0;
  }
  fvs_view FaceView(std::vector<double>& buf, int dir) {
    fvs_view v{};
    int64_t lo[3], hi[3];
    Interior(lo, hi);
    hi[dir] += 1;
    v.ncomp = D + 2;
    v.stride[0] = 1;
    v.stride[1] = hi[0] - lo[0] + 1;
    v.stride[2] = v.stride[1] * (hi[1] - lo[1] + 1);
    v.comp_stride = v.stride[2];
    for (int d = 0; d < 3; ++d) {
      v.lo[d] = lo[d];
      v.hi[d] = hi[d];
    }
    v.data = buf.data();
    return v;
  }

  KbnLevel(const fvs_method_desc* d, const fvs_patch_desc* p, const fvs_cutcell_data* g, int integ)
      : desc(d), geom(*g), integrator(integ) {
    for (int i = 0; i < 2; ++i) {
      n[i] = p->n[i];
      ng[i] = n[i] + 2 * G;
      dx[i] = p->dx[i];
    }
    scratch.assign(ng[0] * ng[1] * (D + 4), 0.0);
    refs.assign(ng[0] * ng[1] * (D + 4), 0.0);
    fb.assign(ng[0] * ng[1] * (D + 2), 0.0);
    for (auto& f : flux) f.assign((n[0] + 1) * (n[1] + 1) * (D + 2), 0.0);
    // shielded fractions of every face with both neighbours in the grown box
    for (int dir = 0; dir < D; ++dir) {
      int64_t lo[3] = {-G, -G, 0}, hi[3] = {n[0] - 1 + G, n[1] - 1 + G, 0};
      hi[dir] += 1;
      lo[dir] += 1;
      hi[dir] -= 1;
      FillCutCellData(&geom, dir, lo, hi);
    }
  }

  void FillGhosts(int dir) {
    int64_t glo[3] = {-G, -G, 0}, ghi[3] = {n[0] - 1 + G, n[1] - 1 + G, 0};
    int64_t ilo[3], ihi[3];
    Interior(ilo, ihi);
    fvs_view v = Grid(scratch, D + 4, glo, ghi);
    TransmissiveFill(&v, ilo, ihi, dir);
  }

  double DirStableDt(int dir) {
    FillGhosts(dir);
    int64_t lo[3], hi[3];
    Interior(lo, hi);
    const int w = desc->reconstruction == FVS_RECON_MUSCL_HANCOCK ? 2 : 1;
    lo[dir] -= w;
    hi[dir] += w;
    fvs_view s = Grid(scratch, D + 4, lo, hi);
    double dt;
    KbnStableDt<D>(desc, &s, &geom, dx[dir], dir, &dt);
    return dt;
  }

  void ReferenceStates() {
    int64_t lo[3], hi[3];
    Interior(lo, hi);
    fvs_view rv = Grid(refs, D + 4, lo, hi);
    fvs_view sv = Grid(scratch, D + 4, lo, hi);
    ::ReferenceStates<D>(desc, &rv, &sv, &geom);
  }

  int SplitStep(int dir, double split_dt, double* level_dt) {
    const double local_dt = DirStableDt(dir);
    *level_dt = local_dt;
    if (local_dt < split_dt) return FVS_TIME_STEP_TOO_LARGE;
    int64_t glo[3] = {-G, -G, 0}, ghi[3] = {n[0] - 1 + G, n[1] - 1 + G, 0};
    int64_t ilo[3], ihi[3];
    Interior(ilo, ihi);
    std::fill(fb.begin(), fb.end(), 0.0);
    fvs_view fbv = Grid(fb, D + 2, glo, ghi);
    fvs_view sg = Grid(scratch, D + 4, glo, ghi);
    fvs_view ri = Grid(refs, D + 4, ilo, ihi);
    BoundaryFluxes<D>(desc, &fbv, &sg, &ri, &geom, dir);
    fvs_view F = FaceView(flux[0], dir), Fs = FaceView(flux[1], dir);
    fvs_view FsL = FaceView(flux[2], dir), FsR = FaceView(flux[3], dir);
    RegularFluxes<D>(desc, &F, &sg, &geom, split_dt, dx[dir], dir);
    CutCellFluxes(D, &Fs, &FsL, &FsR, &fbv, &F, &geom, dir);
    fvs_view ui = Grid(scratch, D + 2, ilo, ihi);
    if (integrator == 0) {
      Update(&ui, &F, &Fs, &FsL, &FsR, &fbv, &geom, split_dt, dx[dir], dir);
    } else {
      Update2(&ui, &ui, &Fs, &fbv, &geom, split_dt, dx[dir], dir);
    }
    fvs_view ci = Grid(scratch, D + 4, ilo, ihi);
    CompleteMasked<D>(desc->gamma, &ci, &ui, &geom);
    return FVS_OK;
  }
};

int KbnLevelRun(const fvs_method_desc* d, const fvs_patch_desc* p, const fvs_cutcell_data* g,
                int integrator, double* cons, double t0, int max_steps, double final_time,
                double* dt_log, int* steps_done, int* failures) {
  KbnLevel L(d, p, g, integrator);
  constexpr int D = 2;
  const int64_t ncell = L.n[0] * L.n[1];
  {
    int64_t ilo[3], ihi[3];
    L.Interior(ilo, ihi);
    fvs_view ci = L.Grid(L.scratch, D + 4, ilo, ihi);
    fvs_view uv{};
    uv.data = cons;
    uv.ncomp = D + 2;
    uv.stride[0] = 1;
    uv.stride[1] = L.n[0];
    uv.stride[2] = ncell;
    uv.comp_stride = ncell;
    for (int i = 0; i < 3; ++i) {
      uv.lo[i] = ilo[i];
      uv.hi[i] = ihi[i];
    }
    CompleteMasked<D>(d->gamma, &ci, &uv, &L.geom);
  }
  double t = t0;
  int64_t cycle = 0;
  int nfail = 0, consecutive = 0;
  bool have_failure = false;
  double failure_dt = 0.0;
  std::vector<double> backup;
  int rc = FVS_OK;
  // RunOptions::smallest_time_step_size (RunSimulation.hpp:61-63, 97, 109)
  const double eps = p->smallest_dt > 0.0 ? p->smallest_dt : 1e-12;
  while (t + eps < final_time && cycle < max_steps) {
    backup = L.scratch;
    double stable = std::numeric_limits<double>::max();
    if (have_failure) {
      stable = failure_dt;
    } else {
      for (int dir = 0; dir < D; ++dir) stable = std::min(stable, L.DirStableDt(dir));
    }
    const double dt = std::min(final_time - t, p->cfl * stable);
    L.ReferenceStates();
    const int is_odd = static_cast<int>(cycle % 2);
    int dirs[2] = {is_odd ? 1 : 0, is_odd ? 0 : 1};
    std::vector<std::pair<int, double>> seq;
    if (p->splitting == FVS_SPLIT_STRANG) {
      seq = {{dirs[0], 0.5 * dt}, {dirs[1], 0.5 * dt}, {dirs[1], 0.5 * dt}, {dirs[0], 0.5 * dt}};
    } else {
      seq = {{dirs[0], dt}, {dirs[1], dt}};
    }
    bool failed = false;
    for (auto& st : seq) {
      double level_dt;
      const int r = L.SplitStep(st.first, st.second, &level_dt);
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
      if (++consecutive > 64) {
        rc = FVS_TIME_STEP_TOO_LARGE;
        break;
      }
      continue;
    }
    consecutive = 0;
    have_failure = false;
    if (dt_log) dt_log[cycle] = dt;
    t += dt;
    ++cycle;
  }
  {
    int64_t ilo[3], ihi[3];
    L.Interior(ilo, ihi);
    fvs_view ci = L.Grid(L.scratch, D + 4, ilo, ihi);
    for (int c = 0; c < D + 2; ++c)
      for (int64_t j = 0; j < L.n[1]; ++j)
        for (int64_t i = 0; i < L.n[0]; ++i) cons[c * ncell + j * L.n[0] + i] = At(ci, i, j, 0, c);
  }
  if (steps_done) *steps_done = static_cast<int>(cycle);
  if (failures) *failures = nfail;
  return rc;
}

} // namespace

extern "C" {

int oracle_kbn_level_run(const fvs_method_desc* desc, const fvs_patch_desc* patch,
                         const fvs_cutcell_data* geom, int integrator, double* cons, double t0,
                         int max_steps, double final_time, double* dt_log, int* steps_done,
                         int* failures) {
  if (!desc || desc->dim != 2) return FVS_BAD_ARGUMENT;
  return KbnLevelRun(desc, patch, geom, integrator, cons, t0, max_steps, final_time, dt_log,
                     steps_done, failures);
}

int oracle_kbn_fill_cutcell_data(const fvs_cutcell_data* geom, int dir, const int64_t lo[3],
                                 const int64_t hi[3]) {
  return FillCutCellData(geom, dir, lo, hi);
}

int oracle_kbn_is_cut_cell(int dim, const fvs_cutcell_data* geom, const int64_t lo[3],
                           const int64_t hi[3], int32_t* flags) {
  int64_t n = 0;
  for (int64_t k = lo[2]; k <= hi[2]; ++k)
    for (int64_t j = lo[1]; j <= hi[1]; ++j)
      for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        const int64_t idx[3] = {i, j, k};
        flags[n++] = dim == 2 ? IsCutCell<2>(geom, idx) : IsCutCell<3>(geom, idx);
      }
  return FVS_OK;
}

int oracle_kbn_reference_states(const fvs_method_desc* desc, const fvs_view* refs,
                                const fvs_view* states, const fvs_cutcell_data* geom) {
  return desc->dim == 2 ? ReferenceStates<2>(desc, refs, states, geom)
                        : ReferenceStates<3>(desc, refs, states, geom);
}

int oracle_kbn_boundary_fluxes(const fvs_method_desc* desc, const fvs_view* bflux,
                               const fvs_view* states, const fvs_view* refs,
                               const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  (void)dt;
  (void)dx;
  return desc->dim == 2 ? BoundaryFluxes<2>(desc, bflux, states, refs, geom, dir)
                        : BoundaryFluxes<3>(desc, bflux, states, refs, geom, dir);
}

int oracle_kbn_regular_fluxes(const fvs_method_desc* desc, const fvs_view* fluxes,
                              const fvs_view* states, const fvs_cutcell_data* geom, double dt,
                              double dx, int dir) {
  return desc->dim == 2 ? RegularFluxes<2>(desc, fluxes, states, geom, dt, dx, dir)
                        : RegularFluxes<3>(desc, fluxes, states, geom, dt, dx, dir);
}

int oracle_kbn_stable_dt(const fvs_method_desc* desc, const fvs_view* states,
                         const fvs_cutcell_data* geom, double dx, int dir, double* out) {
  return desc->dim == 2 ? KbnStableDt<2>(desc, states, geom, dx, dir, out)
                        : KbnStableDt<3>(desc, states, geom, dx, dir, out);
}

int oracle_kbn_cutcell_fluxes(int dim, const fvs_view* stab, const fvs_view* fsl,
                              const fvs_view* fsr, const fvs_view* bflux,
                              const fvs_view* regular, const fvs_cutcell_data* geom, int dir) {
  return CutCellFluxes(dim, stab, fsl, fsr, bflux, regular, geom, dir);
}

int oracle_kbn_update(const fvs_view* scratch, const fvs_view* fus, const fvs_view* fst,
                      const fvs_view* fsl, const fvs_view* fsr, const fvs_view* fb,
                      const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  return Update(scratch, fus, fst, fsl, fsr, fb, geom, dt, dx, dir);
}

int oracle_kbn_update2(const fvs_view* dest, const fvs_view* src, const fvs_view* fst,
                       const fvs_view* fb, const fvs_cutcell_data* geom, double dt, double dx,
                       int dir) {
  return Update2(dest, src, fst, fb, geom, dt, dx, dir);
}

int oracle_kbn_complete_from_cons(int dim, double gamma, const fvs_view* complete,
                                  const fvs_view* cons, const fvs_cutcell_data* geom) {
  return dim == 2 ? CompleteMasked<2>(gamma, complete, cons, geom)
                  : CompleteMasked<3>(gamma, complete, cons, geom);
}

// Pointwise helpers for pinning tests.
int oracle_kbn_rotate(int dim, const double* from, const double* to, const double* mom,
                      double* out) {
  if (dim == 2) {
    const Rot<2> r = MakeRotation<2>(from, to);
    State<2> q{};
    q.mom[0] = mom[0];
    q.mom[1] = mom[1];
    RotateMomentum<2>(q, r);
    out[0] = q.mom[0];
    out[1] = q.mom[1];
  } else {
    const Rot<3> r = MakeRotation<3>(from, to);
    State<3> q{};
    for (int i = 0; i < 3; ++i) q.mom[i] = mom[i];
    RotateMomentum<3>(q, r);
    for (int i = 0; i < 3; ++i) out[i] = q.mom[i];
  }
  return FVS_OK;
}

int oracle_hllem_riemann_state(int dim, double gamma, int dir, int larrouturou, int64_t n,
                               const double* left, const double* right, double* out) {
  const Gas g(gamma);
  const int nc = dim + 4;
  for (int64_t i = 0; i < n; ++i) {
    if (dim == 2) {
      State<2> L, R, S{};
      for (int c = 0; c < nc; ++c) {
        C(L, c) = left[c * n + i];
        C(R, c) = right[c * n + i];
      }
      HllemRiemannState<2>(g, S, L, R, dir, larrouturou != 0);
      for (int c = 0; c < nc; ++c) out[c * n + i] = C(S, c);
    } else {
      State<3> L, R, S{};
      for (int c = 0; c < nc; ++c) {
        C(L, c) = left[c * n + i];
        C(R, c) = right[c * n + i];
      }
      HllemRiemannState<3>(g, S, L, R, dir, larrouturou != 0);
      for (int c = 0; c < nc; ++c) out[c * n + i] = C(S, c);
    }
  }
  return FVS_OK;
}

} // extern "C"
