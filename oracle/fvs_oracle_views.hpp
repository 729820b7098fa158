// TEST INFRASTRUCTURE ONLY: view access and per-face helpers shared by the
// oracle's translation units (fvs_oracle.cpp, fvs_kbn_oracle.cpp).
#pragma once

#include "fvs_oracle.hpp"

#include "../include/fvs.h"

namespace fvs_oracle {
struct ViewAcc {
  const fvs_view* v;
  double* at(int64_t i, int64_t j, int64_t k, int c) const {
    return v->data + (i - v->lo[0]) * v->stride[0] + (j - v->lo[1]) * v->stride[1] +
           (k - v->lo[2]) * v->stride[2] + c * v->comp_stride;
  }
};

inline void Shift(int64_t idx[3], int dir, int64_t by) { idx[dir] += by; }

template <int D>
void LoadState(State<D>& q, const ViewAcc& a, const int64_t idx[3]) {
  for (int c = 0; c < D + 4; ++c) C(q, c) = *a.at(idx[0], idx[1], idx[2], c);
}

struct Method {
  Gas gas;
  int base;
  int recon;
  int limiter;
  bool quirk3d;
  double tol;
  int max_it;
  bool seq; // sequential (scalar) semantics: required for exact Godunov
  explicit Method(const fvs_method_desc* d)
      : gas(d->gamma), base(d->base), recon(d->reconstruction),
        limiter(d->limiter), quirk3d(d->quirks != 0),
        tol(d->newton_tolerance > 0 ? d->newton_tolerance : 1e-7),
        max_it(d->max_newton_iterations > 0 ? d->max_newton_iterations : 5000),
        seq(d->base == FVS_BASE_GODUNOV_EXACT) {}
  int width() const { return recon != FVS_RECON_NONE ? 2 : 1; }
};

// Base two-state flux (Hllem or Hll array path, or Godunov scalar path).
template <int D>
void BaseFlux(const Method& m, Cons<D>& f, const State<D>& L, const State<D>& R,
              int dir, int* newton_fail) {
  if (m.base == FVS_BASE_GODUNOV_EXACT) {
    State<D> sol{};
    MiddleState ms = SolveRiemannProblem<D>(m.gas, sol, L, R, dir, -1.0, m.tol, m.max_it);
    if (!ms.converged && newton_fail) *newton_fail = 1;
    Flux<D>(f, sol, dir);
  } else if (m.base == FVS_BASE_HLL_EINFELDT) {
    HllFlux<D>(m.gas, f, L, R, dir);
  } else {
    HllemFlux<D>(m.gas, f, L, R, dir, m.base == FVS_BASE_HLLEM_LARROUTUROU, m.quirk3d);
  }
}

template <int D>
double BaseStableDt(const Method& m, const State<D>& L, const State<D>& R, double dx,
                    int dir) {
  if (m.base == FVS_BASE_GODUNOV_EXACT) {
    return GodunovStableDt<D>(m.gas, L, R, dx, dir, m.tol, m.max_it);
  }
  if (m.base == FVS_BASE_HLL_EINFELDT) return HllStableDt<D>(m.gas, L, R, dx, dir);
  return HllemStableDt<D>(m.gas, L, R, dx, dir);
}

// MusclHancock::ComputeNumericFlux (MusclHancockMethod.hpp:294-356 scalar,
// :419-485 array).  The only difference between the two on this path is the
// kinetic-energy reduction order inside CompleteFromCons (3D).
template <int D>
void MusclHancockFlux(const Method& m, Cons<D>& flux, const State<D> st[4],
                      double dt, double dx, int dir, int* newton_fail) {
  const double lambda_half = Planted(PLANT_MUSCL_HALF_STEP_SIGN)   ? -(0.5 * dt / dx)
                             : Planted(PLANT_MUSCL_HALF_STEP_DROP) ? 0.0
                                                                   : 0.5 * dt / dx;
  State<D> rec[2];
  for (int side = 0; side < 2; ++side) {
    const State<D>* s3 = st + side; // first<3>() or last<3>()
    State<D> qL{}, qR{};
    for (int k = 0; k < D + 2; ++k) {
      const double slope = LimitedSlope(m.limiter, C(s3[0], k), C(s3[1], k), C(s3[2], k));
      C(qL, k) = C(s3[1], k) - 0.5 * slope;
      C(qR, k) = C(s3[1], k) + 0.5 * slope;
    }
    if (m.seq) {
      CompleteFromConsScalar<D>(m.gas, qL);
      CompleteFromConsScalar<D>(m.gas, qR);
    } else {
      CompleteFromConsArray<D>(m.gas, qL);
      CompleteFromConsArray<D>(m.gas, qR);
    }
    Cons<D> fL, fR;
    Flux<D>(fL, qL, dir);
    Flux<D>(fR, qR, dir);
    State<D>& r = rec[side];
    const State<D>& base = side == 0 ? qR : qL;
    for (int k = 0; k < D + 2; ++k) {
      C(r, k) = C(base, k) + lambda_half * (C(fL, k) - C(fR, k));
    }
    if (m.seq) {
      CompleteFromConsScalar<D>(m.gas, r);
    } else {
      CompleteFromConsArray<D>(m.gas, r);
    }
  }
  BaseFlux<D>(m, flux, rec[0], rec[1], dir, newton_fail);
}

template <int D>
void FaceFlux(const Method& m, Cons<D>& flux, const State<D>* st, double dt,
              double dx, int dir, int* newton_fail) {
  if (m.recon == FVS_RECON_MUSCL_HANCOCK_CHARACTERISTIC) {
    State<D> w[2];
    const double lambda = dt / dx;
    CharacteristicFaceState<D>(m.gas, st, lambda, 0, dir, w[0]);
    CharacteristicFaceState<D>(m.gas, st + 1, lambda, 1, dir, w[1]);
    HllemFlux<D>(m.gas, flux, w[0], w[1], dir, true, m.quirk3d);
  } else if (m.recon == FVS_RECON_MUSCL_HANCOCK_PRIM) {
    State<D> w[2];
    MusclHancockPrimStates<D>(m.gas, st, dt, dx, dir, w);
    HllemFlux<D>(m.gas, flux, w[0], w[1], dir, true, m.quirk3d);
  } else if (m.recon == FVS_RECON_MUSCL_HANCOCK) {
    MusclHancockFlux<D>(m, flux, st, dt, dx, dir, newton_fail);
  } else {
    BaseFlux<D>(m, flux, st[0], st[1], dir, newton_fail);
  }
}

template <int D>
double FaceStableDt(const Method& m, const State<D>* st, double dx, int dir) {
  // MusclHancock::ComputeStableDt uses the middle two cells (MusclHancockMethod.hpp:214-218).
  // MusclHancockPrim::ComputeStableDt: Hllem on the middle two
  // (src/equations/perfect_gas/MusclHancockPrimMethod.hpp:40-46).
  if (m.recon != FVS_RECON_NONE) {
    return BaseStableDt<D>(m, st[1], st[2], dx, dir);
  }
  return BaseStableDt<D>(m, st[0], st[1], dx, dir);
}

inline void BoxOf(const fvs_view* v, int64_t lo[3], int64_t hi[3]) {
  for (int d = 0; d < 3; ++d) {
    lo[d] = v->lo[d];
    hi[d] = v->hi[d];
  }
}

// TransmissiveBoundary::FillBoundary (src/AMReX/boundary_condition/TransmissiveBoundary.cpp:31-93):
// every cell of v outside [dlo, dhi] along dir copies the boundary cell of its row.
inline int TransmissiveFill(const fvs_view* v, const int64_t dlo[3], const int64_t dhi[3],
                     int dir) {
  ViewAcc a{v};
  int64_t lo[3], hi[3];
  BoxOf(v, lo, hi);
  for (int64_t k = lo[2]; k <= hi[2]; ++k)
    for (int64_t j = lo[1]; j <= hi[1]; ++j)
      for (int64_t i = lo[0]; i <= hi[0]; ++i) {
        int64_t idx[3] = {i, j, k};
        int64_t src[3] = {i, j, k};
        if (idx[dir] < dlo[dir]) {
          src[dir] = dlo[dir];
        } else if (idx[dir] > dhi[dir]) {
          src[dir] = dhi[dir];
        } else {
          continue;
        }
        for (int c = 0; c < v->ncomp; ++c)
          *a.at(i, j, k, c) = *a.at(src[0], src[1], src[2], c);
      }
  return FVS_OK;
}

} // namespace fvs_oracle
