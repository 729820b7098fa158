// KBN h-box cut-cell method on strided views: geometry preprocessing, the
// boundary Riemann problems, masked regular fluxes and CFL, the stabilised
// flux components and the cut-cell time integrators.
//
//   FillCutCellData__              src/AMReX/cutcell/FillCutCellData.cpp:28-118
//   IsCutCell / GetBoundaryNormal  src/CutCellData.cpp:90-105,225-256
//   MakeRotation, Rotate, Reflect  include/fub/ext/Eigen.hpp:104-130,
//                                  src/equations/PerfectGas.cpp:230-300
//   Hllem::SolveRiemannProblem     src/equations/perfect_gas/HllemMethod.hpp:25-369
//   KbnCutCellMethod               include/fub/cutcell_method/KbnStabilisation.hpp:188-429
//   ComputeStableFluxComponents    src/cutcell_method/KbnStabilisation.cpp:55-184
//   cutcell::TimeIntegrator(2)     src/AMReX/cutcell/TimeIntegrator.cpp:65-508
//   cutcell::Reconstruction        include/fub/AMReX/cutcell/Reconstruction.hpp:123-157
//
// One thread per face or cell; the boundary problems run one thread per cell
// and return at once for cells that are not cut.  All arithmetic is FP64
// with the reference's operation order (the divisions by small volume
// fractions in particular), except in the FP32-flux form of the fused level
// sweeps: there the masked MUSCL-Hancock states and the HLLEM fluxes are
// float, while the CFL values, the boundary fluxes, the stabilised fluxes,
// the h-box update with its divisions by the volume fraction and the
// accumulation into the state stay FP64.
#include "fvs_internal.hpp"
#include "fvs_lanes.hpp"
#include "fvs_physics.hpp"
#include "fvs_riemann.hpp"
#include "fvs_views.hpp"

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <new>
#include <vector>

namespace fvs {

namespace {

constexpr int kBlock = 256;

struct GeomViews {
  DevView alpha;
  DevView beta[3];
  DevView normals;
  DevView centroids;
};

__device__ __forceinline__ double Val(const DevView& v, const int64_t* c, int k = 0) {
  return *v.at(c, k);
}

template <int D>
__device__ __forceinline__ bool IsCutCell(const GeomViews& g, const int64_t* c) {
  const double alpha = Val(g.alpha, c);
  if (alpha <= 0.0) return false;
  if (alpha < 1.0) return true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    int64_t r[3] = {c[0], c[1], c[2]};
    r[d] += 1;
    if (Val(g.beta[d], c) < 1.0 || Val(g.beta[d], r) < 1.0) return true;
  }
  return false;
}

// MakeRotation(a, b) applied to a momentum vector in place.
template <int D>
__device__ __forceinline__ void RotateMomentum(double* m, const double* a, const double* b) {
  bool opposite = true;
#pragma unroll
  for (int i = 0; i < D; ++i) opposite = opposite && (a[i] == -b[i]);
  double r[D][D];
  if (opposite) {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) r[i][j] = (i == j) ? -1.0 : -0.0;
  } else if constexpr (D == 2) {
    const double s = a[0] * b[1] - a[1] * b[0];
    const double c = a[0] * b[0] + a[1] * b[1];
    r[0][0] = c;
    r[0][1] = -s;
    r[1][0] = s;
    r[1][1] = c;
  } else {
    const double v[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                         a[0] * b[1] - a[1] * b[0]};
    const double c = (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
    const double sk[3][3] = {{0.0, -v[2], +v[1]}, {+v[2], 0.0, -v[0]}, {-v[1], +v[0], 0.0}};
    const double den = 1.0 + c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double sq = (sk[i][0] * sk[0][j] + sk[i][1] * sk[1][j]) + sk[i][2] * sk[2][j];
        r[i][j] = ((i == j ? 1.0 : 0.0) + sk[i][j]) + DivRn(sq, den);
      }
  }
  double out[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double acc = r[i][0] * m[0];
#pragma unroll
    for (int k = 1; k < D; ++k) acc = acc + r[i][k] * m[k];
    out[i] = acc;
  }
#pragma unroll
  for (int i = 0; i < D; ++i) m[i] = out[i];
}

// Hllem::SolveRiemannProblem (scalar) along X between complete states.
template <int D>
__device__ void HllemRiemannState(double* sol, const double* left, const double* right,
                                  bool larrouturou, const Gas<double>& g) {
  constexpr int NQ = D + 4, M = 1, d = 0;
  double L[NQ], R[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    L[k] = left[k];
    R[k] = right[k];
  }
  const double Msq = 1.0;
  const double Minv = DivRn(1.0, sqrt(Msq));
  const double rhoL = L[0], rhoR = R[0];
  double vL_[D], vR_[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    vL_[i] = DivRn(L[M + i], rhoL);
    vR_[i] = DivRn(R[M + i], rhoR);
  }
  double eKinRL = 0.0, eKinRR = 0.0;
  if (larrouturou) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const int mask = (i == d);
      eKinRL += (1 - mask) * vL_[i] * vL_[i];
      eKinRR += (1 - mask) * vR_[i] * vR_[i];
      L[M + i] *= mask;
      R[M + i] *= mask;
    }
    eKinRL *= 0.5 * Msq;
    eKinRR *= 0.5 * Msq;
    L[D + 1] -= rhoL * eKinRL;
    R[D + 1] -= rhoR * eKinRR;
  }
  double fluxL[D + 2], fluxR[D + 2];
  PhysFlux<D, 0>(fluxL, L);
  PhysFlux<D, 0>(fluxR, R);
  const double uL = DivRn(L[M + d], rhoL);
  const double uR = DivRn(R[M + d], rhoR);
  const double aL = L[D + 3], aR = R[D + 3];
  const double rhoEL = L[D + 1], rhoER = R[D + 1];
  const double pL = L[D + 2], pR = R[D + 2];
  const double hL = DivRn(rhoEL + pL, rhoL);
  const double hR = DivRn(rhoER + pR, rhoR);
  const double sqRhoL = sqrt(rhoL);
  const double sqRhoR = sqrt(rhoR);
  const double sqRhoSum = sqRhoL + sqRhoR;
  double roeU[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double vL = DivRn(L[M + i], rhoL);
    const double vR = DivRn(R[M + i], rhoR);
    roeU[i] = DivRn(sqRhoL * vL + sqRhoR * vR, sqRhoSum);
  }
  const double roeU0 = roeU[d];
  const double roeH = DivRn(sqRhoL * hL + sqRhoR * hR, sqRhoSum);
  const double roeGamma = DivRn(sqRhoL * g.gamma + sqRhoR * g.gamma, sqRhoSum);
  const double gm1 = roeGamma - 1.0;
  const double sqU = SqNormScalar<D>(roeU);
  const double roeA2 = gm1 * (roeH - 0.5 * Msq * sqU);
  const double roeA = sqrt(roeA2);
  const double maxA = rmax(aL, aR);
  const double sL1 = uL - maxA * Minv;
  const double sL2 = roeU0 - roeA * Minv;
  const double sR1 = roeU0 + roeA * Minv;
  const double sR2 = uR + maxA * Minv;
  const double sL = rmin(sL1, sL2);
  const double sR = rmax(sR1, sR2);
  const double bL = rmin(sL, 0.0);
  const double bR = rmax(sR, 0.0);
  const double db = bR - bL;
  const double db_positive = int(db <= 0) + int(db > 0) * db;
  double w[D + 2];
#pragma unroll
  for (int k = 0; k < D + 2; ++k) w[k] = DivRn(bR * R[k] - bL * L[k] + fluxL[k] - fluxR[k], db_positive);
  const double squaredNormRoeU_half = 0.5 * (Msq * sqU);
  const double u_bar = 0.5 * (sR + sL);
  const double delta = DivRn(roeA, roeA + fabs(u_bar));
  const double deltaRho = rhoR - rhoL;
  double dRhoU[D];
#pragma unroll
  for (int i = 0; i < D; ++i) dRhoU[i] = R[M + i] - L[M + i];
  const double deltaRhoE = rhoER - rhoEL;
  double wm[D + 2];
#pragma unroll
  for (int k = 0; k < D + 2; ++k) wm[k] = w[k];
  const int upwind = (u_bar > 0.0);
  if constexpr (D == 2) {
    constexpr int ix = 0, iy = 1;
    const double l21 = DivRn(gm1, roeA2) * (roeH - Msq * sqU);
    const double l22 = DivRn(gm1, roeA2) * Msq * roeU[0];
    const double l23 = DivRn(gm1, roeA2) * roeU[1];
    const double l24 = DivRn(gm1, roeA2) * (-1.0);
    const double alpha_2 = l21 * deltaRho + l22 * dRhoU[0] + l23 * dRhoU[1] + l24 * deltaRhoE;
    const double l31 = -roeU[iy];
    const double l33 = 1;
    const double alpha_3 = l31 * deltaRho + l33 * dRhoU[iy];
    const double ubda2 = u_bar * delta * alpha_2;
    const double ubda3 = u_bar * delta * alpha_3;
    wm[0] = w[0] - ubda2 * 1.0;
    wm[M + ix] = w[M + ix] - ubda2 * roeU[ix];
    if (larrouturou) {
      wm[M + iy] = wm[0] * (upwind * vL_[iy] + (1 - upwind) * vR_[iy]);
      wm[D + 1] = w[D + 1] - ubda2 * squaredNormRoeU_half +
                  wm[0] * (upwind * eKinRL + (1 - upwind) * eKinRR);
    } else {
      wm[M + iy] = w[M + iy] - ubda2 * roeU[iy] - ubda3 * 1.0;
      wm[D + 1] = w[D + 1] - ubda2 * squaredNormRoeU_half - ubda3 * roeU[iy];
    }
  } else {
    constexpr int ix = 0, iy = 1, iz = 2;
    const double l21 = DivRn(gm1, roeA2) * (roeH - Msq * sqU);
    const double l22 = DivRn(gm1, roeA2) * Msq * roeU[0];
    const double l23 = DivRn(gm1, roeA2) * roeU[1];
    const double l24 = DivRn(gm1, roeA2) * roeU[2];
    const double l25 = DivRn(gm1, roeA2) * (-1.0);
    const double l31 = -roeU[iy];
    const double l33 = 1;
    const double l41 = -roeU[iz];
    const double l44 = 1;
    const double alpha_2 =
        l21 * deltaRho + l22 * dRhoU[0] + l23 * dRhoU[1] + l24 * dRhoU[2] + l25 * deltaRhoE;
    const double alpha_3 = l31 * deltaRho + l33 * dRhoU[iy];
    const double alpha_4 = l41 * deltaRho + l44 * dRhoU[iz];
    const double ubd = u_bar * delta;
    wm[0] = w[0] - ubd * alpha_2;
    wm[M + ix] = w[M + ix] - ubd * alpha_2 * roeU[ix];
    if (larrouturou) {
      wm[M + iy] = wm[0] * (upwind * vL_[iy] + (1 - upwind) * vR_[iy]);
      wm[M + iz] = wm[0] * (upwind * vL_[iz] + (1 - upwind) * vR_[iz]);
      wm[D + 1] = w[D + 1] - ubd * alpha_2 * squaredNormRoeU_half +
                  wm[0] * (upwind * eKinRL + (1 - upwind) * eKinRR);
    } else {
      wm[M + iy] = w[M + iy] - ubd * (alpha_2 * roeU[iy] + alpha_3 * 1.0);
      wm[M + iz] = w[M + iz] - ubd * (alpha_2 * roeU[iz] + alpha_4 * 1.0);
      wm[D + 1] = w[D + 1] - ubd * (alpha_2 * squaredNormRoeU_half + alpha_3 * roeU[iy] +
                                    alpha_4 * roeU[iz]);
    }
  }
  if (0.0 < sL) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) sol[k] = L[k];
  } else if (sR < 0.0) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) sol[k] = R[k];
  } else {
#pragma unroll
    for (int k = 0; k < D + 2; ++k) sol[k] = wm[k];
    CompleteFromConsScalar<D>(sol, g);
  }
}

struct BoundaryParams {
  Gas<double> g;
  NewtonParams np;
  int hllem;        // boundary solver: 0 exact, 1 HLLEM
  int larrouturou;  // HLLEM variant
};

// The state at the wall: `state` (complete) rotated onto the boundary
// normal, solved against its reflection along X; returns the solution in
// the rotated frame and rotates `state` in place.
template <int D>
__device__ void WallProblem(double* sol, double* state, const double* normal,
                            const BoundaryParams& bp) {
  constexpr int NQ = D + 4;
  double unit[D];
#pragma unroll
  for (int i = 0; i < D; ++i) unit[i] = (i == 0) ? 1.0 : 0.0;
  RotateMomentum<D>(state + 1, normal, unit);
  double reflected[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) reflected[k] = state[k];
  {
    double dot = state[1] * unit[0];
#pragma unroll
    for (int k = 1; k < D; ++k) dot = dot + state[1 + k] * unit[k];
#pragma unroll
    for (int i = 0; i < D; ++i) reflected[1 + i] = state[1 + i] - 2.0 * (dot * unit[i]);
  }
#pragma unroll
  for (int k = 0; k < NQ; ++k) sol[k] = 0.0;
  if (bp.hllem) {
    HllemRiemannState<D>(sol, reflected, state, bp.larrouturou != 0, bp.g);
  } else if (0.0 < reflected[0] && 0.0 < state[0]) {
    const MiddleState ms = ComputeMiddleState<D, 0>(reflected, state, bp.g, bp.np,
                                                    0.5 * (reflected[D + 2] + state[D + 2]));
    SampleRiemann<D, 0>(sol, reflected, state, ms, bp.g);
  }
  RotateMomentum<D>(sol + 1, unit, normal);
}

template <int D>
__device__ __forceinline__ void LoadQ(const DevView& v, const int64_t* c, double* q) {
#pragma unroll
  for (int k = 0; k < D + 4; ++k) q[k] = *v.at(c, k);
}

template <int D>
__device__ __forceinline__ void Normal(const GeomViews& g, const int64_t* c, double* n) {
#pragma unroll
  for (int k = 0; k < D; ++k) n[k] = -*g.normals.at(c, k);
}

template <int D>
__global__ __launch_bounds__(kBlock) void ReferenceStatesKernel(DevView refs, DevView states,
                                                                GeomViews geom,
                                                                BoundaryParams bp) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= states.count()) return;
  int64_t c[3];
  states.index(q, c);
  if (!IsCutCell<D>(geom, c)) return;
  double state[D + 4];
  LoadQ<D>(states, c, state);
  if (!(state[0] > 0.0)) return;
  double normal[D];
  Normal<D>(geom, c, normal);
  double sol[D + 4];
  WallProblem<D>(sol, state, normal, bp);
#pragma unroll
  for (int k = 0; k < D + 4; ++k) *refs.at(c, k) = sol[k];
}

template <int D>
__global__ __launch_bounds__(kBlock) void BoundaryFluxKernel(DevView bflux, DevView states,
                                                             DevView refs, GeomViews geom,
                                                             BoundaryParams bp, int dir) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= refs.count()) return;
  int64_t c[3];
  refs.index(q, c);
  if (!IsCutCell<D>(geom, c)) return;
  double state[D + 4], ref[D + 4];
  LoadQ<D>(states, c, state);
  LoadQ<D>(refs, c, ref);
  double normal[D];
  Normal<D>(geom, c, normal);
  double flux[D + 2];
  if (dir == 0) PhysFlux<D, 0>(flux, ref);
  else if (dir == 1) PhysFlux<D, 1>(flux, ref);
  else PhysFlux<D, (D > 2 ? 2 : 0)>(flux, ref);
  double sol[D + 4];
  WallProblem<D>(sol, state, normal, bp);
  flux[1 + dir] += sol[D + 2] - ref[D + 2];
#pragma unroll
  for (int k = 0; k < D + 2; ++k) *bflux.at(c, k) = flux[k];
}

// PerfectGas::CompleteFromCons, masked array form: the zero state unless
// the lane is valid with positive density.  T / A: the arithmetic of the
// reconstruction (double IEEE, or float for the FP32-flux form).
template <int D, class T = double, class A = IeeeArith<T>>
__device__ __forceinline__ void CompleteMasked(T* q, bool mask, const Gas<T>& g) {
  if (mask && q[0] > T(0)) {
    CompleteFromCons<D, T, A>(q, g);
  } else {
#pragma unroll
    for (int k = 0; k < D + 4; ++k) q[k] = T(0);
  }
}

// Masked PerfectGas::Flux: zero for the zero state (density 1 stands in
// for invalid lanes, so every product vanishes).
template <int D, int DIR, class T = double, class A = IeeeArith<T>>
__device__ __forceinline__ void FluxMasked(T* f, const T* q) {
  if (q[0] > T(0)) {
    PhysFlux<D, DIR, T, A>(f, q);
  } else {
#pragma unroll
    for (int k = 0; k < D + 2; ++k) f[k] = T(0);
  }
}

// Masked MUSCL-Hancock half-step states of one cell: slopes are zero unless
// the three stencil cells are all uncovered.
template <int D, int DIR, int LIM, class T = double, class A = IeeeArith<T>>
__device__ __forceinline__ void MaskedRecon(const T* um, const T* u0, const T* up,
                                            bool slope_mask, T lh, const Gas<T>& g,
                                            T* recL, T* recR) {
  constexpr int NC = D + 2, NQ = D + 4;
  T qL[NQ], qR[NQ];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const T s = slope_mask ? LimitedSlope<LIM, T, A>(um[k], u0[k], up[k]) : T(0);
    qL[k] = u0[k] - T(0.5) * s;
    qR[k] = u0[k] + T(0.5) * s;
  }
  CompleteMasked<D, T, A>(qL, true, g);
  CompleteMasked<D, T, A>(qR, true, g);
  T fL[NC], fR[NC];
  FluxMasked<D, DIR, T, A>(fL, qL);
  FluxMasked<D, DIR, T, A>(fR, qR);
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const T dq = lh * (fL[k] - fR[k]);
    recR[k] = qR[k] + dq;
    recL[k] = qL[k] + dq;
  }
  CompleteMasked<D, T, A>(recR, true, g);
  CompleteMasked<D, T, A>(recL, true, g);
}

struct FluxParams {
  Gas<double> g;
  double dt, dx;
  int limiter, muscl, quirk;
};

template <int D, int DIR, bool LARR>
__global__ __launch_bounds__(kBlock) void RegularFluxKernel(DevView fl, DevView st, DevView alpha,
                                                            DevView beta, FluxParams fp) {
  constexpr int NC = D + 2, NQ = D + 4;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= fl.count()) return;
  int64_t f[3];
  fl.index(q, f);
  double out[NC];
  if (!(Val(beta, f) > 0.0)) {
#pragma unroll
    for (int k = 0; k < NC; ++k) out[k] = 0.0;
  } else {
    double L[NQ], R[NQ];
    if (fp.muscl) {
      double u[4][NQ];
      double a[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        int64_t c[3] = {f[0], f[1], f[2]};
        c[DIR] += s - 2;
        LoadQ<D>(st, c, u[s]);
        a[s] = Val(alpha, c);
      }
      const bool mL = a[0] > 0.0 && a[1] > 0.0 && a[2] > 0.0;
      const bool mR = a[1] > 0.0 && a[2] > 0.0 && a[3] > 0.0;
      const double lh = DivRn(0.5 * fp.dt, fp.dx);
      double dummy[NQ];
      switch (fp.limiter) {
      case kMinMod:
        MaskedRecon<D, DIR, kMinMod>(u[0], u[1], u[2], mL, lh, fp.g, dummy, L);
        MaskedRecon<D, DIR, kMinMod>(u[1], u[2], u[3], mR, lh, fp.g, R, dummy);
        break;
      case kVanLeer:
        MaskedRecon<D, DIR, kVanLeer>(u[0], u[1], u[2], mL, lh, fp.g, dummy, L);
        MaskedRecon<D, DIR, kVanLeer>(u[1], u[2], u[3], mR, lh, fp.g, R, dummy);
        break;
      case kNoLimiter:
        MaskedRecon<D, DIR, kNoLimiter>(u[0], u[1], u[2], mL, lh, fp.g, dummy, L);
        MaskedRecon<D, DIR, kNoLimiter>(u[1], u[2], u[3], mR, lh, fp.g, R, dummy);
        break;
      default:
        MaskedRecon<D, DIR, kNoGradient>(u[0], u[1], u[2], mL, lh, fp.g, dummy, L);
        MaskedRecon<D, DIR, kNoGradient>(u[1], u[2], u[3], mR, lh, fp.g, R, dummy);
        break;
      }
    } else {
      int64_t c[3] = {f[0], f[1], f[2]};
      c[DIR] -= 1;
      LoadQ<D>(st, c, L);
      LoadQ<D>(st, f, R);
    }
    if constexpr (D == 3) {
      if (fp.quirk) HllemFlux<D, DIR, LARR, true>(out, L, R, fp.g);
      else HllemFlux<D, DIR, LARR, false>(out, L, R, fp.g);
    } else {
      HllemFlux<D, DIR, LARR, false>(out, L, R, fp.g);
    }
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) *fl.at(f, k) = out[k];
}

__device__ __forceinline__ unsigned long long KeyMin(unsigned long long a, unsigned long long b) {
  return b < a ? b : a;
}

__device__ __forceinline__ unsigned long long WaveMinKey(unsigned long long k) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(k, off, 64);
    k = o < k ? o : k;
  }
  return k;
}

// Masked HLLEM stable dt of the faces enumerated by `faces` (cells f - 1, f);
// DBL_MAX where beta = 0.  One atomic per workgroup.
template <int D, int DIR>
__global__ __launch_bounds__(kBlock) void MaskedStableDtKernel(DevView st, DevView faces,
                                                               DevView beta, double dx,
                                                               Gas<double> g,
                                                               unsigned long long* key) {
  constexpr int NQ = D + 4;
  __shared__ unsigned long long s[kBlock / 64];
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long k = 0x7ff0000000000000ull;
  if (q < faces.count()) {
    int64_t f[3];
    faces.index(q, f);
    double dt = DBL_MAX;
    if (Val(beta, f) > 0.0) {
      double L[NQ], R[NQ];
      int64_t c[3] = {f[0], f[1], f[2]};
      c[DIR] -= 1;
      LoadQ<D>(st, c, L);
      LoadQ<D>(st, f, R);
      dt = HllemStableDt<D, DIR>(L, R, dx, g);
    }
    k = DtKey(dt);
  }
  k = WaveMinKey(k);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = s[0];
    for (int w = 1; w < kBlock / 64; ++w) m = s[w] < m ? s[w] : m;
    atomicMin(key, m);
  }
}

__device__ __forceinline__ double Clamp01(double x) {
  // Vc: where(x < lo) x = lo; where(hi < x) x = hi;  (std::clamp for non-NaN)
  if (x < 0.0) x = 0.0;
  if (1.0 < x) x = 1.0;
  return x;
}

__global__ __launch_bounds__(kBlock) void CutCellFluxKernel(DevView stab, DevView fsl,
                                                            DevView fsr, DevView bflux,
                                                            DevView regular, DevView us,
                                                            DevView sl, DevView sr,
                                                            DevView centroids, int dir,
                                                            int ncomp) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= regular.count()) return;
  int64_t f[3];
  regular.index(q, f);
  int64_t l[3] = {f[0], f[1], f[2]};
  l[dir] -= 1;
  const double dL = Clamp01(0.5 - *centroids.at(l, dir));
  const double dR = Clamp01(0.5 + *centroids.at(f, dir));
  const double betaUS = *us.at(f, 1);
  const double betaL = *sl.at(f, 1);
  const double betaR = *sr.at(f, 1);
  for (int c = 0; c < ncomp; ++c) {
    const double fr = *regular.at(f, c);
    const double sL = dL * fr + (1.0 - dL) * *bflux.at(l, c);
    const double sR = dR * fr + (1.0 - dR) * *bflux.at(f, c);
    *fsl.at(f, c) = sL;
    *fsr.at(f, c) = sR;
    *stab.at(f, c) = betaUS * fr + betaL * sL + betaR * sR;
  }
}

struct UpdateViews {
  DevView scratch, fus, fst, fsl, fsr, fb;
  DevView alpha, beta, us, sl, sr, centroids;
};

__global__ __launch_bounds__(kBlock) void KbnUpdateKernel(UpdateViews v, double dt_over_dx,
                                                          int dir, int ncomp) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= v.scratch.count()) return;
  int64_t c[3];
  v.scratch.index(q, c);
  int64_t r[3] = {c[0], c[1], c[2]};
  r[dir] += 1;
  const double alpha = *v.alpha.at(c, 0);
  if (alpha == 0.0) {
    for (int k = 0; k < ncomp; ++k) *v.scratch.at(c, k) = 0.0;
    return;
  }
  const double betaL = *v.beta.at(c, 0);
  const double betaR = *v.beta.at(r, 0);
  const double center = *v.centroids.at(c, dir);
  const double betaL_us_over_alpha = DivRn(*v.us.at(c, 0), alpha);
  const double betaL_sL_over_alpha = DivRn(*v.sl.at(c, 0), alpha);
  const double dLg = Clamp01(DivRn(*v.sr.at(c, 0) * (0.5 + center), alpha));
  const double betaR_us_over_alpha = DivRn(*v.us.at(r, 0), alpha);
  const double betaR_sR_over_alpha = DivRn(*v.sr.at(r, 0), alpha);
  const double dRg = Clamp01(DivRn(*v.sl.at(r, 0) * (0.5 - center), alpha));
  for (int k = 0; k < ncomp; ++k) {
    double* next = v.scratch.at(c, k);
    const double prev = *next;
    const double fLs = *v.fst.at(c, k);
    const double fRs = *v.fst.at(r, k);
    double out;
    if (betaL == betaR) {
      out = prev + dt_over_dx * (fLs - fRs);
    } else if (betaL > betaR) {
      const double fLus = *v.fus.at(c, k);
      const double fLsL = *v.fsl.at(c, k);
      const double fB = *v.fb.at(c, k);
      out = prev + dt_over_dx * betaL_us_over_alpha * (fLus - fRs) +
            dt_over_dx * betaL_sL_over_alpha * (fLsL - fRs) + dt_over_dx * dLg * (fLus - fB);
    } else {
      const double fRus = *v.fus.at(r, k);
      const double fRsR = *v.fsr.at(r, k);
      const double fB = *v.fb.at(c, k);
      out = prev + dt_over_dx * betaR_us_over_alpha * (fLs - fRus) +
            dt_over_dx * betaR_sR_over_alpha * (fLs - fRsR) + dt_over_dx * dRg * (fB - fRus);
    }
    *next = out;
  }
}

__global__ __launch_bounds__(kBlock) void KbnUpdate2Kernel(DevView dest, DevView src, DevView fst,
                                                           DevView fb, DevView alpha, DevView beta,
                                                           double dt_over_dx, int dir,
                                                           int ncomp) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= dest.count()) return;
  int64_t c[3];
  dest.index(q, c);
  const double volfrac = *alpha.at(c, 0);
  if (!(volfrac > 0.0)) return;
  int64_t r[3] = {c[0], c[1], c[2]};
  r[dir] += 1;
  const double bL = *beta.at(c, 0);
  const double bR = *beta.at(r, 0);
  for (int k = 0; k < ncomp; ++k) {
    const double fL = *fst.at(c, k);
    const double fR = *fst.at(r, k);
    const double prev = *src.at(c, k);
    double next;
    if (bL == bR) {
      next = prev + dt_over_dx * (fL - fR);
    } else {
      const double dB = bL - bR;
      next = prev + DivRn(dt_over_dx, volfrac) * (bL * fL - bR * fR - dB * *fb.at(c, k));
    }
    *dest.at(c, k) = next;
  }
}

template <int D>
__global__ __launch_bounds__(kBlock) void KbnCompleteKernel(DevView complete, DevView cons,
                                                            DevView alpha, Gas<double> g) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= complete.count()) return;
  int64_t c[3];
  complete.index(q, c);
  double u[D + 4];
#pragma unroll
  for (int k = 0; k < D + 2; ++k) u[k] = *cons.at(c, k);
  CompleteMasked<D>(u, *alpha.at(c, 0) > 0.0, g);
#pragma unroll
  for (int k = 0; k < D + 4; ++k) *complete.at(c, k) = u[k];
}

__global__ __launch_bounds__(kBlock) void FillCutCellKernel(DevView beta, DevView faces,
                                                            DevView us, DevView sl, DevView sr,
                                                            DevView ds, int dir) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= faces.count()) return;
  int64_t f[3];
  faces.index(q, f);
  int64_t l[3] = {f[0], f[1], f[2]}, r[3] = {f[0], f[1], f[2]};
  l[dir] -= 1;
  r[dir] += 1;
  const double betaL = *beta.at(l, 0);
  double betaM = *beta.at(f, 0);
  const double betaR = *beta.at(r, 0);
  const double dBetaL = rmax(0.0, betaM - betaL);
  const double dBetaR = rmax(0.0, betaM - betaR);
  const double unshielded = rmax(0.0, betaM - rmax(dBetaL, dBetaR));
  const double doubly = rmin(dBetaL, dBetaR);
  const double sLeft = dBetaL - doubly;
  const double sRight = dBetaR - doubly;
  betaM = betaM > 0 ? betaM : 1.0;
  *us.at(f, 0) = unshielded;
  *us.at(f, 1) = DivRn(unshielded, betaM);
  *sl.at(f, 0) = sLeft;
  *sl.at(f, 1) = DivRn(sLeft, betaM);
  *sr.at(f, 0) = sRight;
  *sr.at(f, 1) = DivRn(sRight, betaM);
  *ds.at(f, 0) = doubly;
  *ds.at(f, 1) = DivRn(doubly, betaM);
}

template <int D>
__global__ __launch_bounds__(kBlock) void IsCutCellKernel(GeomViews geom, DevView box,
                                                          int32_t* flags) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= box.count()) return;
  int64_t c[3];
  box.index(q, c);
  flags[q] = IsCutCell<D>(geom, c) ? 1 : 0;
}

unsigned Blocks(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// A view covering box [lo, hi] with no data (index enumeration only).
DevView BoxView(const int64_t lo[3], const int64_t hi[3]) {
  DevView v{};
  for (int d = 0; d < 3; ++d) {
    v.lo[d] = lo[d];
    v.ext[d] = hi[d] - lo[d] + 1;
  }
  return v;
}

int GeomOf(const fvs_cutcell_data* g, int D, GeomViews* out) {
  if (!g) {
    SetError("null cut-cell data");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = MakeDevView(&g->volume_fractions, 1, &out->alpha)) return rc;
  for (int d = 0; d < 3; ++d) {
    if (d < D) {
      if (int rc = MakeDevView(&g->face_fractions[d], 1, &out->beta[d])) return rc;
    } else {
      out->beta[d] = DevView{};
    }
  }
  if (g->boundary_normals.data) {
    if (int rc = MakeDevView(&g->boundary_normals, D, &out->normals)) return rc;
  } else {
    out->normals = DevView{};
  }
  if (g->boundary_centroids.data) {
    if (int rc = MakeDevView(&g->boundary_centroids, D, &out->centroids)) return rc;
  } else {
    out->centroids = DevView{};
  }
  return FVS_OK;
}

BoundaryParams BoundaryOf(const fvs_method_desc& d) {
  BoundaryParams bp;
  bp.g.gamma = d.gamma;
  bp.g.gm1inv = 1.0 / (d.gamma - 1.0);
  bp.g.gm1 = d.gamma - 1.0;
  bp.np.tol = d.newton_tolerance > 0 ? d.newton_tolerance : 1e-7;
  bp.np.max_it = d.max_newton_iterations > 0 ? d.max_newton_iterations : 5000;
  bp.hllem = d.boundary_solver == FVS_BOUNDARY_HLLEM;
  bp.larrouturou = d.base == FVS_BASE_HLLEM_LARROUTUROU;
  return bp;
}

int CheckKbn(const fvs_handle* h, int dir, bool needs_hllem) {
  if (!h) {
    SetError("null handle");
    return FVS_BAD_ARGUMENT;
  }
  if (dir < 0 || dir >= h->desc.dim) {
    SetError("direction %d out of range", dir);
    return FVS_BAD_ARGUMENT;
  }
  if (h->desc.base == FVS_BASE_HLL_EINFELDT ||
      h->desc.reconstruction >= FVS_RECON_MUSCL_HANCOCK_PRIM ||
      (needs_hllem && h->desc.base == FVS_BASE_GODUNOV_EXACT)) {
    SetError("the cut-cell regular fluxes use an HLLEM base method");
    return FVS_BAD_ARGUMENT;
  }
  return FVS_OK;
}

} // namespace
} // namespace fvs

using namespace fvs;

extern "C" {

int fvs_kbn_fill_cutcell_data(fvs_handle* h, const fvs_cutcell_data* geom, int dir,
                              const int64_t face_lo[3], const int64_t face_hi[3]) {
  if (int rc = CheckKbn(h, dir, false)) return rc;
  if (!geom || !face_lo || !face_hi) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  DevView beta, us, sl, sr, ds;
  if (int rc = MakeDevView(&geom->face_fractions[dir], 1, &beta)) return rc;
  if (int rc = MakeDevView(&geom->unshielded[dir], 2, &us)) return rc;
  if (int rc = MakeDevView(&geom->shielded_left[dir], 2, &sl)) return rc;
  if (int rc = MakeDevView(&geom->shielded_right[dir], 2, &sr)) return rc;
  if (int rc = MakeDevView(&geom->doubly_shielded[dir], 2, &ds)) return rc;
  int64_t blo[3], bhi[3];
  for (int d = 0; d < 3; ++d) {
    blo[d] = face_lo[d] - (d == dir);
    bhi[d] = face_hi[d] + (d == dir);
    if (face_hi[d] < face_lo[d]) {
      SetError("empty face box");
      return FVS_BAD_ARGUMENT;
    }
  }
  if (!ContainsBox(&geom->face_fractions[dir], blo, bhi)) {
    SetError("face fractions must cover the face box grown by one along dir");
    return FVS_BAD_ARGUMENT;
  }
  const fvs_view* outs[4] = {&geom->unshielded[dir], &geom->shielded_left[dir],
                             &geom->shielded_right[dir], &geom->doubly_shielded[dir]};
  for (const fvs_view* o : outs)
    if (!ContainsBox(o, face_lo, face_hi)) {
      SetError("fraction views must cover the face box");
      return FVS_BAD_ARGUMENT;
    }
  const DevView faces = BoxView(face_lo, face_hi);
  hipLaunchKernelGGL(FillCutCellKernel, dim3(Blocks(faces.count())), dim3(kBlock), 0, h->stream,
                     beta, faces, us, sl, sr, ds, dir);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_kbn_is_cut_cell(fvs_handle* h, const fvs_cutcell_data* geom, const int64_t lo[3],
                        const int64_t hi[3], int32_t* flags) {
  if (int rc = CheckKbn(h, 0, false)) return rc;
  if (!lo || !hi || !flags) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  const int D = h->desc.dim;
  GeomViews g;
  if (int rc = GeomOf(geom, D, &g)) return rc;
  if (!ContainsBox(&geom->volume_fractions, lo, hi)) {
    SetError("volume fractions must cover the box");
    return FVS_BAD_ARGUMENT;
  }
  for (int d = 0; d < D; ++d) {
    int64_t fhi[3] = {hi[0], hi[1], hi[2]};
    fhi[d] += 1;
    if (!ContainsBox(&geom->face_fractions[d], lo, fhi)) {
      SetError("face fractions of direction %d must cover the faces of the box", d);
      return FVS_BAD_ARGUMENT;
    }
  }
  const DevView box = BoxView(lo, hi);
  if (D == 2)
    hipLaunchKernelGGL(IsCutCellKernel<2>, dim3(Blocks(box.count())), dim3(kBlock), 0, h->stream, g, box, flags);
  else
    hipLaunchKernelGGL(IsCutCellKernel<3>, dim3(Blocks(box.count())), dim3(kBlock), 0, h->stream, g, box, flags);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

// Checks that the geometry covers the cells of box and their upper faces
// (IsCutCell reads beta at the cell and its right neighbour face).
static int CheckCellGeometry(const fvs_cutcell_data* geom, int D, const fvs_view* cells,
                             bool normals, bool centroids) {
  if (!ContainsBox(&geom->volume_fractions, cells->lo, cells->hi)) {
    SetError("volume fractions must cover the cell box");
    return FVS_BAD_ARGUMENT;
  }
  for (int d = 0; d < D; ++d) {
    int64_t fhi[3] = {cells->hi[0], cells->hi[1], cells->hi[2]};
    fhi[d] += 1;
    if (!ContainsBox(&geom->face_fractions[d], cells->lo, fhi)) {
      SetError("face fractions of direction %d must cover the cell box's faces", d);
      return FVS_BAD_ARGUMENT;
    }
  }
  if (normals && (geom->boundary_normals.ncomp < D ||
                  !ContainsBox(&geom->boundary_normals, cells->lo, cells->hi))) {
    SetError("boundary normals must cover the cell box");
    return FVS_BAD_ARGUMENT;
  }
  if (centroids && (geom->boundary_centroids.ncomp < D ||
                    !ContainsBox(&geom->boundary_centroids, cells->lo, cells->hi))) {
    SetError("boundary centroids must cover the cell box");
    return FVS_BAD_ARGUMENT;
  }
  return FVS_OK;
}

int fvs_kbn_reference_states(fvs_handle* h, const fvs_view* references, const fvs_view* states,
                             const fvs_cutcell_data* geom) {
  if (int rc = CheckKbn(h, 0, false)) return rc;
  const int D = h->desc.dim;
  DevView rv, sv;
  if (int rc = MakeDevView(references, D + 4, &rv)) return rc;
  if (int rc = MakeDevView(states, D + 4, &sv)) return rc;
  if (!ContainsBox(references, states->lo, states->hi)) {
    SetError("references must cover the states box");
    return FVS_BAD_ARGUMENT;
  }
  GeomViews g;
  if (int rc = GeomOf(geom, D, &g)) return rc;
  if (int rc = CheckCellGeometry(geom, D, states, true, false)) return rc;
  const BoundaryParams bp = BoundaryOf(h->desc);
  if (D == 2)
    hipLaunchKernelGGL(ReferenceStatesKernel<2>, dim3(Blocks(sv.count())), dim3(kBlock), 0, h->stream, rv, sv, g, bp);
  else
    hipLaunchKernelGGL(ReferenceStatesKernel<3>, dim3(Blocks(sv.count())), dim3(kBlock), 0, h->stream, rv, sv, g, bp);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_kbn_boundary_fluxes(fvs_handle* h, const fvs_view* boundary_fluxes,
                            const fvs_view* states, const fvs_view* references,
                            const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  (void)dt;
  (void)dx;
  if (int rc = CheckKbn(h, dir, false)) return rc;
  const int D = h->desc.dim;
  DevView bv, sv, rv;
  if (int rc = MakeDevView(boundary_fluxes, D + 2, &bv)) return rc;
  if (int rc = MakeDevView(states, D + 4, &sv)) return rc;
  if (int rc = MakeDevView(references, D + 4, &rv)) return rc;
  if (!ContainsBox(states, references->lo, references->hi) ||
      !ContainsBox(boundary_fluxes, references->lo, references->hi)) {
    SetError("states and boundary fluxes must cover the references box");
    return FVS_BAD_ARGUMENT;
  }
  GeomViews g;
  if (int rc = GeomOf(geom, D, &g)) return rc;
  if (int rc = CheckCellGeometry(geom, D, references, true, false)) return rc;
  const BoundaryParams bp = BoundaryOf(h->desc);
  if (D == 2)
    hipLaunchKernelGGL(BoundaryFluxKernel<2>, dim3(Blocks(rv.count())), dim3(kBlock), 0, h->stream, bv, sv, rv, g, bp, dir);
  else
    hipLaunchKernelGGL(BoundaryFluxKernel<3>, dim3(Blocks(rv.count())), dim3(kBlock), 0, h->stream, bv, sv, rv, g, bp, dir);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_kbn_regular_fluxes(fvs_handle* h, const fvs_view* fluxes, const fvs_view* states,
                           const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  if (int rc = CheckKbn(h, dir, true)) return rc;
  const int D = h->desc.dim;
  DevView fl, st, alpha, beta;
  if (int rc = MakeDevView(fluxes, D + 2, &fl)) return rc;
  if (int rc = MakeDevView(states, D + 4, &st)) return rc;
  const bool muscl = h->desc.reconstruction == FVS_RECON_MUSCL_HANCOCK;
  const int w = muscl ? 2 : 1;
  int64_t clo[3], chi[3];
  for (int d = 0; d < 3; ++d) {
    clo[d] = fluxes->lo[d] - (d == dir ? w : 0);
    chi[d] = fluxes->hi[d] + (d == dir ? w - 1 : 0);
  }
  if (!ContainsBox(states, clo, chi) || !ContainsBox(&geom->volume_fractions, clo, chi) ||
      !ContainsBox(&geom->face_fractions[dir], fluxes->lo, fluxes->hi)) {
    SetError("states and geometry must cover the stencils of the flux box");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = MakeDevView(&geom->volume_fractions, 1, &alpha)) return rc;
  if (int rc = MakeDevView(&geom->face_fractions[dir], 1, &beta)) return rc;
  FluxParams fp;
  fp.g.gamma = h->desc.gamma;
  fp.g.gm1inv = 1.0 / (h->desc.gamma - 1.0);
  fp.g.gm1 = h->desc.gamma - 1.0;
  fp.dt = dt;
  fp.dx = dx;
  fp.limiter = h->desc.limiter;
  fp.muscl = muscl;
  fp.quirk = h->desc.quirks;
  const bool larr = h->desc.base == FVS_BASE_HLLEM_LARROUTUROU;
  const unsigned nb = Blocks(fl.count());
  hipStream_t s = h->stream;
#define FVS_KF(DD, DR, LA)                                                                     \
  hipLaunchKernelGGL((RegularFluxKernel<DD, DR, LA>), dim3(nb), dim3(kBlock), 0, s, fl, st,   \
                     alpha, beta, fp)
  if (D == 2) {
    if (dir == 0) { if (larr) FVS_KF(2, 0, true); else FVS_KF(2, 0, false); }
    else { if (larr) FVS_KF(2, 1, true); else FVS_KF(2, 1, false); }
  } else {
    if (dir == 0) { if (larr) FVS_KF(3, 0, true); else FVS_KF(3, 0, false); }
    else if (dir == 1) { if (larr) FVS_KF(3, 1, true); else FVS_KF(3, 1, false); }
    else { if (larr) FVS_KF(3, 2, true); else FVS_KF(3, 2, false); }
  }
#undef FVS_KF
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_kbn_stable_dt(fvs_handle* h, const fvs_view* states, const fvs_cutcell_data* geom,
                      double dx, int dir, double* dt_device) {
  if (int rc = CheckKbn(h, dir, true)) return rc;
  if (!dt_device) {
    SetError("null output");
    return FVS_BAD_ARGUMENT;
  }
  const int D = h->desc.dim;
  DevView st, beta;
  if (int rc = MakeDevView(states, D + 4, &st)) return rc;
  const int w = h->desc.reconstruction == FVS_RECON_MUSCL_HANCOCK ? 2 : 1;
  int64_t flo[3], fhi[3];
  for (int d = 0; d < 3; ++d) {
    flo[d] = states->lo[d] + (d == dir ? w : 0);
    fhi[d] = states->hi[d] - (d == dir ? w - 1 : 0);
    if (fhi[d] < flo[d]) {
      SetError("states box too small for the stencil");
      return FVS_BAD_ARGUMENT;
    }
  }
  if (!ContainsBox(&geom->face_fractions[dir], flo, fhi)) {
    SetError("face fractions must cover the faces of the states box");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = MakeDevView(&geom->face_fractions[dir], 1, &beta)) return rc;
  const DevView faces = BoxView(flo, fhi);
  Gas<double> g;
  g.gamma = h->desc.gamma;
  g.gm1inv = 1.0 / (h->desc.gamma - 1.0);
  g.gm1 = h->desc.gamma - 1.0;
  const unsigned nb = Blocks(faces.count());
  hipStream_t s = h->stream;
  unsigned long long* key = h->scratch_keys;
  if (D == 2) {
    if (dir == 0) hipLaunchKernelGGL((MaskedStableDtKernel<2, 0>), dim3(nb), dim3(kBlock), 0, s, st, faces, beta, dx, g, key);
    else hipLaunchKernelGGL((MaskedStableDtKernel<2, 1>), dim3(nb), dim3(kBlock), 0, s, st, faces, beta, dx, g, key);
  } else {
    if (dir == 0) hipLaunchKernelGGL((MaskedStableDtKernel<3, 0>), dim3(nb), dim3(kBlock), 0, s, st, faces, beta, dx, g, key);
    else if (dir == 1) hipLaunchKernelGGL((MaskedStableDtKernel<3, 1>), dim3(nb), dim3(kBlock), 0, s, st, faces, beta, dx, g, key);
    else hipLaunchKernelGGL((MaskedStableDtKernel<3, 2>), dim3(nb), dim3(kBlock), 0, s, st, faces, beta, dx, g, key);
  }
  FVS_LAUNCH_CHECK();
  return LaunchKeysToDt(s, key, 1, dt_device);
}

int fvs_kbn_cutcell_fluxes(fvs_handle* h, const fvs_view* stabilised,
                           const fvs_view* shielded_left, const fvs_view* shielded_right,
                           const fvs_view* boundary_fluxes, const fvs_view* regular_fluxes,
                           const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  (void)dt;
  (void)dx;
  if (int rc = CheckKbn(h, dir, false)) return rc;
  const int D = h->desc.dim;
  DevView stab, fsl, fsr, fb, reg, us, sl, sr, xc;
  if (int rc = MakeDevView(regular_fluxes, 1, &reg)) return rc;
  const int nc = regular_fluxes->ncomp;
  if (int rc = MakeDevView(stabilised, nc, &stab)) return rc;
  if (int rc = MakeDevView(shielded_left, nc, &fsl)) return rc;
  if (int rc = MakeDevView(shielded_right, nc, &fsr)) return rc;
  if (int rc = MakeDevView(boundary_fluxes, nc, &fb)) return rc;
  if (int rc = MakeDevView(&geom->unshielded[dir], 2, &us)) return rc;
  if (int rc = MakeDevView(&geom->shielded_left[dir], 2, &sl)) return rc;
  if (int rc = MakeDevView(&geom->shielded_right[dir], 2, &sr)) return rc;
  if (int rc = MakeDevView(&geom->boundary_centroids, D, &xc)) return rc;
  const int64_t* flo = regular_fluxes->lo;
  const int64_t* fhi = regular_fluxes->hi;
  int64_t clo[3] = {flo[0], flo[1], flo[2]};
  clo[dir] -= 1;
  int64_t chi[3] = {fhi[0], fhi[1], fhi[2]};
  if (!ContainsBox(stabilised, flo, fhi) || !ContainsBox(shielded_left, flo, fhi) ||
      !ContainsBox(shielded_right, flo, fhi) || !ContainsBox(&geom->unshielded[dir], flo, fhi) ||
      !ContainsBox(&geom->shielded_left[dir], flo, fhi) ||
      !ContainsBox(&geom->shielded_right[dir], flo, fhi) ||
      !ContainsBox(boundary_fluxes, clo, chi) || !ContainsBox(&geom->boundary_centroids, clo, chi)) {
    SetError("views must cover the face box (cells on both sides for boundary data)");
    return FVS_BAD_ARGUMENT;
  }
  hipLaunchKernelGGL(CutCellFluxKernel, dim3(Blocks(reg.count())), dim3(kBlock), 0, h->stream,
                     stab, fsl, fsr, fb, reg, us, sl, sr, xc, dir, nc);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_kbn_update_conservatively(fvs_handle* h, const fvs_view* scratch,
                                  const fvs_view* unshielded, const fvs_view* stabilised,
                                  const fvs_view* shielded_left, const fvs_view* shielded_right,
                                  const fvs_view* boundary_fluxes,
                                  const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  if (int rc = CheckKbn(h, dir, false)) return rc;
  const int D = h->desc.dim;
  UpdateViews v;
  if (int rc = MakeDevView(scratch, 1, &v.scratch)) return rc;
  const int nc = scratch->ncomp;
  if (int rc = MakeDevView(unshielded, nc, &v.fus)) return rc;
  if (int rc = MakeDevView(stabilised, nc, &v.fst)) return rc;
  if (int rc = MakeDevView(shielded_left, nc, &v.fsl)) return rc;
  if (int rc = MakeDevView(shielded_right, nc, &v.fsr)) return rc;
  if (int rc = MakeDevView(boundary_fluxes, nc, &v.fb)) return rc;
  if (int rc = MakeDevView(&geom->volume_fractions, 1, &v.alpha)) return rc;
  if (int rc = MakeDevView(&geom->face_fractions[dir], 1, &v.beta)) return rc;
  if (int rc = MakeDevView(&geom->unshielded[dir], 1, &v.us)) return rc;
  if (int rc = MakeDevView(&geom->shielded_left[dir], 1, &v.sl)) return rc;
  if (int rc = MakeDevView(&geom->shielded_right[dir], 1, &v.sr)) return rc;
  if (int rc = MakeDevView(&geom->boundary_centroids, D, &v.centroids)) return rc;
  int64_t fhi[3] = {scratch->hi[0], scratch->hi[1], scratch->hi[2]};
  fhi[dir] += 1;
  const int64_t* lo = scratch->lo;
  const fvs_view* faces[] = {unshielded, stabilised, shielded_left, shielded_right,
                             &geom->face_fractions[dir], &geom->unshielded[dir],
                             &geom->shielded_left[dir], &geom->shielded_right[dir]};
  for (const fvs_view* f : faces)
    if (!ContainsBox(f, lo, fhi)) {
      SetError("face views must cover the faces enclosing the cell box");
      return FVS_BAD_ARGUMENT;
    }
  if (!ContainsBox(boundary_fluxes, lo, scratch->hi) ||
      !ContainsBox(&geom->volume_fractions, lo, scratch->hi) ||
      !ContainsBox(&geom->boundary_centroids, lo, scratch->hi)) {
    SetError("cell views must cover the cell box");
    return FVS_BAD_ARGUMENT;
  }
  hipLaunchKernelGGL(KbnUpdateKernel, dim3(Blocks(v.scratch.count())), dim3(kBlock), 0,
                     h->stream, v, dt / dx, dir, nc);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_kbn_update_conservatively2(fvs_handle* h, const fvs_view* dest, const fvs_view* src,
                                   const fvs_view* stabilised, const fvs_view* boundary_fluxes,
                                   const fvs_cutcell_data* geom, double dt, double dx, int dir) {
  if (int rc = CheckKbn(h, dir, false)) return rc;
  DevView dv, sv, fst, fb, alpha, beta;
  if (int rc = MakeDevView(dest, 1, &dv)) return rc;
  const int nc = dest->ncomp;
  if (int rc = MakeDevView(src, nc, &sv)) return rc;
  if (int rc = MakeDevView(stabilised, nc, &fst)) return rc;
  if (int rc = MakeDevView(boundary_fluxes, nc, &fb)) return rc;
  if (int rc = MakeDevView(&geom->volume_fractions, 1, &alpha)) return rc;
  if (int rc = MakeDevView(&geom->face_fractions[dir], 1, &beta)) return rc;
  int64_t fhi[3] = {dest->hi[0], dest->hi[1], dest->hi[2]};
  fhi[dir] += 1;
  if (!ContainsBox(src, dest->lo, dest->hi) || !ContainsBox(boundary_fluxes, dest->lo, dest->hi) ||
      !ContainsBox(&geom->volume_fractions, dest->lo, dest->hi) ||
      !ContainsBox(stabilised, dest->lo, fhi) ||
      !ContainsBox(&geom->face_fractions[dir], dest->lo, fhi)) {
    SetError("views must cover the cell box and its faces");
    return FVS_BAD_ARGUMENT;
  }
  hipLaunchKernelGGL(KbnUpdate2Kernel, dim3(Blocks(dv.count())), dim3(kBlock), 0, h->stream, dv,
                     sv, fst, fb, alpha, beta, dt / dx, dir, nc);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_kbn_complete_from_cons(fvs_handle* h, const fvs_view* complete, const fvs_view* cons,
                               const fvs_cutcell_data* geom) {
  if (int rc = CheckKbn(h, 0, false)) return rc;
  const int D = h->desc.dim;
  DevView cv, uv, alpha;
  if (int rc = MakeDevView(complete, D + 4, &cv)) return rc;
  if (int rc = MakeDevView(cons, D + 2, &uv)) return rc;
  if (!geom) {
    SetError("null cut-cell data");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = MakeDevView(&geom->volume_fractions, 1, &alpha)) return rc;
  if (!ContainsBox(cons, complete->lo, complete->hi) ||
      !ContainsBox(&geom->volume_fractions, complete->lo, complete->hi)) {
    SetError("cons and volume fractions must cover the box");
    return FVS_BAD_ARGUMENT;
  }
  Gas<double> g;
  g.gamma = h->desc.gamma;
  g.gm1inv = 1.0 / (h->desc.gamma - 1.0);
  g.gm1 = h->desc.gamma - 1.0;
  if (D == 2)
    hipLaunchKernelGGL(KbnCompleteKernel<2>, dim3(Blocks(cv.count())), dim3(kBlock), 0, h->stream, cv, uv, alpha, g);
  else
    hipLaunchKernelGGL(KbnCompleteKernel<3>, dim3(Blocks(cv.count())), dim3(kBlock), 0, h->stream, cv, uv, alpha, g);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Single-patch 2D cut-cell level: the sequence of the reference's
// DimensionalSplitLevelIntegrator with the cut-cell IntegratorContext
// (reference states per level advance, then per split step the boundary
// fill, KBN CFL check, boundary/regular/stabilised fluxes, cut-cell update
// and masked reconstruction).  The state is stored as conservative
// variables with two ghost layers; complete states are recomputed from them
// with the masked reconstruction wherever a kernel needs them (the values
// the reference stores).  The transmissive boundary fill is implicit: a
// kernel reading a ghost cell's state reads the nearest interior cell's
// (LevelGeom::state_cell), which is what the fill would have copied there,
// while the ghost cell's geometry stays its own.  Kernels per split step:
//   boundary fluxes over the list of cut cells (a dense zero-initialised
//     array; only cut cells are ever written)
//   regular fluxes + masked CFL of every face, with per-workgroup minima
//   h-box update of every interior cell, the stabilised fluxes of its two
//     faces recomputed from the regular fluxes and the face fractions
namespace fvs {
namespace {

constexpr int kG = 2;  // ghost layers

struct LevelGeom {
  const double* alpha;      // [ngy][ngx]
  const double* beta[2];    // x: [ngy][ngx + 1], y: [ngy + 1][ngx]
  const double* normals;    // [2][ngy][ngx] (AMReX orientation)
  const double* centroids;  // [2][ngy][ngx]
  // Per direction, per cell ([ngy][ngx], ghost ring included), what the
  // fused sweeps need of the geometry in one word (CellCode).
  const int32_t* code[2];
  double* irr_flux[2];
  int64_t ngx, ngy;
  __device__ __forceinline__ int64_t cell(int64_t i, int64_t j) const {
    return (j + kG) * ngx + (i + kG);
  }
  // Cell holding the state of cell (i, j) under the transmissive boundary:
  // (i, j) clamped into the interior.
  __device__ __forceinline__ int64_t state_cell(int64_t i, int64_t j) const {
    const int64_t nx = ngx - 2 * kG, ny = ngy - 2 * kG;
    i = i < 0 ? 0 : (i < nx ? i : nx - 1);
    j = j < 0 ? 0 : (j < ny ? j : ny - 1);
    return cell(i, j);
  }
  // face f of direction d at transverse index t (interior coordinates)
  __device__ __forceinline__ double beta_at(int d, int64_t i, int64_t j) const {
    return d == 0 ? beta[0][(j + kG) * (ngx + 1) + (i + kG)]
                  : beta[1][(j + kG) * ngx + (i + kG)];
  }
};

// Geometry word of a cell for the fused sweep of one direction:
//   bit 0: volume fraction > 0 (the reconstruction mask),
//   bit 1: fraction of the cell's lower face in the direction > 0 (the face
//          carries a regular flux),
//   bits 2..: 0 covered (volume fraction 0: the update is the zero state),
//          1 regular (volume fraction 1 and the four face fractions around
//          the cell along the direction 1: the plain update), or 2 + the
//          cell's slot in the direction's list of irregular cells, whose two
//          face fluxes the sweep leaves in irr_flux[slot * 8 ..] for
//          KbnIrregularKernel.
struct CellCode {
  static constexpr int32_t kCovered = 0, kRegular = 1, kSlot0 = 2;
  __host__ __device__ static int32_t Make(bool pos, bool face, int32_t kind) {
    return (pos ? 1 : 0) | (face ? 2 : 0) | (kind << 2);
  }
  __device__ static bool Pos(int32_t c) { return (c & 1) != 0; }
  __device__ static bool Face(int32_t c) { return (c & 2) != 0; }
  __device__ static int32_t Kind(int32_t c) { return c >> 2; }
};

struct LevelState {
  double* u[4];
};

__device__ __forceinline__ void LoadCons(const LevelState& s, int64_t c, double* q) {
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = s.u[k][c];
}


__device__ __forceinline__ unsigned long long BlockMinKey(unsigned long long k) {
  __shared__ unsigned long long s[kBlock / 64];
  k = WaveMinKey(k);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = k;
  __syncthreads();
  unsigned long long m = s[0];
  for (int w = 1; w < kBlock / 64; ++w) m = s[w] < m ? s[w] : m;
  return m;
}

// Complete state of interior cell c with the volume mask.
__device__ __forceinline__ void CellComplete(const LevelState& st, const LevelGeom& g, int64_t c,
                                             const Gas<double>& gas, double* q) {
  LoadCons(st, c, q);
  CompleteMasked<2>(q, g.alpha[c] > 0.0, gas);
}

// Complete state of cell (i, j), a ghost cell included (transmissive
// state, the cell's own volume mask).
__device__ __forceinline__ void CellCompleteAt(const LevelState& st, const LevelGeom& g,
                                               int64_t i, int64_t j, const Gas<double>& gas,
                                               double* q) {
  LoadCons(st, g.state_cell(i, j), q);
  CompleteMasked<2>(q, g.alpha[g.cell(i, j)] > 0.0, gas);
}

// Masked HLLEM CFL of the faces 0..n of direction DIR (interior rows); the
// grid is padded to `nblocks` workgroups so that every direction writes the
// same number of partial minima.
template <int DIR>
__global__ __launch_bounds__(kBlock) void LevelDtKernel(LevelState st, LevelGeom g, int64_t nx,
                                                        int64_t ny, double dx, Gas<double> gas,
                                                        unsigned long long* partials,
                                                        const int* halt) {
  if (*halt) return;
  const int64_t fx = nx + (DIR == 0), fy = ny + (DIR == 1);
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long key = 0x7ff0000000000000ull;
  if (q < fx * fy) {
    const int64_t i = q % fx, j = q / fx;
    double dt = DBL_MAX;
    if (g.beta_at(DIR, i, j) > 0.0) {
      double L[6], R[6];
      CellCompleteAt(st, g, i - (DIR == 0), j - (DIR == 1), gas, L);
      CellCompleteAt(st, g, i, j, gas, R);
      dt = HllemStableDt<2, DIR>(L, R, dx, gas);
    }
    key = DtKey(dt);
  }
  const unsigned long long m = BlockMinKey(key);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// Step-start masked HLLEM CFL of both directions in one pass: a workgroup
// is a tile of 256 columns (column x0 - 1 + t; threads 1..255 own theirs)
// marching kDt2Seg rows, each thread completing its cell once per row; the
// x face to its left pairs it with the left neighbour through LDS, the y
// face below with the previous row kept in registers.  Face values are
// those of LevelDtKernel (DBL_MAX where beta is zero); the per-direction
// minima go to partials[b] (x) and partials[nblocks + b] (y); blocks beyond
// the tiles write +inf.
constexpr int kDt2Seg = 8;

__global__ __launch_bounds__(kBlock) void LevelDt2Kernel(LevelState st, LevelGeom g, int64_t nx,
                                                         int64_t ny, double dx0, double dx1,
                                                         Gas<double> gas,
                                                         unsigned long long* partials,
                                                         int64_t nblocks, const int* halt) {
  if (*halt) return;
  constexpr int kOwn = kBlock - 1;
  const int64_t tiles_x = (nx + 1 + kOwn - 1) / kOwn;
  const int64_t segs = (ny + 1 + kDt2Seg - 1) / kDt2Seg;
  const int64_t b = blockIdx.x;
  unsigned long long kx = 0x7ff0000000000000ull, ky = 0x7ff0000000000000ull;
  if (b < tiles_x * segs) {
    __shared__ double s_q[6][kBlock];
    const int t = threadIdx.x;
    const int64_t i = (b % tiles_x) * kOwn - 1 + t;   // column, -1 .. nx (+ tile padding)
    const int64_t j0 = (b / tiles_x) * kDt2Seg;        // first face row of the segment
    const int64_t j1 = j0 + kDt2Seg < ny + 1 ? j0 + kDt2Seg : ny + 1;
    const bool col_ok = i <= nx;
    const int64_t ic = col_ok ? i : nx;
    // The cell words carry the volume mask (bit 0 of either direction's)
    // and the fractions of the cell's lower x and y faces.
    double prev[6];
    {
      LoadCons(st, g.state_cell(ic, j0 - 1), prev);
      CompleteMasked<2>(prev, CellCode::Pos(g.code[0][g.cell(ic, j0 - 1)]), gas);
    }
    for (int64_t j = j0; j < j1; ++j) {
      const int64_t c = g.cell(ic, j);
      const int32_t cx = g.code[0][c], cy = g.code[1][c];
      double cur[6];
      LoadCons(st, g.state_cell(ic, j), cur);
      CompleteMasked<2>(cur, CellCode::Pos(cx), gas);
#pragma unroll
      for (int k = 0; k < 6; ++k) s_q[k][t] = cur[k];
      __syncthreads();
      if (t >= 1 && col_ok) {
        // x face i between cells i - 1 and i (rows 0 .. ny - 1, faces 0 .. nx)
        if (j < ny && i >= 0 && CellCode::Face(cx)) {
          double L[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) L[k] = s_q[k][t - 1];
          kx = KeyMin(kx, DtKey(HllemStableDt<2, 0>(L, cur, dx0, gas)));
        } else if (j < ny && i >= 0) {
          kx = KeyMin(kx, DtKey(DBL_MAX));
        }
        // y face j between cells (i, j - 1) and (i, j) (columns 0 .. nx - 1)
        if (i >= 0 && i < nx) {
          const double v = CellCode::Face(cy) ? HllemStableDt<2, 1>(prev, cur, dx1, gas) : DBL_MAX;
          ky = KeyMin(ky, DtKey(v));
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 6; ++k) prev[k] = cur[k];
    }
  }
  const unsigned long long mx = BlockMinKey(kx);
  __syncthreads();  // BlockMinKey's shared slots are reused
  const unsigned long long my = BlockMinKey(ky);
  if (threadIdx.x == 0) {
    partials[b] = mx;
    partials[nblocks + b] = my;
  }
}

// Reference states of the cut cells (list) from the step-start state.
__global__ __launch_bounds__(kBlock) void LevelRefsKernel(LevelState st, LevelGeom g,
                                                          const int32_t* cut, int ncut,
                                                          BoundaryParams bp, double* refs,
                                                          const int* halt) {
  if (*halt) return;
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= ncut) return;
  const int64_t c = cut[q];
  double state[6];
  CellComplete(st, g, c, bp.g, state);
  if (!(state[0] > 0.0)) return;  // references keep their previous value
  const int64_t plane = g.ngx * g.ngy;
  double normal[2] = {-g.normals[c], -g.normals[plane + c]};
  double sol[6];
  WallProblem<2>(sol, state, normal, bp);
#pragma unroll
  for (int k = 0; k < 6; ++k) refs[k * ncut + q] = sol[k];
}

// Boundary fluxes of the cut cells (list) from the sweep's input state.
__global__ __launch_bounds__(kBlock) void LevelBoundaryFluxKernel(LevelState st, LevelGeom g,
                                                                  const int32_t* cut, int ncut,
                                                                  BoundaryParams bp,
                                                                  const double* refs, double* fb,
                                                                  int dir, const int* halt) {
  if (*halt) return;
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= ncut) return;
  const int64_t c = cut[q];
  double state[6], ref[6];
  CellComplete(st, g, c, bp.g, state);
#pragma unroll
  for (int k = 0; k < 6; ++k) ref[k] = refs[k * ncut + q];
  const int64_t plane = g.ngx * g.ngy;
  double normal[2] = {-g.normals[c], -g.normals[plane + c]};
  double flux[4];
  if (dir == 0) PhysFlux<2, 0>(flux, ref);
  else PhysFlux<2, 1>(flux, ref);
  double sol[6];
  WallProblem<2>(sol, state, normal, bp);
  flux[1 + dir] += sol[4] - ref[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) fb[k * plane + c] = flux[k];
}

// Face fractions of face i (absolute and relative) from the face fractions
// of faces i - 1, i, i + 1 (FillCutCellData__ arithmetic).
struct FaceFrac {
  double us, sl, sr, us_rel, sl_rel, sr_rel;
};
__device__ __forceinline__ FaceFrac Fractions(double bl, double bm, double br) {
  FaceFrac f;
  const double dBetaL = rmax(0.0, bm - bl);
  const double dBetaR = rmax(0.0, bm - br);
  f.us = rmax(0.0, bm - rmax(dBetaL, dBetaR));
  const double doubly = rmin(dBetaL, dBetaR);
  f.sl = dBetaL - doubly;
  f.sr = dBetaR - doubly;
  const double b = bm > 0 ? bm : 1.0;
  f.us_rel = DivRn(f.us, b);
  f.sl_rel = DivRn(f.sl, b);
  f.sr_rel = DivRn(f.sr, b);
  return f;
}

// A regular cell away from the boundary: both faces are unshielded
// (fractions 1, 0, 0), no boundary flux enters, and the general formula of
// KbnUpdateCell reduces to the plain update with the same result bits.
// u: the cell's conservative state.
__device__ __forceinline__ void KbnRegularUpdate(const double* u, const double* Fl,
                                                 const double* Fr, double dt_over_dx,
                                                 double* res) {
#pragma unroll
  for (int k = 0; k < 4; ++k) res[k] = u[k] + dt_over_dx * (Fl[k] - Fr[k]);
  const bool keep = res[0] > 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) res[k] = keep ? res[k] : 0.0;
}

// Update of a cell from its conservative state u and its two face fluxes
// along DIR.  Irregular cells hand their fluxes to KbnIrregularKernel, which
// writes them after the sweep (the h-box formula inline would cost the
// sweeps registers and occupancy for the few cells near the boundary), and
// return false; otherwise res holds the update (the zero state for covered
// cells).
template <int DIR>
__device__ __forceinline__ bool KbnSweepUpdate(const LevelGeom& g, int32_t code, const double* u,
                                               const double* Fl, const double* Fr,
                                               double dt_over_dx, double* res) {
  const int32_t kind = CellCode::Kind(code);
  if (kind >= CellCode::kSlot0) {
    double* f = g.irr_flux[DIR] + 8 * static_cast<int64_t>(kind - CellCode::kSlot0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      f[k] = Fl[k];
      f[4 + k] = Fr[k];
    }
    return false;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) res[k] = 0.0;
  if (kind == CellCode::kRegular) KbnRegularUpdate(u, Fl, Fr, dt_over_dx, res);
  return true;
}

// h-box update (TimeIntegrator, or TimeIntegrator2 when `second`) of the
// uncovered interior cell (i, j) from the regular fluxes Fl / Fr of its two
// faces in DIR; the boundary fluxes fb are read only where a shielded
// fraction or unequal face fractions need them.  Returns the masked result
// (the zero state unless the density is positive).
template <int DIR>
__device__ __forceinline__ void KbnUpdateCell(const LevelState& in, const LevelGeom& g,
                                              const double* fb, int64_t i, int64_t j,
                                              const double* Fl, const double* Fr,
                                              double dt_over_dx, int second, double* res) {
  const int64_t c = g.cell(i, j);
  const double alpha = g.alpha[c];
  const int di = DIR == 0, dj = DIR == 1;
  const double bLL = g.beta_at(DIR, i - di, j - dj);
  const double bL = g.beta_at(DIR, i, j);
  const double bR = g.beta_at(DIR, i + di, j + dj);
  const double bRR = g.beta_at(DIR, i + 2 * di, j + 2 * dj);
  if (alpha == 1.0 && bLL == 1.0 && bL == 1.0 && bR == 1.0 && bRR == 1.0) {
    double u[4];
    LoadCons(in, c, u);
    KbnRegularUpdate(u, Fl, Fr, dt_over_dx, res);
    return;
  }
  const FaceFrac fl = Fractions(bLL, bL, bR);
  const FaceFrac fr = Fractions(bL, bR, bRR);
  const int64_t plane = g.ngx * g.ngy;
  const int64_t cl = g.cell(i - di, j - dj), cr = g.cell(i + di, j + dj);
  const double* xc = g.centroids + DIR * plane;
  const bool affL = fl.sl_rel != 0.0 || fl.sr_rel != 0.0;
  const bool affR = fr.sl_rel != 0.0 || fr.sr_rel != 0.0;
  // boundary-side blend weights of the two faces (ComputeStableFluxes_Row)
  const double dL_l = affL ? Clamp01(0.5 - xc[cl]) : 0.5;
  const double dR_l = affL ? Clamp01(0.5 + xc[c]) : 0.5;
  const double dL_r = affR ? Clamp01(0.5 - xc[c]) : 0.5;
  const double dR_r = affR ? Clamp01(0.5 + xc[cr]) : 0.5;
  const double center = xc[c];
  for (int k = 0; k < 4; ++k) {
    const double fbl = affL ? fb[k * plane + cl] : 0.0;
    const double fbc = (affL || affR || bL != bR) ? fb[k * plane + c] : 0.0;
    const double fbr = affR ? fb[k * plane + cr] : 0.0;
    // stabilised fluxes and the shielded ones of both faces
    const double sL_l = dL_l * Fl[k] + (1.0 - dL_l) * fbl;
    const double sR_l = dR_l * Fl[k] + (1.0 - dR_l) * fbc;
    const double Fs_l = fl.us_rel * Fl[k] + fl.sl_rel * sL_l + fl.sr_rel * sR_l;
    const double sL_r = dL_r * Fr[k] + (1.0 - dL_r) * fbc;
    const double sR_r = dR_r * Fr[k] + (1.0 - dR_r) * fbr;
    const double Fs_r = fr.us_rel * Fr[k] + fr.sl_rel * sL_r + fr.sr_rel * sR_r;
    const double prev = in.u[k][c];
    double next;
    if (second) {
      next = (bL == bR) ? prev + dt_over_dx * (Fs_l - Fs_r)
                        : prev + DivRn(dt_over_dx, alpha) * (bL * Fs_l - bR * Fs_r - (bL - bR) * fbc);
    } else if (bL == bR) {
      next = prev + dt_over_dx * (Fs_l - Fs_r);
    } else if (bL > bR) {
      const double dLg = Clamp01(DivRn(fl.sr * (0.5 + center), alpha));
      next = prev + dt_over_dx * DivRn(fl.us, alpha) * (Fl[k] - Fs_r) +
             dt_over_dx * DivRn(fl.sl, alpha) * (sL_l - Fs_r) + dt_over_dx * dLg * (Fl[k] - fbc);
    } else {
      const double dRg = Clamp01(DivRn(fr.sl * (0.5 - center), alpha));
      next = prev + dt_over_dx * DivRn(fr.us, alpha) * (Fs_l - Fr[k]) +
             dt_over_dx * DivRn(fr.sr, alpha) * (Fs_l - sR_r) + dt_over_dx * dRg * (fbc - Fr[k]);
    }
    res[k] = next;
  }
  // the masked reconstruction turns a cell without positive density into
  // the zero state
  const bool keep = res[0] > 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) res[k] = keep ? res[k] : 0.0;
}

// One fused cut-cell split step in X (the regular masked MUSCL-Hancock +
// HLLEM fluxes, their masked CFL values and the h-box update in one pass
// over the level, with no face-flux arrays in memory): a wave covers 64
// consecutive cells of one row, lanes 2..61 own theirs; each lane completes
// and reconstructs its cell once, neighbours come from DPP lane shifts, the
// flux of a lane's left face is shifted to its left neighbour as that one's
// right-face flux.  Ghost columns read the transmissive state.
constexpr int kKbnOwned = 60;

// Flux arithmetic of the fused cut-cell sweeps: T = double with IEEE
// division and square root (bit-exact), or T = float with the hardware
// estimates (the FP32-flux form).  The masked CFL, the h-box update with its
// divisions by the volume fraction and the accumulation stay FP64.
template <class T> __device__ __forceinline__ Gas<T> GasAs(const Gas<double>& g) {
  Gas<T> r;
  r.gamma = static_cast<T>(g.gamma);
  r.gm1inv = static_cast<T>(g.gm1inv);
  r.gm1 = static_cast<T>(g.gm1);
  return r;
}

template <int LIM, bool LARR, class T, bool FAST>
__global__ __launch_bounds__(kBlock) void KbnSweepXKernel(LevelState in, LevelState out,
                                                          LevelGeom g, int64_t nx, int64_t ny,
                                                          double dx, const double* dt_ptr,
                                                          Gas<double> gas,
                                                          unsigned long long* partials,
                                                          const int* halt) {
  using TA = typename Policies<T, FAST>::TA;
  if (*halt) return;
  const Gas<T> gasT = GasAs<T>(gas);
  const int lane = threadIdx.x & 63;
  const int64_t nseg = (nx + kKbnOwned - 1) / kKbnOwned;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const bool wave_ok = wave < nseg * ny;
  const int64_t w = wave_ok ? wave : 0;
  const int64_t j = w / nseg;
  const int64_t i = (w % nseg) * kKbnOwned + lane - 2;  // -2 .. nx + 61
  const int64_t ic = i < nx + 1 ? i : nx + 1;            // stay inside the ghost ring
  const int64_t c = g.cell(ic, j);
  const int32_t code = g.code[0][c];
  double q[6];
  LoadCons(in, g.state_cell(ic, j), q);
  CompleteMasked<2>(q, CellCode::Pos(code), gas);
  double qm[6], qp[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    qm[k] = ShiftUp(q[k]);
    qp[k] = ShiftDown(q[k]);
  }
  // Every lane must take part in a DPP shift (an inactive source lane yields
  // the reader's own value): shift first, combine without short-circuit.
  const int code_m = DppUp(code), code_p = DppDown(code);
  const bool mask = CellCode::Pos(code_m) & CellCode::Pos(code) & CellCode::Pos(code_p);
  const double dt = *dt_ptr;
  const double lh = DivRn(0.5 * dt, dx);
  T recL[6], recR[6];
  {
    T tm[4], t0[4], tp[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      tm[k] = static_cast<T>(qm[k]);
      t0[k] = static_cast<T>(q[k]);
      tp[k] = static_cast<T>(qp[k]);
    }
    MaskedRecon<2, 0, LIM, T, TA>(tm, t0, tp, mask, static_cast<T>(lh), gasT, recL, recR);
  }
  T recRm[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) recRm[k] = ShiftUp(recR[k]);
  // left face of the lane's cell: faces 0 .. nx, lanes 2 .. 62
  const bool face_lane = wave_ok && lane >= 2 && lane <= 2 + kKbnOwned && i >= 0 && i <= nx;
  double F[4] = {0.0, 0.0, 0.0, 0.0};
  unsigned long long key = 0x7ff0000000000000ull;
  if (face_lane && CellCode::Face(code)) {
    T Ft[4];
    HllemFlux<2, 0, LARR, false, T, TA>(Ft, recRm, recL, gasT);
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k] = static_cast<double>(Ft[k]);
    key = DtKey(HllemStableDt<2, 0>(qm, q, dx, gas));
  }
  double Fr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) Fr[k] = ShiftDown(F[k]);
  if (wave_ok && lane >= 2 && lane < 2 + kKbnOwned && i >= 0 && i < nx) {
    double u[4], res[4];
    LoadCons(in, c, u);
    if (KbnSweepUpdate<0>(g, code, u, F, Fr, DivRn(dt, dx), res)) {
#pragma unroll
      for (int k = 0; k < 4; ++k) out.u[k][c] = res[k];
    }
  }
  const unsigned long long m = BlockMinKey(key);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// The same step in Y through an LDS transpose.  A workgroup stages a tile of
// kYtCols columns by 64 rows (rows j0 - 2 .. j0 + 61, clamped into the ghost
// ring) with coalesced loads: the conservative state and the cell words.
// Each wave then takes one column at a time with lane l on row j0 - 2 + l
// and runs the X kernel's lane-parallel step along it (neighbours by DPP,
// lanes 2 .. 61 own their rows), writes the updates back into its column of
// the tile, and the workgroup stores the owned rows with coalesced stores.
constexpr int kYtCols = 16;   // columns per tile (one 128-byte row segment)
constexpr int kYtPitch = 68;  // LDS doubles per tile column (64 rows + padding:
                              // row-wise staging writes are 2-way conflicted at most)

template <int LIM, bool LARR, class T, bool FAST>
__global__ __launch_bounds__(kBlock) void KbnSweepYKernel(LevelState in, LevelState out,
                                                          LevelGeom g, int64_t nx, int64_t ny,
                                                          double dx, const double* dt_ptr,
                                                          Gas<double> gas,
                                                          unsigned long long* partials,
                                                          const int* halt) {
  using TA = typename Policies<T, FAST>::TA;
  __shared__ double s_u[4][kYtCols][kYtPitch];
  __shared__ int32_t s_code[kYtCols][kYtPitch];
  if (*halt) return;
  const Gas<T> gasT = GasAs<T>(gas);
  // Tiles in x-fastest order; consecutive tiles (which share cache lines at
  // their edges) go to the same XCD (workgroup b runs on XCD b % 8).
  const int64_t ncols = (nx + kYtCols - 1) / kYtCols;
  const int64_t ntiles = ncols * ((ny + kKbnOwned - 1) / kKbnOwned);
  const int64_t nt8 = ntiles / 8 * 8;
  const int64_t b = blockIdx.x;
  const int64_t tile = b < nt8 ? (b % 8) * (nt8 / 8) + b / 8 : b;
  const int64_t i0 = (tile % ncols) * kYtCols;
  const int64_t j0 = (tile / ncols) * kKbnOwned;
  for (int e = threadIdx.x; e < 64 * kYtCols; e += kBlock) {
    const int r = e / kYtCols, cc = e % kYtCols;
    const int64_t jr = j0 - 2 + r;
    const int64_t jj = jr < ny + 1 ? jr : ny + 1;
    const int64_t ii = i0 + cc < nx ? i0 + cc : nx - 1;
    const int64_t cs = g.state_cell(ii, jj);
#pragma unroll
    for (int k = 0; k < 4; ++k) s_u[k][cc][r] = in.u[k][cs];
    s_code[cc][r] = g.code[1][g.cell(ii, jj)];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int64_t j = j0 - 2 + lane;  // the lane's row, -2 .. ny + 61
  const double dt = *dt_ptr;
  const double lh = DivRn(0.5 * dt, dx);
  const double dt_over_dx = DivRn(dt, dx);
  const bool face_row = lane >= 2 && lane <= 2 + kKbnOwned && j >= 0 && j <= ny;
  const bool own_row = lane >= 2 && lane < 2 + kKbnOwned && j < ny;
  unsigned long long key = 0x7ff0000000000000ull;
#pragma unroll 1
  for (int cc = threadIdx.x >> 6; cc < kYtCols; cc += kBlock / 64) {
    const bool col_ok = i0 + cc < nx;
    const int32_t code = s_code[cc][lane];
    double u[4], q[6];
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = u[k] = s_u[k][cc][lane];
    CompleteMasked<2>(q, CellCode::Pos(code), gas);
    double qm[6], qp[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      qm[k] = ShiftUp(q[k]);
      qp[k] = ShiftDown(q[k]);
    }
    const int code_m = DppUp(code), code_p = DppDown(code);
    const bool mask = CellCode::Pos(code_m) & CellCode::Pos(code) & CellCode::Pos(code_p);
    T recL[6], recR[6];
    {
      T tm[4], t0[4], tp[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        tm[k] = static_cast<T>(qm[k]);
        t0[k] = static_cast<T>(q[k]);
        tp[k] = static_cast<T>(qp[k]);
      }
      MaskedRecon<2, 1, LIM, T, TA>(tm, t0, tp, mask, static_cast<T>(lh), gasT, recL, recR);
    }
    T recRm[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) recRm[k] = ShiftUp(recR[k]);
    // lower face of the lane's cell: faces 0 .. ny, lanes 2 .. 62
    double F[4] = {0.0, 0.0, 0.0, 0.0};
    if (face_row && CellCode::Face(code)) {
      T Ft[4];
      HllemFlux<2, 1, LARR, false, T, TA>(Ft, recRm, recL, gasT);
#pragma unroll
      for (int k = 0; k < 4; ++k) F[k] = static_cast<double>(Ft[k]);
      if (col_ok) key = KeyMin(key, DtKey(HllemStableDt<2, 1>(qm, q, dx, gas)));
    }
    double Fr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) Fr[k] = ShiftDown(F[k]);
    // The column's tile entries belong to this wave alone: the update
    // replaces the state it was read from.  Irregular cells keep it (the
    // store below skips them).
    if (own_row && col_ok) {
      double res[4];
      if (KbnSweepUpdate<1>(g, code, u, F, Fr, dt_over_dx, res)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) s_u[k][cc][lane] = res[k];
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * kYtCols; e += kBlock) {
    const int r = e / kYtCols, cc = e % kYtCols;
    const int64_t jr = j0 - 2 + r, ii = i0 + cc;
    if (r < 2 || r >= 2 + kKbnOwned || jr >= ny || ii >= nx) continue;
    if (CellCode::Kind(s_code[cc][r]) >= CellCode::kSlot0) continue;
    const int64_t c = g.cell(ii, jr);
#pragma unroll
    for (int k = 0; k < 4; ++k) out.u[k][c] = s_u[k][cc][r];
  }
  const unsigned long long m = BlockMinKey(key);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// The h-box update (KbnUpdateCell) of the irregular cells of direction DIR,
// one thread per listed cell, from the two face fluxes the fused sweep left
// in irr_flux; runs after that sweep on the same stream.
template <int DIR>
__global__ __launch_bounds__(kBlock) void KbnIrregularKernel(LevelState in, LevelState out,
                                                             LevelGeom g, const int32_t* cells,
                                                             int n, const double* fb,
                                                             const double* dt_ptr, double dx,
                                                             int second, const int* halt) {
  if (*halt) return;
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= n) return;
  const int64_t c = cells[q];
  const int64_t i = c % g.ngx - kG, j = c / g.ngx - kG;
  const double* f = g.irr_flux[DIR] + 8 * static_cast<int64_t>(q);
  double Fl[4], Fr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    Fl[k] = f[k];
    Fr[k] = f[4 + k];
  }
  double res[4];
  KbnUpdateCell<DIR>(in, g, fb, i, j, Fl, Fr, DivRn(*dt_ptr, dx), second, res);
#pragma unroll
  for (int k = 0; k < 4; ++k) out.u[k][c] = res[k];
}

// Workgroups (= partial minima) of a fused cut-cell sweep.
int64_t KbnSweepBlocks(int dir, int64_t nx, int64_t ny) {
  if (dir == 0) {
    const int64_t waves = ((nx + kKbnOwned - 1) / kKbnOwned) * ny;
    return (waves + kBlock / 64 - 1) / (kBlock / 64);
  }
  return ((nx + kYtCols - 1) / kYtCols) * ((ny + kKbnOwned - 1) / kKbnOwned);
}

// Masked reconstruction of the interior at set_state: covered cells and
// cells without positive density become the zero state.
__global__ void LevelMaskKernel(LevelState st, LevelGeom g, int64_t nx, int64_t ny) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= nx * ny) return;
  const int64_t c = g.cell(q % nx, q / nx);
  if (g.alpha[c] > 0.0 && st.u[0][c] > 0.0) return;
#pragma unroll
  for (int k = 0; k < 4; ++k) st.u[k][c] = 0.0;
}

__global__ void CopyGeomKernel(DevView v, int comp, int64_t lo0, int64_t lo1, int64_t nx,
                               int64_t ny, double* dst) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= nx * ny) return;
  const int64_t i = q % nx, j = q / nx;
  dst[q] = *v.at(lo0 + i, lo1 + j, 0, comp);
}

} // namespace
} // namespace fvs

struct fvs_kbn_level {
  fvs_handle* h;
  fvs_patch_desc patch;
  int integrator;
  int64_t nx, ny, ngx, ngy;
  double* buf[3][4];
  double* geom;  // alpha | beta_x | beta_y | normals(2) | centroids(2)
  LevelGeom lg;
  int32_t* cut;
  int ncut;
  double* refs;
  double* fb;
  double* F;
  unsigned long long* partials;
  int64_t nblocks;
  Ctl* ctl;
  int cur;
  int32_t* code[2];       // LevelGeom::code
  int32_t* irr_cells[2];  // cell index of each irregular slot
  double* irr_flux[2];
  int nirr[2];
};

namespace {

LevelState StateOf(fvs_kbn_level* L, int b) {
  LevelState s;
  for (int k = 0; k < 4; ++k) s.u[k] = L->buf[b][k];
  return s;
}

} // namespace

extern "C" {

int fvs_kbn_level_create(fvs_handle* h, const fvs_patch_desc* patch,
                         const fvs_cutcell_data* geom, int integrator, fvs_kbn_level** out) {
  if (!h || !patch || !geom || !out || (integrator != 0 && integrator != 1)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  *out = nullptr;
  if (h->desc.dim != 2 || h->desc.base == FVS_BASE_GODUNOV_EXACT ||
      h->desc.base == FVS_BASE_HLL_EINFELDT) {
    SetError("the cut-cell level is 2D with an HLLEM base method");
    return FVS_BAD_ARGUMENT;
  }
  if (h->desc.reconstruction != FVS_RECON_MUSCL_HANCOCK) {
    SetError("the cut-cell level uses MUSCL-Hancock reconstruction");
    return FVS_BAD_ARGUMENT;
  }
  const int64_t nx = patch->n[0], ny = patch->n[1];
  if (nx < 4 || ny < 4 || nx * ny > (int64_t(1) << 31)) {
    SetError("patch size out of range");
    return FVS_BAD_ARGUMENT;
  }
  const int64_t ngx = nx + 2 * kG, ngy = ny + 2 * kG;
  const int64_t lo[3] = {-kG, -kG, 0}, hi[3] = {nx - 1 + kG, ny - 1 + kG, 0};
  const int64_t bxhi[3] = {nx + kG, ny - 1 + kG, 0}, byhi[3] = {nx - 1 + kG, ny + kG, 0};
  if (!ContainsBox(&geom->volume_fractions, lo, hi) ||
      !ContainsBox(&geom->face_fractions[0], lo, bxhi) ||
      !ContainsBox(&geom->face_fractions[1], lo, byhi) ||
      !ContainsBox(&geom->boundary_normals, lo, hi) || geom->boundary_normals.ncomp < 2 ||
      !ContainsBox(&geom->boundary_centroids, lo, hi) || geom->boundary_centroids.ncomp < 2) {
    SetError("geometry must cover the patch grown by two ghost layers");
    return FVS_BAD_ARGUMENT;
  }
  fvs_kbn_level* L = new (std::nothrow) fvs_kbn_level();
  if (!L) {
    SetError("out of host memory");
    return FVS_BAD_ARGUMENT;
  }
  std::memset(L, 0, sizeof *L);
  L->h = h;
  L->patch = *patch;
  L->integrator = integrator;
  L->nx = nx;
  L->ny = ny;
  L->ngx = ngx;
  L->ngy = ngy;
  hipStream_t s = h->stream;
  const int64_t ncg = ngx * ngy;
  auto fail = [&](int rc) {
    fvs_kbn_level_destroy(L);
    return rc;
  };
  for (int b = 0; b < 3; ++b)
    for (int k = 0; k < 4; ++k)
      if (hipMalloc(&L->buf[b][k], ncg * sizeof(double)) != hipSuccess)
        return fail(FVS_HIP_ERROR);
  const int64_t nbx = ngy * (ngx + 1), nby = (ngy + 1) * ngx;
  const int64_t gsize = ncg + nbx + nby + 4 * ncg;
  if (hipMalloc(&L->geom, gsize * sizeof(double)) != hipSuccess) return fail(FVS_HIP_ERROR);
  double* ga = L->geom;
  double* gbx = ga + ncg;
  double* gby = gbx + nbx;
  double* gn = gby + nby;
  double* gc = gn + 2 * ncg;
  // dense copies of the caller's geometry views
  auto copy = [&](const fvs_view* v, int comp, int64_t cx, int64_t cy, double* dst) {
    DevView dv;
    if (MakeDevView(v, comp + 1, &dv)) return FVS_BAD_ARGUMENT;
    hipLaunchKernelGGL(CopyGeomKernel, dim3(Blocks(cx * cy)), dim3(kBlock), 0, s, dv, comp,
                       -kG, -kG, cx, cy, dst);
    return hipGetLastError() == hipSuccess ? FVS_OK : FVS_HIP_ERROR;
  };
  int rc = copy(&geom->volume_fractions, 0, ngx, ngy, ga);
  if (!rc) rc = copy(&geom->face_fractions[0], 0, ngx + 1, ngy, gbx);
  if (!rc) rc = copy(&geom->face_fractions[1], 0, ngx, ngy + 1, gby);
  if (!rc) rc = copy(&geom->boundary_normals, 0, ngx, ngy, gn);
  if (!rc) rc = copy(&geom->boundary_normals, 1, ngx, ngy, gn + ncg);
  if (!rc) rc = copy(&geom->boundary_centroids, 0, ngx, ngy, gc);
  if (!rc) rc = copy(&geom->boundary_centroids, 1, ngx, ngy, gc + ncg);
  if (rc) return fail(rc);
  L->lg.alpha = ga;
  L->lg.beta[0] = gbx;
  L->lg.beta[1] = gby;
  L->lg.normals = gn;
  L->lg.centroids = gc;
  L->lg.ngx = ngx;
  L->lg.ngy = ngy;
  // geometry words of every cell (CellCode) and the irregular cells per
  // direction
  {
    std::vector<double> ha(ncg), hbx(nbx), hby(nby);
    if (hipMemcpyAsync(ha.data(), ga, ncg * sizeof(double), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(hbx.data(), gbx, nbx * sizeof(double), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(hby.data(), gby, nby * sizeof(double), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(FVS_HIP_ERROR);
    for (int d = 0; d < 2; ++d) {
      // fraction of face f (ghosted index) of the row or column through cell (ci, cj)
      auto beta = [&](int64_t ci, int64_t cj) {
        return d == 0 ? hbx[cj * (ngx + 1) + ci] : hby[cj * ngx + ci];
      };
      std::vector<int32_t> code(ncg), cells;
      for (int64_t cj = 0; cj < ngy; ++cj)
        for (int64_t ci = 0; ci < ngx; ++ci) {
          const int64_t c = cj * ngx + ci;
          const double a = ha[c];
          int32_t kind = CellCode::kCovered;
          const bool interior = ci >= kG && ci < nx + kG && cj >= kG && cj < ny + kG;
          if (interior && a != 0.0) {
            // face fractions at faces i - 1 .. i + 2 (or j - 1 .. j + 2)
            bool regular = a == 1.0;
            for (int64_t f = -1; f <= 2 && regular; ++f)
              regular = (d == 0 ? beta(ci + f, cj) : beta(ci, cj + f)) == 1.0;
            if (regular) {
              kind = CellCode::kRegular;
            } else {
              kind = CellCode::kSlot0 + static_cast<int32_t>(cells.size());
              cells.push_back(static_cast<int32_t>(c));
            }
          }
          code[c] = CellCode::Make(a > 0.0, beta(ci, cj) > 0.0, kind);
        }
      L->nirr[d] = static_cast<int>(cells.size());
      const size_t ncells = cells.empty() ? 1 : cells.size();
      if (hipMalloc(&L->code[d], ncg * sizeof(int32_t)) != hipSuccess ||
          hipMalloc(&L->irr_cells[d], ncells * sizeof(int32_t)) != hipSuccess ||
          hipMalloc(&L->irr_flux[d], 8 * ncells * sizeof(double)) != hipSuccess)
        return fail(FVS_HIP_ERROR);
      if (hipMemcpyAsync(L->code[d], code.data(), ncg * sizeof(int32_t), hipMemcpyHostToDevice,
                         s) != hipSuccess ||
          (!cells.empty() &&
           hipMemcpyAsync(L->irr_cells[d], cells.data(), cells.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, s) != hipSuccess) ||
          hipStreamSynchronize(s) != hipSuccess)
        return fail(FVS_HIP_ERROR);
      L->lg.code[d] = L->code[d];
      L->lg.irr_flux[d] = L->irr_flux[d];
    }
  }
  // list of the interior cut cells
  int32_t* flags = nullptr;
  if (hipMalloc(&flags, nx * ny * sizeof(int32_t)) != hipSuccess) return fail(FVS_HIP_ERROR);
  {
    GeomViews gv;
    if (GeomOf(geom, 2, &gv)) {
      (void)hipFree(flags);
      return fail(FVS_BAD_ARGUMENT);
    }
    const int64_t ilo[3] = {0, 0, 0}, ihi[3] = {nx - 1, ny - 1, 0};
    const DevView box = BoxView(ilo, ihi);
    hipLaunchKernelGGL(IsCutCellKernel<2>, dim3(Blocks(box.count())), dim3(kBlock), 0, s, gv, box,
                       flags);
  }
  std::vector<int32_t> hflags(nx * ny);
  if (hipMemcpyAsync(hflags.data(), flags, nx * ny * sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipFree(flags);
    return fail(FVS_HIP_ERROR);
  }
  (void)hipFree(flags);
  std::vector<int32_t> cut;
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i)
      if (hflags[j * nx + i]) cut.push_back(static_cast<int32_t>((j + kG) * ngx + (i + kG)));
  L->ncut = static_cast<int>(cut.size());
  const int nc = L->ncut > 0 ? L->ncut : 1;
  if (hipMalloc(&L->cut, nc * sizeof(int32_t)) != hipSuccess ||
      hipMalloc(&L->refs, 6 * nc * sizeof(double)) != hipSuccess ||
      hipMalloc(&L->fb, 4 * ncg * sizeof(double)) != hipSuccess ||
      hipMalloc(&L->F, 4 * (nx + 1) * (ny + 1) * sizeof(double)) != hipSuccess ||
      hipMalloc(&L->ctl, sizeof(Ctl)) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  if (L->ncut > 0 && hipMemcpyAsync(L->cut, cut.data(), L->ncut * sizeof(int32_t),
                                    hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  if (hipMemsetAsync(L->fb, 0, 4 * ncg * sizeof(double), s) != hipSuccess ||
      hipMemsetAsync(L->refs, 0, 6 * nc * sizeof(double), s) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  for (int b = 0; b < 3; ++b)
    for (int k = 0; k < 4; ++k)
      if (hipMemsetAsync(L->buf[b][k], 0, ncg * sizeof(double), s) != hipSuccess)
        return fail(FVS_HIP_ERROR);
  L->nblocks = (std::max((nx + 1) * ny, nx * (ny + 1)) + kBlock - 1) / kBlock;
  const int64_t npart = std::max({2 * L->nblocks, KbnSweepBlocks(0, nx, ny), KbnSweepBlocks(1, nx, ny)});
  if (npart > 0x7fffffff) {
    SetError("cut-cell patch too large for the sweep grid");
    return fail(FVS_BAD_ARGUMENT);
  }
  if (hipMalloc(&L->partials, npart * sizeof(unsigned long long)) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  Ctl c{};
  c.cfl = patch->cfl;
  if (hipMemcpyAsync(L->ctl, &c, sizeof c, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  *out = L;
  return FVS_OK;
}

int fvs_kbn_level_destroy(fvs_kbn_level* L) {
  if (!L) return FVS_OK;
  (void)hipStreamSynchronize(L->h->stream);
  for (int b = 0; b < 3; ++b)
    for (int k = 0; k < 4; ++k)
      if (L->buf[b][k]) (void)hipFree(L->buf[b][k]);
  void* ps[] = {L->geom,         L->cut,          L->refs,         L->fb,
                L->F,            L->partials,     L->ctl,          L->code[0],
                L->code[1],      L->irr_cells[0], L->irr_cells[1], L->irr_flux[0],
                L->irr_flux[1]};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete L;
  return FVS_OK;
}

int fvs_kbn_level_ncut(fvs_kbn_level* L) { return L ? L->ncut : -1; }

int fvs_kbn_level_set_state(fvs_kbn_level* L, const double* const* comps, double t0) {
  if (!L || !comps) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  hipStream_t s = L->h->stream;
  L->cur = 0;
  for (int k = 0; k < 4; ++k) {
    if (!comps[k]) {
      SetError("null component");
      return FVS_BAD_ARGUMENT;
    }
    FVS_HIP_CHECK(hipMemcpy2DAsync(L->buf[0][k] + kG * L->ngx + kG, L->ngx * sizeof(double),
                                   comps[k], L->nx * sizeof(double), L->nx * sizeof(double), L->ny,
                                   hipMemcpyDeviceToDevice, s));
  }
  hipLaunchKernelGGL(LevelMaskKernel, dim3(Blocks(L->nx * L->ny)), dim3(kBlock), 0, s,
                     StateOf(L, 0), L->lg, L->nx, L->ny);
  FVS_LAUNCH_CHECK();
  Ctl c{};
  c.t = t0;
  c.cfl = L->patch.cfl;
  FVS_HIP_CHECK(hipMemcpyAsync(L->ctl, &c, sizeof c, hipMemcpyHostToDevice, s));
  FVS_HIP_CHECK(hipStreamSynchronize(s));
  return FVS_OK;
}

int fvs_kbn_level_get_state(fvs_kbn_level* L, double* const* comps) {
  if (!L || !comps) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  hipStream_t s = L->h->stream;
  for (int k = 0; k < 4; ++k)
    FVS_HIP_CHECK(hipMemcpy2DAsync(comps[k], L->nx * sizeof(double), L->buf[L->cur][k] + kG * L->ngx + kG,
                                   L->ngx * sizeof(double), L->nx * sizeof(double), L->ny,
                                   hipMemcpyDeviceToDevice, s));
  FVS_HIP_CHECK(hipStreamSynchronize(s));
  return FVS_OK;
}

int fvs_kbn_level_time(fvs_kbn_level* L, double* t, int64_t* cycle) {
  if (!L) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  Ctl c;
  FVS_HIP_CHECK(hipMemcpyAsync(&c, L->ctl, sizeof c, hipMemcpyDeviceToHost, L->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(L->h->stream));
  if (t) *t = c.t;
  if (cycle) *cycle = c.cycle;
  return FVS_OK;
}

// Enqueues one split sweep (input buffer src, output dst) of direction dir.
static int KbnSweep(fvs_kbn_level* L, int dir, int src, int dst, bool check) {
  fvs_handle* h = L->h;
  hipStream_t s = h->stream;
  const LevelState in = StateOf(L, src), out = StateOf(L, dst);
  const int* halt = &L->ctl->halt;
  Gas<double> gas;
  gas.gamma = h->desc.gamma;
  gas.gm1inv = 1.0 / (h->desc.gamma - 1.0);
  gas.gm1 = h->desc.gamma - 1.0;
  const BoundaryParams bp = BoundaryOf(h->desc);
  if (L->ncut > 0) {
    hipLaunchKernelGGL(LevelBoundaryFluxKernel, dim3(Blocks(L->ncut)), dim3(kBlock), 0, s, in,
                       L->lg, L->cut, L->ncut, bp, L->refs, L->fb, dir, halt);
    FVS_LAUNCH_CHECK();
  }
  const double dx = L->patch.dx[dir];
  const bool larr = h->desc.base == FVS_BASE_HLLEM_LARROUTUROU;
  const double* dtp = &L->ctl->split_dt;
  const int64_t nb = KbnSweepBlocks(dir, L->nx, L->ny);
  const dim3 grid(static_cast<unsigned>(nb));
  const bool fp32 = h->desc.precision == FVS_FP32_FLUX;
#define FVS_KS(K, LM, LA)                                                                     \
  if (fp32)                                                                                   \
    hipLaunchKernelGGL((K<LM, LA, float, true>), grid, dim3(kBlock), 0, s, in, out, L->lg,    \
                       L->nx, L->ny, dx, dtp, gas, L->partials, halt);                        \
  else                                                                                        \
    hipLaunchKernelGGL((K<LM, LA, double, false>), grid, dim3(kBlock), 0, s, in, out, L->lg,  \
                       L->nx, L->ny, dx, dtp, gas, L->partials, halt)
#define FVS_KS_LIM(K, LA)                                  \
  switch (h->desc.limiter) {                               \
  case kMinMod: FVS_KS(K, kMinMod, LA); break;             \
  case kVanLeer: FVS_KS(K, kVanLeer, LA); break;           \
  case kNoLimiter: FVS_KS(K, kNoLimiter, LA); break;       \
  default: FVS_KS(K, kNoGradient, LA); break;              \
  }
  if (dir == 0) {
    if (larr) { FVS_KS_LIM(KbnSweepXKernel, true) } else { FVS_KS_LIM(KbnSweepXKernel, false) }
  } else {
    if (larr) { FVS_KS_LIM(KbnSweepYKernel, true) } else { FVS_KS_LIM(KbnSweepYKernel, false) }
  }
#undef FVS_KS_LIM
#undef FVS_KS
  FVS_LAUNCH_CHECK();
  if (L->nirr[dir] > 0) {
    if (dir == 0)
      hipLaunchKernelGGL(KbnIrregularKernel<0>, dim3(Blocks(L->nirr[0])), dim3(kBlock), 0, s, in,
                         out, L->lg, L->irr_cells[0], L->nirr[0], L->fb, dtp, dx, L->integrator,
                         halt);
    else
      hipLaunchKernelGGL(KbnIrregularKernel<1>, dim3(Blocks(L->nirr[1])), dim3(kBlock), 0, s, in,
                         out, L->lg, L->irr_cells[1], L->nirr[1], L->fb, dtp, dx, L->integrator,
                         halt);
    FVS_LAUNCH_CHECK();
  }
  if (check) {
    if (int rc = LaunchFinalizeCheck(s, L->partials, nb, nullptr, L->ctl)) return rc;
  }
  return FVS_OK;
}

int fvs_kbn_level_advance(fvs_kbn_level* L, int max_steps, double final_time, int batch,
                          double* dt_log, int* steps_done, int* failures) {
  if (!L || max_steps < 0 || batch < 1) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  fvs_handle* h = L->h;
  hipStream_t s = h->stream;
  const bool strang = L->patch.splitting == FVS_SPLIT_STRANG;
  Gas<double> gas;
  gas.gamma = h->desc.gamma;
  gas.gm1inv = 1.0 / (h->desc.gamma - 1.0);
  gas.gm1 = h->desc.gamma - 1.0;
  const BoundaryParams bp = BoundaryOf(h->desc);
  Ctl c;
  FVS_HIP_CHECK(hipMemcpyAsync(&c, L->ctl, sizeof c, hipMemcpyDeviceToHost, s));
  FVS_HIP_CHECK(hipStreamSynchronize(s));
  double* dlog = nullptr;
  if (dt_log && max_steps > 0) FVS_HIP_CHECK(hipMalloc(&dlog, max_steps * sizeof(double)));
  c.final_time = final_time;
  c.eps = SmallestDt(L->patch);
  c.cfl = L->patch.cfl;
  c.halt = 0;
  c.dt_log = dlog;
  c.dt_log_base = c.cycle;
  c.dt_log_len = max_steps;
  FVS_HIP_CHECK(hipMemcpyAsync(L->ctl, &c, sizeof c, hipMemcpyHostToDevice, s));
  const int64_t cycle0 = c.cycle;
  int64_t cycle = c.cycle;
  int nfail = 0;        // restarts in this call
  int consecutive = 0;  // restarts since the last completed step
  int rc = FVS_OK;
  std::vector<int> start_buf;
  const int* halt = &L->ctl->halt;
  while (cycle - cycle0 < max_steps) {
    const int k = static_cast<int>(std::min<int64_t>(batch, max_steps - (cycle - cycle0)));
    start_buf.assign(k + 1, 0);
    int cur = L->cur;
    for (int sidx = 0; sidx < k && rc == FVS_OK; ++sidx) {
      start_buf[sidx] = cur;
      const int64_t cyc = cycle + sidx;
      const int is_odd = static_cast<int>(cyc % 2);
      const int dirs[2] = {is_odd ? 1 : 0, is_odd ? 0 : 1};
      int seq[4];
      int ns = 0;
      seq[ns++] = dirs[0];
      seq[ns++] = dirs[1];
      if (strang) {
        seq[ns++] = dirs[1];
        seq[ns++] = dirs[0];
      }
      // step-start CFL of both directions
      const LevelState st0 = StateOf(L, cur);
      const int64_t dt2_tiles = ((L->nx + 1 + kBlock - 2) / (kBlock - 1)) *
                                ((L->ny + 1 + kDt2Seg - 1) / kDt2Seg);
      int64_t npart = L->nblocks;  // partial minima per direction
      if (dt2_tiles <= L->nblocks) {
        npart = dt2_tiles;
        hipLaunchKernelGGL(LevelDt2Kernel, dim3(unsigned(dt2_tiles)), dim3(kBlock), 0, s, st0,
                           L->lg, L->nx, L->ny, L->patch.dx[0], L->patch.dx[1], gas, L->partials,
                           dt2_tiles, halt);
      } else {  // very thin levels: one pass per direction
        hipLaunchKernelGGL(LevelDtKernel<0>, dim3(unsigned(L->nblocks)), dim3(kBlock), 0, s, st0,
                           L->lg, L->nx, L->ny, L->patch.dx[0], gas, L->partials, halt);
        hipLaunchKernelGGL(LevelDtKernel<1>, dim3(unsigned(L->nblocks)), dim3(kBlock), 0, s, st0,
                           L->lg, L->nx, L->ny, L->patch.dx[1], gas, L->partials + L->nblocks,
                           halt);
      }
      rc = LaunchFinalizeCfl(s, L->partials, npart, 2, nullptr, L->ctl, seq[0], strang);
      if (rc) break;
      if (L->ncut > 0) {
        hipLaunchKernelGGL(LevelRefsKernel, dim3(Blocks(L->ncut)), dim3(kBlock), 0, s, st0, L->lg,
                           L->cut, L->ncut, bp, L->refs, halt);
      }
      const int b0 = cur, b1 = (cur + 1) % 3, b2 = (cur + 2) % 3;
      int src = b0, dst = b1;
      for (int w = 0; w < ns && rc == FVS_OK; ++w) {
        rc = KbnSweep(L, seq[w], src, dst, w > 0);
        src = dst;
        dst = (dst == b1) ? b2 : b1;
      }
      if (rc) break;
      cur = src;
      rc = LaunchStepEnd(s, L->ctl);
    }
    if (rc) break;
    FVS_HIP_CHECK(hipMemcpyAsync(&c, L->ctl, sizeof c, hipMemcpyDeviceToHost, s));
    FVS_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t completed = c.cycle - cycle;
    if (completed > 0) consecutive = 0;
    if (c.halt == 0) {
      L->cur = cur;
      cycle = c.cycle;
      continue;
    }
    L->cur = start_buf[completed];
    cycle = c.cycle;
    if (c.halt == 2) break;
    ++nfail;
    ++consecutive;
    c.halt = 0;
    c.have_failure = 1;
    FVS_HIP_CHECK(hipMemcpyAsync(L->ctl, &c, sizeof c, hipMemcpyHostToDevice, s));
    if (consecutive > 64) {
      SetError("more than 64 consecutive CFL failures");
      rc = FVS_TIME_STEP_TOO_LARGE;
      break;
    }
  }
  if (dlog) {
    if (rc == FVS_OK && cycle > cycle0)
      (void)hipMemcpyAsync(dt_log, dlog, (cycle - cycle0) * sizeof(double), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    (void)hipFree(dlog);
  }
  if (steps_done) *steps_done = static_cast<int>(cycle - cycle0);
  if (failures) *failures = nfail;
  return rc;
}

} // extern "C"
