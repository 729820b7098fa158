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
// fractions in particular).
#include "fvs_internal.hpp"
#include "fvs_physics.hpp"
#include "fvs_riemann.hpp"
#include "fvs_views.hpp"

#include <cfloat>

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
        r[i][j] = ((i == j ? 1.0 : 0.0) + sk[i][j]) + sq / den;
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
  const double Minv = 1.0 / sqrt(Msq);
  const double rhoL = L[0], rhoR = R[0];
  double vL_[D], vR_[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    vL_[i] = L[M + i] / rhoL;
    vR_[i] = R[M + i] / rhoR;
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
  const double uL = L[M + d] / rhoL;
  const double uR = R[M + d] / rhoR;
  const double aL = L[D + 3], aR = R[D + 3];
  const double rhoEL = L[D + 1], rhoER = R[D + 1];
  const double pL = L[D + 2], pR = R[D + 2];
  const double hL = (rhoEL + pL) / rhoL;
  const double hR = (rhoER + pR) / rhoR;
  const double sqRhoL = sqrt(rhoL);
  const double sqRhoR = sqrt(rhoR);
  const double sqRhoSum = sqRhoL + sqRhoR;
  double roeU[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double vL = L[M + i] / rhoL;
    const double vR = R[M + i] / rhoR;
    roeU[i] = (sqRhoL * vL + sqRhoR * vR) / sqRhoSum;
  }
  const double roeU0 = roeU[d];
  const double roeH = (sqRhoL * hL + sqRhoR * hR) / sqRhoSum;
  const double roeGamma = (sqRhoL * g.gamma + sqRhoR * g.gamma) / sqRhoSum;
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
  for (int k = 0; k < D + 2; ++k) w[k] = (bR * R[k] - bL * L[k] + fluxL[k] - fluxR[k]) / db_positive;
  const double squaredNormRoeU_half = 0.5 * (Msq * sqU);
  const double u_bar = 0.5 * (sR + sL);
  const double delta = roeA / (roeA + fabs(u_bar));
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
    const double l21 = gm1 / roeA2 * (roeH - Msq * sqU);
    const double l22 = gm1 / roeA2 * Msq * roeU[0];
    const double l23 = gm1 / roeA2 * roeU[1];
    const double l24 = gm1 / roeA2 * (-1.0);
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
    const double l21 = gm1 / roeA2 * (roeH - Msq * sqU);
    const double l22 = gm1 / roeA2 * Msq * roeU[0];
    const double l23 = gm1 / roeA2 * roeU[1];
    const double l24 = gm1 / roeA2 * roeU[2];
    const double l25 = gm1 / roeA2 * (-1.0);
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
// the lane is valid with positive density.
template <int D>
__device__ __forceinline__ void CompleteMasked(double* q, bool mask, const Gas<double>& g) {
  if (mask && q[0] > 0.0) {
    CompleteFromCons<D>(q, g);
  } else {
#pragma unroll
    for (int k = 0; k < D + 4; ++k) q[k] = 0.0;
  }
}

// Masked PerfectGas::Flux: zero for the zero state (density 1 stands in
// for invalid lanes, so every product vanishes).
template <int D, int DIR>
__device__ __forceinline__ void FluxMasked(double* f, const double* q) {
  if (q[0] > 0.0) {
    PhysFlux<D, DIR>(f, q);
  } else {
#pragma unroll
    for (int k = 0; k < D + 2; ++k) f[k] = 0.0;
  }
}

// Masked MUSCL-Hancock half-step states of one cell: slopes are zero unless
// the three stencil cells are all uncovered.
template <int D, int DIR, int LIM>
__device__ __forceinline__ void MaskedRecon(const double* um, const double* u0, const double* up,
                                            bool slope_mask, double lh, const Gas<double>& g,
                                            double* recL, double* recR) {
  constexpr int NC = D + 2, NQ = D + 4;
  double qL[NQ], qR[NQ];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const double s = slope_mask ? LimitedSlope<LIM>(um[k], u0[k], up[k]) : 0.0;
    qL[k] = u0[k] - 0.5 * s;
    qR[k] = u0[k] + 0.5 * s;
  }
  CompleteMasked<D>(qL, true, g);
  CompleteMasked<D>(qR, true, g);
  double fL[NC], fR[NC];
  FluxMasked<D, DIR>(fL, qL);
  FluxMasked<D, DIR>(fR, qR);
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const double dq = lh * (fL[k] - fR[k]);
    recR[k] = qR[k] + dq;
    recL[k] = qL[k] + dq;
  }
  CompleteMasked<D>(recR, true, g);
  CompleteMasked<D>(recL, true, g);
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
      const double lh = 0.5 * fp.dt / fp.dx;
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
  const double betaL_us_over_alpha = *v.us.at(c, 0) / alpha;
  const double betaL_sL_over_alpha = *v.sl.at(c, 0) / alpha;
  const double dLg = Clamp01((*v.sr.at(c, 0) * (0.5 + center)) / alpha);
  const double betaR_us_over_alpha = *v.us.at(r, 0) / alpha;
  const double betaR_sR_over_alpha = *v.sr.at(r, 0) / alpha;
  const double dRg = Clamp01((*v.sl.at(r, 0) * (0.5 - center)) / alpha);
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
      next = prev + dt_over_dx / volfrac * (bL * fL - bR * fR - dB * *fb.at(c, k));
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
  *us.at(f, 1) = unshielded / betaM;
  *sl.at(f, 0) = sLeft;
  *sl.at(f, 1) = sLeft / betaM;
  *sr.at(f, 0) = sRight;
  *sr.at(f, 1) = sRight / betaM;
  *ds.at(f, 0) = doubly;
  *ds.at(f, 1) = doubly / betaM;
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
  if (needs_hllem && h->desc.base == FVS_BASE_GODUNOV_EXACT) {
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
