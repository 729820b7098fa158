// TEST INFRASTRUCTURE ONLY.  This is the FP64 CPU oracle of the split patch
// update: a restatement of FiniteVolumeSolver's perfect-gas flux methods,
// written as scalar code that follows the reference's 8-lane (SIMD, Eigen
// Array1d) semantics lane by lane, because the reference's production
// adapters run the SIMD paths (include/fub/AMReX/FluxMethodAdapter.hpp:54-56).
// Only tests, the smoke check and the CPU-baseline leg of the benchmark may
// load it.  The product library never includes, links or calls it.
//
// Arithmetic is plain IEEE binary64; the build pins -ffp-contract=off so no
// multiply-add is fused, and every expression keeps the reference's
// operation order (left-to-right, as the Eigen expression templates
// evaluate lane-wise).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

namespace fvs_oracle {

// Planted mistakes.  Tests that pin the oracle switch one of these on
// (oracle_plant) to show that their closed form notices it, then switch it
// off again; 0 (the default) runs the restated algorithm unchanged.
enum Plant {
  PLANT_NONE = 0,
  PLANT_MUSCL_HALF_STEP_SIGN = 1,  // rec = q -/+ (dt/2dx)(F(qL) - F(qR))
  PLANT_MUSCL_HALF_STEP_DROP = 2,  // rec = q (no half step)
  PLANT_HLLEM_NO_ANTIDIFFUSION = 3,  // b delta alpha_2 = 0
  PLANT_HLLEM_L21 = 4,             // l21 with |roeU|^2 / 2 instead of |roeU|^2
  PLANT_HLLEM_L2U_SIGN = 5,        // velocity coefficients of l2 negated
  PLANT_HLLEM_DELTA = 6,           // delta = roeA / (roeA + 2 |u_bar|)
  PLANT_EXACT_FAN_LEFT_U = 7,      // left fan u = 2/(g+1) (aL - (g-1)/2 uL)
  PLANT_EXACT_FAN_RIGHT_RHO = 8,   // right fan density exponent 2/g
  PLANT_HBOX_RIGHT_DISTANCE = 9,   // beta_L < beta_R: 0.5 + centre
  PLANT_HBOX_RIGHT_DB_SIGN = 10,   // beta_L < beta_R: F_R,us - F_B
};
inline int g_plant = PLANT_NONE;
inline bool Planted(int which) { return g_plant == which; }

// PerfectGas<Dim> parameters (include/fub/equations/PerfectGas.hpp:216-222):
// gamma = 1.4, gamma_minus_1_inv = 1 / (gamma - 1) evaluated at run time.
struct Gas {
  double gamma;
  double gm1inv;
  explicit Gas(double g = 1.4) : gamma(g), gm1inv(1.0 / (g - 1.0)) {}
};

template <int D> struct Cons {
  double rho;
  double mom[D];
  double E;
};

// Complete<PerfectGas<D>>: rho, momentum, energy, pressure, speed_of_sound.
template <int D> struct State {
  double rho;
  double mom[D];
  double E;
  double p;
  double a;
};

// Squared momentum norm of the array path: Eigen 3.3 reduces a column block
// of 2 or 3 rows (no packet access for a 3-row block) with its split-half
// unroller, a0 + (a1 + a2) (PerfectGas.cpp:33-38).
template <int D> inline double SquaredNormArray(const double* m) {
  if constexpr (D == 1) {
    return m[0] * m[0];
  } else if constexpr (D == 2) {
    return m[0] * m[0] + m[1] * m[1];
  } else {
    return m[0] * m[0] + (m[1] * m[1] + m[2] * m[2]);
  }
}

// Squared norm of a fixed Vector<D> in the scalar path: Eigen's linear
// vectorised reduction with SSE2 packets of two, (a0 + a1) + a2
// (PerfectGas.cpp:25-29, ExactRiemannSolver.hpp:28-31).
template <int D> inline double SquaredNormScalar(const double* m) {
  if constexpr (D == 1) {
    return m[0] * m[0];
  } else if constexpr (D == 2) {
    return m[0] * m[0] + m[1] * m[1];
  } else {
    return (m[0] * m[0] + m[1] * m[1]) + m[2] * m[2];
  }
}

// PerfectGas::CompleteFromCons, array form (src/equations/PerfectGas.cpp:112-124).
template <int D>
inline void CompleteFromConsArray(const Gas& g, State<D>& q) {
  const double e_kin = 0.5 * SquaredNormArray<D>(q.mom) / q.rho;
  const double e_int = q.E - e_kin;
  q.p = e_int / g.gm1inv;
  q.a = std::sqrt(g.gamma * q.p / q.rho);
}

// PerfectGas::CompleteFromCons, scalar form (src/equations/PerfectGas.cpp:96-110).
template <int D>
inline void CompleteFromConsScalar(const Gas& g, State<D>& q) {
  const double e_kin = 0.5 * SquaredNormScalar<D>(q.mom) / q.rho;
  const double e_int = q.E - e_kin;
  q.p = e_int / g.gm1inv;
  q.a = std::sqrt(g.gamma * q.p / q.rho);
}

// PerfectGas::CompleteFromCons, masked array form
// (src/equations/PerfectGas.cpp:126-144): invalid lanes (mask false or
// rho <= 0) become the zero state.
template <int D>
inline void CompleteFromConsMasked(const Gas& g, State<D>& q, bool mask) {
  mask = mask && (q.rho > 0.0);
  const double sq = SquaredNormArray<D>(q.mom);
  const double safe_rho_ke = mask ? q.rho : 1.0;
  const double e_kin = 0.5 * sq / safe_rho_ke;
  const double e_int = q.E - e_kin;
  State<D> out{};
  out.rho = mask ? q.rho : 0.0;
  for (int d = 0; d < D; ++d) out.mom[d] = mask ? q.mom[d] : 0.0;
  out.E = mask ? q.E : 0.0;
  out.p = mask ? e_int / g.gm1inv : 0.0;
  const double safe_rho = mask ? out.rho : 1.0;
  out.a = mask ? std::sqrt(g.gamma * out.p / safe_rho) : 0.0;
  q = out;
}

// PerfectGas::CompleteFromPrim, scalar (src/equations/PerfectGas.cpp:146-158).
template <int D>
inline State<D> CompleteFromPrim(const Gas& g, double rho, const double* v,
                                 double p) {
  State<D> q{};
  q.rho = rho;
  for (int d = 0; d < D; ++d) q.mom[d] = rho * v[d];
  q.p = p;
  const double e_kin = 0.5 * SquaredNormScalar<D>(q.mom) / q.rho;
  q.E = e_kin + p * g.gm1inv;
  q.a = std::sqrt(g.gamma * q.p / q.rho);
  return q;
}

// PerfectGas::Flux (src/equations/PerfectGas.cpp:52-78); array and scalar
// forms are the same expression.
template <int D>
inline void Flux(Cons<D>& f, const State<D>& q, int dir) {
  const double velocity = q.mom[dir] / q.rho;
  f.rho = q.mom[dir];
  for (int d = 0; d < D; ++d) f.mom[d] = velocity * q.mom[d];
  f.mom[dir] += q.p;
  f.E = velocity * (q.E + q.p);
}

// Masked array flux (src/equations/PerfectGas.cpp:80-93).
template <int D>
inline void FluxMasked(Cons<D>& f, const State<D>& q, bool mask, int dir) {
  mask = mask && (q.rho > 0.0);
  const double density = mask ? q.rho : 1.0;
  const double velocity = q.mom[dir] / density;
  f.rho = q.mom[dir];
  for (int d = 0; d < D; ++d) f.mom[d] = velocity * q.mom[d];
  f.mom[dir] += q.p;
  f.E = velocity * (q.E + q.p);
}

enum Limiter { kMinMod = 0, kVanLeer = 1, kNoLimiter = 2, kNoGradient = 3 };

// Limited slope of one component (include/fub/flux_method/MusclHancockMethod.hpp:31-187),
// array forms.
inline double LimitedSlope(int limiter, double qL, double qM, double qR) {
  const double sL = qM - qL;
  const double sR = qR - qM;
  switch (limiter) {
  case kMinMod:
    // (sR > 0).select(zero.max(sL.min(sR)), zero.min(sL.max(sR)))
    return (sR > 0) ? std::max(0.0, std::min(sL, sR))
                    : std::min(0.0, std::max(sL, sR));
  case kVanLeer: {
    const double r = 2.0 * sL * sR;
    return (r > 0.0) ? r / (sL + sR) : 0.0;
  }
  case kNoLimiter:
    return 0.5 * (sL + sR);
  default:
    return 0.0;
  }
}

template <int D> inline double& C(Cons<D>& c, int k) {
  return (&c.rho)[k];
}
template <int D> inline double C(const Cons<D>& c, int k) {
  return (&c.rho)[k];
}
template <int D> inline double& C(State<D>& c, int k) {
  return (&c.rho)[k];
}
template <int D> inline double C(const State<D>& c, int k) {
  return (&c.rho)[k];
}

// perfect_gas::Hllem<PerfectGas<D>, Larrouturou>::ComputeNumericFlux, array
// form (src/equations/perfect_gas/HllemMethod.hpp:711-1072).  Returns p*.
// quirk3d: the 3D branch subtracts b*delta*alpha_2*roeU[iy] from the normal
// momentum flux (:990-992); with quirk3d = false it uses roeU[ix].
template <int D>
inline double HllemFlux(const Gas& g, Cons<D>& flux, const State<D>& stL,
                        const State<D>& stR, int dir, bool larrouturou,
                        bool quirk3d) {
  State<D> L = stL;
  State<D> R = stR;
  const double Msq = 1.0; // euler::MaSq(PerfectGas) == 1 (EulerEquation.hpp:53-62)
  const double Minv = 1.0 / std::sqrt(Msq);
  const double rhoL = L.rho;
  const double rhoR = R.rho;
  double vL_[D > 0 ? D : 1];
  double vR_[D > 0 ? D : 1];
  double eKinRL = 0.0;
  double eKinRR = 0.0;
  if (larrouturou) {
    for (int i = 0; i < D; ++i) {
      const int dim_mask = (i == dir);
      vL_[i] = L.mom[i] / rhoL;
      vR_[i] = R.mom[i] / rhoR;
      eKinRL += static_cast<double>(1 - dim_mask) * vL_[i] * vL_[i];
      eKinRR += static_cast<double>(1 - dim_mask) * vR_[i] * vR_[i];
      L.mom[i] *= static_cast<double>(dim_mask);
      R.mom[i] *= static_cast<double>(dim_mask);
    }
    eKinRL *= 0.5 * Msq;
    eKinRR *= 0.5 * Msq;
    L.E -= rhoL * eKinRL;
    R.E -= rhoR * eKinRR;
  }
  Cons<D> fluxL, fluxR;
  Flux<D>(fluxL, L, dir);
  Flux<D>(fluxR, R, dir);

  const double aL = L.a;
  const double aR = R.a;
  const double rhoEL = L.E;
  const double rhoER = R.E;
  const double pL = L.p;
  const double pR = R.p;
  const double hL = (rhoEL + pL) / rhoL;
  const double hR = (rhoER + pR) / rhoR;
  const double sqRhoL = std::sqrt(rhoL);
  const double sqRhoR = std::sqrt(rhoR);
  const double sqRhoSum = sqRhoL + sqRhoR;
  const double sqRhoL_over_Sum = sqRhoL / sqRhoSum;
  const double sqRhoR_over_Sum = sqRhoR / sqRhoSum;

  double uL[D], uR[D], roeU[D];
  for (int i = 0; i < D; ++i) {
    uL[i] = L.mom[i] / rhoL;
    uR[i] = R.mom[i] / rhoR;
    roeU[i] = sqRhoL_over_Sum * uL[i] + sqRhoR_over_Sum * uR[i];
  }
  const double roeH = sqRhoL_over_Sum * hL + sqRhoR_over_Sum * hR;
  double squaredNormRoeU = 0.0;
  for (int i = 0; i < D; ++i) squaredNormRoeU += roeU[i] * roeU[i];
  squaredNormRoeU *= Msq;
  const double squaredNormRoeU_half = 0.5 * squaredNormRoeU;

  const double gammaL = g.gamma;
  const double gammaR = g.gamma;
  const double roeGamma = (sqRhoL * gammaL + sqRhoR * gammaR) / sqRhoSum;
  const double gm1 = roeGamma - 1.0;

  const double roeA2 = gm1 * (roeH - squaredNormRoeU_half);
  const double roeA = std::sqrt(roeA2);
  const double maxA = std::max(aL, aR);
  const double sL1 = uL[dir] - maxA * Minv;
  const double sL2 = roeU[dir] - roeA * Minv;
  const double sR1 = roeU[dir] + roeA * Minv;
  const double sR2 = uR[dir] + maxA * Minv;
  const double sL = std::min(sL1, sL2);
  const double sR = std::max(sR1, sR2);

  const double bL = std::min(sL, 0.0);
  const double bR = std::max(sR, 0.0);
  const double bLbR = bL * bR;
  const double db = bR - bL;
  const double db_positive = (db > 0.0) ? db : 1.0;

  Cons<D> hlle;
  for (int k = 0; k < D + 2; ++k) {
    C(hlle, k) = (bR * C(fluxL, k) - bL * C(fluxR, k) +
                  bLbR * (C(R, k) - C(L, k))) /
                 db_positive;
  }

  const double u_bar = 0.5 * (sR + sL);
  const double u_bar_abs = std::abs(u_bar);
  const double delta = Planted(PLANT_HLLEM_DELTA) ? roeA / (roeA + 2.0 * u_bar_abs)
                                                  : roeA / (roeA + u_bar_abs);
  const double b = Planted(PLANT_HLLEM_NO_ANTIDIFFUSION) ? 0.0 : bLbR / db_positive;
  const double l21_norm = Planted(PLANT_HLLEM_L21) ? squaredNormRoeU_half : squaredNormRoeU;
  const double l2u = Planted(PLANT_HLLEM_L2U_SIGN) ? -1.0 : 1.0;

  const double deltaRho = rhoR - rhoL;
  double deltaRhoU[D];
  for (int i = 0; i < D; ++i) deltaRhoU[i] = R.mom[i] - L.mom[i];
  const double deltaRhoE = rhoER - rhoEL;
  const double gm1_over_roeA2 = gm1 / roeA2;

  if constexpr (D == 1) {
    const double l21 = gm1_over_roeA2 * (roeH - l21_norm);
    const double l22 = l2u * gm1_over_roeA2 * Msq * roeU[0];
    const double l23 = gm1_over_roeA2 * (-1.0);
    const double alpha_2 = l21 * deltaRho + l22 * deltaRhoU[0] + l23 * deltaRhoE;
    const double b_delta_alpha_2 = b * delta * alpha_2;
    flux.rho = hlle.rho - b_delta_alpha_2;
    flux.mom[0] = hlle.mom[0] - b_delta_alpha_2 * roeU[0];
    flux.E = hlle.E - b_delta_alpha_2 * squaredNormRoeU_half;
  } else if constexpr (D == 2) {
    const int ix = dir;
    const int iy = (ix == 0);
    const double l21 = gm1_over_roeA2 * (roeH - l21_norm);
    const double l22 = l2u * gm1_over_roeA2 * Msq * roeU[0];
    const double l23 = l2u * gm1_over_roeA2 * roeU[1];
    const double l24 = -gm1_over_roeA2;
    const double alpha_2 = l21 * deltaRho + l22 * deltaRhoU[0] +
                           l23 * deltaRhoU[1] + l24 * deltaRhoE;
    const double b_delta = b * delta;
    const double b_delta_alpha_2 = b_delta * alpha_2;
    flux.rho = hlle.rho - b_delta_alpha_2;
    flux.mom[ix] = hlle.mom[ix] - b_delta_alpha_2 * roeU[ix];
    if (larrouturou) {
      const bool upwind = flux.rho > 0.0;
      flux.mom[iy] = flux.rho * (upwind ? vL_[iy] : vR_[iy]);
      const double f_rhoE = flux.rho * (upwind ? eKinRL : eKinRR);
      flux.E = hlle.E - b_delta_alpha_2 * squaredNormRoeU_half + f_rhoE;
    } else {
      const double l31 = -roeU[iy];
      const double alpha_3 = l31 * deltaRho + deltaRhoU[iy];
      const double b_delta_alpha_3 = b_delta * alpha_3;
      flux.mom[iy] = hlle.mom[iy] - b_delta_alpha_2 * roeU[iy] - b_delta_alpha_3;
      flux.E = hlle.E - b_delta_alpha_2 * squaredNormRoeU_half -
               b_delta_alpha_3 * roeU[iy];
    }
  } else {
    const int ix = dir;
    const int iy = (ix + 1) % 3;
    const int iz = (iy + 1) % 3;
    const double l21 = gm1_over_roeA2 * (roeH - l21_norm);
    const double l22 = l2u * gm1_over_roeA2 * Msq * roeU[0];
    const double l23 = l2u * gm1_over_roeA2 * roeU[1];
    const double l24 = l2u * gm1_over_roeA2 * roeU[2];
    const double l25 = -gm1_over_roeA2;
    const double alpha_2 = l21 * deltaRho + l22 * deltaRhoU[0] +
                           l23 * deltaRhoU[1] + l24 * deltaRhoU[2] +
                           l25 * deltaRhoE;
    const double b_delta = b * delta;
    const double b_delta_alpha_2 = b_delta * alpha_2;
    flux.rho = hlle.rho - b_delta_alpha_2;
    flux.mom[ix] =
        hlle.mom[ix] - b_delta_alpha_2 * roeU[quirk3d ? iy : ix];
    if (larrouturou) {
      const bool upwind = flux.rho > 0.0;
      flux.mom[iy] = flux.rho * (upwind ? vL_[iy] : vR_[iy]);
      flux.mom[iz] = flux.rho * (upwind ? vL_[iz] : vR_[iz]);
      const double f_rhoE = flux.rho * (upwind ? eKinRL : eKinRR);
      flux.E = hlle.E - b_delta_alpha_2 * squaredNormRoeU_half + f_rhoE;
    } else {
      const double l31 = -roeU[iy];
      const double l41 = -roeU[iz];
      const double alpha_3 = l31 * deltaRho + deltaRhoU[iy];
      const double alpha_4 = l41 * deltaRho + deltaRhoU[iz];
      const double b_delta_alpha_3 = b_delta * alpha_3;
      const double b_delta_alpha_4 = b_delta * alpha_4;
      flux.mom[iy] = hlle.mom[iy] - b_delta_alpha_2 * roeU[iy] - b_delta_alpha_3;
      flux.mom[iz] = hlle.mom[iz] - b_delta_alpha_2 * roeU[iz] - b_delta_alpha_4;
      flux.E = hlle.E - b_delta_alpha_2 * squaredNormRoeU_half -
               b_delta_alpha_3 * roeU[iy] - b_delta_alpha_4 * roeU[iz];
    }
  }
  return (bR * pL - bL * pR) / db_positive;
}

// EinfeldtSignalVelocities<PerfectGas<D>>, array form
// (src/equations/perfect_gas/EinfeldtSignalVelocities.hpp:64-97; the scalar
// form at :25-62 performs the same operations in the same order).  Roe
// averages of the normal velocity and of the enthalpy only; the Roe sound
// speed uses the normal velocity alone.
// MusclHancockPrim<Dim>::ComputeNumericFlux, array path
// (src/equations/perfect_gas/MusclHancockPrimMethod.hpp:62-176): the
// limiter lambda (:68-81) on rho, u and p of cells 1 and 2 of the stencil,
// the half-step evolution of the primitives (:114-141), the complete face
// states (:148-168); the caller applies Hllem<PerfectGas<Dim>, true>.
inline double PrimLimiter(double qL, double qM, double qR) {
  const double delta_q = 0.5 * (qR - qL);
  const bool is_relevant = std::fabs(delta_q) > 1e-12;
  const double delta_q_if_relevant = is_relevant ? delta_q : 1.0;
  const double rL = is_relevant ? (qM - qL) / delta_q_if_relevant : 0.0;
  const double rR = is_relevant ? (qR - qM) / delta_q_if_relevant : 0.0;
  const double r = (rL > 0.0 && rR > 0.0) ? std::min(rL, rR) : 0.0;
  const double r_1 = 2 * r / (1.0 + r);
  const double r_2 = 0.5 * (r + 1.0);
  const double sigma = r < 1.0 ? r_1 : r_2;
  return sigma * delta_q;
}

template <int D>
inline void MusclHancockPrimStates(const Gas& g, const State<D>* st, double dt, double dx,
                                   int d, State<D> w[2]) {
  double rho[4], p[4], u[4][D];
  for (int s = 0; s < 4; ++s) {
    rho[s] = st[s].rho;
    p[s] = st[s].p;
    for (int i = 0; i < D; ++i) u[s][i] = st[s].mom[i] / st[s].rho;
  }
  const double lambda = dt / dx;
  // left face state from cell 1
  const double a_l = st[1].a;
  const double drho_l = PrimLimiter(rho[0], rho[1], rho[2]);
  const double dp_l = PrimLimiter(p[0], p[1], p[2]);
  double du_l[D];
  for (int i = 0; i < D; ++i) du_l[i] = PrimLimiter(u[0][i], u[1][i], u[2][i]);
  // right face state from cell 2
  const double a_r = st[2].a;
  const double drho_r = PrimLimiter(rho[1], rho[2], rho[3]);
  const double dp_r = PrimLimiter(p[1], p[2], p[3]);
  double du_r[D];
  for (int i = 0; i < D; ++i) du_r[i] = PrimLimiter(u[1][i], u[2][i], u[3][i]);

  const double rec_rho_l =
      rho[1] + 0.5 * (drho_l - lambda * (u[1][d] * drho_l + rho[1] * du_l[d]));
  const double rec_p_l =
      p[1] + 0.5 * (dp_l - lambda * (rho[1] * a_l * a_l * du_l[d] + u[1][d] * dp_l));
  double rec_u_l[D];
  for (int i = 0; i < D; ++i)
    rec_u_l[i] = u[1][i] + 0.5 * (du_l[i] - lambda * (u[1][d] * du_l[i]));
  rec_u_l[d] -= 0.5 * lambda * dp_l / rho[1];

  const double rec_rho_r =
      rho[2] - 0.5 * (drho_r + lambda * (u[2][d] * drho_r + rho[2] * du_r[d]));
  const double rec_p_r =
      p[2] - 0.5 * (dp_r + lambda * (rho[2] * a_r * a_r * du_r[d] + u[2][d] * dp_r));
  double rec_u_r[D];
  for (int i = 0; i < D; ++i)
    rec_u_r[i] = u[2][i] - 0.5 * (du_r[i] + lambda * (u[2][d] * du_r[i]));
  rec_u_r[d] -= 0.5 * lambda * dp_r / rho[2];

  const double rr[2] = {rec_rho_l, rec_rho_r}, rp[2] = {rec_p_l, rec_p_r};
  const double* ru[2] = {rec_u_l, rec_u_r};
  for (int s = 0; s < 2; ++s) {
    w[s].rho = rr[s];
    for (int i = 0; i < D; ++i) w[s].mom[i] = rr[s] * ru[s][i];
    w[s].p = rp[s];
    w[s].a = std::sqrt(g.gamma * rp[s] / rr[s]);
    w[s].E = 0.5 * rr[s] * SquaredNormArray<D>(ru[s]) + rp[s] * g.gm1inv;
  }
}

// MusclHancockCharacteristic<Rank>, array path
// (src/equations/perfect_gas/MusclHancockCharacteristicMethod.hpp:26-285):
// wave amplitudes of the primitive differences (ComputeAmplitudes, :49-65),
// limited by VanLeer's array operator (include/fub/flux_method/
// MusclHancockMethod.hpp:144-149), advanced by the characteristic speeds
// over half a step (DoTimeStep, :170-180), recombined into primitive
// differences (Recombine, :76-84), added to the cell and completed
// (Reconstruct, :238-259; PerfectGas::CompleteFromPrim array form,
// src/equations/PerfectGas.cpp:160-174).  Velocities have three rows
// (zero beyond the dimension); ix = dir, iy = ix + 1, iz = iy + 1 (mod 3).
inline double VanLeerPair(double sL, double sR) {
  const double r = 2.0 * sL * sR;
  return r > 0.0 ? r / (sL + sR) : 0.0;
}

template <int D>
inline void CharacteristicFaceState(const Gas& g, const State<D>* s3, double lambda,
                                    int is_right_side, int dir, State<D>& rec) {
  const State<D>& q = s3[1];
  const double rho = q.rho;
  const double u = q.mom[dir] / rho;
  const double c = q.a;
  const double rhoc = rho * c;
  const double ooc2 = 1.0 / (c * c);
  const int ix = dir, iy = (ix + 1) % 3, iz = (iy + 1) % 3;
  double pr[3], pp[3], pv[3][3];
  for (int s = 0; s < 3; ++s) {
    pr[s] = s3[s].rho;
    pp[s] = s3[s].p;
    for (int d = 0; d < 3; ++d) pv[s][d] = d < D ? s3[s].mom[d] / s3[s].rho : 0.0;
  }
  double amp[2][5];  // (left, right) x (minus, zero, plus, v, w)
  for (int k = 0; k < 2; ++k) {
    const int a = k, b = k + 1;
    const double dp = pp[b] - pp[a];
    const double du = pv[b][ix] - pv[a][ix];
    const double drho = pr[b] - pr[a];
    amp[k][0] = dp - rhoc * du;
    amp[k][1] = drho - ooc2 * dp;
    amp[k][2] = dp + rhoc * du;
    amp[k][3] = pv[b][iy] - pv[a][iy];
    amp[k][4] = pv[b][iz] - pv[a][iz];
  }
  double slope[5];
  for (int m = 0; m < 5; ++m) slope[m] = VanLeerPair(amp[0][m], amp[1][m]);
  const double sign = 1 - 2 * is_right_side;
  double ampl[5];
  ampl[0] = sign * 0.5 * slope[0] * (1.0 - sign * lambda * (u - c));
  ampl[1] = sign * 0.5 * slope[1] * (1.0 - sign * lambda * u);
  ampl[2] = sign * 0.5 * slope[2] * (1.0 - sign * lambda * (u + c));
  ampl[3] = sign * 0.5 * slope[3] * (1.0 - sign * lambda * u);
  ampl[4] = sign * 0.5 * slope[4] * (1.0 - sign * lambda * u);
  double dpres = 0.5 * (ampl[0] + ampl[2]);
  double drho = ooc2 * dpres + ampl[1];
  double dvel[3];
  dvel[ix] = 0.5 * (ampl[2] - ampl[0]) / rhoc;
  dvel[iy] = ampl[3];
  dvel[iz] = ampl[4];
  drho += rho;
  for (int d = 0; d < D; ++d) dvel[d] += (q.mom[d] / q.rho);
  dpres += q.p;
  rec = State<D>{};
  rec.rho = drho;
  for (int d = 0; d < D; ++d) rec.mom[d] = drho * dvel[d];
  rec.p = dpres;
  const double e_kin = 0.5 * SquaredNormArray<D>(rec.mom) / rec.rho;
  rec.E = e_kin + dpres * g.gm1inv;
  rec.a = std::sqrt(g.gamma * rec.p / rec.rho);
}

template <int D>
inline void EinfeldtSignals(const Gas& g, const State<D>& L, const State<D>& R, int dir,
                            double s[2]) {
  const double rhoL = L.rho;
  const double rhoR = R.rho;
  const double rhoUL = L.mom[dir];
  const double rhoUR = R.mom[dir];
  const double aL = L.a;
  const double aR = R.a;
  const double sqRhoL = std::sqrt(rhoL);
  const double sqRhoR = std::sqrt(rhoR);
  const double uL = rhoUL / rhoL;
  const double uR = rhoUR / rhoR;
  const double oosqRhoSum = 1.0 / (sqRhoL + sqRhoR);
  auto roeAvg = [&](double xL, double xR) { return (xL * sqRhoL + xR * sqRhoR) * oosqRhoSum; };
  const double roeU = roeAvg(uL, uR);
  const double hL = (L.E + L.p) / rhoL;
  const double hR = (R.E + R.p) / rhoR;
  const double roeH = roeAvg(hL, hR);
  const double roeA2 = (g.gamma - 1.0) * (roeH - 0.5 * roeU * roeU);
  const double roeA = std::sqrt(roeA2);
  const double sL1 = uL - aL;
  const double sL2 = roeU - roeA;
  const double sR1 = roeU + roeA;
  const double sR2 = uR + aR;
  s[0] = std::min(sL1, sL2);
  s[1] = std::max(sR1, sR2);
}

// Hll<PerfectGas<D>, EinfeldtSignalVelocities>::ComputeNumericFlux, array
// form (include/fub/flux_method/HllMethod.hpp:260-293).  Returns p*.  The
// scalar form (:191-222) differs only where the array form guards a
// non-positive wave-speed difference (then 1).
template <int D>
inline double HllFlux(const Gas& g, Cons<D>& nf, const State<D>& L, const State<D>& R,
                      int dir) {
  double s[2];
  EinfeldtSignals<D>(g, L, R, dir, s);
  Cons<D> fL, fR;
  Flux<D>(fL, L, dir);
  Flux<D>(fR, R, dir);
  const double sL = std::min(s[0], 0.0);
  const double sR = std::max(s[1], 0.0);
  const double sLsR = sL * sR;
  double ds = sR - sL;
  ds = (ds > 0.0) ? ds : 1.0;
  for (int k = 0; k < D + 2; ++k) {
    C(nf, k) = (sR * C(fL, k) - sL * C(fR, k) + sLsR * (C(R, k) - C(L, k))) / ds;
  }
  return (sR * L.p - sL * R.p) / ds;
}

// Hll::ComputeStableDt, array form (include/fub/flux_method/HllMethod.hpp:337-347):
// dx / max(0, |sL|, |sR|) accumulated left to right.
template <int D>
inline double HllStableDt(const Gas& g, const State<D>& L, const State<D>& R, double dx,
                          int dir) {
  double s[2];
  EinfeldtSignals<D>(g, L, R, dir, s);
  double m = 0.0;
  m = std::max(m, std::abs(s[0]));
  m = std::max(m, std::abs(s[1]));
  return dx / m;
}

// Hllem::ComputeStableDt, array form (src/equations/perfect_gas/HllemMethod.hpp:1152-1216).
template <int D>
inline double HllemStableDt(const Gas& g, const State<D>& L, const State<D>& R,
                            double dx, int dir) {
  const double gm1 = g.gamma - 1.0;
  const double Msq = 1.0;
  const double Minv = 1.0 / std::sqrt(Msq);
  const double rhoL = L.rho;
  const double rhoR = R.rho;
  const double aL = L.a;
  const double aR = R.a;
  const double hL = (L.E + L.p) / rhoL;
  const double hR = (R.E + R.p) / rhoR;
  const double sqRhoL = std::sqrt(rhoL);
  const double sqRhoR = std::sqrt(rhoR);
  const double sqRhoSum = sqRhoL + sqRhoR;
  const double sqRhoL_over_Sum = sqRhoL / sqRhoSum;
  const double sqRhoR_over_Sum = sqRhoR / sqRhoSum;
  double uL[D], uR[D], roeU[D];
  for (int i = 0; i < D; ++i) {
    uL[i] = L.mom[i] / rhoL;
    uR[i] = R.mom[i] / rhoR;
    roeU[i] = sqRhoL_over_Sum * uL[i] + sqRhoR_over_Sum * uR[i];
  }
  const double roeH = sqRhoL_over_Sum * hL + sqRhoR_over_Sum * hR;
  double squaredNormRoeU = 0.0;
  for (int i = 0; i < D; ++i) squaredNormRoeU += roeU[i] * roeU[i];
  squaredNormRoeU *= Msq;
  const double squaredNormRoeU_half = 0.5 * squaredNormRoeU;
  const double roeA2 = gm1 * (roeH - squaredNormRoeU_half);
  const double roeA = std::sqrt(roeA2);
  const double maxA = std::max(aL, aR);
  const double sL1 = uL[dir] - maxA * Minv;
  const double sL2 = roeU[dir] - roeA * Minv;
  const double sR1 = roeU[dir] + roeA * Minv;
  const double sR2 = uR[dir] + maxA * Minv;
  const double sL = std::min(sL1, sL2);
  const double sR = std::max(sR1, sR2);
  const double maxS = std::max(std::abs(sL), std::abs(sR));
  return dx / maxS;
}

// ---------------------------------------------------------------------------
// Exact Riemann solver (src/equations/perfect_gas/ExactRiemannSolver.hpp).

struct NewtonResult {
  double x;
  int iterations;
  bool converged;
};

// NewtonIteration (include/fub/NewtonIteration.hpp:28-52).
template <typename F, typename DF>
inline NewtonResult NewtonIteration(F&& f, DF&& Df, double x0,
                                    double tolerance = 1e-7,
                                    int max_iterations = 5000) {
  double x = x0;
  double fx = f(x);
  int counter = 0;
  double err = std::abs(fx);
  while (err > tolerance && counter < max_iterations) {
    const double dfx = Df(x);
    const double offset = fx / dfx;
    const double next = x - offset;
    x = ((next - next) != (next - next) || next < 0.0) ? 0.5 * x : next;
    fx = f(x);
    err = std::abs(fx);
    counter += 1;
  }
  const bool failed = (counter == max_iterations && fx > tolerance);
  return {x, counter, !failed};
}

struct MiddleState {
  double pM;
  double uM;
  int iterations;
  bool converged;
};

// ExactRiemannSolver::ComputeMiddleState (ExactRiemannSolver.hpp:34-97).
// x0 < 0 selects the reference's initial guess 0.5 * (pL + pR).
template <int D>
inline MiddleState ComputeMiddleState(const Gas& eq, const State<D>& left,
                                      const State<D>& right, int dir,
                                      double x0 = -1.0, double tol = 1e-7,
                                      int max_it = 5000) {
  const double g = eq.gamma;
  const double gp1 = g + 1.0;
  const double gm1 = g - 1.0;
  const double rhoL = left.rho;
  const double pL = left.p;
  const double aL = left.a;
  const double A_l = 2.0 / (gp1 * rhoL);
  const double B_l = gm1 / gp1 * pL;
  const double exponent = 0.5 * gm1 / g;
  const double d_exponent = -0.5 * gp1 / g;
  auto f_left = [=](double p) {
    if (p > pL) {
      return (p - pL) * std::sqrt(A_l / (p + B_l));
    } else {
      return 2.0 * aL / gm1 * (std::pow(p / pL, exponent) - 1.0);
    }
  };
  auto df_left = [=](double p) {
    if (p > pL) {
      return std::sqrt(A_l / (B_l + p)) * (1.0 - 0.5 * (p - pL) / (B_l + p));
    } else {
      return std::pow(p / pL, d_exponent) / (rhoL * aL);
    }
  };
  const double rhoR = right.rho;
  const double pR = right.p;
  const double aR = right.a;
  const double A_r = 2.0 / (gp1 * rhoR);
  const double B_r = gm1 / gp1 * pR;
  auto f_right = [=](double p) {
    if (p > pR) {
      return (p - pR) * std::sqrt(A_r / (p + B_r));
    } else {
      return 2 * aR / gm1 * (std::pow(p / pR, exponent) - 1.0);
    }
  };
  auto df_right = [=](double p) {
    if (p > pR) {
      return std::sqrt(A_r / (B_r + p)) * (1.0 - 0.5 * (p - pR) / (B_r + p));
    } else {
      return std::pow(p / pR, d_exponent) / (rhoR * aR);
    }
  };
  const double uL = left.mom[dir] / left.rho;
  const double uR = right.mom[dir] / right.rho;
  const double du = uR - uL;
  auto f = [=](double h) { return f_right(h) + f_left(h) + du; };
  auto df = [=](double h) { return df_right(h) + df_left(h); };
  const double guess = x0 < 0.0 ? 0.5 * (pL + pR) : x0;
  const NewtonResult r = NewtonIteration(f, df, guess, tol, max_it);
  const double pM = r.x;
  const double fL = f_left(pM);
  const double fR = f_right(pM);
  const double uM = 0.5 * (uL + uR) + 0.5 * (fR - fL);
  return {pM, uM, r.iterations, r.converged};
}

// ExactRiemannSolver::ComputeSignals, scalar (ExactRiemannSolver.hpp:105-148).
template <int D>
inline void ComputeSignals(const Gas& eq, const State<D>& left,
                           const State<D>& right, int dir, double s[2],
                           MiddleState* ms_out = nullptr, double x0 = -1.0,
                           double tol = 1e-7, int max_it = 5000) {
  const MiddleState ms = ComputeMiddleState<D>(eq, left, right, dir, x0, tol, max_it);
  if (ms_out) *ms_out = ms;
  const double pM = ms.pM;
  const double g = eq.gamma;
  const double gp1 = g + 1.0;
  const double gm1 = g - 1.0;
  const double uL = left.mom[dir] / left.rho;
  const double aL = left.a;
  const double pL = left.p;
  const double pL_ratio = pM / pL;
  if (pM > pL) {
    s[0] = uL - aL * std::sqrt((gp1 * pL_ratio + gm1) / (2.0 * g));
  } else {
    s[0] = uL - aL;
  }
  const double uR = right.mom[dir] / right.rho;
  const double aR = right.a;
  const double pR = right.p;
  const double pR_ratio = pM / pR;
  if (pM > pR) {
    s[1] = uR + aR * std::sqrt((gp1 * pR_ratio + gm1) / (2.0 * g));
  } else {
    s[1] = uR + aR;
  }
}

// ExactRiemannSolver::SolveRiemannProblem, scalar (ExactRiemannSolver.hpp:204-353).
// The state at x/t = 0 of the exact solution.
template <int D>
inline MiddleState SolveRiemannProblem(const Gas& eq, State<D>& state,
                                       const State<D>& left,
                                       const State<D>& right, int dir,
                                       double x0 = -1.0, double tol = 1e-7,
                                       int max_it = 5000) {
  auto velocity = [dir](const State<D>& s) { return s.mom[dir] / s.rho; };
  auto set_momentum = [dir](State<D>& s, double momentum, const State<D>& side) {
    for (int d = 0; d < D; ++d) s.mom[d] = s.rho * (side.mom[d] / side.rho);
    s.mom[dir] = momentum;
  };
  auto from_prim = [&eq](State<D>& s) {
    const double rhoE_kin = 0.5 * SquaredNormScalar<D>(s.mom) / s.rho;
    const double rhoE_internal = s.p * eq.gm1inv;
    s.E = rhoE_internal + rhoE_kin;
    s.a = std::sqrt(eq.gamma * s.p / s.rho);
  };
  MiddleState ms{0.0, 0.0, 0, true};
  if (0.0 < left.rho && 0.0 < right.rho) {
    ms = ComputeMiddleState<D>(eq, left, right, dir, x0, tol, max_it);
    const double pM = ms.pM;
    const double uM = ms.uM;
    const double g = eq.gamma;
    const double gp1 = g + 1.0;
    const double gm1 = g - 1.0;
    if (0.0 <= uM) {
      const double uL = velocity(left);
      const double aL = left.a;
      const double pL = left.p;
      const double p_ratio = pM / pL;
      const double g_ratio = gm1 / gp1;
      if (pM > pL) {
        const double shock_speed =
            uL - aL * std::sqrt((gp1 * p_ratio + gm1) / (2.0 * g));
        if (shock_speed < 0.0) {
          state.rho = left.rho * (p_ratio + g_ratio) / (g_ratio * p_ratio + 1.0);
          set_momentum(state, state.rho * uM, left);
          state.p = pM;
          from_prim(state);
        } else {
          state = left;
        }
      } else {
        const double aM = aL * std::pow(p_ratio, gm1 / (2.0 * g));
        const double shock_speed_TL = uM - aM;
        const double shock_speed_HL = uL - aL;
        if (0.0 < shock_speed_HL) {
          state = left;
        } else if (shock_speed_TL < 0.0) {
          state.rho = left.rho * std::pow(p_ratio, 1 / g);
          set_momentum(state, state.rho * uM, left);
          state.p = pM;
          from_prim(state);
        } else {
          state.rho = left.rho * std::pow(2.0 / gp1 + g_ratio * uL / aL, 2.0 / gm1);
          const double u = Planted(PLANT_EXACT_FAN_LEFT_U) ? 2.0 / gp1 * (aL - 0.5 * gm1 * uL)
                                                           : 2.0 / gp1 * (aL + 0.5 * gm1 * uL);
          set_momentum(state, state.rho * u, left);
          state.p = pL * std::pow(2.0 / gp1 + g_ratio * uL / aL, 2.0 * g / gm1);
          from_prim(state);
        }
      }
    } else {
      const double uR = velocity(right);
      const double aR = right.a;
      const double pR = right.p;
      const double p_ratio = pM / pR;
      const double g_ratio = gm1 / gp1;
      if (pM > pR) {
        const double shock_speed =
            uR + aR * std::sqrt((gp1 * p_ratio + gm1) / (2.0 * g));
        if (0.0 < shock_speed) {
          state.rho = right.rho * (p_ratio + g_ratio) / (g_ratio * p_ratio + 1.0);
          set_momentum(state, state.rho * uM, right);
          state.p = pM;
          from_prim(state);
        } else {
          state = right;
        }
      } else {
        const double aM = aR * std::pow(p_ratio, gm1 / (2.0 * g));
        const double shock_speed_TR = uM + aM;
        const double shock_speed_HR = uR + aR;
        if (shock_speed_HR <= 0.0) {
          state = right;
        } else if (shock_speed_TR >= 0.0) {
          state.rho = right.rho * std::pow(p_ratio, 1 / g);
          set_momentum(state, state.rho * uM, right);
          state.p = pM;
          from_prim(state);
        } else {
          state.rho = right.rho * std::pow(2.0 / gp1 - g_ratio * uR / aR,
                                           Planted(PLANT_EXACT_FAN_RIGHT_RHO) ? 2.0 / g : 2.0 / gm1);
          const double u = 2.0 / gp1 * (-aR + 0.5 * gm1 * uR);
          set_momentum(state, state.rho * u, right);
          state.p = pR * std::pow(2.0 / gp1 - g_ratio * uR / aR, 2.0 * g / gm1);
          from_prim(state);
        }
      }
    }
  }
  return ms;
}

// Godunov<Eq, ExactRiemannSolver>::ComputeStableDt, scalar
// (include/fub/flux_method/GodunovMethod.hpp:163-170).
template <int D>
inline double GodunovStableDt(const Gas& eq, const State<D>& L,
                              const State<D>& R, double dx, int dir,
                              double tol = 1e-7, int max_it = 5000) {
  double s[2];
  ComputeSignals<D>(eq, L, R, dir, s, nullptr, -1.0, tol, max_it);
  double s_max = 0.0;
  s_max = std::max(s_max, std::abs(s[0]));
  s_max = std::max(s_max, std::abs(s[1]));
  return dx / s_max;
}

} // namespace fvs_oracle
