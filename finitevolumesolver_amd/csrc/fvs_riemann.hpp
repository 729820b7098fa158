// Exact Riemann solver of the perfect gas on the device (FP64).
//
// Follows ExactRiemannSolver<PerfectGas<Dim>> operation for operation:
//   ComputeMiddleState   src/equations/perfect_gas/ExactRiemannSolver.hpp:34-97
//   NewtonIteration      include/fub/NewtonIteration.hpp:28-52
//   ComputeSignals       src/equations/perfect_gas/ExactRiemannSolver.hpp:105-148
//   SolveRiemannProblem  src/equations/perfect_gas/ExactRiemannSolver.hpp:204-353
//   Godunov flux / CFL   include/fub/flux_method/GodunovMethod.hpp:130-170
// The exact solver forces the reference onto its sequential (scalar) code
// path (the array ComputeSignals throws), so kinetic energies here use the
// scalar reduction order (m0^2 + m1^2) + m2^2.  pow is the device libm's;
// it can differ from the host libm in the last bit, so these results agree
// with the CPU to rounding, not bit for bit.
//
// Mixed form (FAST = true): the Newton iteration stays FP64 but evaluates
// x^e as exp(e log x) (FP64 exp and log, a few ulps instead of pow's
// correctly rounded-like accuracy), and the sampled state and its flux are
// computed in float with x^e = exp2(e log2 x) from the hardware estimates.
#pragma once

#include "fvs_physics.hpp"

namespace fvs {

// x^e for the Newton pressure function (FP64).
template <bool FAST> __device__ __forceinline__ double PowN(double x, double e) {
  if constexpr (FAST) {
    return exp(e * log(x));
  } else {
    return pow(x, e);
  }
}

// x^e of the sampled state (float or double).
__device__ __forceinline__ double PowS(double x, double e) { return pow(x, e); }
__device__ __forceinline__ float PowS(float x, float e) {
  return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}

template <int D, class T> __device__ __forceinline__ T SqNormScalar(const T* m) {
  if constexpr (D == 2) {
    return m[0] * m[0] + m[1] * m[1];
  } else {
    return (m[0] * m[0] + m[1] * m[1]) + m[2] * m[2];
  }
}

// PerfectGas::CompleteFromCons, scalar form (src/equations/PerfectGas.cpp:96-110).
template <int D>
__device__ __forceinline__ void CompleteFromConsScalar(double* q, const Gas<double>& g) {
  using I = Idx<D>;
  const double e_kin = 0.5 * SqNormScalar<D>(q + I::kMom) / q[I::kRho];
  const double e_int = q[I::kE] - e_kin;
  q[I::kP] = e_int / g.gm1inv;
  q[I::kA] = rsqrt_ieee(g.gamma * q[I::kP] / q[I::kRho]);
}

struct NewtonParams {
  double tol;      // 1e-7
  int max_it;      // 5000
};

struct MiddleState {
  double pM;
  double uM;
  int iterations;
  int converged;
};

// Toro's pressure function of one side and its derivative, both at x (the
// Newton iteration needs the derivative at the point whose residual it just
// tested).  Shock branch: the square root sqrt((x + B) / A) is shared, as in
// the reference's two calls (same arguments, same value).  Rarefaction
// branch: the reference evaluates (x/p)^((gamma+1)/(-2 gamma)) with a second
// pow; here it is (x/p)^((gamma-1)/(2 gamma)) / (x/p), the same number to
// rounding (one pow instead of two per side and iteration; at x = 0 the
// reference's pow form is kept).
template <bool FAST = false> struct SideFn {
  double rho, p, a, A, B;
  __device__ __forceinline__ double eval(double x, double gm1, double exponent,
                                         double d_exponent, double* df) const {
    if (x > p) {
      const double r = rsqrt_ieee(A / (x + B));
      *df = r * (1.0 - 0.5 * (x - p) / (B + x));
      return (x - p) * r;
    } else {
      const double t = x / p;
      const double s = PowN<FAST>(t, exponent);
      *df = (t > 0.0 ? s / t : PowN<FAST>(t, d_exponent)) / (rho * a);
      return 2.0 * a / gm1 * (s - 1.0);
    }
  }
};

// ComputeMiddleState with an explicit initial guess x0 (the reference's is
// 0.5 * (pL + pR)).  The side functions at the final iterate, which the
// reference evaluates once more for the middle velocity, are the ones the
// last convergence test computed (same arguments, same values).
template <int D, int DIR, bool FAST = false>
__device__ __forceinline__ MiddleState ComputeMiddleState(const double* L, const double* R,
                                                          const Gas<double>& eq,
                                                          const NewtonParams& np,
                                                          double x0) {
  using I = Idx<D>;
  const double g = eq.gamma;
  const double gp1 = g + 1.0;
  const double gm1 = g - 1.0;
  SideFn<FAST> left, right;
  left.rho = L[I::kRho];
  left.p = L[I::kP];
  left.a = L[I::kA];
  left.A = 2.0 / (gp1 * left.rho);
  left.B = gm1 / gp1 * left.p;
  right.rho = R[I::kRho];
  right.p = R[I::kP];
  right.a = R[I::kA];
  right.A = 2.0 / (gp1 * right.rho);
  right.B = gm1 / gp1 * right.p;
  const double exponent = 0.5 * gm1 / g;
  const double d_exponent = -0.5 * gp1 / g;
  const double uL = L[I::kMom + DIR] / L[I::kRho];
  const double uR = R[I::kMom + DIR] / R[I::kRho];
  const double du = uR - uL;
  // NewtonIteration (NewtonIteration.hpp:28-52).
  double x = x0;
  double dR, dL;
  double fR = right.eval(x, gm1, exponent, d_exponent, &dR);
  double fL = left.eval(x, gm1, exponent, d_exponent, &dL);
  double fx = fR + fL + du;
  int counter = 0;
  double err = rabs(fx);
  while (err > np.tol && counter < np.max_it) {
    const double dfx = dR + dL;
    const double offset = fx / dfx;
    const double next = x - offset;
    x = ((next - next) != (next - next) || next < 0.0) ? 0.5 * x : next;
    fR = right.eval(x, gm1, exponent, d_exponent, &dR);
    fL = left.eval(x, gm1, exponent, d_exponent, &dL);
    fx = fR + fL + du;
    err = rabs(fx);
    counter += 1;
  }
  MiddleState ms;
  ms.pM = x;
  ms.uM = 0.5 * (uL + uR) + 0.5 * (fR - fL);
  ms.iterations = counter;
  ms.converged = !(counter == np.max_it && fx > np.tol);
  return ms;
}

// Wave speeds from the middle pressure (ComputeSignals) and the Godunov
// stable dt dx / max(0, |sL|, |sR|).
template <int D, int DIR>
__device__ __forceinline__ double GodunovDtFromMiddle(const double* L, const double* R,
                                                      double pM, const Gas<double>& eq,
                                                      double dx) {
  using I = Idx<D>;
  const double g = eq.gamma;
  const double gp1 = g + 1.0;
  const double gm1 = g - 1.0;
  double s0, s1;
  {
    const double uL = L[I::kMom + DIR] / L[I::kRho];
    const double aL = L[I::kA];
    const double pL = L[I::kP];
    const double pL_ratio = pM / pL;
    if (pM > pL) {
      s0 = uL - aL * rsqrt_ieee((gp1 * pL_ratio + gm1) / (2.0 * g));
    } else {
      s0 = uL - aL;
    }
  }
  {
    const double uR = R[I::kMom + DIR] / R[I::kRho];
    const double aR = R[I::kA];
    const double pR = R[I::kP];
    const double pR_ratio = pM / pR;
    if (pM > pR) {
      s1 = uR + aR * rsqrt_ieee((gp1 * pR_ratio + gm1) / (2.0 * g));
    } else {
      s1 = uR + aR;
    }
  }
  double s_max = 0.0;
  s_max = rmax(s_max, rabs(s0));
  s_max = rmax(s_max, rabs(s1));
  return dx / s_max;
}

// SolveRiemannProblem: the complete state at x/t = 0 given the middle state.
// T = double reproduces the reference; T = float is the mixed form's sample.
template <int D, int DIR, class T>
__device__ __forceinline__ void SampleRiemannT(T* state, const T* left, const T* right, T pM,
                                               T uM, const Gas<T>& eq) {
  using I = Idx<D>;
  constexpr int NQ = I::kComp;
  constexpr int M = I::kMom;
  auto set_momentum = [&](T momentum, const T* side) {
#pragma unroll
    for (int d = 0; d < D; ++d) state[M + d] = state[I::kRho] * (side[M + d] / side[I::kRho]);
    state[M + DIR] = momentum;
  };
  auto from_prim = [&]() {
    const T rhoE_kin = T(0.5) * SqNormScalar<D>(state + M) / state[I::kRho];
    const T rhoE_internal = state[I::kP] * eq.gm1inv;
    state[I::kE] = rhoE_internal + rhoE_kin;
    state[I::kA] = rsqrt_ieee(eq.gamma * state[I::kP] / state[I::kRho]);
  };
  auto copy = [&](const T* s) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) state[k] = s[k];
  };
  const T g = eq.gamma;
  const T gp1 = g + T(1.0);
  const T gm1 = g - T(1.0);
  if (T(0.0) <= uM) {
    const T uL = left[M + DIR] / left[I::kRho];
    const T aL = left[I::kA];
    const T pL = left[I::kP];
    const T p_ratio = pM / pL;
    const T g_ratio = gm1 / gp1;
    if (pM > pL) {
      const T shock_speed = uL - aL * rsqrt_ieee((gp1 * p_ratio + gm1) / (T(2.0) * g));
      if (shock_speed < T(0.0)) {
        state[I::kRho] = left[I::kRho] * (p_ratio + g_ratio) / (g_ratio * p_ratio + T(1.0));
        set_momentum(state[I::kRho] * uM, left);
        state[I::kP] = pM;
        from_prim();
      } else {
        copy(left);
      }
    } else {
      const T aM = aL * PowS(p_ratio, gm1 / (T(2.0) * g));
      const T shock_speed_TL = uM - aM;
      const T shock_speed_HL = uL - aL;
      if (T(0.0) < shock_speed_HL) {
        copy(left);
      } else if (shock_speed_TL < T(0.0)) {
        state[I::kRho] = left[I::kRho] * PowS(p_ratio, T(1) / g);
        set_momentum(state[I::kRho] * uM, left);
        state[I::kP] = pM;
        from_prim();
      } else {
        state[I::kRho] = left[I::kRho] * PowS(T(2.0) / gp1 + g_ratio * uL / aL, T(2.0) / gm1);
        const T u = T(2.0) / gp1 * (aL + T(0.5) * gm1 * uL);
        set_momentum(state[I::kRho] * u, left);
        state[I::kP] = pL * PowS(T(2.0) / gp1 + g_ratio * uL / aL, T(2.0) * g / gm1);
        from_prim();
      }
    }
  } else {
    const T uR = right[M + DIR] / right[I::kRho];
    const T aR = right[I::kA];
    const T pR = right[I::kP];
    const T p_ratio = pM / pR;
    const T g_ratio = gm1 / gp1;
    if (pM > pR) {
      const T shock_speed = uR + aR * rsqrt_ieee((gp1 * p_ratio + gm1) / (T(2.0) * g));
      if (T(0.0) < shock_speed) {
        state[I::kRho] = right[I::kRho] * (p_ratio + g_ratio) / (g_ratio * p_ratio + T(1.0));
        set_momentum(state[I::kRho] * uM, right);
        state[I::kP] = pM;
        from_prim();
      } else {
        copy(right);
      }
    } else {
      const T aM = aR * PowS(p_ratio, gm1 / (T(2.0) * g));
      const T shock_speed_TR = uM + aM;
      const T shock_speed_HR = uR + aR;
      if (shock_speed_HR <= T(0.0)) {
        copy(right);
      } else if (shock_speed_TR >= T(0.0)) {
        state[I::kRho] = right[I::kRho] * PowS(p_ratio, T(1) / g);
        set_momentum(state[I::kRho] * uM, right);
        state[I::kP] = pM;
        from_prim();
      } else {
        state[I::kRho] = right[I::kRho] * PowS(T(2.0) / gp1 - g_ratio * uR / aR, T(2.0) / gm1);
        const T u = T(2.0) / gp1 * (-aR + T(0.5) * gm1 * uR);
        set_momentum(state[I::kRho] * u, right);
        state[I::kP] = pR * PowS(T(2.0) / gp1 - g_ratio * uR / aR, T(2.0) * g / gm1);
        from_prim();
      }
    }
  }
}

template <int D, int DIR>
__device__ __forceinline__ void SampleRiemann(double* state, const double* left,
                                              const double* right, const MiddleState& ms,
                                              const Gas<double>& eq) {
  SampleRiemannT<D, DIR, double>(state, left, right, ms.pM, ms.uM, eq);
}

// Godunov<PerfectGas<D>, ExactRiemannSolver> numeric flux between complete
// states L and R from a precomputed middle state.
template <int D, int DIR>
__device__ __forceinline__ void GodunovFluxFromMiddle(double* flux, const double* L,
                                                      const double* R, const MiddleState& ms,
                                                      const Gas<double>& eq) {
  double sol[D + 4];
#pragma unroll
  for (int k = 0; k < D + 4; ++k) sol[k] = 0.0;
  if (0.0 < L[0] && 0.0 < R[0]) SampleRiemann<D, DIR>(sol, L, R, ms, eq);
  PhysFlux<D, DIR>(flux, sol);
}

} // namespace fvs
