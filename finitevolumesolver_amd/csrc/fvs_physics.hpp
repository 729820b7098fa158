// Device arithmetic of the perfect-gas split patch update (gfx950).
//
// Every FP64 expression below keeps the operation order of the reference's
// 8-lane Eigen array path so that the FP64 kernels reproduce the reference
// bit for bit when compiled with -ffp-contract=off:
//   PerfectGas::Flux / CompleteFromCons   src/equations/PerfectGas.cpp:52-144
//   MinMod / VanLeer / NoLimiter          include/fub/flux_method/MusclHancockMethod.hpp:31-187
//   MusclHancock::ComputeNumericFlux      include/fub/flux_method/MusclHancockMethod.hpp:419-485
//   Hllem<Eq, Larrouturou>::ComputeNumericFlux  src/equations/perfect_gas/HllemMethod.hpp:711-1072
//   Hllem::ComputeStableDt                src/equations/perfect_gas/HllemMethod.hpp:1152-1216
// Simplifications used here are exact for finite inputs: multiplications by
// MaSq() == 1 and Minv == 1 are dropped, terms that the Larrouturou variant
// multiplies by zero are dropped (they can only change the sign of a zero
// result, never a non-zero value), and per-cell reconstructions are shared
// by the two faces of a cell instead of being recomputed per face.
//
// min/max follow std::min/std::max argument semantics ((b < a) ? b : a),
// which is what Eigen 3.3's scalar and packet min/max reproduce.
//
// Every division and square root goes through an arithmetic policy A:
//   IeeeArith<T>  correctly rounded a / b (DivRn) and sqrt (the reference's
//                 arithmetic; the FP64 kernels are bit-exact with it);
//   FastArith<T>  a shared reciprocal per denominator from the hardware
//                 estimate (v_rcp_f64 refined by one Newton step), quotients
//                 as products, and sqrt from v_rsq_f64 refined by one
//                 Goldschmidt step and one residual correction.  Each result
//                 is within about ten units in the last place of the
//                 correctly rounded one; a zero denominator yields NaN
//                 instead of inf.
//                 The float form uses the single-precision estimates as is.
// Sharing a reciprocal changes nothing under IeeeArith (Inv keeps the
// denominator and every quotient is still a / b).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fvs {

template <class T> __device__ __forceinline__ T rmin(T a, T b) { return (b < a) ? b : a; }
template <class T> __device__ __forceinline__ T rmax(T a, T b) { return (a < b) ? b : a; }
__device__ __forceinline__ double rsqrt_ieee(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float rsqrt_ieee(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double rabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float rabs(float x) { return __builtin_fabsf(x); }

// Correctly rounded a / b in FP64.  The compiler's division sequence
// (v_div_scale, v_rcp_f64, Newton steps, v_div_fmas, v_div_fixup) returns a
// faithful quotient that can miss the nearest double when a / b lies within
// about 2^-53 ulp of a rounding midpoint; e.g. 0x1.6666666666663p+1 /
// 0x1.ffffffffffffbp+0 (a Roe-averaged gamma) gives 0x1.6666666666666p+0
// instead of ...667p+0.  The exact remainders of the quotient and of its
// neighbour towards a / b decide between the two (a quotient of two doubles
// is never exactly a midpoint).  Zero, infinite and NaN quotients pass
// through unchanged.
// q: a quotient within one ulp of a / b.
__device__ __forceinline__ double RoundQuotient(double a, double b, double q) {
  const double r = __builtin_fma(-q, b, a);
  const bool down = (r < 0.0) != (b < 0.0) ? !(q < 0.0) : (q < 0.0);
  const double q2 = __longlong_as_double(__double_as_longlong(q) + (down ? -1ll : 1ll));
  const double r2 = __builtin_fma(-q2, b, a);
  const bool regular = q != 0.0 && __builtin_fabs(q) < __builtin_huge_val();
  return (regular && __builtin_fabs(r2) < __builtin_fabs(r)) ? q2 : q;
}
// Measured on gfx950 with quotients within ~2^-50 ulp of a midpoint, the
// misrounded ones all have denominators whose significand lies within a few
// ulps of 2 (none among denominators farther than 2^-40 from 2, nor near 1 or
// uniformly drawn); the correction therefore runs only for significands of at
// least 2 - 2^-20 (high 20 mantissa bits all set), a branch that waves skip
// when none of their lanes needs it.
__device__ __forceinline__ double DivRn(double a, double b) {
  const double q = a / b;
  const int hi = static_cast<int>(__double_as_longlong(b) >> 32);
  if (__builtin_expect((hi & 0xFFFFF) == 0xFFFFF, 0)) return RoundQuotient(a, b, q);
  return q;
}
__device__ __forceinline__ float DivRn(float a, float b) { return a / b; }

template <class T> struct Gas {
  T gamma;  // 1.4
  T gm1inv; // 1 / (gamma - 1), evaluated in double on the host
  T gm1;    // gamma - 1 (stable-dt formula)
};

template <class T> struct IeeeArith {
  static constexpr bool kFast = false;
  struct Inv {
    T b;
  };
  static __device__ __forceinline__ Inv inv(T b) { return Inv{b}; }
  static __device__ __forceinline__ T div(T a, const Inv& i) { return DivRn(a, i.b); }
  static __device__ __forceinline__ T quot(T a, T b) { return DivRn(a, b); }
  static __device__ __forceinline__ T sqrt(T x) { return rsqrt_ieee(x); }
  // std::min / std::max argument semantics (signed zeros and NaN as the reference)
  static __device__ __forceinline__ T min(T a, T b) { return rmin(a, b); }
  static __device__ __forceinline__ T max(T a, T b) { return rmax(a, b); }
  // e_int / (1 / (gamma - 1))
  static __device__ __forceinline__ T pressure(T e_int, const Gas<T>& g) {
    return DivRn(e_int, g.gm1inv);
  }
};

template <class T> struct FastArith;

template <> struct FastArith<double> {
  static constexpr bool kFast = true;
  struct Inv {
    double y;
  };
  static __device__ __forceinline__ Inv inv(double b) {
    // v_rcp_f64 is good to ~2^-24 (2.5e8 ulp measured); one Newton step
    // brings it to <= 11 ulp (mean 0.6 ulp over 4M random operands)
    double y = __builtin_amdgcn_rcp(b);
    const double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    return Inv{y};
  }
  static __device__ __forceinline__ double div(double a, const Inv& i) { return a * i.y; }
  static __device__ __forceinline__ double quot(double a, double b) { return a * inv(b).y; }
  static __device__ __forceinline__ double sqrt(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double s = x * r;
    double h = 0.5 * r;
    const double e = __builtin_fma(-h, s, 0.5);
    s = __builtin_fma(s, e, s);
    h = __builtin_fma(h, e, h);
    const double d = __builtin_fma(-s, s, x);
    s = __builtin_fma(d, h, s);
    return (x == 0.0) ? x : s;
  }
  // hardware minimum / maximum (one instruction; may differ from std::min in
  // the sign of a zero result or when an operand is NaN)
  static __device__ __forceinline__ double min(double a, double b) { return __builtin_fmin(a, b); }
  static __device__ __forceinline__ double max(double a, double b) { return __builtin_fmax(a, b); }
  static __device__ __forceinline__ double pressure(double e_int, const Gas<double>& g) {
    return e_int * g.gm1;
  }
};

template <> struct FastArith<float> {
  static constexpr bool kFast = true;
  struct Inv {
    float y;
  };
  static __device__ __forceinline__ Inv inv(float b) { return Inv{__builtin_amdgcn_rcpf(b)}; }
  static __device__ __forceinline__ float div(float a, const Inv& i) { return a * i.y; }
  static __device__ __forceinline__ float quot(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  static __device__ __forceinline__ float min(float a, float b) { return __builtin_fminf(a, b); }
  static __device__ __forceinline__ float max(float a, float b) { return __builtin_fmaxf(a, b); }
  static __device__ __forceinline__ float pressure(float e_int, const Gas<float>& g) {
    return e_int * g.gm1;
  }
};

// Arithmetic policies of a kernel: the flux arithmetic in T and the CFL
// arithmetic (always double).
template <class T, bool FAST> struct Policies {
  using TA = typename std::conditional<FAST, FastArith<T>, IeeeArith<T>>::type;
  using DA = typename std::conditional<FAST, FastArith<double>, IeeeArith<double>>::type;
};

// Component layout of a Complete state in registers: rho, mom[0..D-1], E, p, a.
template <int D> struct Idx {
  static constexpr int kRho = 0;
  static constexpr int kMom = 1;
  static constexpr int kE = D + 1;
  static constexpr int kP = D + 2;
  static constexpr int kA = D + 3;
  static constexpr int kCons = D + 2;
  static constexpr int kComp = D + 4;
};

// Squared momentum norm in the order of the reference's array path
// (Eigen column reduction): 2D m0^2 + m1^2, 3D m0^2 + (m1^2 + m2^2).
template <int D, class T> __device__ __forceinline__ T SqNormArray(const T* m) {
  if constexpr (D == 2) {
    return m[0] * m[0] + m[1] * m[1];
  } else {
    return m[0] * m[0] + (m[1] * m[1] + m[2] * m[2]);
  }
}

// PerfectGas::CompleteFromCons (array form): fills p and a of q and returns
// the density's reciprocal (for further quotients by the same density).
template <int D, class T, class A = IeeeArith<T>>
__device__ __forceinline__ typename A::Inv CompleteFromCons(T* q, const Gas<T>& g) {
  using I = Idx<D>;
  const typename A::Inv ir = A::inv(q[I::kRho]);
  const T e_kin = A::div(T(0.5) * SqNormArray<D>(q + I::kMom), ir);
  const T e_int = q[I::kE] - e_kin;
  q[I::kP] = A::pressure(e_int, g);
  q[I::kA] = A::sqrt(A::div(g.gamma * q[I::kP], ir));
  return ir;
}

// PerfectGas::Flux in direction DIR of a complete state; ir is the
// reciprocal of its density.
template <int D, int DIR, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void PhysFlux(T* f, const T* q, const typename A::Inv& ir) {
  using I = Idx<D>;
  const T velocity = A::div(q[I::kMom + DIR], ir);
  f[I::kRho] = q[I::kMom + DIR];
#pragma unroll
  for (int d = 0; d < D; ++d) f[I::kMom + d] = velocity * q[I::kMom + d];
  f[I::kMom + DIR] += q[I::kP];
  f[I::kE] = velocity * (q[I::kE] + q[I::kP]);
}

template <int D, int DIR, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void PhysFlux(T* f, const T* q) {
  PhysFlux<D, DIR, T, A>(f, q, A::inv(q[Idx<D>::kRho]));
}

enum { kMinMod = 0, kVanLeer = 1, kNoLimiter = 2, kNoGradient = 3 };

template <int LIM, class T, class A = IeeeArith<T>>
__device__ __forceinline__ T LimitedSlope(T qL, T qM, T qR) {
  const T sL = qM - qL;
  const T sR = qR - qM;
  if constexpr (LIM == kMinMod) {
    return (sR > T(0)) ? A::max(T(0), A::min(sL, sR)) : A::min(T(0), A::max(sL, sR));
  } else if constexpr (LIM == kVanLeer) {
    const T r = T(2) * sL * sR;
    return (r > T(0)) ? A::quot(r, sL + sR) : T(0);
  } else if constexpr (LIM == kNoLimiter) {
    return T(0.5) * (sL + sR);
  } else {
    return T(0);
  }
}

// Slope of MusclHancockPrim (src/equations/perfect_gas/MusclHancockPrimMethod.hpp:68-81,
// the limiter lambda of the array path): sigma(r) * (qR - qL) / 2 with r the
// smaller positive one-sided ratio, sigma = 2r / (1 + r) below 1 and
// (r + 1) / 2 above; differences below 1e-12 give a zero slope.  FP64, IEEE.
__device__ __forceinline__ double PrimSlope(double qL, double qM, double qR) {
  const double dq = 0.5 * (qR - qL);
  const bool relevant = __builtin_fabs(dq) > 1e-12;
  const double den = relevant ? dq : 1.0;
  const double rL = relevant ? DivRn(qM - qL, den) : 0.0;
  const double rR = relevant ? DivRn(qR - qM, den) : 0.0;
  const double r = (rL > 0.0 && rR > 0.0) ? rmin(rL, rR) : 0.0;
  const double sigma = r < 1.0 ? DivRn(2.0 * r, 1.0 + r) : 0.5 * (r + 1.0);
  return sigma * dq;
}

// MusclHancockPrim::ComputeNumericFlux, array path (:83-176): primitive
// slopes of (rho, u, p) in cells 1 and 2 of the complete stencil u[0..3],
// the characteristic half-step evolution of the primitives with
// lambda = dt / dx, and the complete states L (cell 1, at the face) and R
// (cell 2) that the HLLEM base takes.  FP64, IEEE, in the reference's order.
template <int D, int DIR>
__device__ __forceinline__ void MusclHancockPrimStates(const double (*u)[D + 4], double lambda,
                                                       const Gas<double>& g, double* L,
                                                       double* R) {
  double rho[4], p[4], v[4][D];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    rho[s] = u[s][0];
    p[s] = u[s][D + 2];
#pragma unroll
    for (int i = 0; i < D; ++i) v[s][i] = DivRn(u[s][1 + i], rho[s]);
  }
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const int c = 1 + side;  // the cell whose face state is built
    const double a = u[c][D + 3];
    const double drho = PrimSlope(rho[c - 1], rho[c], rho[c + 1]);
    const double dp = PrimSlope(p[c - 1], p[c], p[c + 1]);
    double du[D];
#pragma unroll
    for (int i = 0; i < D; ++i) du[i] = PrimSlope(v[c - 1][i], v[c][i], v[c + 1][i]);
    double rr, rp, ru[D];
    if (side == 0) {
      rr = rho[c] + 0.5 * (drho - lambda * (v[c][DIR] * drho + rho[c] * du[DIR]));
      rp = p[c] + 0.5 * (dp - lambda * (rho[c] * a * a * du[DIR] + v[c][DIR] * dp));
#pragma unroll
      for (int i = 0; i < D; ++i) ru[i] = v[c][i] + 0.5 * (du[i] - lambda * (v[c][DIR] * du[i]));
    } else {
      rr = rho[c] - 0.5 * (drho + lambda * (v[c][DIR] * drho + rho[c] * du[DIR]));
      rp = p[c] - 0.5 * (dp + lambda * (rho[c] * a * a * du[DIR] + v[c][DIR] * dp));
#pragma unroll
      for (int i = 0; i < D; ++i) ru[i] = v[c][i] - 0.5 * (du[i] + lambda * (v[c][DIR] * du[i]));
    }
    ru[DIR] = ru[DIR] - DivRn(0.5 * lambda * dp, rho[c]);
    double* w = side == 0 ? L : R;
    w[0] = rr;
#pragma unroll
    for (int i = 0; i < D; ++i) w[1 + i] = rr * ru[i];
    w[D + 2] = rp;
    w[D + 3] = __builtin_sqrt(DivRn(g.gamma * rp, rr));
    w[D + 1] = 0.5 * rr * SqNormArray<D>(ru) + rp * g.gm1inv;
  }
}

// MusclHancockCharacteristic, array path
// (src/equations/perfect_gas/MusclHancockCharacteristicMethod.hpp:26-285):
// the face state of cell u[1] of the complete stencil u[0..2] on its right
// (side 0) or left (side 1) face: VanLeer-limited wave amplitudes
// (minus, zero, plus, v, w) of the primitive differences, advanced along
// their characteristic speeds over half a step, recombined, added to the
// cell's primitives and completed.  Velocities have three rows (zero beyond
// D); ix = DIR, iy = ix + 1, iz = iy + 1 (mod 3).  FP64, IEEE.
__device__ __forceinline__ double VanLeerPair(double sL, double sR) {
  const double r = 2.0 * sL * sR;
  return r > 0.0 ? DivRn(r, sL + sR) : 0.0;
}

template <int D, int DIR>
__device__ __forceinline__ void CharacteristicFaceState(const double (*u)[D + 4], double lambda,
                                                        int side, const Gas<double>& g,
                                                        double* w) {
  constexpr int ix = DIR, iy = (ix + 1) % 3, iz = (iy + 1) % 3;
  const double rho = u[1][0];
  const double un = DivRn(u[1][1 + DIR], rho);
  const double c = u[1][D + 3];
  const double rhoc = rho * c;
  const double ooc2 = DivRn(1.0, c * c);
  double pr[3], pp[3], pv[3][3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    pr[s] = u[s][0];
    pp[s] = u[s][D + 2];
#pragma unroll
    for (int d = 0; d < 3; ++d) pv[s][d] = d < D ? DivRn(u[s][1 + d], u[s][0]) : 0.0;
  }
  double amp[2][5];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double dp = pp[k + 1] - pp[k];
    const double du = pv[k + 1][ix] - pv[k][ix];
    const double drho = pr[k + 1] - pr[k];
    amp[k][0] = dp - rhoc * du;
    amp[k][1] = drho - ooc2 * dp;
    amp[k][2] = dp + rhoc * du;
    amp[k][3] = pv[k + 1][iy] - pv[k][iy];
    amp[k][4] = pv[k + 1][iz] - pv[k][iz];
  }
  const double sign = side == 0 ? 1.0 : -1.0;
  double a[5];
  const double speed[5] = {un - c, un, un + c, un, un};
#pragma unroll
  for (int m = 0; m < 5; ++m)
    a[m] = sign * 0.5 * VanLeerPair(amp[0][m], amp[1][m]) * (1.0 - sign * lambda * speed[m]);
  double dpres = 0.5 * (a[0] + a[2]);
  double drho = ooc2 * dpres + a[1];
  double dvel[3];
  dvel[ix] = DivRn(0.5 * (a[2] - a[0]), rhoc);
  dvel[iy] = a[3];
  dvel[iz] = a[4];
  drho += rho;
#pragma unroll
  for (int d = 0; d < D; ++d) dvel[d] += DivRn(u[1][1 + d], u[1][0]);
  dpres += u[1][D + 2];
  w[0] = drho;
#pragma unroll
  for (int d = 0; d < D; ++d) w[1 + d] = drho * dvel[d];
  w[D + 2] = dpres;
  const double e_kin = DivRn(0.5 * SqNormArray<D>(w + 1), drho);
  w[D + 1] = e_kin + dpres * g.gm1inv;
  w[D + 3] = __builtin_sqrt(DivRn(g.gamma * dpres, drho));
}

// MUSCL-Hancock reconstruction of one cell from its conservative neighbours
// um (left), u0 (centre), up (right).  Produces the complete half-step
// states at the cell's left face (recL = rec_[1] of that face) and right
// face (recR = rec_[0] of that face).  lh = 0.5 * dt / dx.
// irL / irR receive the reciprocals of recL's and recR's densities.
template <int D, int DIR, int LIM, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void ReconstructCell(const T* um, const T* u0, const T* up,
                                                T lh, const Gas<T>& g, T* recL, T* recR,
                                                typename A::Inv& irL, typename A::Inv& irR) {
  using I = Idx<D>;
  T qL[I::kComp], qR[I::kComp];
#pragma unroll
  for (int k = 0; k < I::kCons; ++k) {
    const T s = LimitedSlope<LIM, T, A>(um[k], u0[k], up[k]);
    qL[k] = u0[k] - T(0.5) * s;
    qR[k] = u0[k] + T(0.5) * s;
  }
  const typename A::Inv iqL = CompleteFromCons<D, T, A>(qL, g);
  const typename A::Inv iqR = CompleteFromCons<D, T, A>(qR, g);
  T fL[I::kCons], fR[I::kCons];
  PhysFlux<D, DIR, T, A>(fL, qL, iqL);
  PhysFlux<D, DIR, T, A>(fR, qR, iqR);
#pragma unroll
  for (int k = 0; k < I::kCons; ++k) {
    const T dq = lh * (fL[k] - fR[k]);
    recR[k] = qR[k] + dq;
    recL[k] = qL[k] + dq;
  }
  irR = CompleteFromCons<D, T, A>(recR, g);
  irL = CompleteFromCons<D, T, A>(recL, g);
}

template <int D, int DIR, int LIM, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void ReconstructCell(const T* um, const T* u0, const T* up,
                                                T lh, const Gas<T>& g, T* recL,
                                                T* recR) {
  typename A::Inv irL, irR;
  ReconstructCell<D, DIR, LIM, T, A>(um, u0, up, lh, g, recL, recR, irL, irR);
}

// Reconstruction of a ghost cell whose three stencil cells are copies of the
// same boundary state u (transmissive boundary): every limiter returns a zero
// slope, qL = qR = u, the half-step correction vanishes and the result is the
// complete state of u (equal in value to the general formula).
template <int D, class T, class A = IeeeArith<T>>
__device__ __forceinline__ typename A::Inv GhostRec(const T* u, const Gas<T>& g, T* rec) {
#pragma unroll
  for (int k = 0; k < D + 2; ++k) rec[k] = u[k];
  return CompleteFromCons<D, T, A>(rec, g);
}

// Hllem<PerfectGas<D>, Larrouturou> numerical flux in direction DIR between
// complete states L and R; writes the D+2 flux components.
// irL / irR: reciprocals of the two densities.
template <int D, int DIR, bool LARR, bool QUIRK, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void HllemFlux(T* flux, const T* L, const typename A::Inv& irL,
                                          const T* R, const typename A::Inv& irR,
                                          const Gas<T>& g) {
  using I = Idx<D>;
  constexpr int M = I::kMom;
  const T rhoL = L[I::kRho];
  const T rhoR = R[I::kRho];
  const T aL = L[I::kA];
  const T aR = R[I::kA];
  const T pL = L[I::kP];
  const T pR = R[I::kP];
  const T sqRhoL = A::sqrt(rhoL);
  const T sqRhoR = A::sqrt(rhoR);
  const T sqRhoSum = sqRhoL + sqRhoR;
  const typename A::Inv isum = A::inv(sqRhoSum);
  const T sqRhoL_over_Sum = A::div(sqRhoL, isum);
  const T sqRhoR_over_Sum = A::div(sqRhoR, isum);
  const T roeGamma = A::div(sqRhoL * g.gamma + sqRhoR * g.gamma, isum);
  const T gm1 = roeGamma - T(1);
  if constexpr (LARR) {
    // Transverse velocities and kinetic energies (HllemMethod.hpp:729-745);
    // the state's transverse momentum becomes zero.
    T vL[D], vR[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      vL[i] = A::div(L[M + i], irL);
      vR[i] = A::div(R[M + i], irR);
    }
    T eKinRL = T(0), eKinRR = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) {
      if (i != DIR) {
        eKinRL += vL[i] * vL[i];
        eKinRR += vR[i] * vR[i];
      }
    }
    eKinRL *= T(0.5);
    eKinRR *= T(0.5);
    const T EL = L[I::kE] - rhoL * eKinRL;
    const T ER = R[I::kE] - rhoR * eKinRR;
    const T mL = L[M + DIR];
    const T mR = R[M + DIR];
    // Physical fluxes of the modified states (only the components used).
    const T uLd = vL[DIR]; // == mL / rhoL
    const T uRd = vR[DIR];
    const T fL_rho = mL;
    const T fR_rho = mR;
    const T fL_m = uLd * mL + pL;
    const T fR_m = uRd * mR + pR;
    const T fL_E = uLd * (EL + pL);
    const T fR_E = uRd * (ER + pR);
    const T hL = A::div(EL + pL, irL);
    const T hR = A::div(ER + pR, irR);
    const T roeU = sqRhoL_over_Sum * uLd + sqRhoR_over_Sum * uRd;
    const T roeH = sqRhoL_over_Sum * hL + sqRhoR_over_Sum * hR;
    const T squaredNormRoeU = roeU * roeU;
    const T squaredNormRoeU_half = T(0.5) * squaredNormRoeU;
    const T roeA2 = gm1 * (roeH - squaredNormRoeU_half);
    const T roeA = A::sqrt(roeA2);
    const T maxA = A::max(aL, aR);
    const T sL1 = uLd - maxA;
    const T sL2 = roeU - roeA;
    const T sR1 = roeU + roeA;
    const T sR2 = uRd + maxA;
    const T sL = A::min(sL1, sL2);
    const T sR = A::max(sR1, sR2);
    const T bL = A::min(sL, T(0));
    const T bR = A::max(sR, T(0));
    const T bLbR = bL * bR;
    const T db = bR - bL;
    const T db_positive = (db > T(0)) ? db : T(1);
    const typename A::Inv idb = A::inv(db_positive);
    const T hlle_rho = A::div(bR * fL_rho - bL * fR_rho + bLbR * (rhoR - rhoL), idb);
    const T hlle_m = A::div(bR * fL_m - bL * fR_m + bLbR * (mR - mL), idb);
    const T hlle_E = A::div(bR * fL_E - bL * fR_E + bLbR * (ER - EL), idb);
    const T u_bar = T(0.5) * (sR + sL);
    const T delta = A::quot(roeA, roeA + rabs(u_bar));
    const T b = A::div(bLbR, idb);
    const T deltaRho = rhoR - rhoL;
    const T deltaRhoU = mR - mL;
    const T deltaRhoE = ER - EL;
    const T gm1_over_roeA2 = A::quot(gm1, roeA2);
    const T l21 = gm1_over_roeA2 * (roeH - squaredNormRoeU);
    const T lUd = gm1_over_roeA2 * roeU;
    const T lE = -gm1_over_roeA2;
    // alpha_2 summed in component order; the transverse terms are zero.
    T alpha_2;
    if constexpr (DIR == 0) {
      alpha_2 = l21 * deltaRho + lUd * deltaRhoU;
    } else {
      alpha_2 = l21 * deltaRho + lUd * deltaRhoU;
    }
    alpha_2 = alpha_2 + lE * deltaRhoE;
    const T b_delta_alpha_2 = b * delta * alpha_2;
    const T frho = hlle_rho - b_delta_alpha_2;
    flux[I::kRho] = frho;
    if constexpr (D == 3 && QUIRK) {
      // HllemMethod.hpp:990-992 uses roeU.row(iy), which the Larrouturou
      // variant has zeroed: the normal momentum flux keeps the HLLE value.
      flux[M + DIR] = hlle_m;
    } else {
      flux[M + DIR] = hlle_m - b_delta_alpha_2 * roeU;
    }
    const bool upwind = frho > T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) {
      if (i != DIR) flux[M + i] = frho * (upwind ? vL[i] : vR[i]);
    }
    const T f_rhoE = frho * (upwind ? eKinRL : eKinRR);
    flux[I::kE] = hlle_E - b_delta_alpha_2 * squaredNormRoeU_half + f_rhoE;
  } else {
    // Plain HLLEM (Larrouturou = false): every component.
    T fL[I::kCons], fR[I::kCons];
    PhysFlux<D, DIR, T, A>(fL, L, irL);
    PhysFlux<D, DIR, T, A>(fR, R, irR);
    const T hL = A::div(L[I::kE] + pL, irL);
    const T hR = A::div(R[I::kE] + pR, irR);
    T uL[D], uR[D], roeU[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      uL[i] = A::div(L[M + i], irL);
      uR[i] = A::div(R[M + i], irR);
      roeU[i] = sqRhoL_over_Sum * uL[i] + sqRhoR_over_Sum * uR[i];
    }
    const T roeH = sqRhoL_over_Sum * hL + sqRhoR_over_Sum * hR;
    T squaredNormRoeU = roeU[0] * roeU[0];
#pragma unroll
    for (int i = 1; i < D; ++i) squaredNormRoeU += roeU[i] * roeU[i];
    const T squaredNormRoeU_half = T(0.5) * squaredNormRoeU;
    const T roeA2 = gm1 * (roeH - squaredNormRoeU_half);
    const T roeA = A::sqrt(roeA2);
    const T maxA = A::max(aL, aR);
    const T sL = A::min(uL[DIR] - maxA, roeU[DIR] - roeA);
    const T sR = A::max(roeU[DIR] + roeA, uR[DIR] + maxA);
    const T bL = A::min(sL, T(0));
    const T bR = A::max(sR, T(0));
    const T bLbR = bL * bR;
    const T db = bR - bL;
    const T db_positive = (db > T(0)) ? db : T(1);
    const typename A::Inv idb = A::inv(db_positive);
    T hlle[I::kCons];
#pragma unroll
    for (int k = 0; k < I::kCons; ++k)
      hlle[k] = A::div(bR * fL[k] - bL * fR[k] + bLbR * (R[k] - L[k]), idb);
    const T u_bar = T(0.5) * (sR + sL);
    const T delta = A::quot(roeA, roeA + rabs(u_bar));
    const T b = A::div(bLbR, idb);
    const T deltaRho = rhoR - rhoL;
    T deltaRhoU[D];
#pragma unroll
    for (int i = 0; i < D; ++i) deltaRhoU[i] = R[M + i] - L[M + i];
    const T deltaRhoE = R[I::kE] - L[I::kE];
    const T gm1_over_roeA2 = A::quot(gm1, roeA2);
    const T l21 = gm1_over_roeA2 * (roeH - squaredNormRoeU);
    T alpha_2 = l21 * deltaRho;
#pragma unroll
    for (int i = 0; i < D; ++i) alpha_2 = alpha_2 + (gm1_over_roeA2 * roeU[i]) * deltaRhoU[i];
    alpha_2 = alpha_2 + (-gm1_over_roeA2) * deltaRhoE;
    const T b_delta = b * delta;
    const T b_delta_alpha_2 = b_delta * alpha_2;
    flux[I::kRho] = hlle[I::kRho] - b_delta_alpha_2;
    constexpr int IY = (D == 2) ? (DIR == 0 ? 1 : 0) : (DIR + 1) % 3;
    constexpr int IZ = (IY + 1) % 3;
    if constexpr (D == 3 && QUIRK) {
      flux[M + DIR] = hlle[M + DIR] - b_delta_alpha_2 * roeU[IY];
    } else {
      flux[M + DIR] = hlle[M + DIR] - b_delta_alpha_2 * roeU[DIR];
    }
    const T alpha_3 = (-roeU[IY]) * deltaRho + deltaRhoU[IY];
    const T b_delta_alpha_3 = b_delta * alpha_3;
    flux[M + IY] = hlle[M + IY] - b_delta_alpha_2 * roeU[IY] - b_delta_alpha_3;
    if constexpr (D == 3) {
      const T alpha_4 = (-roeU[IZ]) * deltaRho + deltaRhoU[IZ];
      const T b_delta_alpha_4 = b_delta * alpha_4;
      flux[M + IZ] = hlle[M + IZ] - b_delta_alpha_2 * roeU[IZ] - b_delta_alpha_4;
      flux[I::kE] = hlle[I::kE] - b_delta_alpha_2 * squaredNormRoeU_half -
                    b_delta_alpha_3 * roeU[IY] - b_delta_alpha_4 * roeU[IZ];
    } else {
      flux[I::kE] = hlle[I::kE] - b_delta_alpha_2 * squaredNormRoeU_half -
                    b_delta_alpha_3 * roeU[IY];
    }
  }
}

template <int D, int DIR, bool LARR, bool QUIRK, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void HllemFlux(T* flux, const T* L, const T* R, const Gas<T>& g) {
  HllemFlux<D, DIR, LARR, QUIRK, T, A>(flux, L, A::inv(L[Idx<D>::kRho]), R,
                                       A::inv(R[Idx<D>::kRho]), g);
}

// Hllem::ComputeStableDt (array form) between complete states L and R.
template <int D, int DIR, class T, class A = IeeeArith<T>>
__device__ __forceinline__ T HllemStableDt(const T* L, const T* R, T dx,
                                           const Gas<T>& g) {
  using I = Idx<D>;
  constexpr int M = I::kMom;
  const T rhoL = L[I::kRho];
  const T rhoR = R[I::kRho];
  const typename A::Inv irL = A::inv(rhoL);
  const typename A::Inv irR = A::inv(rhoR);
  const T hL = A::div(L[I::kE] + L[I::kP], irL);
  const T hR = A::div(R[I::kE] + R[I::kP], irR);
  const T sqRhoL = A::sqrt(rhoL);
  const T sqRhoR = A::sqrt(rhoR);
  const T sqRhoSum = sqRhoL + sqRhoR;
  const typename A::Inv isum = A::inv(sqRhoSum);
  const T wL = A::div(sqRhoL, isum);
  const T wR = A::div(sqRhoR, isum);
  T uL[D], uR[D], roeU[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    uL[i] = A::div(L[M + i], irL);
    uR[i] = A::div(R[M + i], irR);
    roeU[i] = wL * uL[i] + wR * uR[i];
  }
  const T roeH = wL * hL + wR * hR;
  T sq = roeU[0] * roeU[0];
#pragma unroll
  for (int i = 1; i < D; ++i) sq += roeU[i] * roeU[i];
  const T roeA2 = g.gm1 * (roeH - T(0.5) * sq);
  const T roeA = A::sqrt(roeA2);
  const T maxA = A::max(L[I::kA], R[I::kA]);
  const T sL = A::min(uL[DIR] - maxA, roeU[DIR] - roeA);
  const T sR = A::max(roeU[DIR] + roeA, uR[DIR] + maxA);
  const T maxS = A::max(rabs(sL), rabs(sR));
  return A::quot(dx, maxS);
}

// EinfeldtSignalVelocities<PerfectGas<D>> (array form,
// src/equations/perfect_gas/EinfeldtSignalVelocities.hpp:64-97): signal
// speeds from the Roe averages of the normal velocity and the enthalpy; the
// Roe sound speed uses the normal velocity only.  irL / irR are the
// reciprocals of the densities.
template <int D, int DIR, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void EinfeldtSignals(const T* L, const typename A::Inv& irL,
                                                const T* R, const typename A::Inv& irR,
                                                const Gas<T>& g, T& sL, T& sR) {
  using I = Idx<D>;
  const T sqRhoL = A::sqrt(L[I::kRho]);
  const T sqRhoR = A::sqrt(R[I::kRho]);
  const T uL = A::div(L[I::kMom + DIR], irL);
  const T uR = A::div(R[I::kMom + DIR], irR);
  const T oosqRhoSum = A::quot(T(1), sqRhoL + sqRhoR);
  const T roeU = (uL * sqRhoL + uR * sqRhoR) * oosqRhoSum;
  const T hL = A::div(L[I::kE] + L[I::kP], irL);
  const T hR = A::div(R[I::kE] + R[I::kP], irR);
  const T roeH = (hL * sqRhoL + hR * sqRhoR) * oosqRhoSum;
  const T roeA2 = (g.gamma - T(1)) * (roeH - T(0.5) * roeU * roeU);
  const T roeA = A::sqrt(roeA2);
  sL = A::min(uL - L[I::kA], roeU - roeA);
  sR = A::max(roeU + roeA, uR + R[I::kA]);
}

// Hll<PerfectGas<D>, EinfeldtSignalVelocities>::ComputeNumericFlux (array
// form, include/fub/flux_method/HllMethod.hpp:260-293).
template <int D, int DIR, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void HllFlux(T* flux, const T* L, const typename A::Inv& irL,
                                        const T* R, const typename A::Inv& irR,
                                        const Gas<T>& g) {
  using I = Idx<D>;
  T s0, s1;
  EinfeldtSignals<D, DIR, T, A>(L, irL, R, irR, g, s0, s1);
  T fL[I::kCons], fR[I::kCons];
  PhysFlux<D, DIR, T, A>(fL, L, irL);
  PhysFlux<D, DIR, T, A>(fR, R, irR);
  const T sL = A::min(s0, T(0));
  const T sR = A::max(s1, T(0));
  const T sLsR = sL * sR;
  const T ds = sR - sL;
  const typename A::Inv ids = A::inv((ds > T(0)) ? ds : T(1));
#pragma unroll
  for (int k = 0; k < I::kCons; ++k)
    flux[k] = A::div(sR * fL[k] - sL * fR[k] + sLsR * (R[k] - L[k]), ids);
}

template <int D, int DIR, class T, class A = IeeeArith<T>>
__device__ __forceinline__ void HllFlux(T* flux, const T* L, const T* R, const Gas<T>& g) {
  HllFlux<D, DIR, T, A>(flux, L, A::inv(L[Idx<D>::kRho]), R, A::inv(R[Idx<D>::kRho]), g);
}

// Hll::ComputeStableDt (array form, include/fub/flux_method/HllMethod.hpp:337-347):
// dx / max(0, |sL|, |sR|) accumulated left to right.
template <int D, int DIR, class T, class A = IeeeArith<T>>
__device__ __forceinline__ T HllStableDt(const T* L, const T* R, T dx, const Gas<T>& g) {
  T s0, s1;
  EinfeldtSignals<D, DIR, T, A>(L, A::inv(L[Idx<D>::kRho]), R, A::inv(R[Idx<D>::kRho]), g, s0,
                                s1);
  T m = A::max(T(0), rabs(s0));
  m = A::max(m, rabs(s1));
  return A::quot(dx, m);
}

// Stable dt of a face whose two states are the same state Q (boundary face
// between a cell and its transmissive ghost copy).  With L == R the Roe
// weights are exactly 1/2 and every Roe average equals the state's own value,
// so this equals HllemStableDt(Q, Q) bit for bit.
template <int D, int DIR, class T, class A = IeeeArith<T>>
__device__ __forceinline__ T HllemStableDtUniform(const T* Q, T dx, const Gas<T>& g) {
  using I = Idx<D>;
  constexpr int M = I::kMom;
  const T rho = Q[I::kRho];
  const typename A::Inv ir = A::inv(rho);
  const T h = A::div(Q[I::kE] + Q[I::kP], ir);
  T u[D];
#pragma unroll
  for (int i = 0; i < D; ++i) u[i] = A::div(Q[M + i], ir);
  T sq = u[0] * u[0];
#pragma unroll
  for (int i = 1; i < D; ++i) sq += u[i] * u[i];
  const T roeA2 = g.gm1 * (h - T(0.5) * sq);
  const T roeA = A::sqrt(roeA2);
  const T a = Q[I::kA];
  const T sL = A::min(u[DIR] - a, u[DIR] - roeA);
  const T sR = A::max(u[DIR] + roeA, u[DIR] + a);
  const T maxS = A::max(rabs(sL), rabs(sR));
  return A::quot(dx, maxS);
}

// Order-preserving integer key of a non-negative double (for atomic min).
__device__ __forceinline__ unsigned long long DtKey(double dt) {
  // Stable dts are positive (or +inf); NaN maps above +inf so it never wins.
  if (dt != dt) return 0x7ff8000000000000ull;
  return static_cast<unsigned long long>(__double_as_longlong(dt));
}

// Sufficient test of a sweep's CFL check (the level driver needs only its
// outcome while it passes).  With w = sqrt(rho_L) / (sqrt(rho_L) +
// sqrt(rho_R)) and h = a^2 / (g - 1) + |u|^2 / 2, the Roe sound speed of
// FaceSpeed satisfies a~^2 = w aL^2 + (1 - w) aR^2 + (g - 1) / 2 w (1 - w)
// |uR - uL|^2 <= max(aL^2, aR^2) + (g - 1) / 8 |uR - uL|^2, and |u~_n| <=
// max(|uL_n|, |uR_n|), so the face speed is at most U + sqrt(A^2 + k |du|^2)
// (U the larger normal speed, A^2 the larger squared sound speed, k = (g -
// 1) / 8).  A face passes when that bound stays below dx / dt by a relative
// margin of 2^-20, far above the rounding of either side; then its stable
// dt, dx / speed, is at least dt and the exact check passes too.  A wave
// evaluates the exact speeds when any of its faces does not pass the bound,
// so a failing check reports the exact minimum dt as before.  The bound
// uses the refined hardware reciprocal in every arithmetic mode.
template <int D> struct BoundCell {
  double u[D];
  double a2;  // squared speed of sound, gamma p / rho
};

template <int D>
__device__ __forceinline__ BoundCell<D> MakeBoundCell(const double* q, const Gas<double>& g) {
  using FA = FastArith<double>;
  const FA::Inv ir = FA::inv(q[0]);
  BoundCell<D> b;
#pragma unroll
  for (int i = 0; i < D; ++i) b.u[i] = FA::div(q[1 + i], ir);
  const double e_kin = FA::div(0.5 * SqNormArray<D>(q + 1), ir);
  const double p = (q[D + 1] - e_kin) * g.gm1;
  b.a2 = FA::div(g.gamma * p, ir);
  return b;
}

template <int D, int DIR>
__device__ __forceinline__ bool BoundPasses(const BoundCell<D>& L, const BoundCell<D>& R,
                                            double s_crit, double k) {
  const double U = __builtin_fmax(rabs(L.u[DIR]), rabs(R.u[DIR]));
  const double T = s_crit - U;
  double du2 = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const double d = R.u[i] - L.u[i];
    du2 += d * d;
  }
  const double q = __builtin_fmax(L.a2, R.a2) + k * du2;
  return T >= 0.0 && L.a2 >= 0.0 && R.a2 >= 0.0 && q <= T * T;
}

// dx / dt shrunk by the bound's margin (+inf for dt = 0: every face passes).
__device__ __forceinline__ double BoundSpeedLimit(double dx, double dt) {
  return (dx / dt) * (1.0 - 0x1p-20);
}

} // namespace fvs
