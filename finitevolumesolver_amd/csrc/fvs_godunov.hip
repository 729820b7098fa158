// Exact-Godunov flux method on gfx950: FluxMethod<Godunov<PerfectGas<D>,
// ExactRiemannSolver>> (include/fub/flux_method/GodunovMethod.hpp:130-199).
//
// Fused first-order sweeps (X: slot-mapped lanes like the HLLEM sweep; Y/Z:
// pencil march).  The face flux and the face's CFL signal speeds need the
// same Riemann problem of the same two cell states, so one FP64 Newton
// iteration serves both (the reference solves it twice: once in
// ComputeNumericFlux, once in ComputeSignals).  An optional per-face initial
// guess (the learned star pressure) replaces 0.5 (pL + pR); Newton itself,
// its tolerance and its safeguards are unchanged.  Newton statistics
// (iterations, non-converged faces) are accumulated per workgroup.
#include "fvs_internal.hpp"
#include "fvs_mlp_infer.hpp"
#include "fvs_riemann.hpp"
#include "fvs_views.hpp"

#include <cstring>
#include <type_traits>

namespace fvs {

namespace {

constexpr int kBlock = 256;
constexpr int kXOwned = 62;
constexpr int kTX = 256;
constexpr int kCflSeg = 32;
constexpr unsigned long long kInfKey = 0x7ff0000000000000ull;

__device__ __forceinline__ int64_t Clamp(int64_t i, int64_t n) {
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
__device__ __forceinline__ unsigned long long KMin(unsigned long long a, unsigned long long b) {
  return b < a ? b : a;
}
__device__ __forceinline__ unsigned long long WaveMinKey(unsigned long long k) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) k = KMin(k, __shfl_xor(k, off, 64));
  return k;
}
__device__ __forceinline__ unsigned long long WaveSum(unsigned long long k) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) k += __shfl_xor(k, off, 64);
  return k;
}
__device__ __forceinline__ unsigned long long WaveMaxU(unsigned long long k) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(k, off, 64);
    k = o > k ? o : k;
  }
  return k;
}

// Block reduction of (min dt key, sum iterations, max iterations,
// non-converged faces, faces) to partials[block * 5 + i].
__device__ void BlockReduceGodunov(unsigned long long key, unsigned long long iters,
                                   unsigned long long maxit, unsigned long long fails,
                                   unsigned long long faces, unsigned long long* out) {
  __shared__ unsigned long long s[5][kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  key = WaveMinKey(key);
  iters = WaveSum(iters);
  maxit = WaveMaxU(maxit);
  fails = WaveSum(fails);
  faces = WaveSum(faces);
  if (lane == 0) {
    s[0][wave] = key;
    s[1][wave] = iters;
    s[2][wave] = maxit;
    s[3][wave] = fails;
    s[4][wave] = faces;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long r[5] = {s[0][0], s[1][0], s[2][0], s[3][0], s[4][0]};
    for (int w = 1; w < kBlock / 64; ++w) {
      r[0] = KMin(r[0], s[0][w]);
      r[1] += s[1][w];
      r[2] = s[2][w] > r[2] ? s[2][w] : r[2];
      r[3] += s[3][w];
      r[4] += s[4][w];
    }
    for (int i = 0; i < 5; ++i) out[i] = r[i];
  }
}

struct GodunovArgs {
  const double* in[5];
  double* out[5];
  int64_t n[3];
  double dx;
  const double* dt_ptr;
  double dt_val;
  double gamma;
  double gm1inv;
  double tol;
  int max_it;
  const double* guess;   // per-face initial guess (nullable)
  const short* tiles;    // learned guess computed inline (nullable)
  const float* bias;
  unsigned long long* partials;  // 5 per workgroup
  int* halt;
  int seg;
  int update;            // 0: only fluxes' Newton + CFL (no state write)
  int reuse;             // guess holds the star pressures the step-start CFL
                         // pass found for these faces on the same state
                         // (NaN where its Newton did not converge)
};

__device__ __forceinline__ Gas<double> GasOf(const GodunovArgs& a) {
  Gas<double> g;
  g.gamma = a.gamma;
  g.gm1inv = a.gm1inv;
  g.gm1 = a.gamma - 1.0;
  return g;
}

template <int D>
__device__ __forceinline__ void ShflUpState(const double* q, double* r) {
#pragma unroll
  for (int k = 0; k < D + 4; ++k) r[k] = __shfl_up(q[k], 1, 64);
}

// Mixed form of the flux: the sampled state and its physical flux in float.
template <int D, int DIR>
__device__ __forceinline__ void GodunovFluxMixed(double* flux, const double* L, const double* R,
                                                 const MiddleState& ms, const Gas<double>& g) {
  constexpr int NQ = D + 4;
  float lf[NQ], rf[NQ], sol[NQ], ff[D + 2];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    lf[k] = static_cast<float>(L[k]);
    rf[k] = static_cast<float>(R[k]);
    sol[k] = 0.0f;
  }
  Gas<float> gf;
  gf.gamma = static_cast<float>(g.gamma);
  gf.gm1inv = static_cast<float>(g.gm1inv);
  gf.gm1 = static_cast<float>(g.gm1);
  if (0.0 < L[0] && 0.0 < R[0])
    SampleRiemannT<D, DIR, float>(sol, lf, rf, static_cast<float>(ms.pM),
                                  static_cast<float>(ms.uM), gf);
  PhysFlux<D, DIR, float>(ff, sol);
#pragma unroll
  for (int k = 0; k < D + 2; ++k) flux[k] = static_cast<double>(ff[k]);
}

// Newton from a caller's seed (learned or given), with the reference's
// initial guess, the mean pressure, as the fallback.  A seed that is not a
// finite positive pressure (an overflowed or underflowed network output) is
// not tried; a run from the seed that does not converge, or ends on a
// pressure that is not finite and non-negative, is redone from the mean.
// Both count as a fallback.
template <int D, int DIR, bool MIXED>
__device__ __forceinline__ MiddleState SeededMiddleState(const double* L, const double* R,
                                                         const Gas<double>& g,
                                                         const NewtonParams& np, bool seeded,
                                                         double seed, double mean, int* iters,
                                                         int* fallback) {
  const bool usable = seeded && seed > 0.0 && seed <= 1.7976931348623157e308;
  MiddleState ms = ComputeMiddleState<D, DIR, MIXED>(L, R, g, np, usable ? seed : mean);
  int it = ms.iterations;
  int fb = seeded && !usable;
  if (usable && !(ms.converged && ms.pM >= 0.0 && ms.pM <= 1.7976931348623157e308)) {
    ms = ComputeMiddleState<D, DIR, MIXED>(L, R, g, np, mean);
    it += ms.iterations;
    fb = 1;
  }
  *iters = it;
  *fallback = fb;
  return ms;
}

// One face: Newton (shared by flux and CFL), flux, dt.  MIXED: the Newton
// iteration stays FP64 (with exp/log powers) and the sample and flux are
// float.
//
// With a.reuse the seed is the converged iterate of the step-start CFL pass
// for the same face and state: Newton from it evaluates the side functions
// at that pressure, finds the residual within the tolerance and stops at
// once with the same star pressure and velocity, so the face is not solved
// again (and not counted as a solved face, bit 2 of *conv).  A NaN seed (no
// converged iterate) solves the face from the mean pressure.
template <int D, int DIR, bool MIXED>
__device__ __forceinline__ void GodunovFace(const double* L, const double* R, double guess,
                                            bool has_guess, const GodunovArgs& a,
                                            const Gas<double>& g, double* F, double* dt_face,
                                            int* iters, int* conv) {
  NewtonParams np;
  np.tol = a.tol;
  np.max_it = a.max_it;
  using I = Idx<D>;
  const double mean = 0.5 * (L[I::kP] + R[I::kP]);
  int it, fallback = 0, reused = 0;
  MiddleState ms;
  if (a.reuse) {
    reused = guess > 0.0 && guess <= 1.7976931348623157e308;
    ms = ComputeMiddleState<D, DIR, MIXED>(L, R, g, np, reused ? guess : mean);
    it = ms.iterations;
  } else {
    ms = SeededMiddleState<D, DIR, MIXED>(L, R, g, np, has_guess, guess, mean, &it, &fallback);
  }
  if constexpr (MIXED) {
    GodunovFluxMixed<D, DIR>(F, L, R, ms, g);
  } else {
    GodunovFluxFromMiddle<D, DIR>(F, L, R, ms, g);
  }
  *dt_face = GodunovDtFromMiddle<D, DIR>(L, R, ms.pM, g, a.dx);
  *iters = it;
  *conv = ms.converged | (fallback << 1) | (reused << 2);
}

// LEARNED: the network's guess is inferred inline (a.tiles set); without it
// the kernel carries none of the matrix-core code or its registers.
template <int D, bool MIXED, bool LEARNED>
__global__ __launch_bounds__(kBlock) void GodunovSweepX(GodunovArgs a) {
  constexpr int NC = D + 2, NQ = D + 4;
  if (a.halt && *a.halt) return;
  const int64_t nx = a.n[0];
  const uint32_t S = static_cast<uint32_t>(nx + 2);
  const uint32_t Q = static_cast<uint32_t>(a.n[1] * a.n[2]) * S;
  const int lane = threadIdx.x & 63;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  const int64_t q64 = wave * kXOwned + lane - 1;
  const bool valid = q64 >= 0 && q64 < static_cast<int64_t>(Q);
  const uint32_t q = valid ? static_cast<uint32_t>(q64) : 0u;
  const uint32_t row = q / S;
  const uint32_t sl = q - row * S;
  const int64_t cell = static_cast<int64_t>(sl) - 1;
  const bool real = valid && cell >= 0 && cell < nx;
  const int64_t rbase = static_cast<int64_t>(row) * nx;
  const Gas<double> g = GasOf(a);
  const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
  const double dt_dx = DivRn(dt, a.dx);

  double u[NQ];
  {
    const int64_t off = rbase + Clamp(cell, nx);
#pragma unroll
    for (int k = 0; k < NC; ++k) u[k] = a.in[k][off];
    CompleteFromCons<D>(u, g);
  }
  double L[NQ];
  ShflUpState<D>(u, L);
  const bool face_lane = valid && lane >= 1 && sl >= 1;
  const bool owned = lane >= 1 && lane <= kXOwned;
  double F[NC];
  unsigned long long key = kInfKey, it_sum = 0, it_max = 0, nfail = 0, nfaces = 0;
  // Inline learned guess: every lane takes part in the wave's matrix-core step.
  double inline_guess = 0.0;
  if constexpr (LEARNED) inline_guess = mlp::LearnedGuess<D, 0>(L, u, a.tiles, a.bias);
  if (face_lane) {
    const int64_t f = static_cast<int64_t>(sl) - 1;
    const double guess = LEARNED   ? inline_guess
                         : a.guess ? a.guess[static_cast<int64_t>(row) * (nx + 1) + f]
                                   : 0.0;
    double dtf;
    int it, conv;
    GodunovFace<D, 0, MIXED>(L, u, guess, LEARNED || a.guess != nullptr, a, g, F, &dtf, &it,
                             &conv);
    if (owned) {
      key = DtKey(dtf);
      it_sum = it;
      it_max = it;
      nfail = (conv & 1) ? 0 : 1;
      nfaces = ((conv >> 2) & 1) ? 0 : 1 + (static_cast<unsigned long long>((conv >> 1) & 1) << 32);
    }
  } else {
#pragma unroll
    for (int k = 0; k < NC; ++k) F[k] = 0.0;
  }
  double Fnext[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) Fnext[k] = __shfl_down(F[k], 1, 64);
  if (a.update && owned && real) {
    const int64_t off = rbase + cell;
#pragma unroll
    for (int k = 0; k < NC; ++k) a.out[k][off] = u[k] + dt_dx * (F[k] - Fnext[k]);
  }
  BlockReduceGodunov(key, it_sum, it_max, nfail, nfaces, a.partials + 5 * blockIdx.x);
}

template <int D, int DIR, bool MIXED, bool LEARNED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void GodunovSweepPencil(GodunovArgs a) {
  constexpr int NC = D + 2, NQ = D + 4;
  if (a.halt && *a.halt) return;
  const int64_t nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const int64_t nd = a.n[DIR];
  const int64_t npencil = (DIR == 1) ? nx * nz : nx * ny;
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const bool active = p0 < npencil;
  // Lanes past the last pencil repeat it (all lanes of a wave march together,
  // which the inline matrix-core guess needs); their results are not kept.
  const int64_t p = active ? p0 : npencil - 1;
  const int64_t stride = (DIR == 1) ? nx : nx * ny;
  const int64_t base = (DIR == 1) ? (p % nx) + (p / nx) * nx * ny : p;
  // Face (x, j, z) of direction DIR in the face box (n with n[DIR] + 1).
  const int64_t fstride = stride;
  const int64_t fbase = (DIR == 1) ? (p % nx) + (p / nx) * nx * (ny + 1) : p;
  const int64_t j0 = static_cast<int64_t>(blockIdx.y) * a.seg;
  const int64_t j1 = (j0 + a.seg < nd) ? j0 + a.seg : nd;
  const Gas<double> g = GasOf(a);
  const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
  const double dt_dx = DivRn(dt, a.dx);
  (void)nz;

  unsigned long long key = kInfKey, it_sum = 0, it_max = 0, nfail = 0, nfaces = 0;
  {
    auto load = [&](double* q, int64_t j) {
      const int64_t off = base + Clamp(j, nd) * stride;
#pragma unroll
      for (int k = 0; k < NC; ++k) q[k] = a.in[k][off];
      CompleteFromCons<D>(q, g);
    };
    double qm[NQ], q0[NQ];
    load(qm, j0 - 1);
    double Fprev[NC];
    for (int64_t j = j0; j <= j1; ++j) {
      load(q0, j);
      double F[NC], dtf;
      int it, conv;
      double guess = 0.0;
      if constexpr (LEARNED) {
        guess = mlp::LearnedGuess<D, DIR>(qm, q0, a.tiles, a.bias);
      } else if (a.guess) {
        guess = a.guess[fbase + j * fstride];
      }
      GodunovFace<D, DIR, MIXED>(qm, q0, guess, LEARNED || a.guess != nullptr, a, g, F, &dtf,
                                 &it, &conv);
      if (active && (j < j1 || j1 == nd)) {
        key = KMin(key, DtKey(dtf));
        it_sum += it;
        it_max = (unsigned long long)it > it_max ? it : it_max;
        nfail += (conv & 1) ? 0 : 1;
        nfaces += ((conv >> 2) & 1) ? 0
                                    : 1 + (static_cast<unsigned long long>((conv >> 1) & 1) << 32);
      }
      if (a.update && active && j > j0) {
        const int64_t off = base + (j - 1) * stride;
#pragma unroll
        for (int k = 0; k < NC; ++k) a.out[k][off] = qm[k] + dt_dx * (Fprev[k] - F[k]);
      }
#pragma unroll
      for (int k = 0; k < NQ; ++k) qm[k] = q0[k];
#pragma unroll
      for (int k = 0; k < NC; ++k) Fprev[k] = F[k];
    }
  }
  BlockReduceGodunov(key, it_sum, it_max, nfail, nfaces,
                     a.partials + 5 * (blockIdx.x + static_cast<int64_t>(blockIdx.y) * gridDim.x));
}

// Reduction of the per-workgroup Godunov partials: out[0] = min dt,
// stats[0..3] = total iterations, max iterations, non-converged, faces.
constexpr int kRedBlock = 1024;
__global__ __launch_bounds__(kRedBlock) void GodunovFinalize(const unsigned long long* p,
                                                             int64_t n, double* out,
                                                             unsigned long long* stats,
                                                             Ctl* ctl) {
  if (ctl && ctl->halt) return;
  __shared__ unsigned long long s[5][kRedBlock / 64];
  unsigned long long r[5] = {kInfKey, 0, 0, 0, 0};
  for (int64_t i = threadIdx.x; i < n; i += kRedBlock) {
    r[0] = KMin(r[0], p[5 * i]);
    r[1] += p[5 * i + 1];
    r[2] = p[5 * i + 2] > r[2] ? p[5 * i + 2] : r[2];
    r[3] += p[5 * i + 3];
    r[4] += p[5 * i + 4];
  }
  r[0] = WaveMinKey(r[0]);
  r[1] = WaveSum(r[1]);
  r[2] = WaveMaxU(r[2]);
  r[3] = WaveSum(r[3]);
  r[4] = WaveSum(r[4]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int i = 0; i < 5; ++i) s[i][wave] = r[i];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < kRedBlock / 64; ++w) {
    r[0] = KMin(r[0], s[0][w]);
    r[1] += s[1][w];
    r[2] = s[2][w] > r[2] ? s[2][w] : r[2];
    r[3] += s[3][w];
    r[4] += s[4][w];
  }
  r[0] = s[0][0] < r[0] ? s[0][0] : r[0];
  const double v = __longlong_as_double(static_cast<long long>(r[0]));
  if (out) *out = v;
  if (stats) {
    // r[4] packs faces (low 32 bits per face) and fallbacks (high bits).
    stats[0] += r[1];
    stats[1] = r[2] > stats[1] ? r[2] : stats[1];
    stats[2] += r[3];
    stats[3] += r[4] & 0xffffffffull;
    stats[4] += r[4] >> 32;
  }
}

// Per-face Godunov flux on strided views (FluxMethod seq driver): first
// order, or MUSCL-Hancock with the sequential (scalar) CompleteFromCons.
template <int D, int DIR, int LIM>
__device__ __forceinline__ void ReconScalarKE(const double* um, const double* u0,
                                              const double* up, double lh,
                                              const Gas<double>& g, double* recL,
                                              double* recR) {
  constexpr int NC = D + 2, NQ = D + 4;
  double qL[NQ], qR[NQ];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const double s = LimitedSlope<LIM>(um[k], u0[k], up[k]);
    qL[k] = u0[k] - 0.5 * s;
    qR[k] = u0[k] + 0.5 * s;
  }
  CompleteFromConsScalar<D>(qL, g);
  CompleteFromConsScalar<D>(qR, g);
  double fL[NC], fR[NC];
  PhysFlux<D, DIR>(fL, qL);
  PhysFlux<D, DIR>(fR, qR);
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const double dq = lh * (fL[k] - fR[k]);
    recR[k] = qR[k] + dq;
    recL[k] = qL[k] + dq;
  }
  CompleteFromConsScalar<D>(recR, g);
  CompleteFromConsScalar<D>(recL, g);
}

template <int D, int DIR>
__global__ __launch_bounds__(kBlock) void GodunovViewFlux(DevView fl, DevView st, double dt,
                                                          double dx, Gas<double> g, int limiter,
                                                          int muscl, double tol, int max_it,
                                                          unsigned long long* nfail) {
  constexpr int NC = D + 2, NQ = D + 4;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= fl.count()) return;
  int64_t f[3];
  fl.index(q, f);
  double L[NQ], R[NQ];
  auto loadc = [&](int64_t off, double* out) {
    int64_t c[3] = {f[0], f[1], f[2]};
    c[DIR] += off;
    for (int k = 0; k < NQ; ++k) out[k] = *st.at(c, k);
  };
  if (muscl) {
    double u[4][NQ];
    for (int s = 0; s < 4; ++s) loadc(s - 2, u[s]);
    const double lh = 0.5 * dt / dx;
    double dummy[NQ];
    switch (limiter) {
    case kMinMod:
      ReconScalarKE<D, DIR, kMinMod>(u[0], u[1], u[2], lh, g, dummy, L);
      ReconScalarKE<D, DIR, kMinMod>(u[1], u[2], u[3], lh, g, R, dummy);
      break;
    case kVanLeer:
      ReconScalarKE<D, DIR, kVanLeer>(u[0], u[1], u[2], lh, g, dummy, L);
      ReconScalarKE<D, DIR, kVanLeer>(u[1], u[2], u[3], lh, g, R, dummy);
      break;
    case kNoLimiter:
      ReconScalarKE<D, DIR, kNoLimiter>(u[0], u[1], u[2], lh, g, dummy, L);
      ReconScalarKE<D, DIR, kNoLimiter>(u[1], u[2], u[3], lh, g, R, dummy);
      break;
    default:
      ReconScalarKE<D, DIR, kNoGradient>(u[0], u[1], u[2], lh, g, dummy, L);
      ReconScalarKE<D, DIR, kNoGradient>(u[1], u[2], u[3], lh, g, R, dummy);
      break;
    }
  } else {
    loadc(-1, L);
    loadc(0, R);
  }
  NewtonParams np;
  np.tol = tol;
  np.max_it = max_it;
  const MiddleState ms = ComputeMiddleState<D, DIR>(L, R, g, np, 0.5 * (L[D + 2] + R[D + 2]));
  if (!ms.converged) atomicAdd(nfail, 1ull);
  double F[NC];
  GodunovFluxFromMiddle<D, DIR>(F, L, R, ms, g);
  for (int k = 0; k < NC; ++k) *fl.at(f, k) = F[k];
}

// Godunov::ComputeStableDt through the sequential driver: faces c in
// [lo, hi - S] with S = 2 * width; the base sees cells c + w - 1, c + w.
template <int D, int DIR>
__global__ __launch_bounds__(kBlock) void GodunovViewStableDt(DevView st, DevView faces, int w,
                                                              double dx, Gas<double> g,
                                                              double tol, int max_it,
                                                              unsigned long long* partials) {
  constexpr int NQ = D + 4;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long key = kInfKey;
  if (q < faces.count()) {
    int64_t c[3];
    faces.index(q, c);
    c[DIR] += w - 1;
    double L[NQ], R[NQ];
    for (int k = 0; k < NQ; ++k) L[k] = *st.at(c, k);
    c[DIR] += 1;
    for (int k = 0; k < NQ; ++k) R[k] = *st.at(c, k);
    NewtonParams np;
    np.tol = tol;
    np.max_it = max_it;
    const MiddleState ms = ComputeMiddleState<D, DIR>(L, R, g, np, 0.5 * (L[D + 2] + R[D + 2]));
    key = DtKey(GodunovDtFromMiddle<D, DIR>(L, R, ms.pM, g, dx));
  }
  __shared__ unsigned long long s[kBlock / 64];
  key = WaveMinKey(key);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long k = s[0];
    for (int i = 1; i < kBlock / 64; ++i) k = KMin(k, s[i]);
    partials[blockIdx.x] = k;
  }
}

// Step-start CFL of all directions for the exact solver (one Newton per
// face; faces as in the HLLEM pass: x through LDS, y recomputed in 3D,
// march along the last dimension).  guess[d] (nullable) holds per-face
// initial guesses in direction d's face box (n with n[d] + 1).
struct GodunovCflArgs {
  const double* in[5];
  int64_t n[3];
  double dx[3];
  double gamma, gm1inv, tol;
  int max_it;
  const double* guess[3];
  const short* tiles;  // learned guess computed inline (nullable)
  const float* bias;
  unsigned long long* partials;  // 5 per workgroup and direction
  int* halt;
  double* pm_out;  // star pressures of direction pm_dir's faces (nullable;
  int pm_dir;      // NaN where Newton did not converge), for the first sweep
  // Speed floor (nullable: none; the level driver's ctl->cfl_floor): faces of
  // directions other than pm_dir whose certified speed bound stays below it
  // skip the Newton solve; blocks whose evaluated faces reach it set hit[d].
  const double* floor;
  int* hit;
};

// kGodunovFloorCap: Newton iterations within which Newton from the mean
// pressure converges on every face that GodunovSpeedBound certifies (at most
// 6 were observed over millions of sampled pairs of that class); with a
// smaller iteration cap the pass solves every face, so the cap is reported
// where the reference's would be.

// Certified upper bound of the exact solver's signal speeds max(|s0|, |s1|)
// (GodunovDtFromMiddle) of a face, without Newton, for the faces of the
// class on which Newton from the mean pressure converges quickly: positive
// densities and pressures, pressure ratio at most 4, sound speeds at most
// 1e6 and |uR - uL| at most the smaller sound speed (no vacuum).  Toro's f(p)
// = fL(p) + fR(p) + (uR - uL) increases, so any P with f(P) >= 0 bounds the
// star pressure p*: P = max(pL, pR) when uR >= uL (fK(max p) >= 0 on both
// sides), else P = max(pL, pR) + 2 (uL - uR) / S with S the two shock
// slopes sqrt(AK / (max p + BK)), checked on the shock branches.  Newton
// stops at |f(pM)| <= tol, so pM <= P + tol / Dlo with Dlo the smaller of the
// two branches' slopes on [0, 2P] summed over the sides (valid when that stays
// below 2P).  The signal speeds grow with the star pressure, so |uK| + aK
// w(P_eff / pK), w(r) = sqrt(1 + (g + 1) / (2 g) (r - 1)) for r > 1 and 1
// otherwise, bounds them.  Returns false (no bound) outside the class.
template <int D, int DIR>
__device__ __forceinline__ bool GodunovSpeedBound(const double* L, const double* R,
                                                  const Gas<double>& eq, double tol,
                                                  double* bound) {
  using I = Idx<D>;
  const double rhoL = L[I::kRho], rhoR = R[I::kRho];
  const double pL = L[I::kP], pR = R[I::kP], aL = L[I::kA], aR = R[I::kA];
  const double uL = L[I::kMom + DIR] / rhoL, uR = R[I::kMom + DIR] / rhoR;
  const double pmin = fmin(pL, pR), pmax = fmax(pL, pR);
  const double amin = fmin(aL, aR), amax = fmax(aL, aR);
  const double du = uR - uL;
  if (!(rhoL > 0.0 && rhoR > 0.0 && pmin > 0.0 && pmax <= 4.0 * pmin && amax <= 1e6 &&
        fabs(du) <= amin))
    return false;  // NaN fails too
  const double g = eq.gamma, gp1 = g + 1.0, gm1 = g - 1.0;
  const double AL = 2.0 / (gp1 * rhoL), AR = 2.0 / (gp1 * rhoR);
  const double BL = gm1 / gp1 * pL, BR = gm1 / gp1 * pR;
  double P = pmax;
  if (du < 0.0) {
    const double S = sqrt(AL / (pmax + BL)) + sqrt(AR / (pmax + BR));
    P = pmax + 2.0 * (-du) / S;
    const double fP = (P - pL) * sqrt(AL / (P + BL)) + (P - pR) * sqrt(AR / (P + BR));
    if (!(fP >= -du * (1.0 + 0x1p-30))) return false;
  }
  const double Dlo = fmin(1.0 / (rhoL * aL), 0.5 * sqrt(AL / (2.0 * P + BL))) +
                     fmin(1.0 / (rhoR * aR), 0.5 * sqrt(AR / (2.0 * P + BR)));
  const double Peff = P + tol / Dlo;
  if (!(Peff <= 2.0 * P)) return false;
  const double c = gp1 / (2.0 * g);
  const double wL = Peff > pL ? sqrt(1.0 + c * (Peff / pL - 1.0)) : 1.0;
  const double wR = Peff > pR ? sqrt(1.0 + c * (Peff / pR - 1.0)) : 1.0;
  *bound = fmax(fabs(uL) + aL * wL, fabs(uR) + aR * wR);
  return true;
}

// Newton seed of a face: the inline learned guess (lg, computed by the whole
// wave), else the face's guess array entry, else the mean pressure.
template <int D, int DIR, bool MIXED, bool LEARNED>
__device__ __forceinline__ double CflFace(const double* L, const double* R,
                                          const GodunovCflArgs& a, const Gas<double>& g,
                                          int64_t fidx, double lg, unsigned long long* acc) {
  NewtonParams np;
  np.tol = a.tol;
  np.max_it = a.max_it;
  const double mean = 0.5 * (L[D + 2] + R[D + 2]);
  const bool seeded = LEARNED || a.guess[DIR] != nullptr;
  const double x0 = LEARNED ? lg : a.guess[DIR] ? a.guess[DIR][fidx] : mean;
  int iters, fallback;
  const MiddleState ms =
      SeededMiddleState<D, DIR, MIXED>(L, R, g, np, seeded, x0, mean, &iters, &fallback);
  const unsigned long long it = static_cast<unsigned long long>(iters);
  const unsigned long long fb = static_cast<unsigned long long>(fallback);
  if (a.pm_out && DIR == a.pm_dir) a.pm_out[fidx] = ms.converged ? ms.pM : __builtin_nan("");
  const double dt = GodunovDtFromMiddle<D, DIR>(L, R, ms.pM, g, a.dx[DIR]);
  acc[0] = KMin(acc[0], DtKey(dt));
  acc[1] += it;
  acc[2] = it > acc[2] ? it : acc[2];
  acc[3] += ms.converged ? 0 : 1;
  acc[4] += 1 + (fb << 32);
  return dt;
}

template <int D, bool MIXED, bool LEARNED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void GodunovCflKernel(GodunovCflArgs a) {
  constexpr int NC = D + 2, NQ = D + 4;
  constexpr int M = D - 1;
  if (a.halt && *a.halt) return;
  Gas<double> g;
  g.gamma = a.gamma;
  g.gm1inv = a.gm1inv;
  g.gm1 = a.gamma - 1.0;
  const int64_t nx = a.n[0], ny = a.n[1];
  const int64_t tiles = (nx + kTX - 1) / kTX;
  const int64_t x0 = (blockIdx.x % tiles) * kTX;
  const int64_t y = (D == 3) ? static_cast<int64_t>(blockIdx.x / tiles) : 0;
  const int tx = threadIdx.x;
  const int64_t x = x0 + tx;
  const bool active = x < nx;
  const int64_t xs = active ? x : nx - 1;
  const int64_t nm = (D == 3) ? a.n[2] : ny;
  const int64_t ms = (D == 3) ? nx * ny : nx;
  const int64_t m0 = static_cast<int64_t>(blockIdx.y) * kCflSeg;
  const int64_t m1 = (m0 + kCflSeg < nm) ? m0 + kCflSeg : nm;
  const int64_t rowoff = (D == 3) ? y * nx : 0;
  // Speed floor of the directions other than pm_dir (0: every face solved).
  double F[3] = {0.0, 0.0, 0.0}, dtF[3] = {0.0, 0.0, 0.0};
  if (a.floor) {
    for (int d = 0; d < D; ++d) {
      if (d != a.pm_dir && a.floor[d] > 0.0) {
        F[d] = a.floor[d];
        dtF[d] = a.dx[d] / F[d];
      }
    }
  }
  bool hit[3] = {false, false, false};
  // A face skips its Newton solve when its certified speed bound stays below
  // the floor by a relative 2^-30 (far above the rounding of either side).
  auto skipped = [&](auto dir_tag, const double* l, const double* r) -> bool {
    constexpr int DD = decltype(dir_tag)::value;
    if (!(F[DD] > 0.0)) return false;
    double sb;
    return GodunovSpeedBound<D, DD>(l, r, g, a.tol, &sb) && sb * (1.0 + 0x1p-30) < F[DD];
  };
  __shared__ double s_q[NQ][kTX];
  unsigned long long acc[3][5];
  for (int d = 0; d < 3; ++d) {
    acc[d][0] = kInfKey;
    acc[d][1] = acc[d][2] = acc[d][3] = acc[d][4] = 0;
  }
  auto load = [&](int64_t off, double* q) {
#pragma unroll
    for (int k = 0; k < NC; ++k) q[k] = a.in[k][off];
    CompleteFromCons<D>(q, g);
  };
  // Index in direction d's face box of the face at cell coordinates (cx, cy, cz).
  auto fidx = [&](int d, int64_t cx, int64_t cy, int64_t cz) -> int64_t {
    const int64_t fx = nx + (d == 0), fy = ny + (d == 1);
    return cx + fx * (cy + fy * cz);
  };
  // Waves with an active lane run the inline learned guess with all lanes.
  const bool wave_live = __ballot(active) != 0;
  auto learned = [&](auto dir_tag, const double* l, const double* r) -> double {
    constexpr int DD = decltype(dir_tag)::value;
    if constexpr (LEARNED) {
      if (wave_live) return mlp::LearnedGuess<D, DD>(l, r, a.tiles, a.bias);
    }
    return 0.0;
  };
  using X0 = std::integral_constant<int, 0>;
  using X1 = std::integral_constant<int, 1>;
  using XM = std::integral_constant<int, M>;
  double qprev[NQ];
  load(xs + rowoff + Clamp(m0 - 1, nm) * ms, qprev);
  for (int64_t m = m0; m < m1; ++m) {
    const int64_t off = xs + rowoff + m * ms;
    const int64_t cy = (D == 3) ? y : m, cz = (D == 3) ? m : 0;
    double qc[NQ];
    load(off, qc);
    // March direction.
    {
      const double* l = (m == 0) ? qc : qprev;
      const bool skip = active && skipped(XM{}, l, qc);
      const double lg = __all(skip || !active) ? 0.0 : learned(XM{}, l, qc);
      if (active && !skip) {
        const double dt = CflFace<D, M, MIXED, LEARNED>(
            l, qc, a, g, (M == 1) ? fidx(1, xs, m, 0) : fidx(2, xs, cy, m), lg, acc[M]);
        hit[M] = hit[M] || dt <= dtF[M];
      }
      if (active && m == nm - 1 && !skipped(XM{}, qc, qc)) {
        const double dt = CflFace<D, M, MIXED, LEARNED>(
            qc, qc, a, g, (M == 1) ? fidx(1, xs, m + 1, 0) : fidx(2, xs, cy, m + 1),
            0.5 * (qc[D + 2] + qc[D + 2]), acc[M]);
        hit[M] = hit[M] || dt <= dtF[M];
      }
    }
    if constexpr (D == 3) {
      double qy[NQ];
      if (y == 0) {
        for (int k = 0; k < NQ; ++k) qy[k] = qc[k];
      } else {
        load(off - nx, qy);
      }
      const bool skip = active && skipped(X1{}, qy, qc);
      const double lg = __all(skip || !active) ? 0.0 : learned(X1{}, qy, qc);
      if (active && !skip) {
        const double dt = CflFace<D, 1, MIXED, LEARNED>(qy, qc, a, g, fidx(1, xs, y, cz), lg, acc[1]);
        hit[1] = hit[1] || dt <= dtF[1];
      }
      if (active && y == ny - 1 && !skipped(X1{}, qc, qc)) {
        const double dt = CflFace<D, 1, MIXED, LEARNED>(qc, qc, a, g, fidx(1, xs, y + 1, cz),
                                                        0.5 * (qc[D + 2] + qc[D + 2]), acc[1]);
        hit[1] = hit[1] || dt <= dtF[1];
      }
    }
#pragma unroll
    for (int k = 0; k < NQ; ++k) s_q[k][tx] = qc[k];
    __syncthreads();
    {
      double ql[NQ];
      if (x == 0) {
        for (int k = 0; k < NQ; ++k) ql[k] = qc[k];
      } else if (tx > 0) {
        for (int k = 0; k < NQ; ++k) ql[k] = s_q[k][tx - 1];
      } else {
        load(off - 1, ql);
      }
      const bool skip = active && skipped(X0{}, ql, qc);
      const double lg = __all(skip || !active) ? 0.0 : learned(X0{}, ql, qc);
      if (active && !skip) {
        const double dt = CflFace<D, 0, MIXED, LEARNED>(ql, qc, a, g, fidx(0, xs, cy, cz), lg, acc[0]);
        hit[0] = hit[0] || dt <= dtF[0];
      }
      if (active && x == nx - 1 && !skipped(X0{}, qc, qc)) {
        const double dt = CflFace<D, 0, MIXED, LEARNED>(qc, qc, a, g, fidx(0, xs + 1, cy, cz),
                                                        0.5 * (qc[D + 2] + qc[D + 2]), acc[0]);
        hit[0] = hit[0] || dt <= dtF[0];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; ++k) qprev[k] = qc[k];
  }
  const int64_t nblocks = static_cast<int64_t>(gridDim.x) * gridDim.y;
  const int64_t bid = blockIdx.x + static_cast<int64_t>(blockIdx.y) * gridDim.x;
  for (int d = 0; d < D; ++d)
    BlockReduceGodunov(acc[d][0], acc[d][1], acc[d][2], acc[d][3], acc[d][4],
                       a.partials + 5 * (d * nblocks + bid));
  for (int d = 0; d < D; ++d) {
    if (__syncthreads_or(hit[d] ? 1 : 0) && threadIdx.x == 0 && a.hit) a.hit[d] = 1;
  }
}

// ComputeMiddleState on arrays of complete state pairs (SoA (D+4) x n) with
// optional initial guesses; pM, uM, iterations and convergence per pair.
template <int D>
__global__ __launch_bounds__(kBlock) void MiddleStateKernel(const double* __restrict__ L,
                                                            const double* __restrict__ R,
                                                            int64_t n, int dir, double gamma,
                                                            double tol, int max_it,
                                                            const double* __restrict__ guess,
                                                            double* pM, double* uM, int* iters,
                                                            int* conv) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  double l[D + 4], r[D + 4];
  for (int k = 0; k < D + 4; ++k) {
    l[k] = L[k * n + i];
    r[k] = R[k * n + i];
  }
  Gas<double> g;
  g.gamma = gamma;
  g.gm1inv = 1.0 / (gamma - 1.0);
  g.gm1 = gamma - 1.0;
  NewtonParams np;
  np.tol = tol;
  np.max_it = max_it;
  const double x0 = guess ? guess[i] : 0.5 * (l[D + 2] + r[D + 2]);
  MiddleState ms;
  if (dir == 0) ms = ComputeMiddleState<D, 0>(l, r, g, np, x0);
  else if (dir == 1) ms = ComputeMiddleState<D, 1>(l, r, g, np, x0);
  else ms = ComputeMiddleState<D, (D > 2 ? 2 : 0)>(l, r, g, np, x0);
  if (pM) pM[i] = ms.pM;
  if (uM) uM[i] = ms.uM;
  if (iters) iters[i] = ms.iterations;
  if (conv) conv[i] = ms.converged;
}

} // namespace

int LaunchMiddleState(int dim, hipStream_t s, const double* L, const double* R, int64_t n,
                      int dir, double gamma, double tol, int max_it, const double* guess,
                      double* pM, double* uM, int* iters, int* conv) {
  const unsigned blocks = unsigned((n + kBlock - 1) / kBlock);
  if (dim == 2)
    hipLaunchKernelGGL(MiddleStateKernel<2>, dim3(blocks), dim3(kBlock), 0, s, L, R, n, dir, gamma,
                       tol, max_it, guess, pM, uM, iters, conv);
  else
    hipLaunchKernelGGL(MiddleStateKernel<3>, dim3(blocks), dim3(kBlock), 0, s, L, R, n, dir, gamma,
                       tol, max_it, guess, pM, uM, iters, conv);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int64_t GodunovSweepBlocks(int dim, int dir, const int64_t n[3], int seg) {
  const int64_t nz = dim == 3 ? n[2] : 1;
  if (dir == 0) {
    const int64_t slots = (n[0] + 2) * n[1] * nz;
    const int64_t waves = (slots + kXOwned - 1) / kXOwned;
    return (waves + kBlock / 64 - 1) / (kBlock / 64);
  }
  const int64_t npencil = (dir == 1) ? n[0] * nz : n[0] * n[1];
  return ((npencil + kBlock - 1) / kBlock) * ((n[dir] + seg - 1) / seg);
}

int64_t GodunovCflBlocks(int dim, const int64_t n[3]) {
  const int64_t tiles = (n[0] + kTX - 1) / kTX;
  const int64_t rows = dim == 3 ? n[1] : 1;
  const int64_t nm = dim == 3 ? n[2] : n[1];
  return tiles * rows * ((nm + kCflSeg - 1) / kCflSeg);
}

int LaunchGodunovSweep(const fvs_method_desc& desc, hipStream_t s, int dir,
                       const double* const* in, double* const* out, const int64_t n[3],
                       double dx, const double* dt_ptr, double dt_val, const double* guess,
                       unsigned long long* partials, int* halt, int update,
                       const fvs_mlp* net, int reuse) {
  GodunovArgs a;
  std::memset(&a, 0, sizeof a);
  const int D = desc.dim;
  for (int c = 0; c < D + 2; ++c) {
    a.in[c] = in[c];
    a.out[c] = out ? out[c] : nullptr;
  }
  for (int d = 0; d < 3; ++d) a.n[d] = (d < D) ? n[d] : 1;
  a.dx = dx;
  a.dt_ptr = dt_ptr;
  a.dt_val = dt_val;
  a.gamma = desc.gamma;
  a.gm1inv = 1.0 / (desc.gamma - 1.0);
  a.tol = desc.newton_tolerance;
  a.max_it = desc.max_newton_iterations;
  a.guess = guess;
  if (net) MlpOperands(net, &a.tiles, &a.bias);
  a.partials = partials;
  a.halt = halt;
  a.seg = PencilSeg(D, dir, a.n);
  a.update = update;
  a.reuse = reuse && guess != nullptr;
  const bool mixed = desc.precision == FVS_FP32_FLUX;
  const bool learned = a.tiles != nullptr;
  const int64_t blocks = GodunovSweepBlocks(D, dir, a.n, a.seg);
  if (blocks <= 0) return FVS_BAD_ARGUMENT;
  if (dir == 0) {
    if ((a.n[0] + 2) * a.n[1] * a.n[2] >= (int64_t(1) << 32) || blocks > 0x7fffffff) {
      SetError("Godunov X sweep: patch too large for 32-bit slot indices");
      return FVS_BAD_ARGUMENT;
    }
    using K = void (*)(GodunovArgs);
    const K ks[2][2][2] = {{{GodunovSweepX<2, false, false>, GodunovSweepX<2, false, true>},
                            {GodunovSweepX<2, true, false>, GodunovSweepX<2, true, true>}},
                           {{GodunovSweepX<3, false, false>, GodunovSweepX<3, false, true>},
                            {GodunovSweepX<3, true, false>, GodunovSweepX<3, true, true>}}};
    hipLaunchKernelGGL(ks[D == 3][mixed][learned], dim3(unsigned(blocks)), dim3(kBlock), 0, s, a);
  } else {
    const int64_t npencil = (dir == 1) ? a.n[0] * a.n[2] : a.n[0] * a.n[1];
    const int64_t bx = (npencil + kBlock - 1) / kBlock;
    const int64_t nseg = (a.n[dir] + a.seg - 1) / a.seg;
    if (bx > 0x7fffffff || nseg > 65535) return FVS_BAD_ARGUMENT;
    const dim3 grid{static_cast<unsigned>(bx), static_cast<unsigned>(nseg), 1u};
    using K = void (*)(GodunovArgs);
    const K ks[3][2][2] = {
        {{GodunovSweepPencil<2, 1, false, false>, GodunovSweepPencil<2, 1, false, true>},
         {GodunovSweepPencil<2, 1, true, false>, GodunovSweepPencil<2, 1, true, true>}},
        {{GodunovSweepPencil<3, 1, false, false>, GodunovSweepPencil<3, 1, false, true>},
         {GodunovSweepPencil<3, 1, true, false>, GodunovSweepPencil<3, 1, true, true>}},
        {{GodunovSweepPencil<3, 2, false, false>, GodunovSweepPencil<3, 2, false, true>},
         {GodunovSweepPencil<3, 2, true, false>, GodunovSweepPencil<3, 2, true, true>}}};
    const int which = D == 2 ? 0 : (dir == 1 ? 1 : 2);
    hipLaunchKernelGGL(ks[which][mixed][learned], grid, dim3(kBlock), 0, s, a);
  }
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchGodunovFinalize(hipStream_t s, const unsigned long long* partials, int64_t nblocks,
                          double* out, unsigned long long* stats, Ctl* ctl) {
  hipLaunchKernelGGL(GodunovFinalize, dim3(1), dim3(kRedBlock), 0, s, partials, nblocks, out,
                     stats, ctl);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchGodunovCfl(const fvs_method_desc& desc, hipStream_t s, const double* const* cons,
                     const int64_t n[3], const double dx[3], const double* const* guess,
                     unsigned long long* partials, int* halt, const fvs_mlp* net,
                     double* pm_out, int pm_dir, Ctl* floor_ctl) {
  GodunovCflArgs a;
  std::memset(&a, 0, sizeof a);
  const int D = desc.dim;
  for (int c = 0; c < D + 2; ++c) a.in[c] = cons[c];
  for (int d = 0; d < 3; ++d) {
    a.n[d] = (d < D) ? n[d] : 1;
    a.dx[d] = (d < D) ? dx[d] : 1.0;
    a.guess[d] = (guess && d < D) ? guess[d] : nullptr;
  }
  a.gamma = desc.gamma;
  a.gm1inv = 1.0 / (desc.gamma - 1.0);
  a.tol = desc.newton_tolerance;
  a.max_it = desc.max_newton_iterations;
  if (net) MlpOperands(net, &a.tiles, &a.bias);
  a.partials = partials;
  a.halt = halt;
  a.pm_out = pm_out;
  a.pm_dir = pm_dir;
  if (floor_ctl && desc.max_newton_iterations >= kGodunovFloorCap) {
    a.floor = floor_ctl->cfl_floor;
    a.hit = floor_ctl->floor_hit;
  }
  const bool mixed = desc.precision == FVS_FP32_FLUX;
  const int64_t tiles = (a.n[0] + kTX - 1) / kTX;
  const int64_t gx = tiles * (D == 3 ? a.n[1] : 1);
  const int64_t gy = ((D == 3 ? a.n[2] : a.n[1]) + kCflSeg - 1) / kCflSeg;
  if (gx > 0x7fffffff || gy > 65535) return FVS_BAD_ARGUMENT;
  using K = void (*)(GodunovCflArgs);
  const K ks[2][2][2] = {{{GodunovCflKernel<2, false, false>, GodunovCflKernel<2, false, true>},
                          {GodunovCflKernel<2, true, false>, GodunovCflKernel<2, true, true>}},
                         {{GodunovCflKernel<3, false, false>, GodunovCflKernel<3, false, true>},
                          {GodunovCflKernel<3, true, false>, GodunovCflKernel<3, true, true>}}};
  const bool learned = a.tiles != nullptr;
  hipLaunchKernelGGL(ks[D == 3][mixed][learned], dim3(unsigned(gx), unsigned(gy)), dim3(kBlock),
                     0, s, a);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

static int MakeGView(const fvs_view* v, DevView* out) { return MakeDevView(v, 1, out); }

int LaunchGodunovViewFlux(const fvs_method_desc& desc, hipStream_t s, const fvs_view* fluxes,
                          const fvs_view* states, double dt, double dx, int dir,
                          unsigned long long* nfail) {
  DevView fl, st;
  if (MakeGView(fluxes, &fl) || MakeGView(states, &st)) return FVS_BAD_ARGUMENT;
  Gas<double> g;
  g.gamma = desc.gamma;
  g.gm1inv = 1.0 / (desc.gamma - 1.0);
  g.gm1 = desc.gamma - 1.0;
  const int muscl = desc.reconstruction == FVS_RECON_MUSCL_HANCOCK;
  const unsigned nb = unsigned((fl.count() + kBlock - 1) / kBlock);
#define FVS_GVF(DD, DR)                                                                      \
  hipLaunchKernelGGL((GodunovViewFlux<DD, DR>), dim3(nb), dim3(kBlock), 0, s, fl, st, dt, dx, g, \
                     desc.limiter, muscl, desc.newton_tolerance, desc.max_newton_iterations,  \
                     nfail)
  if (desc.dim == 2) {
    if (dir == 0) FVS_GVF(2, 0); else FVS_GVF(2, 1);
  } else {
    if (dir == 0) FVS_GVF(3, 0); else if (dir == 1) FVS_GVF(3, 1); else FVS_GVF(3, 2);
  }
#undef FVS_GVF
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int64_t GodunovViewStableDtBlocks(const fvs_method_desc& desc, const fvs_view* states, int dir) {
  const int S = desc.reconstruction == FVS_RECON_MUSCL_HANCOCK ? 4 : 2;
  int64_t cnt = 1;
  for (int d = 0; d < 3; ++d) {
    int64_t e = states->hi[d] - states->lo[d] + 1;
    if (d == dir) e -= S;
    if (e <= 0) return 0;
    cnt *= e;
  }
  return (cnt + kBlock - 1) / kBlock;
}

int LaunchGodunovViewStableDt(const fvs_method_desc& desc, hipStream_t s, const fvs_view* states,
                              double dx, int dir, unsigned long long* partials) {
  DevView st;
  if (MakeGView(states, &st)) return FVS_BAD_ARGUMENT;
  const int w = desc.reconstruction == FVS_RECON_MUSCL_HANCOCK ? 2 : 1;
  DevView faces = st;
  faces.ext[dir] -= 2 * w;  // sequential driver: Shrink(box, dir, {0, StencilSize})
  if (faces.ext[dir] <= 0) return FVS_BAD_ARGUMENT;
  Gas<double> g;
  g.gamma = desc.gamma;
  g.gm1inv = 1.0 / (desc.gamma - 1.0);
  g.gm1 = desc.gamma - 1.0;
  const unsigned nb = unsigned((faces.count() + kBlock - 1) / kBlock);
#define FVS_GVD(DD, DR)                                                                       \
  hipLaunchKernelGGL((GodunovViewStableDt<DD, DR>), dim3(nb), dim3(kBlock), 0, s, st, faces, w, \
                     dx, g, desc.newton_tolerance, desc.max_newton_iterations, partials)
  if (desc.dim == 2) {
    if (dir == 0) FVS_GVD(2, 0); else FVS_GVD(2, 1);
  } else {
    if (dir == 0) FVS_GVD(3, 0); else if (dir == 1) FVS_GVD(3, 1); else FVS_GVD(3, 2);
  }
#undef FVS_GVD
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

} // namespace fvs
