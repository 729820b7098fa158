// Fused split sweeps for the perfect-gas Euler equations on gfx950.
//
// One launch performs, for one direction, what one split step of
// DimensionalSplitLevelIntegrator::AdvanceLevelNonRecursively does in three
// passes (include/fub/solver/DimensionalSplitLevelIntegrator.hpp:240-256):
//   the CFL check of the input state   (FluxMethod::ComputeStableDt),
//   the numeric fluxes                 (FluxMethod::ComputeNumericFluxes),
//   the conservative update            (HyperbolicPatchIntegrator),
// and writes only the conservative variables; pressure and speed of sound
// are recomputed from them where needed, which equals CompleteFromCons.
// Boundaries are transmissive (TransmissiveBoundary.cpp:31-93): a ghost cell
// holds its row's boundary value, i.e. indices are clamped to the patch.
//
// X sweeps (unit stride): one cell per lane over the flattened sequence of
// row slots (cells plus the two ghost cells of each row), neighbour states
// moving between lanes with wave shuffles; waves are independent.  Y/Z: one
// thread per (x, transverse) pencil, marching along the sweep direction with
// a register window of three cells and a one-cell load prefetch, so every
// load is coalesced along x and no LDS or barrier is needed.
//
// Reductions (CFL minima) write one partial per workgroup and a one-block
// finalize kernel reduces them: no contended device-scope atomics, and the
// result does not depend on workgroup scheduling.
//
// This file is compiled twice.  The plain build holds the IEEE FP64 kernels
// (no fused multiply-adds: bit equality with the reference) and every host
// entry point; fvs_sweep_fast.hip defines FVS_SWEEP_FAST_TU and compiles the
// kernels of the tolerance-checked modes (refined-reciprocal FP64, FP32 flux)
// with multiply-add contraction, reached through LaunchSweepFast and
// LaunchPatchCflFast.
#include "fvs_internal.hpp"
#include "fvs_lanes.hpp"
#include "fvs_physics.hpp"
#include "fvs_xslots.hpp"

namespace fvs {

namespace {

constexpr int kTX = 256;                // CFL-pass x tile (cells per workgroup)
constexpr int kBlock = 256;             // workgroup size of every sweep kernel
constexpr unsigned long long kInfKey = 0x7ff0000000000000ull;

__device__ __forceinline__ int64_t Clamp(int64_t i, int64_t n) {
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

__device__ __forceinline__ unsigned long long KMin(unsigned long long a,
                                                   unsigned long long b) {
  return b < a ? b : a;
}

__device__ __forceinline__ unsigned long long WaveMinKey(unsigned long long k) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) k = KMin(k, __shfl_xor(k, off, 64));
  return k;
}


// Loads and stores as a wave-uniform base pointer plus a 32-bit byte offset:
// the base lives in scalar registers, the offset in one vector register shared
// by every component, and the global accesses take the scalar-base form with
// no 64-bit address arithmetic per component.
__device__ __forceinline__ double LdOff(const double* base, uint32_t byte_off) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + byte_off);
}
__device__ __forceinline__ void StOff(double* base, uint32_t byte_off, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + byte_off) = v;
}
// Streaming store (nontemporal hint) for the pencil sweeps, whose waves write
// whole 512-byte rows: the lines go to memory without being kept in the
// caches (about 4% faster pencils, and the next pass's loads find the caches
// free).  The X sweep keeps StOff: its 60-cell chunks end inside cache lines,
// and streaming partial lines made it 10% slower.
__device__ __forceinline__ void StOffStream(double* base, uint32_t byte_off, double v) {
  __builtin_nontemporal_store(v, reinterpret_cast<double*>(reinterpret_cast<char*>(base) + byte_off));
}

// Largest of two signal speeds; a NaN speed never wins (as a NaN face dt
// never wins the minimum).
__device__ __forceinline__ double SMax(double a, double b) { return __builtin_fmax(a, b); }

__device__ __forceinline__ double WaveMaxSpeed(double s) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s = SMax(s, __shfl_xor(s, off, 64));
  return s;
}

// Block-wide largest signal speed of NK directions (speeds are >= 0, the
// neutral value is 0); thread 0 writes the key of the block's smallest face
// dt, dx[d] / speed, to out[d * stride] (+inf when every speed is 0).  A
// face's stable dt is dx / speed, and a correctly rounded quotient falls
// monotonically with its denominator, so the smallest face dt is the
// quotient by the largest speed, bit for bit: one division per block and
// direction instead of one per face.  Every thread of the block must call it.
// With a speed floor (floor[d] > 0), a block whose largest speed reaches it
// sets hit[d].
template <int NK, int NWAVES, class A>
__device__ __forceinline__ void BlockSpeedToPartials(const double* speed, const double* dx,
                                                     unsigned long long* out, int64_t stride,
                                                     const double* floor = nullptr,
                                                     int* hit = nullptr) {
  __shared__ double s_speed[NK][NWAVES];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < NK; ++d) {
    const double v = WaveMaxSpeed(speed[d]);
    if (lane == 0) s_speed[d][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < NK; ++d) {
      double v = s_speed[d][0];
      for (int w = 1; w < NWAVES; ++w) v = SMax(v, s_speed[d][w]);
      out[d * stride] = v > 0.0 ? DtKey(A::quot(dx[d], v)) : kInfKey;
      if (hit && floor[d] > 0.0 && v >= floor[d]) hit[d] = 1;
    }
  }
}

// Per-cell ingredients of Hllem::ComputeStableDt
// (src/equations/perfect_gas/HllemMethod.hpp:1152-1216): sqrt(rho), the
// velocity, the enthalpy (E + p) / rho and the speed of sound.  A face's dt
// combines two of them with exactly the reference's operations.
template <int D> struct DtCell {
  double sq;
  double u[D];
  double h;
  double a;
};

template <int D, class A>
__device__ __forceinline__ DtCell<D> MakeDtCell(const double* u, const Gas<double>& g) {
  double q[D + 4];
#pragma unroll
  for (int c = 0; c < D + 2; ++c) q[c] = u[c];
  const typename A::Inv ir = CompleteFromCons<D, double, A>(q, g);
  DtCell<D> r;
  r.sq = A::sqrt(q[0]);
#pragma unroll
  for (int i = 0; i < D; ++i) r.u[i] = A::div(q[1 + i], ir);
  r.h = A::div(q[D + 1] + q[D + 2], ir);
  r.a = q[D + 3];
  return r;
}

// The face's largest signal speed max(|sL|, |sR|); its stable dt is
// dx / speed (reduced per block by BlockSpeedToPartials).
template <int D, int DIR, class A>
__device__ __forceinline__ double FaceSpeed(const DtCell<D>& L, const DtCell<D>& R,
                                            double gm1) {
  const double sqRhoSum = L.sq + R.sq;
  const typename A::Inv isum = A::inv(sqRhoSum);
  const double wL = A::div(L.sq, isum);
  const double wR = A::div(R.sq, isum);
  double roeU[D];
#pragma unroll
  for (int i = 0; i < D; ++i) roeU[i] = wL * L.u[i] + wR * R.u[i];
  const double roeH = wL * L.h + wR * R.h;
  double sq = roeU[0] * roeU[0];
#pragma unroll
  for (int i = 1; i < D; ++i) sq += roeU[i] * roeU[i];
  const double roeA2 = gm1 * (roeH - 0.5 * sq);
  const double roeA = A::sqrt(roeA2);
  const double maxA = A::max(L.a, R.a);
  const double sL = A::min(L.u[DIR] - maxA, roeU[DIR] - roeA);
  const double sR = A::max(roeU[DIR] + roeA, R.u[DIR] + maxA);
  return A::max(rabs(sL), rabs(sR));
}

template <int D>
__device__ __forceinline__ BoundCell<D> ShflUp(const BoundCell<D>& v) {
  BoundCell<D> r;
#pragma unroll
  for (int i = 0; i < D; ++i) r.u[i] = ShiftUp(v.u[i]);
  r.a2 = ShiftUp(v.a2);
  return r;
}


template <int D>
__device__ __forceinline__ DtCell<D> ShflUp(const DtCell<D>& v) {
  DtCell<D> r;
  r.sq = ShiftUp(v.sq);
#pragma unroll
  for (int i = 0; i < D; ++i) r.u[i] = ShiftUp(v.u[i]);
  r.h = ShiftUp(v.h);
  r.a = ShiftUp(v.a);
  return r;
}

template <int D, int DIR, class T, class A>
__device__ __forceinline__ void ReconAny(const T* um, const T* u0, const T* up, T lh,
                                         const Gas<T>& g, int limiter, int muscl,
                                         T* recL, T* recR, typename A::Inv& irL,
                                         typename A::Inv& irR) {
  if (!muscl) {
    irL = GhostRec<D, T, A>(u0, g, recL);
    irR = irL;
#pragma unroll
    for (int k = 0; k < D + 4; ++k) recR[k] = recL[k];
    return;
  }
  switch (limiter) {
  case kMinMod:
    ReconstructCell<D, DIR, kMinMod, T, A>(um, u0, up, lh, g, recL, recR, irL, irR);
    break;
  case kVanLeer:
    ReconstructCell<D, DIR, kVanLeer, T, A>(um, u0, up, lh, g, recL, recR, irL, irR);
    break;
  case kNoLimiter:
    ReconstructCell<D, DIR, kNoLimiter, T, A>(um, u0, up, lh, g, recL, recR, irL, irR);
    break;
  default:
    ReconstructCell<D, DIR, kNoGradient, T, A>(um, u0, up, lh, g, recL, recR, irL, irR);
    break;
  }
}

template <int D, int DIR, bool LARR, class T, class A>
__device__ __forceinline__ void FaceFlux(T* f, const T* L, const typename A::Inv& irL,
                                         const T* R, const typename A::Inv& irR,
                                         const Gas<T>& g, int quirk) {
  if constexpr (D == 3) {
    if (quirk) {
      HllemFlux<D, DIR, LARR, true, T, A>(f, L, irL, R, irR, g);
    } else {
      HllemFlux<D, DIR, LARR, false, T, A>(f, L, irL, R, irR, g);
    }
  } else {
    HllemFlux<D, DIR, LARR, false, T, A>(f, L, irL, R, irR, g);
  }
}


template <class T> __device__ __forceinline__ Gas<T> MakeGas(const SweepArgs& a) {
  Gas<T> g;
  g.gamma = static_cast<T>(a.gamma);
  g.gm1inv = static_cast<T>(a.gm1inv);
  g.gm1 = static_cast<T>(a.gamma - 1.0);
  return g;
}

// ---------------------------------------------------------------------------
// X sweep.  Every row of nx cells contributes nx + 2 "slots": the ghost cell
// -1, the cells 0 .. nx-1 and the ghost cell nx.  The slots of all rows form
// one sequence cut into chunks of 64 consecutive slots: lanes 2..61 own
// their slots, lanes 0, 1 and 62, 63 overlap the neighbouring chunks.  A lane
// reconstructs its cell from its neighbours one DPP shift away, computes the
// flux of the face on the cell's left (lanes 2..62) and updates its cell with
// its right neighbour's flux; the outer lanes' reconstructions are never
// used, so no lane reads memory beyond its own slot.  Waves are persistent
// (a few per SIMD) and walk the chunks with a stride, each loading the next
// chunk while it computes the current one, so every wave keeps a load in
// flight behind its FP64 work; no LDS, no barrier until the final CFL
// reduction.
template <int D, bool LARR, class T, bool FAST>
__global__ __launch_bounds__(kBlock) void SweepXKernel(SweepArgs a) {
  using TA = typename Policies<T, FAST>::TA;
  using DA = typename Policies<T, FAST>::DA;
  constexpr int NC = D + 2;
  constexpr int NQ = D + 4;
  if (a.halt && *a.halt) return;
  const int64_t nx = a.n[0];
  const uint32_t S = static_cast<uint32_t>(nx + 2);
  const uint32_t Q = static_cast<uint32_t>(a.n[1] * a.n[2]) * S;
  const int64_t nchunks = (static_cast<int64_t>(Q) + kXOwned - 1) / kXOwned;
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (kBlock / 64);
  const bool owned = lane >= 2 && lane < 2 + kXOwned;

  const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
  const T lh = static_cast<T>(DivRn(0.5 * dt, a.dx));
  const double dt_dx = DivRn(dt, a.dx);
  const Gas<T> g = MakeGas<T>(a);
  const Gas<double> gd = MakeGas<double>(a);
  const double s_crit = BoundSpeedLimit(a.dx, dt);
  const double kbound = 0.125 * gd.gm1;

  // Slot of this lane in chunk c (fvs_xslots.hpp): the row base is lane 0's
  // row, read across the wave.
  XSlotWalker walk(nx, Q, nwaves, lane);
  using Slot = XSlot;
  auto locate = [&](int64_t c) {
    bool valid;
    uint32_t row, sl;
    walk.RowSlot(c, &valid, &row, &sl);
    return walk.Make(valid, row, sl, __builtin_amdgcn_readfirstlane(row));
  };
  auto load = [&](const Slot& s, double* u) {
    const uint32_t off = 8u * (s.rbase + static_cast<uint32_t>(Clamp(s.cell, nx)));
#pragma unroll
    for (int k = 0; k < NC; ++k) u[k] = LdOff(a.in[k] + s.wbase, off);
  };

  double smax = 0.0;
  // The chunk index is wave-uniform (a scalar register), so the chunk loop is
  // a uniform branch and its DPP shifts always run with the whole wave.
  int64_t c = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) +
              __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  walk.Place(c < nchunks ? c : 0);
  Slot cur = locate(c < nchunks ? c : 0);
  double u[NC];
  if (c < nchunks) load(cur, u);
  for (; c < nchunks; c += nwaves) {
    // The next chunk's load is in flight while this chunk is computed.
    const int64_t cn = c + nwaves;
    walk.Next(c, cn, nchunks);
    const Slot nxt = locate(cn < nchunks ? cn : c);
    double un[NC];
    if (cn < nchunks) load(nxt, un);

    const bool real = cur.valid && cur.cell >= 0 && cur.cell < nx;
    double um[NC], up[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      um[k] = ShiftUp(u[k]);
      up[k] = ShiftDown(u[k]);
    }
    // A ghost slot reconstructs with a zero slope, so its half-step states
    // equal the ghost's complete state (up to the sign of zero components)
    // and every lane runs the same reconstruction.  The ghost's clamped value
    // equals its inner neighbour's (slot 0 and slot 1, slot nx and nx + 1
    // hold the same cell), so one of its one-sided differences is exactly
    // zero, and MinMod and VanLeer return a zero slope whatever the other
    // neighbour holds (even NaN or inf).  The central slope of NoLimiter
    // does not vanish: there the outer neighbour is replaced by the cell.
    if (a.muscl && a.limiter == kNoLimiter) {
      if (cur.cell < 0) {
#pragma unroll
        for (int k = 0; k < NC; ++k) um[k] = u[k];
      }
      if (cur.cell >= nx) {
#pragma unroll
        for (int k = 0; k < NC; ++k) up[k] = u[k];
      }
    }

    T recL[NQ], recR[NQ];
    typename TA::Inv irL, irR;
    {
      T tm[NC], t0[NC], tp[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        tm[k] = static_cast<T>(um[k]);
        t0[k] = static_cast<T>(u[k]);
        tp[k] = static_cast<T>(up[k]);
      }
      ReconAny<D, 0, T, TA>(tm, t0, tp, lh, g, a.limiter, a.muscl, recL, recR, irL, irR);
    }
    T recRprev[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) recRprev[k] = ShiftUp(recR[k]);
    typename TA::Inv irRprev;
    if constexpr (TA::kFast) {
      irRprev.y = ShiftUp(irR.y);
    } else {
      irRprev = TA::inv(recRprev[0]);
    }

    // Face sl - 1 (left face of the slot's cell) for slots 1 .. nx + 1.
    const bool face_lane = cur.valid && lane >= 2 && lane <= 2 + kXOwned && cur.sl >= 1;
    T F[NC];
    if (face_lane) {
      FaceFlux<D, 0, LARR, T, TA>(F, recRprev, irRprev, recL, irL, g, a.quirk3d);
    } else {
#pragma unroll
      for (int k = 0; k < NC; ++k) F[k] = T(0);
    }
    T Fnext[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) Fnext[k] = ShiftDown(F[k]);

    if (owned && real) {
      const uint32_t off = 8u * (cur.rbase + static_cast<uint32_t>(cur.cell));
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        StOff(a.out[k] + cur.wbase, off,
              u[k] + dt_dx * (static_cast<double>(F[k]) - static_cast<double>(Fnext[k])));
      }
    }
    if (a.check) {
      // chunk loop is wave-uniform, so the vote and both paths are too
      bool exact = true;
      if (a.bound_check) {
        const BoundCell<D> bc = MakeBoundCell<D>(u, gd);
        const BoundCell<D> bp = ShflUp<D>(bc);
        const bool pass = !(owned && face_lane) || BoundPasses<D, 0>(bp, bc, s_crit, kbound);
        exact = !__all(pass);
      }
      if (exact) {
        const DtCell<D> dc = MakeDtCell<D, DA>(u, gd);
        const DtCell<D> dp = ShflUp<D>(dc);
        if (owned && face_lane) smax = SMax(smax, FaceSpeed<D, 0, DA>(dp, dc, gd.gm1));
      }
    }
    cur = nxt;
#pragma unroll
    for (int k = 0; k < NC; ++k) u[k] = un[k];
  }
  if (a.check) BlockSpeedToPartials<1, kBlock / 64, DA>(&smax, &a.dx, a.partials + blockIdx.x, 0);
}

// ---------------------------------------------------------------------------
// Y / Z sweep: pencil march.
template <int D, int DIR, bool LARR, class T, bool FAST>
__global__ __launch_bounds__(kBlock) void SweepPencilKernel(SweepArgs a) {
  using TA = typename Policies<T, FAST>::TA;
  using DA = typename Policies<T, FAST>::DA;
  constexpr int NC = D + 2;
  constexpr int NQ = D + 4;
  if (a.halt && *a.halt) return;
  const int64_t nx = a.n[0], ny = a.n[1];
  const int64_t nd = a.n[DIR];
  const int64_t npencil = (DIR == 1) ? nx * a.n[2] : nx * ny;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const bool active = p < npencil;
  const int64_t stride = (DIR == 1) ? nx : nx * ny;
  const int64_t base = (DIR == 1) ? (p % nx) + (p / nx) * nx * ny : p;
  // Uniform part of the addresses: the plane of the block's first pencil (Y
  // sweeps; a block spans at most 256 / nx + 2 planes) and the marching index.
  const int64_t ubase =
      (DIR == 1) ? ((static_cast<int64_t>(blockIdx.x) * kBlock) / nx) * nx * ny : 0;
  const uint32_t toff = 8u * static_cast<uint32_t>(base - ubase);
  const int64_t j0 = static_cast<int64_t>(blockIdx.y) * a.seg;
  const int64_t j1 = (j0 + a.seg < nd) ? j0 + a.seg : nd;

  const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
  const T lh = static_cast<T>(DivRn(0.5 * dt, a.dx));
  const double dt_dx = DivRn(dt, a.dx);
  const Gas<T> g = MakeGas<T>(a);
  const Gas<double> gd = MakeGas<double>(a);
  const double s_crit = BoundSpeedLimit(a.dx, dt);
  const double kbound = 0.125 * gd.gm1;

  double smax = 0.0;
  if (active) {
    double Um[NC], U0[NC], Up[NC], Unext[NC];
    auto load = [&](double* u, int64_t j) {
      const int64_t uoff = ubase + Clamp(j, nd) * stride;
#pragma unroll
      for (int c = 0; c < NC; ++c) u[c] = LdOff(a.in[c] + uoff, toff);
    };
    load(Um, j0 - 1);
    load(U0, j0);
    load(Unext, j0 + 1);
    T recRprev[NQ];
    typename TA::Inv irRprev;
    {
      T t0[NC];
      if (j0 == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) t0[c] = static_cast<T>(U0[c]);
        irRprev = GhostRec<D, T, TA>(t0, g, recRprev);
      } else {
        double Umm[NC];
        load(Umm, j0 - 2);
        T tm[NC], tp[NC], dummy[NQ];
        typename TA::Inv idummy;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          tm[c] = static_cast<T>(Umm[c]);
          t0[c] = static_cast<T>(Um[c]);
          tp[c] = static_cast<T>(U0[c]);
        }
        ReconAny<D, DIR, T, TA>(tm, t0, tp, lh, g, a.limiter, a.muscl, dummy, recRprev, idummy,
                                irRprev);
      }
    }
    BoundCell<D> bprev;
    if (a.check && a.bound_check) bprev = MakeBoundCell<D>(Um, gd);
    T Fprev[NC];
    for (int64_t j = j0; j <= j1; ++j) {
#pragma unroll
      for (int c = 0; c < NC; ++c) Up[c] = Unext[c];
      if (j + 1 < nd) load(Unext, j + 2);
      T recL[NQ], recR[NQ];
      typename TA::Inv irL, irR;
      if (j < nd) {
        T tm[NC], t0[NC], tp[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          tm[c] = static_cast<T>(Um[c]);
          t0[c] = static_cast<T>(U0[c]);
          tp[c] = static_cast<T>(Up[c]);
        }
        ReconAny<D, DIR, T, TA>(tm, t0, tp, lh, g, a.limiter, a.muscl, recL, recR, irL, irR);
      } else {
        T t0[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) t0[c] = static_cast<T>(U0[c]);
        irL = GhostRec<D, T, TA>(t0, g, recL);
        irR = irL;
#pragma unroll
        for (int k = 0; k < NQ; ++k) recR[k] = recL[k];
      }
      T F[NC];
      FaceFlux<D, DIR, LARR, T, TA>(F, recRprev, irRprev, recL, irL, g, a.quirk3d);
      if (j > j0) {
        const int64_t uoff = ubase + (j - 1) * stride;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          StOffStream(a.out[c] + uoff, toff,
                      Um[c] + dt_dx * (static_cast<double>(Fprev[c]) - static_cast<double>(F[c])));
        }
      }
      if (a.check && (j < j1 || j1 == nd)) {
        // The face index is uniform over the workgroup, so the vote covers
        // one face per active pencil.
        bool exact = true;
        if (a.bound_check) {
          const BoundCell<D> b0 = MakeBoundCell<D>(U0, gd);
          exact = !__all(BoundPasses<D, DIR>(bprev, b0, s_crit, kbound));
          bprev = b0;
        }
        if (exact) {
          const DtCell<D> dm = MakeDtCell<D, DA>(Um, gd);
          const DtCell<D> d0 = MakeDtCell<D, DA>(U0, gd);
          smax = SMax(smax, FaceSpeed<D, DIR, DA>(dm, d0, gd.gm1));
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        Um[c] = U0[c];
        U0[c] = Up[c];
        Fprev[c] = F[c];
      }
#pragma unroll
      for (int k = 0; k < NQ; ++k) recRprev[k] = recR[k];
      irRprev = irR;
    }
  }
  if (a.check) {
    BlockSpeedToPartials<1, kBlock / 64, DA>(
        &smax, &a.dx, a.partials + blockIdx.x + static_cast<int64_t>(blockIdx.y) * gridDim.x, 0);
  }
}

// ---------------------------------------------------------------------------
// 2D step-start CFL: stable dt of both directions in one pass.  A workgroup
// is a 256-cell x tile marching along y over a segment; x neighbours come
// through LDS and the y neighbour is kept in registers.  Boundary faces pair
// a cell with its transmissive ghost copy.
constexpr int kCflSeg = 32;

template <bool FAST>
__global__ __launch_bounds__(kBlock) void PatchCfl2Kernel(
    const double* __restrict__ u0, const double* __restrict__ u1,
    const double* __restrict__ u2, const double* __restrict__ u3, int64_t nx, int64_t ny,
    double dx0, double dx1, double gamma, double gm1inv, unsigned long long* partials,
    int* halt) {
  using DA = typename Policies<double, FAST>::DA;
  constexpr int D = 2;
  constexpr int NC = D + 2;
  constexpr int M = 1;  // march dimension
  if (halt && *halt) return;
  const double* in[4] = {u0, u1, u2, u3};
  Gas<double> g;
  g.gamma = gamma;
  g.gm1inv = gm1inv;
  g.gm1 = gamma - 1.0;
  const int64_t x0 = static_cast<int64_t>(blockIdx.x) * kTX;
  const int tx = threadIdx.x;
  const int64_t x = x0 + tx;
  const bool active = x < nx;
  const int64_t xs = active ? x : nx - 1;
  const int64_t nm = ny;
  const int64_t ms = nx;
  // rows per segment follow the launch (fewer for small patches)
  const int64_t seg = (nm + gridDim.y - 1) / gridDim.y;
  const int64_t m0 = static_cast<int64_t>(blockIdx.y) * seg;
  const int64_t m1 = (m0 + seg < nm) ? m0 + seg : nm;
  const int64_t rowoff = 0;
  const double dxs[2] = {dx0, dx1};

  __shared__ DtCell<D> s_cells[kTX];
  double smax[2] = {0.0, 0.0};

  auto load_cell = [&](int64_t off) {
    double u[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) u[c] = in[c][off];
    return MakeDtCell<D, DA>(u, g);
  };
  auto load = [&](double* u, int64_t m) {
    const int64_t off = xs + rowoff + Clamp(m, nm) * ms;
#pragma unroll
    for (int c = 0; c < NC; ++c) u[c] = in[c][off];
  };

  // Row m0 - 1 is clamped to row 0 at the lower edge, so there dprev is the
  // row's own transmissive ghost copy and the face pairs the cell with it.
  double ucur[NC], unext[NC];
  load(unext, m0 - 1);
  DtCell<D> dprev = MakeDtCell<D, DA>(unext, g);
  load(ucur, m0);
  for (int64_t m = m0; m < m1; ++m) {
    const int64_t off = xs + rowoff + m * ms;
    // The next row's load is in flight while this row is reduced.
    if (m + 1 < m1) load(unext, m + 1);
    const DtCell<D> dc = MakeDtCell<D, DA>(ucur, g);
#pragma unroll
    for (int c = 0; c < NC; ++c) ucur[c] = unext[c];
    // March direction y.
    smax[M] = SMax(smax[M], FaceSpeed<D, M, DA>(dprev, dc, g.gm1));
    if (m == nm - 1) smax[M] = SMax(smax[M], FaceSpeed<D, M, DA>(dc, dc, g.gm1));
    // x direction through LDS.
    s_cells[tx] = dc;
    __syncthreads();
    if (active) {
      DtCell<D> dl;
      if (x == 0) {
        dl = dc;
      } else if (tx > 0) {
        dl = s_cells[tx - 1];
      } else {
        dl = load_cell(off - 1);
      }
      smax[0] = SMax(smax[0], FaceSpeed<D, 0, DA>(dl, dc, g.gm1));
      if (x == nx - 1) smax[0] = SMax(smax[0], FaceSpeed<D, 0, DA>(dc, dc, g.gm1));
    }
    __syncthreads();
    dprev = dc;
  }
  if (!active) smax[0] = smax[1] = 0.0;
  const int64_t nblocks = static_cast<int64_t>(gridDim.x) * gridDim.y;
  const int64_t bid = blockIdx.x + static_cast<int64_t>(blockIdx.y) * gridDim.x;
  BlockSpeedToPartials<D, kBlock / 64, DA>(smax, dxs, partials + bid, nblocks);
}

// 3D step-start CFL.  A workgroup is a tile of 8 x rows of 64 cells (one
// wave each) marching along z over a segment.  Lanes 1..62 of a wave own
// their cells and lane 0 holds the x - 1 neighbours; row 0 is a halo row (the
// y - 1 neighbours of row 1).  Each cell yields a BoundCell (u and a^2, no
// square root) once per visit, 64 * 8 / (62 * 7) = 1.18 times; x and y faces
// read the neighbour's from LDS, z faces pair the plane with the previous one
// kept in registers.  Two workgroups of eight waves fill a CU at four waves
// per SIMD.  Tiles are grouped per XCD (workgroups are dispatched round-robin
// over the 8 XCDs): XCD k gets a contiguous range of tiles, so vertically
// neighbouring tiles, whose halo rows overlap, share one L2.
constexpr int kC3Rows = 8;
constexpr int kC3OwnX = 62;
constexpr int kC3OwnY = kC3Rows - 1;
constexpr int kXcds = 8;

//
// A face takes the exact FaceSpeed only when its upper bound U + sqrt(A^2 +
// k |du|^2) (see BoundPasses) reaches the speed floor shrunk by 2^-20 (the
// level driver's ctl->cfl_floor; without a control block, or with a zero
// floor, every face with a nonzero bound is evaluated).  A skipped face is
// then slower than the floor, and the floor is at most the largest evaluated
// speed whenever floor_hit is set, so the maximum, and every dt derived from
// it, is that of evaluating every face, bit for bit: the evaluated faces run
// the same instructions with or without a floor.  FinalizeCflKernel rejects
// the step (halt 3) when no evaluated face reached the floor.  Exact speeds
// take the x neighbour from the wave (DPP) and the y and z neighbours from
// memory (cached reloads of the row below and the plane before).
template <bool FAST>
__global__ __launch_bounds__(64 * kC3Rows) void PatchCfl3Kernel(
    const double* __restrict__ u0, const double* __restrict__ u1,
    const double* __restrict__ u2, const double* __restrict__ u3,
    const double* __restrict__ u4, int64_t nx, int64_t ny, int64_t nz, double dx0,
    double dx1, double dx2, double gamma, double gm1inv, int64_t ntiles,
    unsigned long long* partials, int* halt, Ctl* ctl) {
  using DA = typename Policies<double, FAST>::DA;
  constexpr int NC = 5;
  if (halt && *halt) return;
  const double* in[NC] = {u0, u1, u2, u3, u4};
  Gas<double> g;
  g.gamma = gamma;
  g.gm1inv = gm1inv;
  g.gm1 = gamma - 1.0;
  const double kbound = 0.125 * g.gm1;
  double floor[3] = {0.0, 0.0, 0.0};
  if (ctl) {
#pragma unroll
    for (int d = 0; d < 3; ++d) floor[d] = ctl->cfl_floor[d];
  }
  double s_crit[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) s_crit[d] = floor[d] * (1.0 - 0x1p-20);
  const int lane = threadIdx.x & 63;
  const int row = threadIdx.x >> 6;
  const int64_t per_xcd = (ntiles + kXcds - 1) / kXcds;
  const int64_t tile = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  const int64_t tiles_x = (nx + kC3OwnX - 1) / kC3OwnX;
  const bool live = tile < ntiles;
  const int64_t bx = live ? tile % tiles_x : 0;
  const int64_t by = live ? tile / tiles_x : 0;
  const int64_t x = bx * kC3OwnX + lane - 1;
  const int64_t y = by * kC3OwnY + row - 1;
  const int64_t xs = Clamp(x, nx);
  const int64_t ys = Clamp(y, ny);
  const bool own = live && lane >= 1 && lane <= kC3OwnX && x < nx && row >= 1 && y < ny;
  const int64_t m0 = static_cast<int64_t>(blockIdx.y) * kCflSeg;
  const int64_t m1 = (m0 + kCflSeg < nz) ? m0 + kCflSeg : nz;
  const int64_t plane = nx * ny;
  const int64_t col = xs + ys * nx;
  // the y - 1 neighbour's column (clamped: at y = 0 the cell itself)
  const int64_t col_below = xs + Clamp(ys - 1, ny) * nx;

  __shared__ BoundCell<3> s_b[2][kC3Rows][64];
  double smax[3] = {0.0, 0.0, 0.0};
  const double dxs[3] = {dx0, dx1, dx2};
  auto load = [&](double* u, int64_t c, int64_t m) {
    const int64_t off = c + Clamp(m, nz) * plane;
#pragma unroll
    for (int k = 0; k < NC; ++k) u[k] = in[k][off];
  };
  double ucur[NC], un1[NC], un2[NC];
  load(ucur, col, m0 - 1);
  BoundCell<3> bprev = MakeBoundCell<3>(ucur, g);
  load(ucur, col, m0);
  load(un1, col, m0 + 1);
  int buf = 0;
  for (int64_t m = m0; m < m1; ++m) {
    load(un2, col, m + 2);
    const BoundCell<3> bc = MakeBoundCell<3>(ucur, g);
    s_b[buf][row][lane] = bc;
    __syncthreads();
    bool px = true, py = true, pz = true;
    if (own) {
      const BoundCell<3> bl = s_b[buf][row][lane - 1];
      px = BoundPasses<3, 0>(bl, bc, s_crit[0], kbound) &&
           (x != nx - 1 || BoundPasses<3, 0>(bc, bc, s_crit[0], kbound));
      const BoundCell<3> bb = s_b[buf][row - 1][lane];
      py = BoundPasses<3, 1>(bb, bc, s_crit[1], kbound) &&
           (y != ny - 1 || BoundPasses<3, 1>(bc, bc, s_crit[1], kbound));
      pz = BoundPasses<3, 2>(bprev, bc, s_crit[2], kbound) &&
           (m != nz - 1 || BoundPasses<3, 2>(bc, bc, s_crit[2], kbound));
    }
    const bool all_x = __all(px), all_y = __all(py), all_z = __all(pz);
    if (!(all_x && all_y && all_z)) {
      // wave-uniform: every lane computes its DtCell, so the x neighbour's
      // is one lane away
      const DtCell<3> dc = MakeDtCell<3, DA>(ucur, g);
      if (!all_x) {
        const DtCell<3> dl = ShflUp<3>(dc);
        if (!px) {
          smax[0] = SMax(smax[0], FaceSpeed<3, 0, DA>(dl, dc, g.gm1));
          if (x == nx - 1) smax[0] = SMax(smax[0], FaceSpeed<3, 0, DA>(dc, dc, g.gm1));
        }
      }
      if (!py) {
        double ub[NC];
        load(ub, col_below, m);
        const DtCell<3> db = MakeDtCell<3, DA>(ub, g);
        smax[1] = SMax(smax[1], FaceSpeed<3, 1, DA>(db, dc, g.gm1));
        if (y == ny - 1) smax[1] = SMax(smax[1], FaceSpeed<3, 1, DA>(dc, dc, g.gm1));
      }
      if (!pz) {
        double up[NC];
        load(up, col, m - 1);
        const DtCell<3> dp = MakeDtCell<3, DA>(up, g);
        smax[2] = SMax(smax[2], FaceSpeed<3, 2, DA>(dp, dc, g.gm1));
        if (m == nz - 1) smax[2] = SMax(smax[2], FaceSpeed<3, 2, DA>(dc, dc, g.gm1));
      }
    }
    bprev = bc;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      ucur[k] = un1[k];
      un1[k] = un2[k];
    }
    buf ^= 1;
  }
  const int64_t nblocks = static_cast<int64_t>(gridDim.x) * gridDim.y;
  const int64_t bid = blockIdx.x + static_cast<int64_t>(blockIdx.y) * gridDim.x;
  BlockSpeedToPartials<3, kC3Rows, DA>(smax, dxs, partials + bid, nblocks, floor,
                                       ctl ? ctl->floor_hit : nullptr);
}

int64_t Cfl3Tiles(int64_t nx, int64_t ny) {
  return ((nx + kC3OwnX - 1) / kC3OwnX) * ((ny + kC3OwnY - 1) / kC3OwnY);
}

int64_t Cfl3GridX(int64_t nx, int64_t ny) {
  return ((Cfl3Tiles(nx, ny) + kXcds - 1) / kXcds) * kXcds;
}

// One-workgroup reduction of partial minima.
constexpr int kRedBlock = 1024;

__device__ __forceinline__ double KeyToDt(unsigned long long k) {
  return __longlong_as_double(static_cast<long long>(k));
}

__device__ unsigned long long ReduceKeys(const unsigned long long* p, int64_t n) {
  __shared__ unsigned long long s[kRedBlock / 64];
  unsigned long long k = kInfKey;
  for (int64_t i = threadIdx.x; i < n; i += kRedBlock) k = KMin(k, p[i]);
  k = WaveMinKey(k);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = k;
  __syncthreads();
  unsigned long long r = s[0];
  for (int w = 1; w < kRedBlock / 64; ++w) r = KMin(r, s[w]);
  __syncthreads();
  return r;
}

// Minima of up to N key arrays (array r: p[r * stride .. + n[r]]) in one
// pass, every thread's loads of all arrays in flight together; out[r] for
// r < count.  Every thread of the block must call it.
template <int N>
__device__ void ReduceKeysMulti(const unsigned long long* p, int64_t stride, const int64_t* n,
                                int count, unsigned long long* out) {
  __shared__ unsigned long long s[N][kRedBlock / 64];
  unsigned long long k[N];
  int64_t nmax = 0;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    k[r] = kInfKey;
    if (r < count && n[r] > nmax) nmax = n[r];
  }
#pragma unroll 4
  for (int64_t i = threadIdx.x; i < nmax; i += kRedBlock) {
#pragma unroll
    for (int r = 0; r < N; ++r)
      if (r < count && i < n[r]) k[r] = KMin(k[r], p[r * stride + i]);
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const unsigned long long w = WaveMinKey(k[r]);
    if ((threadIdx.x & 63) == 0) s[r][threadIdx.x >> 6] = w;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < N; ++r) {
    unsigned long long m = s[r][0];
    for (int w = 1; w < kRedBlock / 64; ++w) m = KMin(m, s[r][w]);
    out[r] = m;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kRedBlock) void FinalizeCheckKernel(
    const unsigned long long* partials, int64_t n, double* out, Ctl* ctl) {
  if (ctl && ctl->halt) return;
  const double v = KeyToDt(ReduceKeys(partials, n));
  if (threadIdx.x == 0) {
    if (out) *out = v;
    if (ctl) {
      ctl->last_check = v;
      if (v < ctl->split_dt) {
        // TimeStepTooLarge{level_dt} (DimensionalSplitLevelIntegrator.hpp:243-245)
        ctl->halt = 1;
        ctl->failure_dt = v;
      }
    }
  }
}

// The CFL pass's speed floor after a pass: a floor that no evaluated face
// reached may have hidden the fastest face, so the step is repeated without
// a floor (halt 3) and the margin quadruples; otherwise the next floor is
// this pass's largest speed per direction (dx / dt) less a margin that
// follows how fast that speed falls.  Directions equal to solved_dir had
// every face evaluated.  Returns false after setting halt 3.
__device__ bool RenewSpeedFloor(Ctl* ctl, const double* v, const double* fdx, int dim,
                                int solved_dir) {
  bool floor_ok = true;
  for (int d = 0; d < 3; ++d) {
    if (d != solved_dir && ctl->cfl_floor[d] > 0.0 && !ctl->floor_hit[d]) floor_ok = false;
    ctl->floor_hit[d] = 0;
  }
  const double old_margin = ctl->floor_margin > 0.0 ? ctl->floor_margin : kFloorMarginFirst;
  if (!floor_ok) {
    for (int d = 0; d < 3; ++d) ctl->cfl_floor[d] = 0.0;
    ctl->floor_margin = fmin(4.0 * old_margin, kFloorMarginMax);
    ctl->halt = 3;
    return false;
  }
  double speed[3] = {0.0, 0.0, 0.0}, drop = 0.0;
  for (int d = 0; d < dim; ++d) {
    speed[d] = fdx[d] / v[d];
    if (ctl->floor_speed[d] > 0.0 && speed[d] > 0.0)
      drop = fmax(drop, (ctl->floor_speed[d] - speed[d]) / speed[d]);
  }
  const double margin =
      fmin(kFloorMarginMax, fmax(fmax(kFloorMarginMin, 4.0 * drop), 0.5 * old_margin));
  ctl->floor_margin = margin;
  for (int d = 0; d < 3; ++d) {
    const double f = speed[d] * (1.0 - margin);
    const bool finite = d < dim && f > 0.0 && f < 1.7976931348623157e308;
    ctl->cfl_floor[d] = finite ? f : 0.0;
    ctl->floor_speed[d] = finite ? speed[d] : 0.0;
  }
  return true;
}

// With use_floor (the level's 3D pass, fdx the cell widths), the floor of
// the pass is validated first and the next step's floor set from this pass's
// largest speeds.
__global__ __launch_bounds__(kRedBlock) void FinalizeCflKernel(
    const unsigned long long* partials, int64_t n, int dim, double* out, Ctl* ctl,
    int first_dir, int strang, double fdx0, double fdx1, double fdx2, int use_floor) {
  if (ctl && ctl->halt) return;
  double v[3];
  {
    const int64_t ns[3] = {n, n, n};
    unsigned long long keys[3];
    ReduceKeysMulti<3>(partials, n, ns, dim, keys);
    for (int d = 0; d < dim; ++d) v[d] = KeyToDt(keys[d]);
  }
  if (threadIdx.x != 0) return;
  if (out) {
    for (int d = 0; d < dim; ++d) out[d] = v[d];
  }
  if (!ctl) return;
  if (use_floor) {
    const double fdx[3] = {fdx0, fdx1, fdx2};
    if (!RenewSpeedFloor(ctl, v, fdx, dim, -1)) return;
  }
  // DimensionalSplitLevelIntegrator::ComputeStableDt (:204-219) starts from
  // DBL_MAX; RunSimulation's limited_dt (RunSimulation.hpp:121-131).
  double stable = 1.7976931348623157e308;
  for (int d = 0; d < dim; ++d) {
    ctl->dt_dir[d] = v[d];
    stable = (v[d] < stable) ? v[d] : stable;
  }
  if (!(ctl->t + ctl->eps < ctl->final_time)) {
    ctl->halt = 2;
    return;
  }
  const double sdt = ctl->have_failure ? ctl->failure_dt : stable;
  const double rem = ctl->final_time - ctl->t;
  const double cdt = ctl->cfl * sdt;
  const double dt = (cdt < rem) ? cdt : rem;
  ctl->dt = dt;
  ctl->split_dt = strang ? 0.5 * dt : dt;
  // The first sweep's CFL check is the step-start stable dt of its
  // direction: the same faces of the same state.
  const double first = v[first_dir];
  ctl->last_check = first;
  if (first < ctl->split_dt) {
    ctl->halt = 1;
    ctl->failure_dt = first;
  }
}

// Step dt from per-direction stable dts already in ctl->dt_dir (used by
// methods whose CFL pass reduces into ctl->dt_dir itself).
__global__ void StepDtKernel(Ctl* ctl, int dim, int first_dir, int strang, double fdx0,
                             double fdx1, double fdx2, int use_floor) {
  if (ctl->halt) return;
  if (use_floor) {
    const double fdx[3] = {fdx0, fdx1, fdx2};
    if (!RenewSpeedFloor(ctl, ctl->dt_dir, fdx, dim, first_dir)) return;
  }
  double stable = 1.7976931348623157e308;
  for (int d = 0; d < dim; ++d) stable = (ctl->dt_dir[d] < stable) ? ctl->dt_dir[d] : stable;
  if (!(ctl->t + ctl->eps < ctl->final_time)) {
    ctl->halt = 2;
    return;
  }
  const double sdt = ctl->have_failure ? ctl->failure_dt : stable;
  const double rem = ctl->final_time - ctl->t;
  const double cdt = ctl->cfl * sdt;
  const double dt = (cdt < rem) ? cdt : rem;
  ctl->dt = dt;
  ctl->split_dt = strang ? 0.5 * dt : dt;
  const double first = ctl->dt_dir[first_dir];
  ctl->last_check = first;
  if (first < ctl->split_dt) {
    ctl->halt = 1;
    ctl->failure_dt = first;
  }
}

// CFL check of a split step whose stable dt was reduced into ctl->last_check.
__global__ void CheckValueKernel(Ctl* ctl, const double* value) {
  if (ctl->halt) return;
  const double v = *value;
  ctl->last_check = v;
  if (v < ctl->split_dt) {
    ctl->halt = 1;
    ctl->failure_dt = v;
  }
}

__device__ __forceinline__ void EndStep(Ctl* ctl) {
  const int64_t slot = ctl->cycle - ctl->dt_log_base;
  if (ctl->dt_log && slot >= 0 && slot < ctl->dt_log_len) ctl->dt_log[slot] = ctl->dt;
  ctl->t = ctl->t + ctl->dt;
  ctl->cycle += 1;
  ctl->have_failure = 0;
}

__global__ void StepEndKernel(Ctl* ctl) {
  if (ctl->halt) return;
  EndStep(ctl);
}

// The split sweeps' CFL checks of one level step, reduced after the last
// sweep in sweep order (the first failing check gives failure_dt, as if the
// step had stopped there: the sweeps after it only wrote buffers the restart
// discards), then the step end.  One launch instead of a finalize per sweep
// and a step-end kernel.
__global__ __launch_bounds__(kRedBlock) void FinalizeStepKernel(
    const unsigned long long* partials, CheckList cl, Ctl* ctl) {
  if (ctl->halt) return;
  unsigned long long keys[kMaxChecks];
  ReduceKeysMulti<kMaxChecks>(partials, cl.stride, cl.n, cl.count, keys);
  for (int r = 0; r < cl.count; ++r) {
    const double v = KeyToDt(keys[r]);
    const bool fail = v < ctl->split_dt;
    if (threadIdx.x == 0) {
      ctl->last_check = v;
      if (fail) {
        // TimeStepTooLarge{level_dt} (DimensionalSplitLevelIntegrator.hpp:243-245)
        ctl->halt = 1;
        ctl->failure_dt = v;
      }
    }
    if (fail) return;
  }
  if (threadIdx.x == 0) EndStep(ctl);
}

__global__ void KeysToDtKernel(unsigned long long* keys, int n, double* out) {
  const int i = threadIdx.x;
  if (i < n) {
    out[i] = KeyToDt(keys[i]);
    keys[i] = kInfKey;
  }
}

__global__ void ResetKeysKernel(unsigned long long* keys, int n, unsigned int* counters,
                                int nc) {
  const int i = threadIdx.x;
  if (i < n) keys[i] = kInfKey;
  if (i < nc) counters[i] = 0u;
}

template <int D, bool LARR, class T, bool FAST>
int LaunchTyped(hipStream_t s, int dir, const SweepArgs& a) {
  const int64_t blocks = SweepBlocks(D, dir, a.n, a.seg);
  if (blocks <= 0) return FVS_BAD_ARGUMENT;
  // Per-thread byte offsets from the uniform base are 32-bit: a Z sweep spans
  // one x-y plane, a Y sweep 256 / nx + 2 planes.
  const int64_t span = dir == 0 ? 0
                       : dir == 1 ? (kBlock / a.n[0] + 2) * a.n[0] * a.n[1]
                                  : a.n[0] * a.n[1];
  if (span * 8 >= (int64_t(1) << 32)) {
    SetError("fused sweep: a %lld x %lld plane exceeds the 32-bit offset range",
             static_cast<long long>(a.n[0]), static_cast<long long>(a.n[1]));
    return FVS_BAD_ARGUMENT;
  }
  if (dir == 0) {
    if (blocks > 0x7fffffff) return FVS_BAD_ARGUMENT;
    const int64_t slots = (a.n[0] + 2) * a.n[1] * a.n[2];
    if (slots >= (int64_t(1) << 32)) {
      SetError("X sweep: %lld slots exceed the 32-bit slot index", static_cast<long long>(slots));
      return FVS_BAD_ARGUMENT;
    }
    hipLaunchKernelGGL((SweepXKernel<D, LARR, T, FAST>), dim3(static_cast<unsigned>(blocks)),
                       dim3(kBlock), 0, s, a);
  } else {
    const int64_t npencil = (dir == 1) ? a.n[0] * (D == 3 ? a.n[2] : 1) : a.n[0] * a.n[1];
    const int64_t bx = (npencil + kBlock - 1) / kBlock;
    const int64_t nseg = (a.n[dir] + a.seg - 1) / a.seg;
    if (bx > 0x7fffffff || nseg > 65535) return FVS_BAD_ARGUMENT;
    if (dir == 1) {
      hipLaunchKernelGGL((SweepPencilKernel<D, 1, LARR, T, FAST>),
                         dim3(static_cast<unsigned>(bx), static_cast<unsigned>(nseg)),
                         dim3(kBlock), 0, s, a);
    } else {
      if constexpr (D == 3) {
        hipLaunchKernelGGL((SweepPencilKernel<D, 2, LARR, T, FAST>),
                           dim3(static_cast<unsigned>(bx), static_cast<unsigned>(nseg)),
                           dim3(kBlock), 0, s, a);
      } else {
        return FVS_BAD_ARGUMENT;
      }
    }
  }
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

// FP64: IEEE arithmetic (bit-exact with the reference); FP64 reciprocal:
// double with refined hardware reciprocals; FP32 flux: float flux arithmetic
// with the single-precision estimates.
#ifdef FVS_SWEEP_FAST_TU
template <int D>
int LaunchPrecision(int precision, bool larr, hipStream_t s, int dir, const SweepArgs& a) {
  if (precision == FVS_FP32_FLUX) {
    return larr ? LaunchTyped<D, true, float, true>(s, dir, a)
                : LaunchTyped<D, false, float, true>(s, dir, a);
  }
  return larr ? LaunchTyped<D, true, double, true>(s, dir, a)
              : LaunchTyped<D, false, double, true>(s, dir, a);
}
#else
template <int D>
int LaunchPrecision(int precision, bool larr, hipStream_t s, int dir, const SweepArgs& a) {
  if (precision != FVS_FP64) return LaunchSweepFast(D, precision, larr, s, dir, a);
  return larr ? LaunchTyped<D, true, double, false>(s, dir, a)
              : LaunchTyped<D, false, double, false>(s, dir, a);
}
#endif

// Compute units of the current device (queried once per device).
int DeviceCus() {
  static int cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

} // namespace

#ifdef FVS_SWEEP_FAST_TU

int LaunchSweepFast(int dim, int precision, bool larr, hipStream_t s, int dir,
                    const SweepArgs& a) {
  return dim == 2 ? LaunchPrecision<2>(precision, larr, s, dir, a)
                  : LaunchPrecision<3>(precision, larr, s, dir, a);
}

int LaunchPatchCflFast(int dim, unsigned gx, unsigned gy, hipStream_t s,
                       const double* const* cons, const int64_t n[3], const double dx[3],
                       double gamma, double gm1inv, unsigned long long* partials, int* halt,
                       Ctl* floor_ctl) {
  const dim3 grid(gx, gy);
  if (dim == 2) {
    hipLaunchKernelGGL(PatchCfl2Kernel<true>, grid, dim3(kBlock), 0, s, cons[0], cons[1],
                       cons[2], cons[3], n[0], n[1], dx[0], dx[1], gamma, gm1inv, partials,
                       halt);
  } else {
    hipLaunchKernelGGL(PatchCfl3Kernel<true>, grid, dim3(64 * kC3Rows), 0, s, cons[0], cons[1],
                       cons[2], cons[3], cons[4], n[0], n[1], n[2], dx[0], dx[1], dx[2], gamma,
                       gm1inv, Cfl3Tiles(n[0], n[1]), partials, halt, floor_ctl);
  }
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

#else // FVS_SWEEP_FAST_TU

int64_t SweepBlocks(int dim, int dir, const int64_t n[3], int seg) {
  const int64_t nz = dim == 3 ? n[2] : 1;
  if (dir == 0) {
    const int64_t slots = (n[0] + 2) * n[1] * nz;
    const int64_t waves = (slots + kXOwned - 1) / kXOwned;
    const int64_t blocks = (waves + kBlock / 64 - 1) / (kBlock / 64);
    const int64_t persistent = static_cast<int64_t>(DeviceCus()) * kXWavesPerCu / (kBlock / 64);
    return blocks < persistent ? blocks : persistent;
  }
  const int64_t npencil = (dir == 1) ? n[0] * nz : n[0] * n[1];
  const int64_t bx = (npencil + kBlock - 1) / kBlock;
  const int64_t nseg = (n[dir] + seg - 1) / seg;
  return bx * nseg;
}

// Row segments of the 2D CFL pass: 32 rows, fewer when the patch is too
// small to give the device about 1,024 workgroups (down to 1 row).
int64_t Cfl2Segments(int64_t nx, int64_t ny) {
  const int64_t tiles = (nx + kTX - 1) / kTX;
  int64_t seg = (ny * tiles + 1023) / 1024;
  seg = seg < 1 ? 1 : (seg > kCflSeg ? kCflSeg : seg);
  return (ny + seg - 1) / seg;
}

int PencilSeg(int dim, int dir, const int64_t n[3]) {
  if (dir == 0) return 32;
  const int64_t npencil = dir == 1 ? n[0] * (dim == 3 ? n[2] : 1) : n[0] * n[1];
  const int64_t bx = (npencil + kBlock - 1) / kBlock;
  const int64_t nseg = (1024 + bx - 1) / bx;
  int64_t seg = (n[dir] + nseg - 1) / nseg;
  seg = seg < 1 ? 1 : (seg > 32 ? 32 : seg);
  return static_cast<int>(seg);
}

int64_t CflBlocks(int dim, const int64_t n[3]) {
  if (dim == 3) return Cfl3GridX(n[0], n[1]) * ((n[2] + kCflSeg - 1) / kCflSeg);
  const int64_t tiles = (n[0] + kTX - 1) / kTX;
  return tiles * Cfl2Segments(n[0], n[1]);
}

int LaunchSweep(const fvs_method_desc& desc, hipStream_t s, int dir, const SweepArgs& a) {
  if (desc.base != FVS_BASE_HLLEM_LARROUTUROU && desc.base != FVS_BASE_HLLEM) {
    SetError("fused sweep: base method %d has no fused kernel", desc.base);
    return FVS_BAD_ARGUMENT;
  }
  const bool larr = desc.base == FVS_BASE_HLLEM_LARROUTUROU;
  if (desc.dim == 2) return LaunchPrecision<2>(desc.precision, larr, s, dir, a);
  return LaunchPrecision<3>(desc.precision, larr, s, dir, a);
}

int LaunchFinalizeCheck(hipStream_t s, const unsigned long long* partials, int64_t nblocks,
                        double* out, Ctl* ctl) {
  hipLaunchKernelGGL(FinalizeCheckKernel, dim3(1), dim3(kRedBlock), 0, s, partials, nblocks,
                     out, ctl);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchFinalizeCfl(hipStream_t s, const unsigned long long* partials, int64_t nblocks,
                      int dim, double* out, Ctl* ctl, int first_dir, int strang) {
  hipLaunchKernelGGL(FinalizeCflKernel, dim3(1), dim3(kRedBlock), 0, s, partials, nblocks, dim,
                     out, ctl, first_dir, strang, 1.0, 1.0, 1.0, 0);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchPatchCfl(const fvs_method_desc& desc, hipStream_t s, const double* const* cons,
                   const int64_t n[3], const double dx[3], unsigned long long* partials,
                   double* out, Ctl* ctl, int* halt, int first_dir, int strang) {
  const int D = desc.dim;
  const int64_t gx = D == 3 ? Cfl3GridX(n[0], n[1]) : (n[0] + kTX - 1) / kTX;
  const int64_t nm = D == 3 ? n[2] : n[1];
  const int64_t gy = D == 3 ? (nm + kCflSeg - 1) / kCflSeg : Cfl2Segments(n[0], n[1]);
  if (gx <= 0 || gx > 0x7fffffff || gy <= 0 || gy > 65535) return FVS_BAD_ARGUMENT;
  const double gm1inv = 1.0 / (desc.gamma - 1.0);
  const dim3 grid(static_cast<unsigned>(gx), static_cast<unsigned>(gy));
  // The level driver's 3D pass (a control block) uses the speed floor.
  Ctl* floor_ctl = (D == 3 && ctl) ? ctl : nullptr;
  if (desc.precision != FVS_FP64) {
    if (int rc = LaunchPatchCflFast(D, grid.x, grid.y, s, cons, n, dx, desc.gamma, gm1inv,
                                    partials, halt, floor_ctl))
      return rc;
  } else if (D == 2) {
    hipLaunchKernelGGL(PatchCfl2Kernel<false>, grid, dim3(kBlock), 0, s, cons[0], cons[1],
                       cons[2], cons[3], n[0], n[1], dx[0], dx[1], desc.gamma, gm1inv,
                       partials, halt);
  } else {
    hipLaunchKernelGGL(PatchCfl3Kernel<false>, grid, dim3(64 * kC3Rows), 0, s, cons[0],
                       cons[1], cons[2], cons[3], cons[4], n[0], n[1], n[2], dx[0], dx[1],
                       dx[2], desc.gamma, gm1inv, Cfl3Tiles(n[0], n[1]), partials, halt,
                       floor_ctl);
  }
  FVS_LAUNCH_CHECK();
  hipLaunchKernelGGL(FinalizeCflKernel, dim3(1), dim3(kRedBlock), 0, s, partials, gx * gy, D,
                     out, ctl, first_dir, strang, dx[0], dx[1], dx[2], floor_ctl ? 1 : 0);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchStepEnd(hipStream_t s, Ctl* ctl) {
  hipLaunchKernelGGL(StepEndKernel, dim3(1), dim3(1), 0, s, ctl);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchFinalizeStep(hipStream_t s, const unsigned long long* partials, const CheckList& cl,
                       Ctl* ctl) {
  if (cl.count < 0 || cl.count > kMaxChecks) return FVS_BAD_ARGUMENT;
  hipLaunchKernelGGL(FinalizeStepKernel, dim3(1), dim3(kRedBlock), 0, s, partials, cl, ctl);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchStepDt(hipStream_t s, Ctl* ctl, int dim, int first_dir, int strang, const double* dx,
                 int use_floor) {
  const double f0 = dx ? dx[0] : 1.0, f1 = dx ? dx[1] : 1.0, f2 = dx ? dx[2] : 1.0;
  hipLaunchKernelGGL(StepDtKernel, dim3(1), dim3(1), 0, s, ctl, dim, first_dir, strang, f0, f1,
                     f2, (dx && use_floor) ? 1 : 0);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchCheckValue(hipStream_t s, Ctl* ctl, const double* value) {
  hipLaunchKernelGGL(CheckValueKernel, dim3(1), dim3(1), 0, s, ctl, value);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchKeysToDt(hipStream_t s, unsigned long long* keys, int n, double* out) {
  hipLaunchKernelGGL(KeysToDtKernel, dim3(1), dim3(64), 0, s, keys, n, out);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchResetKeys(hipStream_t s, unsigned long long* keys, int n, unsigned int* counters,
                    int nc) {
  hipLaunchKernelGGL(ResetKeysKernel, dim3(1), dim3(64), 0, s, keys, n, counters, nc);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

#endif // FVS_SWEEP_FAST_TU

} // namespace fvs
