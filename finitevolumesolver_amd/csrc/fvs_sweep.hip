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
// X sweeps (unit stride): one workgroup per row tile of 256 cells; the
// conservative states, half-step states and fluxes are exchanged through
// LDS.  Y/Z sweeps: one thread per (x, transverse) pencil, marching along
// the sweep direction with a register window of three cells, so every load
// is coalesced along x and no LDS or barrier is needed.
#include "fvs_internal.hpp"
#include "fvs_physics.hpp"

namespace fvs {

namespace {

constexpr int kTX = 256;      // X-sweep tile (cells per workgroup)
constexpr int kBlock = 256;   // threads per workgroup

__device__ __forceinline__ int64_t Clamp(int64_t i, int64_t n) {
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

__device__ __forceinline__ unsigned long long WaveMinKey(unsigned long long k) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(k, off, 64);
    k = o < k ? o : k;
  }
  return k;
}

// Block-wide min of keys; thread 0 performs one atomic min.  Then the last
// block to finish (if a control block is given) finalises the CFL check.
__device__ void BlockCheckMin(unsigned long long key, const SweepArgs& a) {
  __shared__ unsigned long long s_keys[kBlock / 64];
  key = WaveMinKey(key);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) s_keys[wave] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long k = s_keys[0];
    for (int w = 1; w < kBlock / 64; ++w) k = s_keys[w] < k ? s_keys[w] : k;
    atomicMin(a.check_key, k);
    if (a.ctl) {
      __threadfence();
      const unsigned int total = gridDim.x * gridDim.y * gridDim.z;
      const unsigned int prev = atomicAdd(&a.ctl->check_counter, 1u);
      if (prev == total - 1) {
        __threadfence();
        const unsigned long long fk = atomicAdd(a.check_key, 0ull);
        const double v = __longlong_as_double(static_cast<long long>(fk));
        a.ctl->last_check = v;
        if (v < a.ctl->split_dt) {
          // TimeStepTooLarge{level_dt} (DimensionalSplitLevelIntegrator.hpp:243-245)
          a.ctl->halt = 1;
          a.ctl->failure_dt = v;
        }
        atomicExch(a.check_key, 0x7ff0000000000000ull);
        atomicExch(&a.ctl->check_counter, 0u);
      }
    }
  }
}

template <int D, int DIR, class T>
__device__ __forceinline__ void ReconAny(const T* um, const T* u0, const T* up, T lh,
                                         const Gas<T>& g, int limiter, int muscl,
                                         T* recL, T* recR) {
  if (!muscl) {
    GhostRec<D>(u0, g, recL);
#pragma unroll
    for (int k = 0; k < D + 4; ++k) recR[k] = recL[k];
    return;
  }
  switch (limiter) {
  case kMinMod:
    ReconstructCell<D, DIR, kMinMod>(um, u0, up, lh, g, recL, recR);
    break;
  case kVanLeer:
    ReconstructCell<D, DIR, kVanLeer>(um, u0, up, lh, g, recL, recR);
    break;
  case kNoLimiter:
    ReconstructCell<D, DIR, kNoLimiter>(um, u0, up, lh, g, recL, recR);
    break;
  default:
    ReconstructCell<D, DIR, kNoGradient>(um, u0, up, lh, g, recL, recR);
    break;
  }
}

template <int D, int DIR, bool LARR, class T>
__device__ __forceinline__ void FaceFlux(T* f, const T* L, const T* R, const Gas<T>& g,
                                         int quirk) {
  if constexpr (D == 3) {
    if (quirk) {
      HllemFlux<D, DIR, LARR, true>(f, L, R, g);
    } else {
      HllemFlux<D, DIR, LARR, false>(f, L, R, g);
    }
  } else {
    HllemFlux<D, DIR, LARR, false>(f, L, R, g);
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
// X sweep: workgroup per row tile.
template <int D, bool LARR, class T>
__global__ __launch_bounds__(kBlock) void SweepXKernel(SweepArgs a) {
  constexpr int NC = D + 2;
  constexpr int NQ = D + 4;
  if (a.halt && *a.halt) return;
  const int64_t nx = a.n[0];
  const int64_t tiles = (nx + kTX - 1) / kTX;
  const int64_t row = blockIdx.x / tiles;
  const int64_t x0 = (blockIdx.x % tiles) * kTX;
  const int cnt = static_cast<int>(nx - x0 < kTX ? nx - x0 : kTX);
  const int t = threadIdx.x;
  const int64_t rbase = row * nx;

  __shared__ double su[NC][kTX + 4];   // cells x0-2 .. x0+cnt+1 (clamped)
  __shared__ T srecL[NQ][kTX + 2];     // recL of cells x0-1 .. x0+cnt
  __shared__ T srecR[NQ][kTX + 2];     // recR, later overwritten by fluxes
  __shared__ double spa[2][kTX + 4];   // p, a of the input cells (check)

  const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
  const T lh = static_cast<T>(0.5 * dt / a.dx);
  const double dt_dx = dt / a.dx;
  const Gas<T> g = MakeGas<T>(a);

  for (int i = t; i < cnt + 4; i += kBlock) {
    const int64_t x = Clamp(x0 - 2 + i, nx);
#pragma unroll
    for (int c = 0; c < NC; ++c) su[c][i] = a.in[c][rbase + x];
  }
  __syncthreads();

  for (int i = t; i < cnt + 2; i += kBlock) {
    const int64_t cell = x0 - 1 + i;
    T um[NC], u0[NC], up[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      um[c] = static_cast<T>(su[c][i]);
      u0[c] = static_cast<T>(su[c][i + 1]);
      up[c] = static_cast<T>(su[c][i + 2]);
    }
    T recL[NQ], recR[NQ];
    if (cell < 0 || cell >= nx) {
      GhostRec<D>(u0, g, recL);
#pragma unroll
      for (int k = 0; k < NQ; ++k) recR[k] = recL[k];
    } else {
      ReconAny<D, 0>(um, u0, up, lh, g, a.limiter, a.muscl, recL, recR);
    }
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      srecL[k][i] = recL[k];
      srecR[k][i] = recR[k];
    }
    if (a.check) {
      double q[NQ];
#pragma unroll
      for (int c = 0; c < NC; ++c) q[c] = su[c][i + 1];
      CompleteFromCons<D>(q, MakeGas<double>(a));
      spa[0][i + 1] = q[D + 2];
      spa[1][i + 1] = q[D + 3];
    }
  }
  __syncthreads();

  unsigned long long key = 0x7ff0000000000000ull;
  for (int i = t; i < cnt + 1; i += kBlock) {
    T L[NQ], R[NQ], f[NC];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      L[k] = srecR[k][i];
      R[k] = srecL[k][i + 1];
    }
    FaceFlux<D, 0, LARR>(f, L, R, g, a.quirk3d);
#pragma unroll
    for (int k = 0; k < NC; ++k) srecR[k][i] = f[k];
    if (a.check) {
      double qL[NQ], qR[NQ];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        qL[c] = su[c][i + 1];
        qR[c] = su[c][i + 2];
      }
      qL[D + 2] = spa[0][i + 1];
      qL[D + 3] = spa[1][i + 1];
      qR[D + 2] = spa[0][i + 2];
      qR[D + 3] = spa[1][i + 2];
      const double dtf = HllemStableDt<D, 0>(qL, qR, a.dx, MakeGas<double>(a));
      const unsigned long long k2 = DtKey(dtf);
      key = k2 < key ? k2 : key;
    }
  }
  __syncthreads();

  for (int i = t; i < cnt; i += kBlock) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double fl = static_cast<double>(srecR[c][i]);
      const double fr = static_cast<double>(srecR[c][i + 1]);
      a.out[c][rbase + x0 + i] = su[c][i + 2] + dt_dx * (fl - fr);
    }
  }
  if (a.check) BlockCheckMin(key, a);
}

// ---------------------------------------------------------------------------
// Y / Z sweep: pencil march.
template <int D, int DIR, bool LARR, class T>
__global__ __launch_bounds__(kBlock) void SweepPencilKernel(SweepArgs a) {
  constexpr int NC = D + 2;
  constexpr int NQ = D + 4;
  if (a.halt && *a.halt) return;
  const int64_t nx = a.n[0], ny = a.n[1];
  const int64_t nd = a.n[DIR];
  const int64_t npencil = (DIR == 1) ? nx * a.n[2] : nx * ny;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const bool active = p < npencil;
  const int64_t stride = (DIR == 1) ? nx : nx * ny;
  int64_t base = 0;
  if (DIR == 1) {
    base = (p % nx) + (p / nx) * nx * ny;
  } else {
    base = p;
  }
  const int64_t j0 = static_cast<int64_t>(blockIdx.y) * a.seg;
  const int64_t j1 = (j0 + a.seg < nd) ? j0 + a.seg : nd;

  const double dt = a.dt_ptr ? *a.dt_ptr : a.dt_val;
  const T lh = static_cast<T>(0.5 * dt / a.dx);
  const double dt_dx = dt / a.dx;
  const Gas<T> g = MakeGas<T>(a);
  const Gas<double> gd = MakeGas<double>(a);

  unsigned long long key = 0x7ff0000000000000ull;
  if (active) {
    double Um[NC], U0[NC], Up[NC];
    auto load = [&](double* u, int64_t j) {
      const int64_t off = base + Clamp(j, nd) * stride;
#pragma unroll
      for (int c = 0; c < NC; ++c) u[c] = a.in[c][off];
    };
    load(Um, j0 - 1);
    load(U0, j0);
    T recRprev[NQ];
    {
      T tm[NC], t0[NC], tp[NC];
      T dummy[NQ];
      if (j0 == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) t0[c] = static_cast<T>(U0[c]);
        GhostRec<D>(t0, g, recRprev);
      } else {
        double Umm[NC];
        load(Umm, j0 - 2);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          tm[c] = static_cast<T>(Umm[c]);
          t0[c] = static_cast<T>(Um[c]);
          tp[c] = static_cast<T>(U0[c]);
        }
        ReconAny<D, DIR>(tm, t0, tp, lh, g, a.limiter, a.muscl, dummy, recRprev);
      }
    }
    double cprev[NQ];
    if (a.check) {
#pragma unroll
      for (int c = 0; c < NC; ++c) cprev[c] = Um[c];
      CompleteFromCons<D>(cprev, gd);
    }
    T Fprev[NC];
    for (int64_t j = j0; j <= j1; ++j) {
      T recL[NQ], recR[NQ];
      if (j < nd) {
        load(Up, j + 1);
        T tm[NC], t0[NC], tp[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          tm[c] = static_cast<T>(Um[c]);
          t0[c] = static_cast<T>(U0[c]);
          tp[c] = static_cast<T>(Up[c]);
        }
        ReconAny<D, DIR>(tm, t0, tp, lh, g, a.limiter, a.muscl, recL, recR);
      } else {
        T t0[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          t0[c] = static_cast<T>(U0[c]);
          Up[c] = U0[c];
        }
        GhostRec<D>(t0, g, recL);
#pragma unroll
        for (int k = 0; k < NQ; ++k) recR[k] = recL[k];
      }
      T F[NC];
      FaceFlux<D, DIR, LARR>(F, recRprev, recL, g, a.quirk3d);
      if (j > j0) {
        const int64_t off = base + (j - 1) * stride;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          a.out[c][off] = Um[c] + dt_dx * (static_cast<double>(Fprev[c]) -
                                           static_cast<double>(F[c]));
        }
      }
      if (a.check && (j < j1 || j1 == nd)) {
        double c0[NQ];
#pragma unroll
        for (int c = 0; c < NC; ++c) c0[c] = U0[c];
        CompleteFromCons<D>(c0, gd);
        const double dtf = HllemStableDt<D, DIR>(cprev, c0, a.dx, gd);
        const unsigned long long k2 = DtKey(dtf);
        key = k2 < key ? k2 : key;
#pragma unroll
        for (int k = 0; k < NQ; ++k) cprev[k] = c0[k];
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        Um[c] = U0[c];
        U0[c] = Up[c];
        Fprev[c] = F[c];
      }
#pragma unroll
      for (int k = 0; k < NQ; ++k) recRprev[k] = recR[k];
    }
  }
  if (a.check) BlockCheckMin(key, a);
}

// ---------------------------------------------------------------------------
// Step-start CFL: stable dt of every direction in one pass.  Each thread
// owns a cell c and evaluates the faces (c - e_d, c) for every direction d,
// plus the boundary face after the last cell of each row.
template <int D>
__global__ __launch_bounds__(kBlock) void PatchCflKernel(
    const double* __restrict__ u0, const double* __restrict__ u1,
    const double* __restrict__ u2, const double* __restrict__ u3,
    const double* __restrict__ u4, int64_t nx, int64_t ny, int64_t nz,
    double dx0, double dx1, double dx2, double gamma, double gm1inv,
    unsigned long long* keys, unsigned int* counter, Ctl* ctl, int* halt,
    int first_dir, int strang) {
  constexpr int NC = D + 2;
  constexpr int NQ = D + 4;
  if (halt && *halt) return;
  const double* in[5] = {u0, u1, u2, u3, u4};
  Gas<double> g;
  g.gamma = gamma;
  g.gm1inv = gm1inv;
  g.gm1 = gamma - 1.0;
  const int64_t ncell = nx * ny * nz;
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long key[3] = {0x7ff0000000000000ull, 0x7ff0000000000000ull,
                               0x7ff0000000000000ull};
  if (c < ncell) {
    const int64_t x = c % nx;
    const int64_t y = (c / nx) % ny;
    const int64_t z = c / (nx * ny);
    double q[NQ];
#pragma unroll
    for (int k = 0; k < NC; ++k) q[k] = in[k][c];
    CompleteFromCons<D>(q, g);
    const int64_t pos[3] = {x, y, z};
    const int64_t len[3] = {nx, ny, nz};
    const int64_t str[3] = {1, nx, nx * ny};
    const double dxs[3] = {dx0, dx1, dx2};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double dtf;
      if (pos[d] == 0) {
        // Face 0: the ghost copy and the first cell.
        if (d == 0) dtf = HllemStableDtUniform<D, 0>(q, dxs[0], g);
        else if (d == 1) dtf = HllemStableDtUniform<D, 1>(q, dxs[1], g);
        else dtf = HllemStableDtUniform<D, (D > 2 ? 2 : 0)>(q, dxs[2], g);
      } else {
        double qm[NQ];
#pragma unroll
        for (int k = 0; k < NC; ++k) qm[k] = in[k][c - str[d]];
        CompleteFromCons<D>(qm, g);
        if (d == 0) dtf = HllemStableDt<D, 0>(qm, q, dxs[0], g);
        else if (d == 1) dtf = HllemStableDt<D, 1>(qm, q, dxs[1], g);
        else dtf = HllemStableDt<D, (D > 2 ? 2 : 0)>(qm, q, dxs[2], g);
      }
      unsigned long long k2 = DtKey(dtf);
      key[d] = k2 < key[d] ? k2 : key[d];
      if (pos[d] == len[d] - 1) {
        // Face n: the last cell and its ghost copy.
        if (d == 0) dtf = HllemStableDtUniform<D, 0>(q, dxs[0], g);
        else if (d == 1) dtf = HllemStableDtUniform<D, 1>(q, dxs[1], g);
        else dtf = HllemStableDtUniform<D, (D > 2 ? 2 : 0)>(q, dxs[2], g);
        k2 = DtKey(dtf);
        key[d] = k2 < key[d] ? k2 : key[d];
      }
    }
  }
  __shared__ unsigned long long s_keys[3][kBlock / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const unsigned long long k = WaveMinKey(key[d]);
    if (lane == 0) s_keys[d][wave] = k;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int d = 0; d < D; ++d) {
      unsigned long long k = s_keys[d][0];
      for (int w = 1; w < kBlock / 64; ++w) k = s_keys[d][w] < k ? s_keys[d][w] : k;
      atomicMin(&keys[d], k);
    }
    if (ctl) {
      __threadfence();
      const unsigned int prev = atomicAdd(counter, 1u);
      if (prev == gridDim.x - 1) {
        __threadfence();
        // Last block: RunSimulation's limited_dt (RunSimulation.hpp:121-131).
        double stable = 1.7976931348623157e308; // DimensionalSplitLevelIntegrator min_dt start
        for (int d = 0; d < D; ++d) {
          const unsigned long long k = atomicAdd(&keys[d], 0ull);
          const double v = __longlong_as_double(static_cast<long long>(k));
          ctl->dt_dir[d] = v;
          stable = (v < stable) ? v : stable;
          atomicExch(&keys[d], 0x7ff0000000000000ull);
        }
        atomicExch(counter, 0u);
        if (!(ctl->t < ctl->final_time)) {
          ctl->halt = 2;
        } else {
          const double sdt = ctl->have_failure ? ctl->failure_dt : stable;
          const double rem = ctl->final_time - ctl->t;
          const double cdt = ctl->cfl * sdt;
          const double dt = (cdt < rem) ? cdt : rem;
          ctl->dt = dt;
          ctl->split_dt = strang ? 0.5 * dt : dt;
          // The first sweep's CFL check is the step-start stable dt of its
          // direction: the same faces of the same state.
          const double first = ctl->dt_dir[first_dir];
          ctl->last_check = first;
          if (first < ctl->split_dt) {
            ctl->halt = 1;
            ctl->failure_dt = first;
          }
        }
      }
    }
  }
}

__global__ void StepEndKernel(Ctl* ctl, double* dt_log, int64_t dt_log_base,
                              int dt_log_len) {
  if (ctl->halt) return;
  const int64_t slot = ctl->cycle - dt_log_base;
  if (dt_log && slot >= 0 && slot < dt_log_len) dt_log[slot] = ctl->dt;
  ctl->t = ctl->t + ctl->dt;
  ctl->cycle += 1;
  ctl->have_failure = 0;
}

__global__ void KeysToDtKernel(unsigned long long* keys, int n, double* out) {
  const int i = threadIdx.x;
  if (i < n) {
    out[i] = __longlong_as_double(static_cast<long long>(keys[i]));
    keys[i] = 0x7ff0000000000000ull;
  }
}

__global__ void ResetKeysKernel(unsigned long long* keys, int n, unsigned int* counters,
                                int nc) {
  const int i = threadIdx.x;
  if (i < n) keys[i] = 0x7ff0000000000000ull;
  if (i < nc) counters[i] = 0u;
}

template <int D, bool LARR, class T>
int LaunchTyped(hipStream_t s, int dir, const SweepArgs& a) {
  if (dir == 0) {
    const int64_t tiles = (a.n[0] + kTX - 1) / kTX;
    const int64_t rows = a.n[1] * (D == 3 ? a.n[2] : 1);
    const int64_t blocks = tiles * rows;
    if (blocks <= 0 || blocks > 0x7fffffff) return FVS_BAD_ARGUMENT;
    hipLaunchKernelGGL((SweepXKernel<D, LARR, T>), dim3(static_cast<unsigned>(blocks)),
                       dim3(kBlock), 0, s, a);
  } else {
    const int64_t npencil = (dir == 1) ? a.n[0] * (D == 3 ? a.n[2] : 1) : a.n[0] * a.n[1];
    const int64_t bx = (npencil + kBlock - 1) / kBlock;
    const int64_t nseg = (a.n[dir] + a.seg - 1) / a.seg;
    if (bx <= 0 || bx > 0x7fffffff || nseg <= 0 || nseg > 65535) return FVS_BAD_ARGUMENT;
    if (dir == 1) {
      hipLaunchKernelGGL((SweepPencilKernel<D, 1, LARR, T>),
                         dim3(static_cast<unsigned>(bx), static_cast<unsigned>(nseg)),
                         dim3(kBlock), 0, s, a);
    } else {
      if constexpr (D == 3) {
        hipLaunchKernelGGL((SweepPencilKernel<D, 2, LARR, T>),
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

} // namespace

int LaunchSweep(const fvs_method_desc& desc, hipStream_t s, int dir, const SweepArgs& a) {
  if (desc.base != FVS_BASE_HLLEM_LARROUTUROU && desc.base != FVS_BASE_HLLEM) {
    SetError("fused sweep: base method %d has no fused kernel", desc.base);
    return FVS_BAD_ARGUMENT;
  }
  const bool larr = desc.base == FVS_BASE_HLLEM_LARROUTUROU;
  const bool fp32 = desc.precision == FVS_FP32_FLUX;
  if (desc.dim == 2) {
    if (larr) return fp32 ? LaunchTyped<2, true, float>(s, dir, a) : LaunchTyped<2, true, double>(s, dir, a);
    return fp32 ? LaunchTyped<2, false, float>(s, dir, a) : LaunchTyped<2, false, double>(s, dir, a);
  }
  if (larr) return fp32 ? LaunchTyped<3, true, float>(s, dir, a) : LaunchTyped<3, true, double>(s, dir, a);
  return fp32 ? LaunchTyped<3, false, float>(s, dir, a) : LaunchTyped<3, false, double>(s, dir, a);
}

int LaunchPatchCfl(const fvs_method_desc& desc, hipStream_t s, const double* const* cons,
                   const int64_t n[3], const double dx[3], unsigned long long* keys,
                   unsigned int* counter, Ctl* ctl, int* halt, int first_dir,
                   int strang) {
  const int64_t ncell = n[0] * n[1] * (desc.dim == 3 ? n[2] : 1);
  const int64_t blocks = (ncell + kBlock - 1) / kBlock;
  if (blocks <= 0 || blocks > 0x7fffffff) return FVS_BAD_ARGUMENT;
  const double gm1inv = 1.0 / (desc.gamma - 1.0);
  if (desc.dim == 2) {
    hipLaunchKernelGGL(PatchCflKernel<2>, dim3(static_cast<unsigned>(blocks)), dim3(kBlock),
                       0, s, cons[0], cons[1], cons[2], cons[3], nullptr, n[0], n[1],
                       int64_t(1), dx[0], dx[1], dx[0], desc.gamma, gm1inv, keys, counter,
                       ctl, halt, first_dir, strang);
  } else {
    hipLaunchKernelGGL(PatchCflKernel<3>, dim3(static_cast<unsigned>(blocks)), dim3(kBlock),
                       0, s, cons[0], cons[1], cons[2], cons[3], cons[4], n[0], n[1], n[2],
                       dx[0], dx[1], dx[2], desc.gamma, gm1inv, keys, counter, ctl, halt,
                       first_dir, strang);
  }
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int LaunchStepEnd(hipStream_t s, Ctl* ctl, double* dt_log, int64_t dt_log_base,
                  int dt_log_len) {
  hipLaunchKernelGGL(StepEndKernel, dim3(1), dim3(1), 0, s, ctl, dt_log, dt_log_base,
                     dt_log_len);
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

} // namespace fvs
