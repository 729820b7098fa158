// Batched split sweeps over all patches of an AMR level in one launch.
//
// The reference sweeps a level patch by patch (ForEachFab over the
// MultiFab, include/fub/AMReX/ForEachFab.hpp:48-57, each patch running
// FluxMethod::ComputeNumericFluxes and HyperbolicPatchIntegrator::
// UpdateConservatively, include/fub/flux_method/FluxMethod.hpp:246-293,
// src/HyperbolicPatchIntegrator.cpp:35-154).  Here every patch of uneven
// size is cut into tiles of one shape, and a persistent grid walks the tile
// list: each tile reconstructs its cells once (MUSCL-Hancock half-step
// states into LDS), computes its faces' HLLEM fluxes into LDS and updates
// its owned cells, all in registers/LDS; the CFL minimum of the faces is
// reduced per workgroup.  The tile shape per direction is picked by a
// timing autotuner; the level's sweep sequence (ghost fill, X sweep, ghost
// fill, Y sweep) is captured once into a hipGraph and replayed.
//
// Reconstruction and flux arithmetic run in T (double, or float for the
// mixed-precision form); the conservative update and the CFL values are
// FP64.  Patches are 2D with two ghost layers; states are conservative
// (rho, rho u, rho v, rho E), one pool per buffer.
//
// This file is compiled twice: the plain build holds the IEEE FP64 kernels
// and the C entry points; fvs_batch_fast.hip defines FVS_BATCH_FAST_TU and
// compiles the kernels of the tolerance-checked modes with multiply-add
// contraction.
#include "fvs_batch_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

namespace fvs {
using batch::BatchArgs;
using batch::PatchInfo;
using batch::kG;

namespace {

template <class T>
__device__ __forceinline__ Gas<T> GasT(const Gas<double>& g) {
  Gas<T> r;
  r.gamma = static_cast<T>(g.gamma);
  r.gm1inv = static_cast<T>(g.gm1inv);
  r.gm1 = static_cast<T>(g.gm1);
  return r;
}

__device__ __forceinline__ unsigned long long KMinKey(unsigned long long a, unsigned long long b) {
  return b < a ? b : a;
}

__device__ __forceinline__ unsigned long long WaveMinK(unsigned long long k) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(k, off, 64);
    k = o < k ? o : k;
  }
  return k;
}

// Tile = TX x TY threads; along the sweep direction the T_d threads hold
// cells start-1 .. start+T_d-2, of which start .. start+T_d-3 are owned.
template <int DIR, int LIM, class T, int TX, int TY, bool FAST>
__global__ __launch_bounds__(TX * TY) void BatchSweepKernel(BatchArgs a) {
  using TA = typename Policies<T, FAST>::TA;
  using DA = typename Policies<T, FAST>::DA;
  constexpr int NQ = 6, NC = 4;
  constexpr int TD = DIR == 0 ? TX : TY;
  __shared__ T recL[NQ][TY][TX];
  __shared__ T recR[NQ][TY][TX];
  __shared__ T flx[NC][TY][TX];
  __shared__ unsigned long long red[(TX * TY + 63) / 64];
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int td = DIR == 0 ? tx : ty;   // index along the sweep
  const int tt = DIR == 0 ? ty : tx;   // index across
  const Gas<T> gT = GasT<T>(a.g);
  const double dt = *a.dt;
  const double kbound = 0.125 * a.g.gm1;
  unsigned long long kmin = 0x7ff0000000000000ull;
  for (int tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int4 t = a.tiles[tile];
    const PatchInfo P = a.patches[t.x];
    const int nd = DIR == 0 ? P.nx : P.ny;
    const int nt = DIR == 0 ? P.ny : P.nx;
    const int c = t.y - 1 + td;   // cell along the sweep
    const int r = t.z + tt;       // index across (ghost rows when grown)
    const int64_t ldx = P.nx + 2 * kG;
    const int64_t cs = ldx * (P.ny + 2 * kG);
    const int64_t step = DIR == 0 ? 1 : ldx;
    const bool row_ok = r >= -t.w && r < nt + t.w;
    const bool row_inner = r >= 0 && r < nt;
    const bool recon_ok = row_ok && c >= -1 && c <= nd && c <= t.y + TD - 2;
    const double s_crit = BoundSpeedLimit(P.dx, dt);
    int64_t at = 0;
    if (row_ok) {
      const int i = DIR == 0 ? c : r, j = DIR == 0 ? r : c;
      at = P.offset + (int64_t(j) + kG) * ldx + (int64_t(i) + kG);
    }
    double u0[NC];
    if (recon_ok) {
      double um[NC], up[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        um[k] = a.in[at + k * cs - step];
        u0[k] = a.in[at + k * cs];
        up[k] = a.in[at + k * cs + step];
      }
      // CFL of face c (cells c - 1, c), FP64, interior rows.  With a time
      // step set (dt > 0) the sweep's minimum only has to decide the check
      // dt <= stable dt exactly: a wave whose faces all pass the certified
      // speed bound (BoundPasses: stable dt above dt) reports dt for them
      // instead of evaluating the exact speeds.  The reported minimum is then
      // the exact one whenever it lies below dt (the check fails, and the
      // slowest face's bound fails too, so its wave evaluates exactly), and
      // a value of at least dt otherwise.  dt = 0 (a stable-dt pass)
      // evaluates every face.
      const bool face = row_inner && c >= 0 && c <= nd && td >= 1;
      bool exact = true;
      if (dt > 0.0) {
        const BoundCell<2> bl = MakeBoundCell<2>(um, a.g), br = MakeBoundCell<2>(u0, a.g);
        exact = !__all(!face || BoundPasses<2, DIR>(bl, br, s_crit, kbound));
        if (!exact && face) kmin = KMinKey(kmin, DtKey(dt));
      }
      if (exact && face) {
        double L[NQ], R[NQ];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          L[k] = um[k];
          R[k] = u0[k];
        }
        CompleteFromCons<2, double, DA>(L, a.g);
        CompleteFromCons<2, double, DA>(R, a.g);
        const double sdt = HllemStableDt<2, DIR, double, DA>(L, R, P.dx, a.g);
        kmin = KMinKey(kmin, DtKey(sdt));
      }
      T m[NQ], z[NQ], p[NQ], rl[NQ], rr[NQ];
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        m[k] = static_cast<T>(um[k]);
        z[k] = static_cast<T>(u0[k]);
        p[k] = static_cast<T>(up[k]);
      }
      const T lh = static_cast<T>(DivRn(0.5 * dt, P.dx));
      ReconstructCell<2, DIR, LIM, T, TA>(m, z, p, lh, gT, rl, rr);
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        recL[k][ty][tx] = rl[k];
        recR[k][ty][tx] = rr[k];
      }
    }
    __syncthreads();
    // face c between cells c - 1 (thread td - 1) and c (thread td)
    if (recon_ok && td >= 1 && c >= 0) {
      T L[NQ], R[NQ], f[NC];
      const int ptx = DIR == 0 ? tx - 1 : tx, pty = DIR == 0 ? ty : ty - 1;
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        L[k] = recR[k][pty][ptx];
        R[k] = recL[k][ty][tx];
      }
      HllemFlux<2, DIR, true, false, T, TA>(f, L, R, gT);
#pragma unroll
      for (int k = 0; k < NC; ++k) flx[k][ty][tx] = f[k];
    }
    __syncthreads();
    // update of owned cell c from faces c (this thread) and c + 1 (next)
    if (recon_ok && td >= 1 && td <= TD - 2 && c >= 0 && c < nd) {
      const double dt_dx = DivRn(dt, P.dx);
      const int ntx = DIR == 0 ? tx + 1 : tx, nty = DIR == 0 ? ty : ty + 1;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const double fl = static_cast<double>(flx[k][ty][tx]);
        const double fr = static_cast<double>(flx[k][nty][ntx]);
        a.out[at + k * cs] = u0[k] + dt_dx * (fl - fr);
      }
    }
    __syncthreads();
  }
  kmin = WaveMinK(kmin);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = kmin;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red[0];
    for (int i = 1; i < (TX * TY + 63) / 64; ++i) m = red[i] < m ? red[i] : m;
    a.partials[blockIdx.x] = m;
  }
}

// Transmissive ghost fill of every patch (both directions, corners from the
// filled rows): one thread per ghost cell of a patch, grid-stride.
__global__ void BatchGhostKernel(const PatchInfo* patches, int npatch, double* pool) {
  for (int p = blockIdx.x; p < npatch; p += gridDim.x) {
    const PatchInfo P = patches[p];
    const int64_t ldx = P.nx + 2 * kG, ldy = P.ny + 2 * kG, cs = ldx * ldy;
    // x ghosts of the interior rows
    for (int q = threadIdx.x; q < P.ny * 2 * kG; q += blockDim.x) {
      const int j = q / (2 * kG), g = q % (2 * kG);
      const int i = g < kG ? -(g + 1) : P.nx + (g - kG);
      const int si = g < kG ? 0 : P.nx - 1;
      for (int k = 0; k < 4; ++k)
        pool[P.offset + k * cs + (j + kG) * ldx + (i + kG)] =
            pool[P.offset + k * cs + (j + kG) * ldx + (si + kG)];
    }
    __syncthreads();
    // y ghosts of the grown columns (corners included)
    for (int q = threadIdx.x; q < ldx * 2 * kG; q += blockDim.x) {
      const int i = q / (2 * kG) - kG, g = q % (2 * kG);
      const int j = g < kG ? -(g + 1) : P.ny + (g - kG);
      const int sj = g < kG ? 0 : P.ny - 1;
      for (int k = 0; k < 4; ++k)
        pool[P.offset + k * cs + (j + kG) * ldx + (i + kG)] =
            pool[P.offset + k * cs + (sj + kG) * ldx + (i + kG)];
    }
    __syncthreads();
  }
}

struct Shape {
  int tx, ty;
};
// Candidate tile shapes (threads along x, along y) per sweep direction.
constexpr Shape kShapesX[] = {{64, 4}, {32, 8}, {128, 2}, {64, 2}};
constexpr Shape kShapesY[] = {{16, 16}, {32, 8}, {64, 4}, {8, 32}};
constexpr int kNumShapes = 4;

template <int DIR, class T, int TX, int TY, bool FAST>
void LaunchShape(hipStream_t s, const BatchArgs& a, int limiter, unsigned grid) {
  switch (limiter) {
  case kMinMod:
    hipLaunchKernelGGL((BatchSweepKernel<DIR, kMinMod, T, TX, TY, FAST>), dim3(grid), dim3(TX * TY), 0, s, a);
    break;
  case kVanLeer:
    hipLaunchKernelGGL((BatchSweepKernel<DIR, kVanLeer, T, TX, TY, FAST>), dim3(grid), dim3(TX * TY), 0, s, a);
    break;
  case kNoLimiter:
    hipLaunchKernelGGL((BatchSweepKernel<DIR, kNoLimiter, T, TX, TY, FAST>), dim3(grid), dim3(TX * TY), 0, s, a);
    break;
  default:
    hipLaunchKernelGGL((BatchSweepKernel<DIR, kNoGradient, T, TX, TY, FAST>), dim3(grid), dim3(TX * TY), 0, s, a);
    break;
  }
}

template <class T, bool FAST>
void LaunchSweepShape(hipStream_t s, int dir, int shape, const BatchArgs& a, int limiter,
                      unsigned grid) {
  if (dir == 0) {
    switch (shape) {
    case 0: LaunchShape<0, T, 64, 4, FAST>(s, a, limiter, grid); break;
    case 1: LaunchShape<0, T, 32, 8, FAST>(s, a, limiter, grid); break;
    case 2: LaunchShape<0, T, 128, 2, FAST>(s, a, limiter, grid); break;
    default: LaunchShape<0, T, 64, 2, FAST>(s, a, limiter, grid); break;
    }
  } else {
    switch (shape) {
    case 0: LaunchShape<1, T, 16, 16, FAST>(s, a, limiter, grid); break;
    case 1: LaunchShape<1, T, 32, 8, FAST>(s, a, limiter, grid); break;
    case 2: LaunchShape<1, T, 64, 4, FAST>(s, a, limiter, grid); break;
    default: LaunchShape<1, T, 8, 32, FAST>(s, a, limiter, grid); break;
    }
  }
}


// Flux-register contributions: one thread per register entry, the entry's
// faces in order (the fine faces along the coarse face, or the coarse face),
// each face flux computed exactly as BatchSweepKernel computes it (the two
// adjacent cells' MUSCL-Hancock half-step states, then HLLEM) from the same
// pool and time step, so the register holds the fluxes the sweeps applied.
template <int DIR, int LIM, class T, bool FAST>
__device__ __forceinline__ void RegFluxEntry(const batch::RegArgs& a, int q0) {
  using TA = typename Policies<T, FAST>::TA;
  constexpr int NQ = 6, NC = 4;
  if (q0 >= a.count) return;
  const int q = a.first + q0;
  const batch::RegFace F = a.faces[q];
  const PatchInfo P = a.patches[a.fine ? F.fp : F.cp];
  const int i = a.fine ? F.fi : F.ci, j = a.fine ? F.fj : F.cj;
  const int64_t ldx = P.nx + 2 * kG, cs = ldx * (P.ny + 2 * kG);
  const int64_t step = DIR == 0 ? 1 : ldx, across = DIR == 0 ? ldx : 1;
  const double dt = *a.dt;
  const T lh = static_cast<T>(DivRn(0.5 * dt, P.dx));
  const double w = P.dx * dt;
  const Gas<T> gT = GasT<T>(a.g);
  double acc[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) acc[k] = a.reg[k * a.nfaces + q];
  for (int m = 0; m < a.ratio; ++m) {
    const int64_t b = P.offset + (int64_t(j) + kG) * ldx + (int64_t(i) + kG) + m * across;
    T s0[NQ], s1[NQ], s2[NQ], s3[NQ];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      s0[k] = static_cast<T>(a.pool[b - 2 * step + k * cs]);
      s1[k] = static_cast<T>(a.pool[b - step + k * cs]);
      s2[k] = static_cast<T>(a.pool[b + k * cs]);
      s3[k] = static_cast<T>(a.pool[b + step + k * cs]);
    }
    T rlA[NQ], rrA[NQ], rlB[NQ], rrB[NQ], f[NC];
    ReconstructCell<2, DIR, LIM, T, TA>(s0, s1, s2, lh, gT, rlA, rrA);
    ReconstructCell<2, DIR, LIM, T, TA>(s1, s2, s3, lh, gT, rlB, rrB);
    HllemFlux<2, DIR, true, false, T, TA>(f, rrA, rlB, gT);
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = acc[k] + a.sign * (w * static_cast<double>(f[k]));
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) a.reg[k * a.nfaces + q] = acc[k];
}

// Two independent register contributions of one sweep (different
// registers) in one launch: the first nb0 workgroups serve a0, the rest a1.
template <int DIR, int LIM, class T, bool FAST>
__global__ __launch_bounds__(256) void RegFluxKernel(batch::RegArgs a0, batch::RegArgs a1,
                                                     int nb0) {
  const int b = blockIdx.x;
  if (b < nb0) {
    RegFluxEntry<DIR, LIM, T, FAST>(a0, b * 256 + static_cast<int>(threadIdx.x));
  } else {
    RegFluxEntry<DIR, LIM, T, FAST>(a1, (b - nb0) * 256 + static_cast<int>(threadIdx.x));
  }
}

template <int DIR, class T, bool FAST>
void LaunchRegFluxT(hipStream_t s, const batch::RegArgs& a0, const batch::RegArgs& a1) {
  const int nb0 = (a0.count + 255) / 256, nb1 = (a1.count + 255) / 256;
  const dim3 grid(nb0 + nb1);
  switch (a0.count > 0 ? a0.limiter : a1.limiter) {
  case kMinMod: hipLaunchKernelGGL((RegFluxKernel<DIR, kMinMod, T, FAST>), grid, dim3(256), 0, s, a0, a1, nb0); break;
  case kVanLeer: hipLaunchKernelGGL((RegFluxKernel<DIR, kVanLeer, T, FAST>), grid, dim3(256), 0, s, a0, a1, nb0); break;
  case kNoLimiter: hipLaunchKernelGGL((RegFluxKernel<DIR, kNoLimiter, T, FAST>), grid, dim3(256), 0, s, a0, a1, nb0); break;
  default: hipLaunchKernelGGL((RegFluxKernel<DIR, kNoGradient, T, FAST>), grid, dim3(256), 0, s, a0, a1, nb0); break;
  }
}

} // namespace

#ifdef FVS_BATCH_FAST_TU

void batch::LaunchRegFluxFast(hipStream_t s, int dir, const batch::RegArgs& a0,
                              const batch::RegArgs& a1, int precision) {
  if (precision == FVS_FP32_FLUX) {
    if (dir == 0) LaunchRegFluxT<0, float, true>(s, a0, a1); else LaunchRegFluxT<1, float, true>(s, a0, a1);
  } else {
    if (dir == 0) LaunchRegFluxT<0, double, true>(s, a0, a1); else LaunchRegFluxT<1, double, true>(s, a0, a1);
  }
}

void batch::LaunchSweepFast(hipStream_t s, int dir, int shape, const BatchArgs& a,
                            unsigned grid, int precision) {
  if (precision == FVS_FP32_FLUX)
    LaunchSweepShape<float, true>(s, dir, shape, a, a.limiter, grid);
  else
    LaunchSweepShape<double, true>(s, dir, shape, a, a.limiter, grid);
}

} // namespace fvs

#else // FVS_BATCH_FAST_TU

void batch::LaunchRegFlux(hipStream_t s, int dir, const batch::RegArgs& a0,
                          const batch::RegArgs& a1, int precision) {
  if (a0.count <= 0 && a1.count <= 0) return;
  if (precision != FVS_FP64) {
    batch::LaunchRegFluxFast(s, dir, a0, a1, precision);
  } else if (dir == 0) {
    LaunchRegFluxT<0, double, false>(s, a0, a1);
  } else {
    LaunchRegFluxT<1, double, false>(s, a0, a1);
  }
}

} // namespace fvs

using namespace fvs;


namespace {

// Tiles of every patch; grow > 0 extends the rows (dir 0) / columns (dir 1)
// across the sweep to -grow .. n + grow - 1.
void BuildTiles(const std::vector<PatchInfo>& ps, int dir, Shape sh, int grow,
                std::vector<int4>& out) {
  out.clear();
  const int td = dir == 0 ? sh.tx : sh.ty;
  const int tt = dir == 0 ? sh.ty : sh.tx;
  const int own = td - 2;
  for (size_t p = 0; p < ps.size(); ++p) {
    const int nd = dir == 0 ? ps[p].nx : ps[p].ny;
    const int nt = dir == 0 ? ps[p].ny : ps[p].nx;
    for (int r = -grow; r < nt + grow; r += tt)
      for (int s = 0; s < nd; s += own) out.push_back(make_int4(int(p), s, r, grow));
  }
}

int UploadTiles(fvs_batch* b, const std::vector<int4>& tiles, int4** dst, int* n, unsigned* grid) {
  if (*dst) {
    FVS_HIP_CHECK(hipStreamSynchronize(b->h->stream));
    FVS_HIP_CHECK(hipFree(*dst));
    *dst = nullptr;
  }
  FVS_HIP_CHECK(hipMalloc(dst, std::max<size_t>(tiles.size(), 1) * sizeof(int4)));
  FVS_HIP_CHECK(hipMemcpy(*dst, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
  *n = static_cast<int>(tiles.size());
  // persistent grid: a few workgroups per CU, never more than tiles
  const unsigned want = static_cast<unsigned>(b->cus * 8);
  *grid = std::max(1u, std::min<unsigned>(want, static_cast<unsigned>(tiles.size())));
  return FVS_OK;
}

int Upload(fvs_batch* b, int dir, int shape_idx) {
  const Shape sh = dir == 0 ? kShapesX[shape_idx] : kShapesY[shape_idx];
  std::vector<int4> tiles;
  BuildTiles(b->host_patches, dir, sh, 0, tiles);
  if (int rc = UploadTiles(b, tiles, &b->tiles[dir], &b->ntiles[dir], &b->grid[dir])) return rc;
  BuildTiles(b->host_patches, dir, sh, kG, tiles);
  if (int rc = UploadTiles(b, tiles, &b->tiles_grown[dir], &b->ntiles_grown[dir],
                           &b->grid_grown[dir]))
    return rc;
  b->gen += 1;
  b->shape[dir] = shape_idx;
  return FVS_OK;
}

// The sweep kernel alone, its per-workgroup minima written to `partials`
// (b->grid[dir] entries).
int LaunchSweepOnly(fvs_batch* b, hipStream_t s, int dir, int src, unsigned long long* partials,
                    int precision, bool grown = false) {
  BatchArgs a;
  a.patches = b->patches;
  a.tiles = grown ? b->tiles_grown[dir] : b->tiles[dir];
  a.ntiles = grown ? b->ntiles_grown[dir] : b->ntiles[dir];
  const unsigned grid = grown ? b->grid_grown[dir] : b->grid[dir];
  a.in = b->pool[src];
  a.out = b->pool[1 - src];
  a.dt = b->dts;
  a.g.gamma = b->h->desc.gamma;
  a.g.gm1inv = 1.0 / (b->h->desc.gamma - 1.0);
  a.g.gm1 = b->h->desc.gamma - 1.0;
  a.limiter = b->h->desc.limiter;
  a.partials = partials;
  if (precision != FVS_FP64)
    batch::LaunchSweepFast(s, dir, b->shape[dir], a, grid, precision);
  else
    LaunchSweepShape<double, false>(s, dir, b->shape[dir], a, a.limiter, grid);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int Launch(fvs_batch* b, hipStream_t s, int dir, int src, double* dt_out, int precision) {
  if (int rc = LaunchSweepOnly(b, s, dir, src, b->partials, precision)) return rc;
  return LaunchFinalizeCheck(s, b->partials, b->grid[dir], dt_out, nullptr);
}

} // namespace

int fvs::batch::Sweep(fvs_batch* b, hipStream_t s, int dir, int src, double* dt_out) {
  return Launch(b, s, dir, src, dt_out, b->h->desc.precision);
}

namespace {

// Block q reduces the `cap` partial minima (keys: bit patterns of positive
// doubles) of slot q into out[q].
__global__ __launch_bounds__(256) void SlotMinKernel(const unsigned long long* partials, int cap,
                                                     double* out) {
  __shared__ unsigned long long s[4];
  const unsigned long long* p = partials + static_cast<int64_t>(blockIdx.x) * cap;
  unsigned long long k = 0x7ff0000000000000ull;
  for (int i = threadIdx.x; i < cap; i += blockDim.x) k = p[i] < k ? p[i] : k;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(k, off, 64);
    k = o < k ? o : k;
  }
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = s[0];
    for (int w = 1; w < 4; ++w) m = s[w] < m ? s[w] : m;
    out[blockIdx.x] = __longlong_as_double(static_cast<long long>(m));
  }
}

__global__ void FillKeysKernel(unsigned long long* p, int64_t n) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < n) p[q] = 0x7ff0000000000000ull;
}

} // namespace

int fvs::batch::LaunchSlotMin(hipStream_t s, unsigned long long* partials, int nslots, int cap,
                              double* out, bool clear) {
  if (clear) {
    const int64_t n = static_cast<int64_t>(nslots) * cap;
    hipLaunchKernelGGL(FillKeysKernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                       s, partials, n);
  } else {
    hipLaunchKernelGGL(SlotMinKernel, dim3(nslots), dim3(256), 0, s, partials, cap, out);
  }
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs::batch::SweepPartials(fvs_batch* b, hipStream_t s, int dir, int src,
                              unsigned long long* partials, bool grown) {
  return LaunchSweepOnly(b, s, dir, src, partials, b->h->desc.precision, grown);
}

extern "C" {

int fvs_batch_create(fvs_handle* h, int64_t count, const int32_t* n, const double* dx,
                     fvs_batch** out) {
  if (!h || count <= 0 || !n || !dx || !out) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  *out = nullptr;
  if (h->desc.dim != 2 || h->desc.base != FVS_BASE_HLLEM_LARROUTUROU ||
      h->desc.reconstruction != FVS_RECON_MUSCL_HANCOCK) {
    SetError("batched sweeps: 2D MUSCL-Hancock with the Larrouturou HLLEM base");
    return FVS_BAD_ARGUMENT;
  }
  fvs_batch* b = new (std::nothrow) fvs_batch();
  if (!b) {
    SetError("out of host memory");
    return FVS_BAD_ARGUMENT;
  }
  b->h = h;
  b->count = count;
  int64_t off = 0;
  b->host_patches.resize(count);
  for (int64_t p = 0; p < count; ++p) {
    if (n[2 * p] < 2 || n[2 * p + 1] < 2 || !(dx[p] > 0.0)) {
      delete b;
      SetError("patch %lld: bad size or dx", (long long)p);
      return FVS_BAD_ARGUMENT;
    }
    PatchInfo& P = b->host_patches[p];
    P.offset = off;
    P.nx = n[2 * p];
    P.ny = n[2 * p + 1];
    P.dx = dx[p];
    off += 4 * int64_t(P.nx + 2 * kG) * (P.ny + 2 * kG);
  }
  b->pool_size = off;
  hipDeviceProp_t prop;
  int dev = 0;
  (void)hipGetDevice(&dev);
  b->cus = (hipGetDeviceProperties(&prop, dev) == hipSuccess) ? prop.multiProcessorCount : 256;
  auto fail = [&](int rc) {
    fvs_batch_destroy(b);
    return rc;
  };
  if (hipMalloc(&b->patches, count * sizeof(PatchInfo)) != hipSuccess ||
      hipMalloc(&b->pool[0], off * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->pool[1], off * sizeof(double)) != hipSuccess ||
      hipMalloc(&b->dts, 4 * sizeof(double)) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  if (hipMemcpy(b->patches, b->host_patches.data(), count * sizeof(PatchInfo),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(b->pool[0], 0, off * sizeof(double)) != hipSuccess ||
      hipMemset(b->pool[1], 0, off * sizeof(double)) != hipSuccess ||
      hipMemset(b->dts, 0, 4 * sizeof(double)) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  b->partials_cap = int64_t(b->cus) * 8 * 2;  // two slots for the captured step
  if (hipMalloc(&b->partials, b->partials_cap * sizeof(unsigned long long)) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  for (int d = 0; d < 2; ++d)
    if (int rc = Upload(b, d, 0)) return fail(rc);
  if (hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(FVS_HIP_ERROR);
  *out = b;
  return FVS_OK;
}

int fvs_batch_destroy(fvs_batch* b) {
  if (!b) return FVS_OK;
  if (b->h) (void)hipStreamSynchronize(b->h->stream);
  if (b->cap_stream) (void)hipStreamSynchronize(b->cap_stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  void* ps[] = {b->patches, b->pool[0], b->pool[1], b->tiles[0], b->tiles[1],
                b->tiles_grown[0], b->tiles_grown[1], b->partials, b->dts};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete b;
  return FVS_OK;
}

double* fvs_batch_pool(fvs_batch* b, int which) {
  return (b && (which == 0 || which == 1)) ? b->pool[which] : nullptr;
}

int64_t fvs_batch_offset(fvs_batch* b, int64_t patch) {
  if (!b || patch < 0 || patch >= b->count) return -1;
  return b->host_patches[patch].offset;
}

// Device-to-device copy of a whole pool from / to caller memory
// (fvs_batch_pool_size doubles).
int64_t fvs_batch_pool_size(fvs_batch* b) { return b ? b->pool_size : -1; }

int fvs_batch_upload(fvs_batch* b, int which, const double* src) {
  if (!b || !src || (which != 0 && which != 1)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  FVS_HIP_CHECK(hipMemcpyAsync(b->pool[which], src, b->pool_size * sizeof(double),
                               hipMemcpyDeviceToDevice, b->h->stream));
  return FVS_OK;
}

int fvs_batch_download(fvs_batch* b, int which, double* dst) {
  if (!b || !dst || (which != 0 && which != 1)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  FVS_HIP_CHECK(hipMemcpyAsync(dst, b->pool[which], b->pool_size * sizeof(double),
                               hipMemcpyDeviceToDevice, b->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(b->h->stream));
  return FVS_OK;
}

int fvs_batch_set_dt(fvs_batch* b, double dt) {
  if (!b || !(dt >= 0.0)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  FVS_HIP_CHECK(hipMemcpyAsync(b->dts, &dt, sizeof dt, hipMemcpyHostToDevice, b->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(b->h->stream));
  return FVS_OK;
}

int fvs_batch_fill_ghosts(fvs_batch* b, int which) {
  if (!b || (which != 0 && which != 1)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  const unsigned grid = static_cast<unsigned>(std::min<int64_t>(b->count, int64_t(b->cus) * 4));
  hipLaunchKernelGGL(BatchGhostKernel, dim3(grid), dim3(256), 0, b->h->stream, b->patches,
                     static_cast<int>(b->count), b->pool[which]);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_batch_sweep(fvs_batch* b, int dir, int src, double* dt_device) {
  if (!b || (dir != 0 && dir != 1) || (src != 0 && src != 1)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  return Launch(b, b->h->stream, dir, src, dt_device, b->h->desc.precision);
}

int fvs_batch_tune(fvs_batch* b, int dir, int reps, int* shape_out, float* ms_out) {
  if (!b || (dir != 0 && dir != 1) || reps < 1) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  hipStream_t s = b->h->stream;
  hipEvent_t e0, e1;
  FVS_HIP_CHECK(hipEventCreate(&e0));
  FVS_HIP_CHECK(hipEventCreate(&e1));
  int best = 0;
  float best_ms = 1e30f;
  for (int sh = 0; sh < kNumShapes; ++sh) {
    if (int rc = Upload(b, dir, sh)) return rc;
    if (int rc = Launch(b, s, dir, 0, nullptr, b->h->desc.precision)) return rc;  // warm
    FVS_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r)
      if (int rc = Launch(b, s, dir, 0, nullptr, b->h->desc.precision)) return rc;
    FVS_HIP_CHECK(hipEventRecord(e1, s));
    FVS_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    FVS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    ms /= reps;
    if (ms_out) ms_out[sh] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = sh;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (int rc = Upload(b, dir, best)) return rc;
  if (shape_out) *shape_out = best;
  // a captured graph refers to the old tile list
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return FVS_OK;
}

// Captures one split step of the level (ghost fill, X sweep pool 0 -> 1,
// ghost fill, Y sweep pool 1 -> 0, each sweep's CFL minimum into
// dts[1 + dir]) into a hipGraph on the batch's own stream; the sweeps read
// the step's dt from device memory (fvs_batch_set_dt).
int fvs_batch_capture(fvs_batch* b) {
  if (!b) {
    SetError("null batch");
    return FVS_BAD_ARGUMENT;
  }
  if (b->exec) return FVS_OK;
  hipStream_t cs = b->cap_stream;
  FVS_HIP_CHECK(hipStreamSynchronize(b->h->stream));
  FVS_HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  const unsigned ggrid = static_cast<unsigned>(std::min<int64_t>(b->count, int64_t(b->cus) * 4));
  int rc = FVS_OK;
  hipLaunchKernelGGL(BatchGhostKernel, dim3(ggrid), dim3(256), 0, cs, b->patches,
                     static_cast<int>(b->count), b->pool[0]);
  // the two sweeps' minima go to two slots of the partials buffer, reduced
  // by one launch at the end of the step
  const int cap = b->cus * 8;
  rc = batch::LaunchSlotMin(cs, b->partials, 2, cap, nullptr, true);
  if (rc == FVS_OK) rc = LaunchSweepOnly(b, cs, 0, 0, b->partials, b->h->desc.precision);
  if (rc == FVS_OK) {
    hipLaunchKernelGGL(BatchGhostKernel, dim3(ggrid), dim3(256), 0, cs, b->patches,
                       static_cast<int>(b->count), b->pool[1]);
    rc = LaunchSweepOnly(b, cs, 1, 1, b->partials + cap, b->h->desc.precision);
  }
  if (rc == FVS_OK) rc = batch::LaunchSlotMin(cs, b->partials, 2, cap, b->dts + 1, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(cs, &g);
  if (rc != FVS_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc == FVS_OK) SetError("graph capture failed: %s", hipGetErrorString(e));
    return rc != FVS_OK ? rc : FVS_HIP_ERROR;
  }
  b->graph = g;
  FVS_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return FVS_OK;
}

// Replays the captured step `steps` times, ordered after the handle's
// stream and before its later work.
int fvs_batch_replay(fvs_batch* b, int steps) {
  if (!b || steps < 0) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  if (!b->exec)
    if (int rc = fvs_batch_capture(b)) return rc;
  hipEvent_t ev;
  FVS_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  FVS_HIP_CHECK(hipEventRecord(ev, b->h->stream));
  FVS_HIP_CHECK(hipStreamWaitEvent(b->cap_stream, ev, 0));
  for (int i = 0; i < steps; ++i) FVS_HIP_CHECK(hipGraphLaunch(b->exec, b->cap_stream));
  FVS_HIP_CHECK(hipEventRecord(ev, b->cap_stream));
  FVS_HIP_CHECK(hipStreamWaitEvent(b->h->stream, ev, 0));
  FVS_HIP_CHECK(hipEventDestroy(ev));
  return FVS_OK;
}

// Device pointer of the CFL minima of the last captured step's sweeps
// ([0] X, [1] Y).
double* fvs_batch_cfl(fvs_batch* b) { return b ? b->dts + 1 : nullptr; }

int fvs_batch_read_cfl(fvs_batch* b, double* host2) {
  if (!b || !host2) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  FVS_HIP_CHECK(hipStreamSynchronize(b->h->stream));
  FVS_HIP_CHECK(hipMemcpyAsync(host2, b->dts + 1, 2 * sizeof(double), hipMemcpyDeviceToHost,
                               b->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(b->h->stream));
  return FVS_OK;
}

int fvs_batch_set_shape(fvs_batch* b, int dir, int shape) {
  if (!b || (dir != 0 && dir != 1) || shape < 0 || shape >= kNumShapes) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = Upload(b, dir, shape)) return rc;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  return FVS_OK;
}

int fvs_batch_tile_shape(fvs_batch* b, int dir, int* tx, int* ty, int* ntiles) {
  if (!b || (dir != 0 && dir != 1)) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  const Shape sh = dir == 0 ? kShapesX[b->shape[dir]] : kShapesY[b->shape[dir]];
  if (tx) *tx = sh.tx;
  if (ty) *ty = sh.ty;
  if (ntiles) *ntiles = b->ntiles[dir];
  return FVS_OK;
}

} // extern "C"

#endif // FVS_BATCH_FAST_TU
