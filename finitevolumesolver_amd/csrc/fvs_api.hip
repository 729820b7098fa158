// C ABI of libfvs_hip.so: handles, the reference's patch-level operators on
// strided views, and the single-patch level driver built on the fused
// sweeps.
#include "fvs_internal.hpp"
#include "fvs_physics.hpp"
#include "fvs_views.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace fvs {

namespace {
thread_local std::string g_last_error = "no error";
}

void SetError(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int MakeDevView(const fvs_view* v, int min_comp, DevView* out) {
  if (!v || !v->data) {
    SetError("null view or data pointer");
    return FVS_BAD_ARGUMENT;
  }
  if (v->ncomp < min_comp) {
    SetError("view has %d components, need %d", v->ncomp, min_comp);
    return FVS_BAD_ARGUMENT;
  }
  out->data = v->data;
  for (int d = 0; d < 3; ++d) {
    if (v->hi[d] < v->lo[d]) {
      SetError("empty box in dimension %d", d);
      return FVS_BAD_ARGUMENT;
    }
    out->lo[d] = v->lo[d];
    out->ext[d] = v->hi[d] - v->lo[d] + 1;
    out->stride[d] = v->stride[d];
  }
  out->comp_stride = v->comp_stride;
  return FVS_OK;
}

bool ContainsBox(const fvs_view* v, const int64_t lo[3], const int64_t hi[3]) {
  for (int d = 0; d < 3; ++d)
    if (lo[d] < v->lo[d] || hi[d] > v->hi[d]) return false;
  return true;
}

namespace {

constexpr int kBlock = 256;

int CheckDesc(const fvs_method_desc* d) {
  if (!d) {
    SetError("null method description");
    return FVS_BAD_ARGUMENT;
  }
  if (d->dim != 2 && d->dim != 3) {
    SetError("dim must be 2 or 3 (got %d)", d->dim);
    return FVS_BAD_ARGUMENT;
  }
  if (!(d->gamma > 1.0)) {
    SetError("gamma must exceed 1");
    return FVS_BAD_ARGUMENT;
  }
  if (d->base < 0 || d->base > FVS_BASE_HLL_EINFELDT || d->reconstruction < 0 ||
      d->reconstruction > FVS_RECON_MUSCL_HANCOCK_CHARACTERISTIC || d->limiter < 0 ||
      d->limiter > FVS_LIMITER_NO_GRADIENT || d->precision < 0 ||
      d->precision > FVS_FP64_RCP || d->boundary_solver < FVS_BOUNDARY_EXACT ||
      d->boundary_solver > FVS_BOUNDARY_HLLEM) {
    SetError("method enum out of range");
    return FVS_BAD_ARGUMENT;
  }
  if ((d->reconstruction == FVS_RECON_MUSCL_HANCOCK_PRIM ||
       d->reconstruction == FVS_RECON_MUSCL_HANCOCK_CHARACTERISTIC) &&
      (d->base != FVS_BASE_HLLEM_LARROUTUROU || d->precision != FVS_FP64)) {
    SetError("MusclHancockPrim and MusclHancockCharacteristic take the Larrouturou HLLEM base in FP64");
    return FVS_BAD_ARGUMENT;
  }
  return FVS_OK;
}

int CheckDir(const fvs_handle* h, int dir) {
  if (dir < 0 || dir >= h->desc.dim) {
    SetError("direction %d out of range for dim %d", dir, h->desc.dim);
    return FVS_BAD_ARGUMENT;
  }
  return FVS_OK;
}

Gas<double> GasOf(const fvs_method_desc& d) {
  Gas<double> g;
  g.gamma = d.gamma;
  g.gm1inv = 1.0 / (d.gamma - 1.0);
  g.gm1 = d.gamma - 1.0;
  return g;
}

// ---- view kernels ----------------------------------------------------------

template <int D>
__device__ __forceinline__ void LoadComplete(const DevView& v, const int64_t* idx, double* q) {
  for (int c = 0; c < D + 4; ++c) q[c] = *v.at(idx[0], idx[1], idx[2], c);
}

// FluxMethod::ComputeNumericFluxes on one face per thread.  For MUSCL-Hancock
// the stencil is the four cells face-2 .. face+1; rec_[0] is the right
// half-step state of cell face-1 and rec_[1] the left one of cell face.
// BASE: 0 plain HLLEM, 1 HLLEM (Larrouturou), 3 HLL with Einfeldt signals
// (the fvs_base_method values).
template <int D, int DIR, int BASE>
__global__ __launch_bounds__(kBlock) void ViewFluxKernel(DevView fl, DevView st, double dt,
                                                         double dx, Gas<double> g,
                                                         int limiter, int recon, int quirk) {
  constexpr int NC = D + 2, NQ = D + 4;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= fl.count()) return;
  int64_t f[3];
  fl.index(q, f);
  double L[NQ], R[NQ];
  if (recon == FVS_RECON_MUSCL_HANCOCK_PRIM || recon == FVS_RECON_MUSCL_HANCOCK_CHARACTERISTIC) {
    double u[4][NQ];
    for (int s = 0; s < 4; ++s) {
      int64_t c[3] = {f[0], f[1], f[2]};
      c[DIR] += s - 2;
      LoadComplete<D>(st, c, u[s]);
    }
    if (recon == FVS_RECON_MUSCL_HANCOCK_PRIM) {
      MusclHancockPrimStates<D, DIR>(u, DivRn(dt, dx), g, L, R);
    } else {
      const double lambda = DivRn(dt, dx);
      CharacteristicFaceState<D, DIR>(u, lambda, 0, g, L);
      CharacteristicFaceState<D, DIR>(u + 1, lambda, 1, g, R);
    }
  } else if (recon == FVS_RECON_MUSCL_HANCOCK) {
    double u[4][NQ];
    for (int s = 0; s < 4; ++s) {
      int64_t c[3] = {f[0], f[1], f[2]};
      c[DIR] += s - 2;
      LoadComplete<D>(st, c, u[s]);
    }
    const double lh = DivRn(0.5 * dt, dx);
    double dummy[NQ];
    switch (limiter) {
    case kMinMod:
      ReconstructCell<D, DIR, kMinMod>(u[0], u[1], u[2], lh, g, dummy, L);
      ReconstructCell<D, DIR, kMinMod>(u[1], u[2], u[3], lh, g, R, dummy);
      break;
    case kVanLeer:
      ReconstructCell<D, DIR, kVanLeer>(u[0], u[1], u[2], lh, g, dummy, L);
      ReconstructCell<D, DIR, kVanLeer>(u[1], u[2], u[3], lh, g, R, dummy);
      break;
    case kNoLimiter:
      ReconstructCell<D, DIR, kNoLimiter>(u[0], u[1], u[2], lh, g, dummy, L);
      ReconstructCell<D, DIR, kNoLimiter>(u[1], u[2], u[3], lh, g, R, dummy);
      break;
    default:
      ReconstructCell<D, DIR, kNoGradient>(u[0], u[1], u[2], lh, g, dummy, L);
      ReconstructCell<D, DIR, kNoGradient>(u[1], u[2], u[3], lh, g, R, dummy);
      break;
    }
  } else {
    int64_t c[3] = {f[0], f[1], f[2]};
    c[DIR] -= 1;
    LoadComplete<D>(st, c, L);
    LoadComplete<D>(st, f, R);
  }
  double out[NC];
  constexpr bool LARR = BASE == FVS_BASE_HLLEM_LARROUTUROU;
  if constexpr (BASE == FVS_BASE_HLL_EINFELDT) {
    HllFlux<D, DIR>(out, L, R, g);
  } else if constexpr (D == 3) {
    if (quirk) HllemFlux<D, DIR, LARR, true>(out, L, R, g);
    else HllemFlux<D, DIR, LARR, false>(out, L, R, g);
  } else {
    HllemFlux<D, DIR, LARR, false>(out, L, R, g);
  }
  for (int c = 0; c < NC; ++c) *fl.at(f[0], f[1], f[2], c) = out[c];
}

__device__ __forceinline__ unsigned long long WaveMin(unsigned long long k) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(k, off, 64);
    k = o < k ? o : k;
  }
  return k;
}

__device__ __forceinline__ void BlockAtomicMin(unsigned long long key,
                                               unsigned long long* dst) {
  __shared__ unsigned long long s[kBlock / 64];
  key = WaveMin(key);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long k = s[0];
    for (int w = 1; w < kBlock / 64; ++w) k = s[w] < k ? s[w] : k;
    atomicMin(dst, k);
  }
}

// FluxMethod::ComputeStableDt: `faces` enumerates the first stencil cell of
// every stencil inside the states box; the base method sees the two middle
// cells (MUSCL) or both cells (first order).
template <int D, int DIR, bool HLL>
__global__ __launch_bounds__(kBlock) void ViewStableDtKernel(DevView st, DevView faces,
                                                             int offset, double dx,
                                                             Gas<double> g,
                                                             unsigned long long* key) {
  constexpr int NQ = D + 4;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long k = 0x7ff0000000000000ull;
  if (q < faces.count()) {
    int64_t c[3];
    faces.index(q, c);
    c[DIR] += offset;
    double L[NQ], R[NQ];
    LoadComplete<D>(st, c, L);
    c[DIR] += 1;
    LoadComplete<D>(st, c, R);
    k = DtKey(HLL ? HllStableDt<D, DIR>(L, R, dx, g) : HllemStableDt<D, DIR>(L, R, dx, g));
  }
  BlockAtomicMin(k, key);
}

template <int DIR>
__global__ __launch_bounds__(kBlock) void ViewUpdateKernel(DevView nx, DevView pv, DevView fl,
                                                           double dt_dx, int ncomp) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= nx.count()) return;
  int64_t c[3];
  nx.index(q, c);
  int64_t r[3] = {c[0], c[1], c[2]};
  r[DIR] += 1;
  for (int k = 0; k < ncomp; ++k) {
    const double U = *pv.at(c[0], c[1], c[2], k);
    const double fl_ = *fl.at(c[0], c[1], c[2], k);
    const double fr_ = *fl.at(r[0], r[1], r[2], k);
    *nx.at(c[0], c[1], c[2], k) = U + dt_dx * (fl_ - fr_);
  }
}

template <int D>
__global__ __launch_bounds__(kBlock) void ViewCompleteKernel(DevView cv, DevView uv,
                                                             Gas<double> g) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= cv.count()) return;
  int64_t c[3];
  cv.index(q, c);
  double s[D + 4];
  for (int k = 0; k < D + 2; ++k) s[k] = *uv.at(c[0], c[1], c[2], k);
  CompleteFromCons<D>(s, g);
  for (int k = 0; k < D + 4; ++k) *cv.at(c[0], c[1], c[2], k) = s[k];
}

__global__ __launch_bounds__(kBlock) void ViewFillKernel(DevView v, int ncomp, int dir,
                                                         int64_t dlo, int64_t dhi) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (q >= v.count()) return;
  int64_t c[3];
  v.index(q, c);
  int64_t s[3] = {c[0], c[1], c[2]};
  if (c[dir] < dlo) {
    s[dir] = dlo;
  } else if (c[dir] > dhi) {
    s[dir] = dhi;
  } else {
    return;
  }
  for (int k = 0; k < ncomp; ++k) *v.at(c[0], c[1], c[2], k) = *v.at(s[0], s[1], s[2], k);
}

unsigned Blocks(int64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

int64_t Count(const DevView& v) { return v.ext[0] * v.ext[1] * v.ext[2]; }

} // namespace
} // namespace fvs

using namespace fvs;

// Partial minima of fvs_patch_stable_dt_noalloc: a fixed device array that
// every call on a device shares (2 MiB per device, held until the process
// ends).  Every workgroup of the CFL pass writes its keys before the
// finalize kernel reads them, so a call never reads an earlier call's
// contents and nothing is cleared.  Host code never passes the array's name
// to a kernel: it uses the device address hipGetSymbolAddress returns for
// the device, looked up once per device when the first handle on it is
// created.
static constexpr int64_t kNoAllocKeys = 262144;
static __device__ __align__(256) unsigned long long scratch[kNoAllocKeys];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current device.
int NoAllocScratchInit(int device) {
  if (device < 0 || device >= kMaxDevices) {
    SetError("device %d outside the %d devices the scratch table holds", device, kMaxDevices);
    return FVS_BAD_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return FVS_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    SetError("hipGetSymbolAddress of the scratch array failed on device %d: %s", device,
             hipGetErrorString(e));
    return FVS_HIP_ERROR;
  }
  g_scratch_ptr[device] = static_cast<unsigned long long*>(ptr);
  return FVS_OK;
}

unsigned long long* NoAllocScratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}
} // namespace

static int EnsurePartials(fvs_handle* h, int64_t n) {
  if (n <= h->partials_cap) return FVS_OK;
  if (h->partials) {
    FVS_HIP_CHECK(hipStreamSynchronize(h->stream));
    FVS_HIP_CHECK(hipFree(h->partials));
    h->partials = nullptr;
    h->partials_cap = 0;
  }
  FVS_HIP_CHECK(hipMalloc(&h->partials, n * sizeof(unsigned long long)));
  h->partials_cap = n;
  return FVS_OK;
}

extern "C" {

const char* fvs_last_error(void) { return g_last_error.c_str(); }

const char* fvs_version(void) { return "fvs-mi355x 0.1 (gfx950)"; }

int fvs_create(const fvs_method_desc* desc, void* hip_stream, fvs_handle** out) {
  if (!out) {
    SetError("null output handle pointer");
    return FVS_BAD_ARGUMENT;
  }
  *out = nullptr;
  if (int rc = CheckDesc(desc)) return rc;
  fvs_handle* h = new (std::nothrow) fvs_handle();
  if (!h) {
    SetError("out of host memory");
    return FVS_BAD_ARGUMENT;
  }
  h->desc = *desc;
  if (h->desc.max_newton_iterations <= 0) h->desc.max_newton_iterations = 5000;
  if (!(h->desc.newton_tolerance > 0.0)) h->desc.newton_tolerance = 1e-7;
  h->stream = static_cast<hipStream_t>(hip_stream);
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    SetError("hipGetDevice failed: no usable HIP device");
    return FVS_HIP_ERROR;
  }
  void* mem = nullptr;
  if (hipMalloc(&mem, 512) != hipSuccess) {
    delete h;
    SetError("hipMalloc of the handle scratch failed");
    return FVS_HIP_ERROR;
  }
  h->partials = nullptr;
  h->partials_cap = 0;
  h->mlp = nullptr;
  h->scratch_dt = static_cast<double*>(mem);
  h->scratch_keys = reinterpret_cast<unsigned long long*>(static_cast<char*>(mem) + 64);
  h->scratch_counter = reinterpret_cast<unsigned int*>(static_cast<char*>(mem) + 192);
  h->newton_stats = reinterpret_cast<unsigned long long*>(static_cast<char*>(mem) + 256);
  if (hipMemsetAsync(h->newton_stats, 0, 5 * sizeof(unsigned long long), h->stream) != hipSuccess) {
    (void)hipFree(mem);
    delete h;
    SetError("hipMemsetAsync of the Newton counters failed");
    return FVS_HIP_ERROR;
  }
  if (int rc = LaunchResetKeys(h->stream, h->scratch_keys, 8, h->scratch_counter, 4)) {
    (void)hipFree(mem);
    delete h;
    return rc;
  }
  if (int rc = NoAllocScratchInit(h->device)) {
    (void)hipFree(mem);
    delete h;
    return rc;
  }
  *out = h;
  return FVS_OK;
}

int fvs_destroy(fvs_handle* h) {
  if (!h) return FVS_OK;
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(h->scratch_dt);
  if (h->partials) (void)hipFree(h->partials);
  delete h;
  return FVS_OK;
}

int fvs_stable_dt(fvs_handle* h, const fvs_view* states, double dx, int dir,
                  double* dt_device) {
  if (!h || !dt_device) {
    SetError("null handle or output");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = CheckDir(h, dir)) return rc;
  const int D = h->desc.dim;
  if (h->desc.base == FVS_BASE_GODUNOV_EXACT) {
    if (!states || !states->data || states->ncomp < D + 4) {
      SetError("states view needs %d components", D + 4);
      return FVS_BAD_ARGUMENT;
    }
    const int64_t nb = GodunovViewStableDtBlocks(h->desc, states, dir);
    if (nb <= 0) {
      SetError("states box too small for the stencil");
      return FVS_BAD_ARGUMENT;
    }
    if (int rc = EnsurePartials(h, nb)) return rc;
    if (int rc = LaunchGodunovViewStableDt(h->desc, h->stream, states, dx, dir, h->partials))
      return rc;
    return LaunchFinalizeCheck(h->stream, h->partials, nb, dt_device, nullptr);
  }
  DevView st;
  if (int rc = MakeDevView(states, D + 4, &st)) return rc;
  const int S = h->desc.reconstruction != FVS_RECON_NONE ? 4 : 2;
  DevView faces = st;
  faces.ext[dir] -= (S - 1);
  if (faces.ext[dir] <= 0) {
    SetError("states box too small for the stencil");
    return FVS_BAD_ARGUMENT;
  }
  const int offset = S == 4 ? 1 : 0;
  const Gas<double> g = GasOf(h->desc);
  const unsigned nb = Blocks(Count(faces));
  hipStream_t s = h->stream;
  unsigned long long* key = h->scratch_keys;
  const bool hll = h->desc.base == FVS_BASE_HLL_EINFELDT;
#define FVS_VD(DD, DIR_)                                                                 \
  hipLaunchKernelGGL((ViewStableDtKernel<DD, DIR_, false>), dim3(nb), dim3(kBlock), 0, s, st, \
                     faces, offset, dx, g, key)
#define FVS_VDH(DD, DIR_)                                                               \
  hipLaunchKernelGGL((ViewStableDtKernel<DD, DIR_, true>), dim3(nb), dim3(kBlock), 0, s, st, \
                     faces, offset, dx, g, key)
  if (D == 2) {
    if (dir == 0) { if (hll) FVS_VDH(2, 0); else FVS_VD(2, 0); }
    else { if (hll) FVS_VDH(2, 1); else FVS_VD(2, 1); }
  } else {
    if (dir == 0) { if (hll) FVS_VDH(3, 0); else FVS_VD(3, 0); }
    else if (dir == 1) { if (hll) FVS_VDH(3, 1); else FVS_VD(3, 1); }
    else { if (hll) FVS_VDH(3, 2); else FVS_VD(3, 2); }
  }
#undef FVS_VD
#undef FVS_VDH
  FVS_LAUNCH_CHECK();
  return LaunchKeysToDt(s, key, 1, dt_device);
}

int fvs_numeric_fluxes(fvs_handle* h, const fvs_view* fluxes, const fvs_view* states,
                       double dt, double dx, int dir) {
  if (!h) {
    SetError("null handle");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = CheckDir(h, dir)) return rc;
  const int D = h->desc.dim;
  DevView fl, st;
  if (int rc = MakeDevView(fluxes, D + 2, &fl)) return rc;
  if (int rc = MakeDevView(states, D + 4, &st)) return rc;
  const int recon = h->desc.reconstruction;
  const int w = recon != FVS_RECON_NONE ? 2 : 1;
  int64_t need_lo[3], need_hi[3];
  for (int d = 0; d < 3; ++d) {
    need_lo[d] = fluxes->lo[d];
    need_hi[d] = fluxes->hi[d];
  }
  need_lo[dir] -= w;
  need_hi[dir] += w - 1;
  if (!ContainsBox(states, need_lo, need_hi)) {
    SetError("states box does not cover the flux stencils");
    return FVS_BAD_ARGUMENT;
  }
  if (h->desc.base == FVS_BASE_GODUNOV_EXACT) {
    return LaunchGodunovViewFlux(h->desc, h->stream, fluxes, states, dt, dx, dir,
                                 h->newton_stats + 2);
  }
  const Gas<double> g = GasOf(h->desc);
  const int base = h->desc.base;
  const unsigned nb = Blocks(Count(fl));
  hipStream_t s = h->stream;
  const int lim = h->desc.limiter, q = h->desc.quirks;
#define FVS_VF1(DD, DIR_, B)                                                             \
  hipLaunchKernelGGL((ViewFluxKernel<DD, DIR_, B>), dim3(nb), dim3(kBlock), 0, s, fl, st, \
                     dt, dx, g, lim, recon, q)
#define FVS_VF(DD, DIR_)                                                    \
  do {                                                                      \
    if (base == FVS_BASE_HLLEM_LARROUTUROU)                                 \
      FVS_VF1(DD, DIR_, FVS_BASE_HLLEM_LARROUTUROU);                        \
    else if (base == FVS_BASE_HLLEM)                                        \
      FVS_VF1(DD, DIR_, FVS_BASE_HLLEM);                                    \
    else                                                                    \
      FVS_VF1(DD, DIR_, FVS_BASE_HLL_EINFELDT);                             \
  } while (0)
  if (D == 2) {
    if (dir == 0) FVS_VF(2, 0);
    else FVS_VF(2, 1);
  } else {
    if (dir == 0) FVS_VF(3, 0);
    else if (dir == 1) FVS_VF(3, 1);
    else FVS_VF(3, 2);
  }
#undef FVS_VF
#undef FVS_VF1
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_update_conservatively(fvs_handle* h, const fvs_view* next, const fvs_view* prev,
                              const fvs_view* fluxes, double dt, double dx, int dir) {
  if (!h) {
    SetError("null handle");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = CheckDir(h, dir)) return rc;
  DevView nv, pv, fv;
  if (int rc = MakeDevView(next, 1, &nv)) return rc;
  if (int rc = MakeDevView(prev, next->ncomp, &pv)) return rc;
  if (int rc = MakeDevView(fluxes, next->ncomp, &fv)) return rc;
  int64_t flo[3], fhi[3];
  for (int d = 0; d < 3; ++d) {
    flo[d] = next->lo[d];
    fhi[d] = next->hi[d];
  }
  fhi[dir] += 1;
  if (!ContainsBox(fluxes, flo, fhi) || !ContainsBox(prev, next->lo, next->hi)) {
    SetError("flux or prev box does not cover the update box");
    return FVS_BAD_ARGUMENT;
  }
  const double dt_dx = dt / dx;
  const unsigned nb = Blocks(Count(nv));
  hipStream_t s = h->stream;
  if (dir == 0) hipLaunchKernelGGL(ViewUpdateKernel<0>, dim3(nb), dim3(kBlock), 0, s, nv, pv, fv, dt_dx, next->ncomp);
  else if (dir == 1) hipLaunchKernelGGL(ViewUpdateKernel<1>, dim3(nb), dim3(kBlock), 0, s, nv, pv, fv, dt_dx, next->ncomp);
  else hipLaunchKernelGGL(ViewUpdateKernel<2>, dim3(nb), dim3(kBlock), 0, s, nv, pv, fv, dt_dx, next->ncomp);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_complete_from_cons(fvs_handle* h, const fvs_view* complete, const fvs_view* cons) {
  if (!h) {
    SetError("null handle");
    return FVS_BAD_ARGUMENT;
  }
  const int D = h->desc.dim;
  DevView cv, uv;
  if (int rc = MakeDevView(complete, D + 4, &cv)) return rc;
  if (int rc = MakeDevView(cons, D + 2, &uv)) return rc;
  if (!ContainsBox(cons, complete->lo, complete->hi)) {
    SetError("cons box does not cover the complete box");
    return FVS_BAD_ARGUMENT;
  }
  const Gas<double> g = GasOf(h->desc);
  const unsigned nb = Blocks(Count(cv));
  if (D == 2) hipLaunchKernelGGL(ViewCompleteKernel<2>, dim3(nb), dim3(kBlock), 0, h->stream, cv, uv, g);
  else hipLaunchKernelGGL(ViewCompleteKernel<3>, dim3(nb), dim3(kBlock), 0, h->stream, cv, uv, g);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_transmissive_fill(fvs_handle* h, const fvs_view* states, const int64_t domain_lo[3],
                          const int64_t domain_hi[3], int dir) {
  if (!h || !domain_lo || !domain_hi) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = CheckDir(h, dir)) return rc;
  DevView v;
  if (int rc = MakeDevView(states, 1, &v)) return rc;
  if (domain_lo[dir] < states->lo[dir] || domain_hi[dir] > states->hi[dir] ||
      domain_hi[dir] < domain_lo[dir]) {
    SetError("domain boundary cells lie outside the view");
    return FVS_BAD_ARGUMENT;
  }
  hipLaunchKernelGGL(ViewFillKernel, dim3(Blocks(Count(v))), dim3(kBlock), 0, h->stream, v,
                     states->ncomp, dir, domain_lo[dir], domain_hi[dir]);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

// ---- fused single-patch operators ------------------------------------------


static int FillSweepArgs(fvs_handle* h, const double* const* in, double* const* out,
                         const int64_t n[3], double dx, int dir, SweepArgs* a) {
  const int D = h->desc.dim;
  std::memset(a, 0, sizeof *a);
  for (int c = 0; c < D + 2; ++c) {
    if (!in || !in[c] || (out && !out[c])) {
      SetError("null component pointer %d", c);
      return FVS_BAD_ARGUMENT;
    }
    a->in[c] = in[c];
    if (out) a->out[c] = out[c];
  }
  for (int d = 0; d < 3; ++d) a->n[d] = (d < D) ? n[d] : 1;
  for (int d = 0; d < D; ++d) {
    if (n[d] < 1) {
      SetError("patch extent %d must be positive", d);
      return FVS_BAD_ARGUMENT;
    }
  }
  a->dx = dx;
  a->gamma = h->desc.gamma;
  a->gm1inv = 1.0 / (h->desc.gamma - 1.0);
  a->limiter = h->desc.limiter;
  a->muscl = h->desc.reconstruction == FVS_RECON_MUSCL_HANCOCK;
  a->quirk3d = h->desc.quirks;
  a->seg = PencilSeg(h->desc.dim, dir, a->n);
  return FVS_OK;
}

int fvs_split_sweep(fvs_handle* h, const double* const* cons_in, double* const* cons_out,
                    const int64_t n[3], double dt, double dx, int dir,
                    double* stable_dt_device) {
  if (!h || !n) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = CheckDir(h, dir)) return rc;
  if (h->desc.reconstruction >= FVS_RECON_MUSCL_HANCOCK_PRIM) {
    SetError("MusclHancockPrim runs through the patch operators only");
    return FVS_BAD_ARGUMENT;
  }
  SweepArgs a;
  if (int rc = FillSweepArgs(h, cons_in, cons_out, n, dx, dir, &a)) return rc;
  for (int c = 0; c < h->desc.dim + 2; ++c) {
    if (cons_in[c] == cons_out[c]) {
      SetError("fused sweep needs distinct input and output arrays");
      return FVS_BAD_ARGUMENT;
    }
  }
  if (h->desc.base == FVS_BASE_GODUNOV_EXACT) {
    return fvs_godunov_sweep(h, cons_in, cons_out, n, dt, dx, dir, nullptr, stable_dt_device);
  }
  a.dt_val = dt;
  a.check = stable_dt_device != nullptr;
  const int64_t nb = SweepBlocks(h->desc.dim, dir, a.n, a.seg);
  if (a.check) {
    if (int rc = EnsurePartials(h, nb)) return rc;
    a.partials = h->partials;
  }
  if (int rc = LaunchSweep(h->desc, h->stream, dir, a)) return rc;
  if (stable_dt_device) return LaunchFinalizeCheck(h->stream, h->partials, nb, stable_dt_device, nullptr);
  return FVS_OK;
}

// Step-start CFL of a ghost-free patch with the partial minima in `scratch`
// (capacity `cap` keys) when it is given, else in the handle's buffer, grown
// as needed.
static int PatchStableDt(fvs_handle* h, const double* const* cons, const int64_t n[3],
                         const double dx[3], double* dt_dir_device,
                         unsigned long long* scratch_keys, int64_t cap) {
  if (!h || !cons || !n || !dx || !dt_dir_device) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  const int D = h->desc.dim;
  for (int c = 0; c < D + 2; ++c) {
    if (!cons[c]) {
      SetError("null component pointer %d", c);
      return FVS_BAD_ARGUMENT;
    }
  }
  int64_t nn[3] = {n[0], n[1], D == 3 ? n[2] : 1};
  for (int d = 0; d < D; ++d) {
    if (nn[d] < 1) {
      SetError("patch extent %d must be positive", d);
      return FVS_BAD_ARGUMENT;
    }
  }
  const bool godunov = h->desc.base == FVS_BASE_GODUNOV_EXACT;
  if (!godunov && (h->desc.base == FVS_BASE_HLL_EINFELDT ||
                   h->desc.reconstruction >= FVS_RECON_MUSCL_HANCOCK_PRIM)) {
    SetError("fvs_patch_stable_dt has no HLL or MusclHancockPrim kernel (use fvs_stable_dt on views)");
    return FVS_BAD_ARGUMENT;
  }
  const int64_t nb = godunov ? GodunovCflBlocks(D, nn) : CflBlocks(D, nn);
  const int64_t keys = (godunov ? 5 : 1) * D * nb;
  unsigned long long* partials = scratch_keys;
  if (partials) {
    if (keys > cap) {
      SetError("patch needs %lld reduction keys, the allocation-free entry point holds %lld",
               static_cast<long long>(keys), static_cast<long long>(cap));
      return FVS_BAD_ARGUMENT;
    }
  } else {
    if (int rc = EnsurePartials(h, keys)) return rc;
    partials = h->partials;
  }
  if (godunov) {
    if (int rc = LaunchGodunovCfl(h->desc, h->stream, cons, nn, dx, nullptr, partials, nullptr))
      return rc;
    for (int d = 0; d < D; ++d) {
      if (int rc = LaunchGodunovFinalize(h->stream, partials + 5 * d * nb, nb,
                                         dt_dir_device + d, h->newton_stats))
        return rc;
    }
    return FVS_OK;
  }
  return LaunchPatchCfl(h->desc, h->stream, cons, nn, dx, partials, dt_dir_device, nullptr,
                        nullptr, 0, 0);
}

int fvs_patch_stable_dt(fvs_handle* h, const double* const* cons, const int64_t n[3],
                        const double dx[3], double* dt_dir_device) {
  return PatchStableDt(h, cons, n, dx, dt_dir_device, nullptr, 0);
}

int fvs_patch_stable_dt_noalloc(fvs_handle* h, const double* const* cons, const int64_t n[3],
                                const double dx[3], double* dt_dir_device) {
  if (!h) {
    SetError("null handle");
    return FVS_BAD_ARGUMENT;
  }
  unsigned long long* ptr = NoAllocScratch(h->device);
  if (!ptr) {
    SetError("no scratch address for device %d (handle not created on it)", h->device);
    return FVS_BAD_ARGUMENT;
  }
  return PatchStableDt(h, cons, n, dx, dt_dir_device, ptr, kNoAllocKeys);
}

}  // extern "C"

namespace {

// Streams ncomp arrays of n doubles through the chip, 16 bytes per lane and
// array per access, grid-stride: the traffic of a fused sweep (every
// component read and written once) without its arithmetic.  Loads and stores
// carry the nontemporal hint, the fastest of the copy forms measured on
// MI355X (plain accesses and deeper unrolling move about 5% less).
typedef double StreamD2 __attribute__((ext_vector_type(2)));
struct StreamArgs {
  const double* src[5];
  double* dst[5];
  int ncomp;
  int64_t n2;  // double2 elements per array
};

__global__ __launch_bounds__(256) void StreamCopyKernel(StreamArgs a) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n2;
       i += stride) {
    StreamD2 v[5];
#pragma unroll
    for (int c = 0; c < 5; ++c)
      if (c < a.ncomp)
        v[c] = __builtin_nontemporal_load(reinterpret_cast<const StreamD2*>(a.src[c]) + i);
#pragma unroll
    for (int c = 0; c < 5; ++c)
      if (c < a.ncomp)
        __builtin_nontemporal_store(v[c], reinterpret_cast<StreamD2*>(a.dst[c]) + i);
  }
}

}  // namespace

extern "C" {

int fvs_stream_copy(fvs_handle* h, const double* const* src, double* const* dst, int ncomp,
                    int64_t n) {
  if (!h || !src || !dst || ncomp < 1 || ncomp > 5 || n < 0 || n % 2 != 0) {
    SetError("bad argument (1..5 arrays of an even number of doubles)");
    return FVS_BAD_ARGUMENT;
  }
  StreamArgs a{};
  for (int c = 0; c < ncomp; ++c) {
    if (!src[c] || !dst[c] || (reinterpret_cast<uintptr_t>(src[c]) | reinterpret_cast<uintptr_t>(dst[c])) % 16) {
      SetError("array %d is null or not 16-byte aligned", c);
      return FVS_BAD_ARGUMENT;
    }
    a.src[c] = src[c];
    a.dst[c] = dst[c];
  }
  a.ncomp = ncomp;
  a.n2 = n / 2;
  if (a.n2 == 0) return FVS_OK;
  // eight waves per SIMD
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess ||
      cus <= 0)
    cus = 256;
  const int64_t blocks = std::min<int64_t>((a.n2 + 255) / 256, static_cast<int64_t>(cus) * 8);
  hipLaunchKernelGGL(StreamCopyKernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                     h->stream, a);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_godunov_sweep(fvs_handle* h, const double* const* cons_in, double* const* cons_out,
                      const int64_t n[3], double dt, double dx, int dir,
                      const double* guess_device, double* stable_dt_device) {
  if (!h || !n || !cons_in) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = CheckDir(h, dir)) return rc;
  if (h->desc.base != FVS_BASE_GODUNOV_EXACT || h->desc.reconstruction != FVS_RECON_NONE) {
    SetError("fvs_godunov_sweep needs a first-order exact-Godunov handle");
    return FVS_BAD_ARGUMENT;
  }
  const int D = h->desc.dim;
  int64_t nn[3] = {n[0], n[1], D == 3 ? n[2] : 1};
  for (int c = 0; c < D + 2; ++c) {
    if (!cons_in[c] || (cons_out && !cons_out[c]) || (cons_out && cons_out[c] == cons_in[c])) {
      SetError("bad component pointer %d (null, or output aliasing input)", c);
      return FVS_BAD_ARGUMENT;
    }
  }
  for (int d = 0; d < D; ++d) {
    if (nn[d] < 1) {
      SetError("patch extent %d must be positive", d);
      return FVS_BAD_ARGUMENT;
    }
  }
  const int64_t nb = GodunovSweepBlocks(D, dir, nn, PencilSeg(D, dir, nn));
  if (int rc = EnsurePartials(h, 5 * nb)) return rc;
  if (int rc = LaunchGodunovSweep(h->desc, h->stream, dir, cons_in, cons_out, nn, dx, nullptr, dt,
                                  guess_device, h->partials, nullptr, cons_out != nullptr))
    return rc;
  return LaunchGodunovFinalize(h->stream, h->partials, nb, stable_dt_device, h->newton_stats);
}

int fvs_exact_middle_state(fvs_handle* h, int dir, const double* left, const double* right,
                           int64_t n, const double* guess, double* pM, double* uM,
                           int* iterations, int* converged) {
  if (!h || !left || !right || n <= 0) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = CheckDir(h, dir)) return rc;
  return LaunchMiddleState(h->desc.dim, h->stream, left, right, n, dir, h->desc.gamma,
                           h->desc.newton_tolerance, h->desc.max_newton_iterations, guess, pM, uM,
                           iterations, converged);
}

int fvs_set_mlp(fvs_handle* h, fvs_mlp* mlp) {
  if (!h) {
    SetError("null handle");
    return FVS_BAD_ARGUMENT;
  }
  if (mlp && (h->desc.base != FVS_BASE_GODUNOV_EXACT || h->desc.reconstruction != FVS_RECON_NONE)) {
    SetError("a learned guess needs a first-order exact-Godunov handle");
    return FVS_BAD_ARGUMENT;
  }
  h->mlp = mlp;
  h->desc.newton_guess = mlp ? FVS_GUESS_LEARNED : FVS_GUESS_MEAN_PRESSURE;
  return FVS_OK;
}

int fvs_newton_stats(fvs_handle* h, uint64_t stats[5], int reset) {
  if (!h || !stats) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  unsigned long long tmp[5];
  FVS_HIP_CHECK(hipMemcpyAsync(tmp, h->newton_stats, sizeof tmp, hipMemcpyDeviceToHost, h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(h->stream));
  for (int i = 0; i < 5; ++i) stats[i] = tmp[i];
  if (reset) FVS_HIP_CHECK(hipMemsetAsync(h->newton_stats, 0, sizeof tmp, h->stream));
  if (tmp[2] > 0) {
    SetError("%llu Riemann problems hit the Newton iteration cap", tmp[2]);
    return FVS_NEWTON_NOT_CONVERGED;
  }
  return FVS_OK;
}

} // extern "C"

// ---- level driver ----------------------------------------------------------

struct fvs_level {
  fvs_handle* h;
  fvs_patch_desc patch;
  fvs_composed_level* composed = nullptr;  // methods without a fused sweep
  int64_t n[3];
  int64_t ncell;
  double* block;          // 3 state buffers
  double* buf[3][5];
  Ctl* ctl;               // device control block
  unsigned long long* partials;  // per-workgroup minima
  int64_t cfl_redos = 0;  // steps repeated because a speed floor was wrong
  double* pm = nullptr;   // exact Godunov: the step-start CFL pass's star
                          // pressures of the first sweep's faces
  int cur;                // buffer holding the current state
  // Kernel timing (off by default): HIP events recorded on the level's stream
  // around every launch, reduced per kernel class when read.
  int timing = 0;
  struct Mark {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> spare;
  double t_ms[FVS_LEVEL_TIMERS] = {};
  int64_t t_n[FVS_LEVEL_TIMERS] = {};
  // Captured steps (HIP graphs), one per (start buffer, cycle parity, network).
  int use_graphs = 1;
  hipStream_t cap = nullptr;  // capture stream (default priority)
  // A graph holds nsteps consecutive steps (1, or kChainSteps for runs of
  // steps: one graph launch instead of one per step); curs[i] is the buffer
  // step i starts from, curs[nsteps] the buffer holding the result.
  struct StepGraph {
    int cur, parity, nsteps;
    std::vector<int> curs;
    const fvs_mlp* net;
    hipGraphExec_t exec;
  };
  std::vector<StepGraph> graphs;
};

// Steps per chained graph (the step sequence depends only on the start
// buffer and the cycle parity, so a chain is captured once per pair).
constexpr int kChainSteps = 8;

namespace {

hipEvent_t TakeEvent(fvs_level* L) {
  hipEvent_t e = nullptr;
  if (!L->spare.empty()) {
    e = L->spare.back();
    L->spare.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  return e;
}

// Runs launch() and, when timing is on, brackets it with two events.
template <class F> int Timed(fvs_level* L, hipStream_t s, int cls, F&& launch) {
  if (!L->timing) return launch();
  hipEvent_t a = TakeEvent(L), b = TakeEvent(L);
  if (!a || !b) {
    SetError("hipEventCreate failed");
    return FVS_HIP_ERROR;
  }
  FVS_HIP_CHECK(hipEventRecord(a, s));
  const int rc = launch();
  FVS_HIP_CHECK(hipEventRecord(b, s));
  L->marks.push_back({cls, a, b});
  return rc;
}

int CollectMarks(fvs_level* L) {
  if (L->marks.empty()) return FVS_OK;
  FVS_HIP_CHECK(hipStreamSynchronize(L->h->stream));
  for (const auto& m : L->marks) {
    float ms = 0.0f;
    FVS_HIP_CHECK(hipEventElapsedTime(&ms, m.a, m.b));
    L->t_ms[m.cls] += ms;
    L->t_n[m.cls] += 1;
    L->spare.push_back(m.a);
    L->spare.push_back(m.b);
  }
  L->marks.clear();
  return FVS_OK;
}

} // namespace

extern "C" {

int fvs_level_create(fvs_handle* h, const fvs_patch_desc* patch, fvs_level** out) {
  if (!h || !patch || !out) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  *out = nullptr;
  const int D = h->desc.dim;
  const bool godunov = h->desc.base == FVS_BASE_GODUNOV_EXACT;
  int64_t n[3] = {patch->n[0], patch->n[1], D == 3 ? patch->n[2] : 1};
  for (int d = 0; d < D; ++d) {
    if (n[d] < 1 || !(patch->dx[d] > 0)) {
      SetError("bad patch extent or cell width in dimension %d", d);
      return FVS_BAD_ARGUMENT;
    }
  }
  if (patch->splitting != FVS_SPLIT_GODUNOV && patch->splitting != FVS_SPLIT_STRANG) {
    SetError("unknown splitting %d", patch->splitting);
    return FVS_BAD_ARGUMENT;
  }
  fvs_level* L = new (std::nothrow) fvs_level();
  if (!L) {
    SetError("out of host memory");
    return FVS_BAD_ARGUMENT;
  }
  L->h = h;
  L->patch = *patch;
  if (ComposedLevelNeeded(h->desc)) {
    // MUSCL-Hancock over exact Godunov, HLL, MusclHancockPrim and
    // MusclHancockCharacteristic: the level composed from the patch operators
    for (int d = 0; d < 3; ++d) L->n[d] = n[d];
    L->ncell = n[0] * n[1] * n[2];
    if (int rc = ComposedCreate(h, patch, &L->composed)) {
      delete L;
      return rc;
    }
    *out = L;
    return FVS_OK;
  }
  for (int d = 0; d < 3; ++d) L->n[d] = n[d];
  // The launch geometry queries the device's CU count once; do it here, not
  // inside a step capture.
  (void)SweepBlocks(D, 0, n, 32);
  L->ncell = n[0] * n[1] * n[2];
  const size_t bytes = static_cast<size_t>(L->ncell) * (D + 2) * 3 * sizeof(double);
  if (hipMalloc(&L->block, bytes) != hipSuccess) {
    delete L;
    SetError("hipMalloc of %zu bytes for the level state failed", bytes);
    return FVS_HIP_ERROR;
  }
  if (hipMalloc(&L->ctl, sizeof(Ctl)) != hipSuccess) {
    (void)hipFree(L->block);
    delete L;
    SetError("hipMalloc of the control block failed");
    return FVS_HIP_ERROR;
  }
  int64_t np = D * CflBlocks(D, n);
  int64_t sweep_blocks = 0;
  for (int d = 0; d < D; ++d)
    sweep_blocks = std::max<int64_t>(sweep_blocks, SweepBlocks(D, d, n, PencilSeg(D, d, n)));
  // the checked sweeps of a step keep their partial minima side by side
  // (at most 2D - 1 of them, Strang)
  np = std::max<int64_t>(np, (2 * D - 1) * sweep_blocks);
  if (godunov) {
    np = std::max<int64_t>(np, 5 * D * GodunovCflBlocks(D, n));
    for (int d = 0; d < D; ++d) np = std::max<int64_t>(np, 5 * GodunovSweepBlocks(D, d, n, PencilSeg(D, d, n)));
  }
  if (hipMalloc(&L->partials, np * sizeof(unsigned long long)) != hipSuccess) {
    (void)hipFree(L->block);
    (void)hipFree(L->ctl);
    delete L;
    SetError("hipMalloc of the reduction partials failed");
    return FVS_HIP_ERROR;
  }
  if (godunov) {
    int64_t nf = 0;
    for (int d = 0; d < D; ++d) nf = std::max<int64_t>(nf, L->ncell / n[d] * (n[d] + 1));
    if (hipMalloc(&L->pm, nf * sizeof(double)) != hipSuccess) {
      (void)hipFree(L->block);
      (void)hipFree(L->ctl);
      (void)hipFree(L->partials);
      delete L;
      SetError("hipMalloc of the star-pressure faces failed");
      return FVS_HIP_ERROR;
    }
  }
  for (int b = 0; b < 3; ++b)
    for (int c = 0; c < D + 2; ++c) L->buf[b][c] = L->block + (static_cast<int64_t>(b) * (D + 2) + c) * L->ncell;
  L->cur = 0;
  (void)hipMemsetAsync(L->block, 0, bytes, h->stream);
  Ctl init{};
  init.cfl = patch->cfl;
  (void)hipMemcpyAsync(L->ctl, &init, sizeof init, hipMemcpyHostToDevice, h->stream);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    SetError("level initialisation failed");
    return FVS_HIP_ERROR;
  }
  *out = L;
  return FVS_OK;
}

int fvs_level_destroy(fvs_level* L) {
  if (!L) return FVS_OK;
  if (L->composed) {
    ComposedDestroy(L->composed);
    delete L;
    return FVS_OK;
  }
  (void)hipStreamSynchronize(L->h->stream);
  (void)hipFree(L->block);
  (void)hipFree(L->ctl);
  (void)hipFree(L->partials);
  if (L->pm) (void)hipFree(L->pm);
  for (const auto& m : L->marks) {
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
  for (hipEvent_t e : L->spare) (void)hipEventDestroy(e);
  for (const auto& g : L->graphs) (void)hipGraphExecDestroy(g.exec);
  if (L->cap) (void)hipStreamDestroy(L->cap);
  delete L;
  return FVS_OK;
}

int fvs_level_set_graphs(fvs_level* L, int on) {
  if (!L) {
    SetError("null level");
    return FVS_BAD_ARGUMENT;
  }
  L->use_graphs = on ? 1 : 0;
  return FVS_OK;
}

int fvs_level_set_timing(fvs_level* L, int on) {
  if (!L) {
    SetError("null level");
    return FVS_BAD_ARGUMENT;
  }
  L->timing = on ? 1 : 0;
  return FVS_OK;
}

int fvs_level_kernel_times(fvs_level* L, double* ms, int64_t* count) {
  if (!L || !ms || !count) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = CollectMarks(L)) return rc;
  for (int k = 0; k < FVS_LEVEL_TIMERS; ++k) {
    ms[k] = L->t_ms[k];
    count[k] = L->t_n[k];
    L->t_ms[k] = 0.0;
    L->t_n[k] = 0;
  }
  return FVS_OK;
}

double* fvs_level_state(fvs_level* L, int c) {
  if (!L || c < 0 || c >= L->h->desc.dim + 2) return nullptr;
  if (L->composed) {
    SetError("a composed level keeps its state inside a ghosted scratch (fvs_level_get_state)");
    return nullptr;
  }
  return L->buf[L->cur][c];
}

int fvs_level_set_state(fvs_level* L, const double* const* comps, double t0) {
  if (!L || !comps) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  const int D = L->h->desc.dim;
  for (int c = 0; c < D + 2; ++c) {
    if (!comps[c]) {
      SetError("null component pointer %d", c);
      return FVS_BAD_ARGUMENT;
    }
  }
  if (L->composed) return ComposedSetState(L->composed, comps, t0);
  for (int c = 0; c < D + 2; ++c) {
    FVS_HIP_CHECK(hipMemcpyAsync(L->buf[L->cur][c], comps[c], L->ncell * sizeof(double),
                                 hipMemcpyDeviceToDevice, L->h->stream));
  }
  Ctl init{};
  init.t = t0;
  init.cfl = L->patch.cfl;
  FVS_HIP_CHECK(hipMemcpyAsync(L->ctl, &init, sizeof init, hipMemcpyHostToDevice, L->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(L->h->stream));
  return FVS_OK;
}

int fvs_level_get_state(fvs_level* L, double* const* comps) {
  if (!L || !comps) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  const int D = L->h->desc.dim;
  for (int c = 0; c < D + 2; ++c) {
    if (!comps[c]) {
      SetError("null component pointer %d", c);
      return FVS_BAD_ARGUMENT;
    }
  }
  if (L->composed) return ComposedGetState(L->composed, comps);
  for (int c = 0; c < D + 2; ++c) {
    FVS_HIP_CHECK(hipMemcpyAsync(comps[c], L->buf[L->cur][c], L->ncell * sizeof(double),
                                 hipMemcpyDeviceToDevice, L->h->stream));
  }
  return FVS_OK;
}

int fvs_level_time(fvs_level* L, double* t, int64_t* cycles) {
  if (!L) {
    SetError("null level");
    return FVS_BAD_ARGUMENT;
  }
  if (L->composed) {
    ComposedTime(L->composed, t, cycles);
    return FVS_OK;
  }
  Ctl c;
  FVS_HIP_CHECK(hipMemcpyAsync(&c, L->ctl, sizeof c, hipMemcpyDeviceToHost, L->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(L->h->stream));
  if (t) *t = c.t;
  if (cycles) *cycles = c.cycle;
  return FVS_OK;
}

int fvs_level_hint_max_speed(fvs_level* L, const double speed[3]) {
  if (!L || !speed) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  for (int d = 0; d < 3; ++d) {
    if (!(speed[d] >= 0.0) || !(speed[d] < 1.7976931348623157e308)) {
      SetError("speed hint %d must be finite and non-negative", d);
      return FVS_BAD_ARGUMENT;
    }
  }
  // the exact-Godunov passes and the 3D HLLEM pass have a floor
  if (L->composed) return FVS_OK;
  if (L->h->desc.dim != 3 && L->h->desc.base != FVS_BASE_GODUNOV_EXACT) return FVS_OK;
  Ctl c;
  FVS_HIP_CHECK(hipMemcpyAsync(&c, L->ctl, sizeof c, hipMemcpyDeviceToHost, L->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(L->h->stream));
  for (int d = 0; d < 3; ++d) c.cfl_floor[d] = speed[d];
  FVS_HIP_CHECK(hipMemcpyAsync(L->ctl, &c, sizeof c, hipMemcpyHostToDevice, L->h->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(L->h->stream));
  return FVS_OK;
}

int fvs_level_cfl_redos(fvs_level* L, int64_t* redos) {
  if (!L || !redos) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  *redos = L->cfl_redos;
  return FVS_OK;
}

// One time step of the level (CFL pass, the split sweeps with their CFL
// checks, step end) launched on the handle's stream from buffer `cur`;
// *next receives the buffer holding the result.  The launch sequence
// depends only on (cur, cycle parity, network), so it can be captured once
// per combination into a HIP graph.
static int LaunchLevelStep(fvs_level* L, hipStream_t s, int cur, int64_t cyc,
                           const fvs_mlp* net, int* next) {
  fvs_handle* h = L->h;
  const int D = h->desc.dim;
  const bool strang = L->patch.splitting == FVS_SPLIT_STRANG;
  const bool godunov = h->desc.base == FVS_BASE_GODUNOV_EXACT;
  int rc = FVS_OK;
  // MakeSplitDirections (SplittingMethod.hpp:115-126), subcycle {0, 1}.
  const int is_odd = static_cast<int>(cyc % 2);
  int dirs[3];
  for (int i = 0; i < D; ++i) dirs[i] = is_odd ? (D - 1 - i) : i;
  int seq[6];
  int ns = 0;
  for (int i = 0; i < D; ++i) seq[ns++] = dirs[i];
  if (strang)
    for (int i = D - 1; i >= 0; --i) seq[ns++] = dirs[i];
  const double* cin[5];
  for (int q = 0; q < D + 2; ++q) cin[q] = L->buf[cur][q];
  if (godunov) {
    const int64_t nb = GodunovCflBlocks(D, L->n);
    rc = Timed(L, s, FVS_TIMER_CFL, [&] {
      return LaunchGodunovCfl(h->desc, s, cin, L->n, L->patch.dx, nullptr, L->partials,
                              &L->ctl->halt, net, L->pm, seq[0], L->ctl);
    });
    for (int d = 0; d < D && rc == FVS_OK; ++d) {
      rc = LaunchGodunovFinalize(s, L->partials + 5 * d * nb, nb, &L->ctl->dt_dir[d],
                                 h->newton_stats, L->ctl);
    }
    const int use_floor = h->desc.max_newton_iterations >= kGodunovFloorCap;
    if (rc == FVS_OK) rc = LaunchStepDt(s, L->ctl, D, seq[0], strang, L->patch.dx, use_floor);
  } else {
    rc = Timed(L, s, FVS_TIMER_CFL, [&] {
      return LaunchPatchCfl(h->desc, s, cin, L->n, L->patch.dx, L->partials, nullptr, L->ctl,
                            &L->ctl->halt, seq[0], strang);
    });
  }
  if (rc) return rc;
  const int b0 = cur, b1 = (cur + 1) % 3, b2 = (cur + 2) % 3;
  int src = b0, dst = b1;
  CheckList checks{};
  checks.stride = 0;
  for (int d = 0; d < D; ++d)
    checks.stride = std::max<int64_t>(checks.stride, SweepBlocks(D, d, L->n, PencilSeg(D, d, L->n)));
  for (int w = 0; w < ns; ++w) {
    SweepArgs a;
    double* outp[5];
    for (int q = 0; q < D + 2; ++q) outp[q] = L->buf[dst][q];
    const double* inp[5];
    for (int q = 0; q < D + 2; ++q) inp[q] = L->buf[src][q];
    if (godunov) {
      const int64_t nb = GodunovSweepBlocks(D, seq[w], L->n, PencilSeg(D, seq[w], L->n));
      // The first sweep runs on the step-start state, whose star pressures
      // in its direction the CFL pass kept: its faces are not solved again.
      rc = Timed(L, s, FVS_TIMER_SWEEP_X + seq[w], [&] {
        return LaunchGodunovSweep(h->desc, s, seq[w], inp, outp, L->n, L->patch.dx[seq[w]],
                                  &L->ctl->split_dt, 0.0, w == 0 ? L->pm : nullptr,
                                  L->partials, &L->ctl->halt, 1, w == 0 ? nullptr : net,
                                  w == 0 ? 1 : 0);
      });
      if (rc == FVS_OK)
        rc = LaunchGodunovFinalize(s, L->partials, nb, &L->ctl->last_check, h->newton_stats,
                                   L->ctl);
      if (rc == FVS_OK && w > 0) rc = LaunchCheckValue(s, L->ctl, &L->ctl->last_check);
      if (rc) break;
      src = dst;
      dst = (dst == b1) ? b2 : b1;
      continue;
    }
    rc = FillSweepArgs(h, inp, outp, L->n, L->patch.dx[seq[w]], seq[w], &a);
    if (rc) break;
    a.dt_ptr = &L->ctl->split_dt;
    a.halt = &L->ctl->halt;
    a.check = (w > 0);
    a.bound_check = 1;
    if (a.check) {
      if (checks.count >= kMaxChecks) return FVS_BAD_ARGUMENT;
      a.partials = L->partials + checks.count * checks.stride;
      checks.n[checks.count++] = SweepBlocks(D, seq[w], L->n, a.seg);
    }
    rc = Timed(L, s, FVS_TIMER_SWEEP_X + seq[w], [&] { return LaunchSweep(h->desc, s, seq[w], a); });
    if (rc) break;
    src = dst;
    dst = (dst == b1) ? b2 : b1;
  }
  if (rc) return rc;
  *next = src;
  if (godunov) return LaunchStepEnd(s, L->ctl);
  return LaunchFinalizeStep(s, L->partials, checks, L->ctl);
}

// The step as a replayed HIP graph, captured on first use per key on the
// level's own capture stream (the handle's stream may be the null stream,
// which cannot capture) and launched on the handle's stream.
// Captures (without running) the graph of nsteps steps from buffer cur at a
// cycle of the given parity, unless it exists; returns its entry.
static int CaptureLevelSteps(fvs_level* L, int cur, int parity, const fvs_mlp* net, int nsteps,
                             const fvs_level::StepGraph** out) {
  for (const auto& g : L->graphs) {
    if (g.cur == cur && g.parity == parity && g.net == net && g.nsteps == nsteps) {
      *out = &g;
      return FVS_OK;
    }
  }
  fvs_level::StepGraph g;
  g.cur = cur;
  g.parity = parity;
  g.net = net;
  g.nsteps = nsteps;
  g.curs.assign(1, cur);
  if (!L->cap) FVS_HIP_CHECK(hipStreamCreateWithFlags(&L->cap, hipStreamNonBlocking));
  FVS_HIP_CHECK(hipStreamBeginCapture(L->cap, hipStreamCaptureModeThreadLocal));
  int rc = FVS_OK;
  for (int i = 0; i < nsteps && rc == FVS_OK; ++i) {
    int next = g.curs.back();
    rc = LaunchLevelStep(L, L->cap, g.curs.back(), parity + i, net, &next);
    g.curs.push_back(next);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(L->cap, &graph);
  if (rc != FVS_OK || ec != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != FVS_OK) return rc;
    SetError("step capture failed: %s", hipGetErrorString(ec));
    return FVS_HIP_ERROR;
  }
  const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    SetError("step graph instantiation failed: %s", hipGetErrorString(ei));
    return FVS_HIP_ERROR;
  }
  L->graphs.push_back(std::move(g));
  *out = &L->graphs.back();
  return FVS_OK;
}

// Every single-step and chained graph of the level for this network (six
// start states: three buffers, two cycle parities), captured before the
// first replay so that no capture falls inside a run of steps.
static int EnsureLevelGraphs(fvs_level* L, const fvs_mlp* net) {
  const fvs_level::StepGraph* g = nullptr;
  for (int cur = 0; cur < 3; ++cur)
    for (int parity = 0; parity < 2; ++parity)
      for (int nsteps : {1, kChainSteps})
        if (int rc = CaptureLevelSteps(L, cur, parity, net, nsteps, &g)) return rc;
  return FVS_OK;
}

// Launches nsteps (1 or kChainSteps) steps from buffer cur as one replayed
// graph; curs receives the start buffer of each step and, last, the result's.
static int LaunchLevelStepGraph(fvs_level* L, int cur, int64_t cyc, const fvs_mlp* net,
                                int nsteps, std::vector<int>* curs) {
  const fvs_level::StepGraph* g = nullptr;
  if (int rc = CaptureLevelSteps(L, cur, static_cast<int>(cyc % 2), net, nsteps, &g)) return rc;
  *curs = g->curs;
  FVS_HIP_CHECK(hipGraphLaunch(g->exec, L->h->stream));
  return FVS_OK;
}

int fvs_level_advance(fvs_level* L, int max_steps, double final_time, int batch,
                      double* dt_log, int* steps_done, int* failures) {
  if (!L || max_steps < 0 || batch < 1) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  if (L->composed)
    return ComposedAdvance(L->composed, max_steps, final_time, dt_log, steps_done, failures);
  fvs_handle* h = L->h;
  hipStream_t s = h->stream;
  const bool godunov = h->desc.base == FVS_BASE_GODUNOV_EXACT;
  // With a network attached, the Godunov kernels infer each face's guess
  // inline (no guess arrays, no extra pass over the state).
  const fvs_mlp* net = godunov ? h->mlp : nullptr;
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
  // Newton non-convergence (exact Godunov): the reference throws from
  // NewtonIteration (include/fub/NewtonIteration.hpp:44-50) and the run
  // ends; here the run ends at the batch whose steps hit the cap.
  unsigned long long newton0[5] = {}, newton1[5] = {};
  if (godunov) {
    FVS_HIP_CHECK(hipMemcpyAsync(newton0, h->newton_stats, sizeof newton0,
                                 hipMemcpyDeviceToHost, s));
    FVS_HIP_CHECK(hipStreamSynchronize(s));
  }
  if (!L->timing && L->use_graphs && max_steps > 0) {
    if (int rc2 = EnsureLevelGraphs(L, net)) {
      if (dlog) (void)hipFree(dlog);
      return rc2;
    }
  }
  while (cycle - cycle0 < max_steps) {
    const int k = static_cast<int>(std::min<int64_t>(batch, max_steps - (cycle - cycle0)));
    start_buf.assign(k + 1, 0);
    int cur = L->cur;
    std::vector<int> curs;
    for (int sidx = 0; sidx < k && rc == FVS_OK;) {
      const int64_t cyc = cycle + sidx;
      if (L->timing || !L->use_graphs) {
        start_buf[sidx] = cur;
        int next = cur;
        rc = LaunchLevelStep(L, s, cur, cyc, net, &next);
        if (rc) break;
        cur = next;
        ++sidx;
        continue;
      }
      const int nsteps = (k - sidx >= kChainSteps) ? kChainSteps : 1;
      rc = LaunchLevelStepGraph(L, cur, cyc, net, nsteps, &curs);
      if (rc) break;
      for (int i = 0; i < nsteps; ++i) start_buf[sidx + i] = curs[i];
      cur = curs[nsteps];
      sidx += nsteps;
    }
    if (rc) break;
    FVS_HIP_CHECK(hipMemcpyAsync(&c, L->ctl, sizeof c, hipMemcpyDeviceToHost, s));
    if (godunov)
      FVS_HIP_CHECK(hipMemcpyAsync(newton1, h->newton_stats, sizeof newton1,
                                   hipMemcpyDeviceToHost, s));
    FVS_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t completed = c.cycle - cycle;
    if (completed > 0) consecutive = 0;
    if (godunov && newton1[2] > newton0[2]) {
      L->cur = c.halt == 0 ? cur : start_buf[completed];
      cycle = c.cycle;
      SetError("%llu Riemann problems hit the Newton iteration cap (%d iterations)",
               newton1[2] - newton0[2], h->desc.max_newton_iterations);
      rc = FVS_NEWTON_NOT_CONVERGED;
      break;
    }
    if (c.halt == 0) {
      L->cur = cur;
      cycle = c.cycle;
      continue;
    }
    // The step `completed` of this batch stopped; its start buffer holds the
    // state after the last completed step.
    L->cur = start_buf[completed];
    cycle = c.cycle;
    if (c.halt == 2) break;
    if (c.halt == 3) {
      // The CFL pass's speed floor (already cleared) was not reached: the
      // step runs again with every face evaluated.  Not a CFL failure.
      ++L->cfl_redos;
      c.halt = 0;
      FVS_HIP_CHECK(hipMemcpyAsync(L->ctl, &c, sizeof c, hipMemcpyHostToDevice, s));
      continue;
    }
    // TimeStepTooLarge: restart the step with cfl * failure_dt
    // (RunSimulation.hpp:141-151).
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
    if (rc == FVS_OK) {
      const int64_t nd = cycle - cycle0;
      if (nd > 0) {
        (void)hipMemcpyAsync(dt_log, dlog, nd * sizeof(double), hipMemcpyDeviceToHost, s);
      }
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(dlog);
  }
  if (steps_done) *steps_done = static_cast<int>(cycle - cycle0);
  if (failures) *failures = nfail;
  return rc;
}

} // extern "C"
