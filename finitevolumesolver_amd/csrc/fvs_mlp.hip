// Learned initial guess for the exact Riemann solver's Newton iteration.
//
// A 6 -> 64 -> 64 -> 1 ReLU MLP maps a state pair to the log of the star
// pressure relative to sqrt(pL pR):
//   features  x = (ln rhoL s, uL / 5, ln pL s, ln rhoR s, uR / 5, ln pR s),
//             s = 1 / ln(1000)
//   output    r,  initial guess p0 = exp(0.5 (ln pL + ln pR) + r)
// Inference runs on the matrix cores: BF16 operands, FP32 accumulation,
// v_mfma_f32_16x16x16_bf16.  Activations are kept transposed (hidden units x
// faces) so that a layer's 16x16 accumulator tile has exactly the lane and
// element layout of the next layer's B operand (no data movement between
// layers).  Training (forward, backward, Adam) is FP32 on the VALU; labels
// come from the FP64 Newton iteration itself.  The Newton iteration that
// refines the guess is unchanged (include/fub/NewtonIteration.hpp:28-52).
#include "fvs_internal.hpp"
#include "fvs_mlp_infer.hpp"
#include "fvs_riemann.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

namespace fvs {

namespace {

using namespace mlp;

constexpr int kBlock = 256;

// Packs the FP32 master parameters into bf16 A-operand tiles and a bias
// vector (b1, b2, b3).
__global__ void MlpPackKernel(const float* __restrict__ p, short* __restrict__ tiles,
                              float* __restrict__ bias) {
  const int t = blockIdx.x;      // tile
  const int lane = threadIdx.x;  // 0..63
  const int i = lane & 15;
  for (int e = 0; e < 4; ++e) {
    const int k = (lane >> 4) * 4 + e;
    float v = 0.0f;
    if (t < 4) {
      const int out = t * 16 + i;
      v = (k < kInPad) ? p[kOffW1 + out * kInPad + k] : 0.0f;
    } else if (t < 20) {
      const int m = (t - 4) / 4, kb = (t - 4) % 4;
      v = p[kOffW2 + (m * 16 + i) * kH + kb * 16 + k];
    } else {
      const int kb = t - 20;
      v = (i == 0) ? p[kOffW3 + kb * 16 + k] : 0.0f;
    }
    tiles[(t * 64 + lane) * 4 + e] = ToBf16(v);
  }
  if (t == 0) {
    for (int j = lane; j < 2 * kH + 1; j += 64) {
      bias[j] = (j < kH) ? p[kOffB1 + j] : (j < 2 * kH ? p[kOffB2 + j - kH] : p[kOffB3]);
    }
  }
}

// Inference on arrays of complete state pairs (SoA (D+4) x n).  Writes the
// network output r and the initial guess.
template <int D>
__global__ __launch_bounds__(kBlock) void MlpInferPairs(const double* __restrict__ L,
                                                        const double* __restrict__ R, int64_t n,
                                                        int dir, const short* __restrict__ tiles,
                                                        const float* __restrict__ bias,
                                                        float* __restrict__ r_out,
                                                        double* __restrict__ guess) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t j = i < n ? i : n - 1;
  double l[D + 4], r[D + 4];
  for (int k = 0; k < D + 4; ++k) {
    l[k] = L[k * n + j];
    r[k] = R[k * n + j];
  }
  float x[kIn];
  Features<D>(l, r, dir, x);
  const float out = MlpWave(x, tiles, bias);
  if (i < n) {
    if (r_out) r_out[i] = out;
    if (guess) {
      const bool weak = WeakPair(l[0], l[1 + dir] / l[0], l[D + 2], l[D + 3], r[0],
                                 r[1 + dir] / r[0], r[D + 2], r[D + 3]);
      guess[i] = weak ? 0.5 * (l[D + 2] + r[D + 2]) : GuessFromOutput(l[D + 2], r[D + 2], out);
    }
  }
}

// Guesses for every face of direction DIR of a ghost-free transmissive patch
// (face box n with n[DIR] + 1, x fastest): the first-order Godunov states of
// face f are the (clamped) cells f - 1 and f.
template <int D, int DIR>
__global__ __launch_bounds__(kBlock) void MlpInferFaces(
    const double* __restrict__ u0, const double* __restrict__ u1, const double* __restrict__ u2,
    const double* __restrict__ u3, const double* __restrict__ u4, int64_t nx, int64_t ny,
    int64_t nz, double gamma, double gm1inv, const short* __restrict__ tiles,
    const float* __restrict__ bias, double* __restrict__ guess) {
  const double* in[5] = {u0, u1, u2, u3, u4};
  const int64_t fx = nx + (DIR == 0), fy = ny + (DIR == 1), fz = nz + (DIR == 2);
  const int64_t nf = fx * fy * fz;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t qq = q < nf ? q : nf - 1;
  int64_t c[3] = {qq % fx, (qq / fx) % fy, qq / (fx * fy)};
  const int64_t n3[3] = {nx, ny, nz};
  Gas<double> g;
  g.gamma = gamma;
  g.gm1inv = gm1inv;
  g.gm1 = gamma - 1.0;
  double L[D + 4], R[D + 4];
  int64_t cl[3] = {c[0], c[1], c[2]}, cr[3] = {c[0], c[1], c[2]};
  cl[DIR] = c[DIR] - 1 < 0 ? 0 : c[DIR] - 1;
  cr[DIR] = c[DIR] >= n3[DIR] ? n3[DIR] - 1 : c[DIR];
  const int64_t ol = cl[0] + nx * (cl[1] + ny * cl[2]);
  const int64_t orr = cr[0] + nx * (cr[1] + ny * cr[2]);
  for (int k = 0; k < D + 2; ++k) {
    L[k] = in[k][ol];
    R[k] = in[k][orr];
  }
  CompleteFromCons<D>(L, g);
  CompleteFromCons<D>(R, g);
  const bool weak = WeakPair(L[0], L[1 + DIR] / L[0], L[D + 2], L[D + 3], R[0],
                             R[1 + DIR] / R[0], R[D + 2], R[D + 3]);
  double p0 = 0.5 * (L[D + 2] + R[D + 2]);
  if (__ballot(!weak) != 0) {  // wave-uniform: skip the network when every face is weak
    float x[kIn];
    Features<D>(L, R, DIR, x);
    const float out = MlpWave(x, tiles, bias);
    if (!weak) p0 = GuessFromOutput(L[D + 2], R[D + 2], out);
  }
  if (q < nf) guess[q] = p0;
}

// ---- training ---------------------------------------------------------------

__device__ __forceinline__ uint64_t SplitMix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double Uniform(uint64_t seed, uint64_t i, int k) {
  return static_cast<double>(SplitMix(seed ^ SplitMix(i * 8 + k)) >> 11) * (1.0 / 9007199254740992.0);
}

// Sample i: rho, p log-uniform in [1e-3, 1e3], u uniform in [-5, 5] on both
// sides; vacuum-generating pairs (2 aL / (g-1) + 2 aR / (g-1) <= uR - uL)
// are flagged; the label is the FP64 Newton star pressure (default guess).
// prim: 6 x n doubles (rhoL, uL, pL, rhoR, uR, pR); x: 6 x n floats;
// target: r = ln p* - 0.5 (ln pL + ln pR); flag: 0 ok, 1 vacuum, 2 no convergence.
__global__ __launch_bounds__(kBlock) void MlpSampleKernel(int64_t n, uint64_t seed, double gamma,
                                                          int max_it, double tol,
                                                          double* __restrict__ prim,
                                                          float* __restrict__ x,
                                                          float* __restrict__ target,
                                                          int* __restrict__ flag,
                                                          int* __restrict__ iters) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const double lo = log(1e-3), hi = log(1e3);
  const double rhoL = exp(lo + (hi - lo) * Uniform(seed, i, 0));
  const double uL = -5.0 + 10.0 * Uniform(seed, i, 1);
  const double pL = exp(lo + (hi - lo) * Uniform(seed, i, 2));
  const double rhoR = exp(lo + (hi - lo) * Uniform(seed, i, 3));
  const double uR = -5.0 + 10.0 * Uniform(seed, i, 4);
  const double pR = exp(lo + (hi - lo) * Uniform(seed, i, 5));
  const double pr[6] = {rhoL, uL, pL, rhoR, uR, pR};
  for (int k = 0; k < 6; ++k) prim[k * n + i] = pr[k];
  Gas<double> g;
  g.gamma = gamma;
  g.gm1inv = 1.0 / (gamma - 1.0);
  g.gm1 = gamma - 1.0;
  // Complete states (1D momentum along x; 2D layout with zero transverse).
  double L[6], R[6];
  auto complete = [&](double rho, double u, double p, double* q) {
    q[0] = rho;
    q[1] = rho * u;
    q[2] = 0.0;
    const double e_kin = 0.5 * (q[1] * q[1] + q[2] * q[2]) / rho;
    q[3] = e_kin + p * g.gm1inv;
    q[4] = p;
    q[5] = sqrt(gamma * p / rho);
  };
  complete(rhoL, uL, pL, L);
  complete(rhoR, uR, pR, R);
  float xf[kIn];
  Features<2>(L, R, 0, xf);
  for (int k = 0; k < kIn; ++k) x[k * n + i] = xf[k];
  const double gm1 = gamma - 1.0;
  if (2.0 * L[5] / gm1 + 2.0 * R[5] / gm1 <= uR - uL) {
    flag[i] = 1;
    target[i] = 0.0f;
    iters[i] = 0;
    return;
  }
  NewtonParams np;
  np.tol = tol;
  np.max_it = max_it;
  const MiddleState ms = ComputeMiddleState<2, 0>(L, R, g, np, 0.5 * (pL + pR));
  iters[i] = ms.iterations;
  if (!ms.converged || !(ms.pM > 0.0)) {
    flag[i] = 2;
    target[i] = 0.0f;
    return;
  }
  flag[i] = 0;
  target[i] = static_cast<float>(log(ms.pM) - 0.5 * (log(pL) + log(pR)));
}

// FP32 forward of one sample per thread; stores activations for backward.
// x: 6 x n, h1/h2: 64 x n, y: n.
__global__ __launch_bounds__(kBlock) void MlpForwardKernel(const float* __restrict__ p,
                                                           const float* __restrict__ x, int64_t n,
                                                           float* __restrict__ h1,
                                                           float* __restrict__ h2,
                                                           float* __restrict__ y) {
  __shared__ float sp[kNParams];
  for (int k = threadIdx.x; k < kNParams; k += kBlock) sp[k] = p[k];
  __syncthreads();
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  float xi[kIn];
  for (int k = 0; k < kIn; ++k) xi[k] = x[k * n + i];
  float a1[kH];
  for (int o = 0; o < kH; ++o) {
    float s = sp[kOffB1 + o];
    for (int k = 0; k < kIn; ++k) s += sp[kOffW1 + o * kInPad + k] * xi[k];
    a1[o] = fmaxf(s, 0.0f);
    if (h1) h1[o * n + i] = a1[o];
  }
  float out = sp[kOffB3];
  for (int o = 0; o < kH; ++o) {
    float s = sp[kOffB2 + o];
    for (int k = 0; k < kH; ++k) s += sp[kOffW2 + o * kH + k] * a1[k];
    const float a2 = fmaxf(s, 0.0f);
    if (h2) h2[o * n + i] = a2;
    out += sp[kOffW3 + o] * a2;
  }
  y[i] = out;
}

// Backward part 1 (per sample): loss e^2 (or the Huber form 2 h |e| - h^2
// beyond |e| > h when huber h > 0) over valid samples, e = y - t;
// dy = 2 clamp(e, -h, h) / N (mask 0 otherwise), dh2 = dy W3 [h2 > 0],
// dh1 = W2^T dh2 [h1 > 0] (ReLU derivative).
// dh2 / dh1 overwrite h2 / h1's storage in separate arrays.
__global__ __launch_bounds__(kBlock) void MlpBackwardSamples(
    const float* __restrict__ p, const float* __restrict__ y, const float* __restrict__ t,
    const int* __restrict__ flag, int64_t n, float inv_count, float huber,
    const float* __restrict__ h1,
    const float* __restrict__ h2, float* __restrict__ dy, float* __restrict__ dh1,
    float* __restrict__ dh2, float* __restrict__ loss) {
  __shared__ float sp[kNParams];
  __shared__ float sl[kBlock / 64];
  for (int k = threadIdx.x; k < kNParams; k += kBlock) sp[k] = p[k];
  __syncthreads();
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  float li = 0.0f;
  if (i < n) {
    const bool ok = flag[i] == 0;
    const float e = ok ? (y[i] - t[i]) : 0.0f;
    const float ae = fabsf(e);
    const bool quad = !(huber > 0.0f) || ae <= huber;
    li = (quad ? e * e : 2.0f * huber * ae - huber * huber) * inv_count;
    const float ec = quad ? e : copysignf(huber, e);
    const float d = 2.0f * ec * inv_count;
    dy[i] = d;
    float g2[kH];
    for (int o = 0; o < kH; ++o) {
      const float a = h2[o * n + i];
      g2[o] = a > 0.0f ? d * sp[kOffW3 + o] : 0.0f;
      dh2[o * n + i] = g2[o];
    }
    for (int k = 0; k < kH; ++k) {
      float s = 0.0f;
      for (int o = 0; o < kH; ++o) s += sp[kOffW2 + o * kH + k] * g2[o];
      const float a = h1[k * n + i];
      dh1[k * n + i] = a > 0.0f ? s : 0.0f;
    }
  }
  for (int off = 32; off > 0; off >>= 1) li += __shfl_xor(li, off, 64);
  if ((threadIdx.x & 63) == 0) sl[threadIdx.x >> 6] = li;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, sl[0] + sl[1] + sl[2] + sl[3]);
}

// Backward part 2: parameter gradients as reductions over the samples of a
// chunk; one workgroup per chunk adds its partial sums to grad (FP32).
// Thread t handles parameters t, t + 256, ...
__global__ __launch_bounds__(kBlock) void MlpWeightGrads(
    const float* __restrict__ x, const float* __restrict__ h1, const float* __restrict__ h2,
    const float* __restrict__ dy, const float* __restrict__ dh1, const float* __restrict__ dh2,
    int64_t n, int64_t chunk, float* __restrict__ grad) {
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * chunk;
  const int64_t b1 = (b0 + chunk < n) ? b0 + chunk : n;
  for (int k = threadIdx.x; k < kNParams; k += kBlock) {
    float s = 0.0f;
    if (k < kOffB1) {
      const int o = k / kInPad, in = k % kInPad;
      if (in < kIn)
        for (int64_t b = b0; b < b1; ++b) s += dh1[o * n + b] * x[in * n + b];
    } else if (k < kOffW2) {
      const int o = k - kOffB1;
      for (int64_t b = b0; b < b1; ++b) s += dh1[o * n + b];
    } else if (k < kOffB2) {
      const int o = (k - kOffW2) / kH, in = (k - kOffW2) % kH;
      for (int64_t b = b0; b < b1; ++b) s += dh2[o * n + b] * h1[in * n + b];
    } else if (k < kOffW3) {
      const int o = k - kOffB2;
      for (int64_t b = b0; b < b1; ++b) s += dh2[o * n + b];
    } else if (k < kOffB3) {
      const int o = k - kOffW3;
      for (int64_t b = b0; b < b1; ++b) s += dy[b] * h2[o * n + b];
    } else {
      for (int64_t b = b0; b < b1; ++b) s += dy[b];
    }
    atomicAdd(&grad[k], s);
  }
}

__global__ void MlpAdamKernel(float* __restrict__ p, const float* __restrict__ grad,
                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                              float b2, float eps, float bc1, float bc2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNParams) return;
  if (k < kOffB1 && (k % kInPad) >= kIn) return;  // padding columns stay zero
  const float gk = grad[k];
  m[k] = b1 * m[k] + (1.0f - b1) * gk;
  v[k] = b2 * v[k] + (1.0f - b2) * gk * gk;
  const float mh = m[k] / bc1, vh = v[k] / bc2;
  p[k] -= lr * mh / (sqrtf(vh) + eps);
}

// Gathers a minibatch (indices idx) of features / targets / flags.
__global__ void MlpGatherKernel(const float* __restrict__ x, const float* __restrict__ t,
                                const int* __restrict__ flag, int64_t n,
                                const int64_t* __restrict__ idx, int64_t nb,
                                float* __restrict__ xb, float* __restrict__ tb,
                                int* __restrict__ fb) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int64_t j = idx[i];
  for (int k = 0; k < kIn; ++k) xb[k * nb + i] = x[k * n + j];
  tb[i] = t[j];
  fb[i] = flag[j];
}

} // namespace
} // namespace fvs

// ---- host side -----------------------------------------------------------------

using namespace fvs;

struct fvs_mlp {
  hipStream_t stream;
  float* params;  // FP32 master (device)
  float* adam_m;
  float* adam_v;
  float* grad;
  float* loss;
  short* tiles;   // packed bf16 A operands
  float* bias;    // b1, b2, b3
  int64_t adam_t;
  // dataset
  int64_t n;
  double* prim;
  float* x;
  float* target;
  int* flag;
  int* iters;
  // scratch for a minibatch
  int64_t cap;
  float *xb, *tb, *yb, *h1, *h2, *dy, *dh1, *dh2;
  int* fb;
  int64_t* idx;
};

namespace {

int Alloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) {
    SetError("hipMalloc of %zu bytes failed", bytes);
    return FVS_HIP_ERROR;
  }
  return FVS_OK;
}

int Pack(fvs_mlp* m) {
  hipLaunchKernelGGL(MlpPackKernel, dim3(kNTiles), dim3(64), 0, m->stream, m->params, m->tiles,
                     m->bias);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

void FreeBatch(fvs_mlp* m) {
  void* ps[] = {m->xb, m->tb, m->yb, m->h1, m->h2, m->dy, m->dh1, m->dh2, m->fb, m->idx};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  m->xb = m->tb = m->yb = m->h1 = m->h2 = m->dy = m->dh1 = m->dh2 = nullptr;
  m->fb = nullptr;
  m->idx = nullptr;
  m->cap = 0;
}

int EnsureBatch(fvs_mlp* m, int64_t nb) {
  if (nb <= m->cap) return FVS_OK;
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  FreeBatch(m);
  const size_t f = sizeof(float);
  if (Alloc((void**)&m->xb, nb * kIn * f) || Alloc((void**)&m->tb, nb * f) ||
      Alloc((void**)&m->yb, nb * f) || Alloc((void**)&m->h1, nb * kH * f) ||
      Alloc((void**)&m->h2, nb * kH * f) || Alloc((void**)&m->dy, nb * f) ||
      Alloc((void**)&m->dh1, nb * kH * f) || Alloc((void**)&m->dh2, nb * kH * f) ||
      Alloc((void**)&m->fb, nb * sizeof(int)) || Alloc((void**)&m->idx, nb * sizeof(int64_t)))
    return FVS_HIP_ERROR;
  m->cap = nb;
  return FVS_OK;
}

// Forward pass of one minibatch already gathered into xb/tb/fb (activations
// kept for Backward); zeroes the gradient and loss accumulators.
int Forward(fvs_mlp* m, int64_t nb) {
  hipStream_t s = m->stream;
  const unsigned blocks = unsigned((nb + kBlock - 1) / kBlock);
  FVS_HIP_CHECK(hipMemsetAsync(m->grad, 0, kNParams * sizeof(float), s));
  FVS_HIP_CHECK(hipMemsetAsync(m->loss, 0, sizeof(float), s));
  hipLaunchKernelGGL(MlpForwardKernel, dim3(blocks), dim3(kBlock), 0, s, m->params, m->xb, nb,
                     m->h1, m->h2, m->yb);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int Backward(fvs_mlp* m, int64_t nb, float inv_count, float huber) {
  hipStream_t s = m->stream;
  const unsigned blocks = unsigned((nb + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(MlpBackwardSamples, dim3(blocks), dim3(kBlock), 0, s, m->params, m->yb,
                     m->tb, m->fb, nb, inv_count, huber, m->h1, m->h2, m->dy, m->dh1, m->dh2,
                     m->loss);
  FVS_LAUNCH_CHECK();
  const int64_t chunk = 64;
  const unsigned gb = unsigned((nb + chunk - 1) / chunk);
  hipLaunchKernelGGL(MlpWeightGrads, dim3(gb), dim3(kBlock), 0, s, m->xb, m->h1, m->h2, m->dy,
                     m->dh1, m->dh2, nb, chunk, m->grad);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

uint64_t HostSplitMix(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

} // namespace

// Initial guesses for every face of direction dir of a ghost-free
// transmissive patch (face box n with n[dir] + 1, x fastest), on stream s.
void fvs::MlpOperands(const fvs_mlp* m, const short** tiles, const float** bias) {
  *tiles = m->tiles;
  *bias = m->bias;
}

int fvs::MlpFaceGuess(fvs_mlp* m, hipStream_t s, int dim, double gamma,
                      const double* const* cons, const int64_t n[3], int dir, double* guess) {
  if (!m || !cons || !n || !guess || (dim != 2 && dim != 3) || dir < 0 || dir >= dim) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  const int64_t nx = n[0], ny = n[1], nz = dim == 3 ? n[2] : 1;
  const int64_t nf = (nx + (dir == 0)) * (ny + (dir == 1)) * (nz + (dir == 2));
  const unsigned blocks = unsigned((nf + kBlock - 1) / kBlock);
  const double gm1inv = 1.0 / (gamma - 1.0);
  const double* u4 = dim == 3 ? cons[4] : nullptr;
#define FVS_MIF(DD, DR)                                                                          \
  hipLaunchKernelGGL((MlpInferFaces<DD, DR>), dim3(blocks), dim3(kBlock), 0, s, cons[0],         \
                     cons[1], cons[2], cons[3], u4, nx, ny, nz, gamma, gm1inv, m->tiles, m->bias,  \
                     guess)
  if (dim == 2) {
    if (dir == 0) FVS_MIF(2, 0); else FVS_MIF(2, 1);
  } else {
    if (dir == 0) FVS_MIF(3, 0); else if (dir == 1) FVS_MIF(3, 1); else FVS_MIF(3, 2);
  }
#undef FVS_MIF
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

extern "C" {

int fvs_mlp_create(void* hip_stream, uint64_t seed, fvs_mlp** out) {
  if (!out) {
    SetError("null output pointer");
    return FVS_BAD_ARGUMENT;
  }
  *out = nullptr;
  fvs_mlp* m = new (std::nothrow) fvs_mlp();
  if (!m) {
    SetError("out of host memory");
    return FVS_BAD_ARGUMENT;
  }
  std::memset(m, 0, sizeof *m);
  m->stream = static_cast<hipStream_t>(hip_stream);
  const size_t pb = kNParams * sizeof(float);
  if (Alloc((void**)&m->params, pb) || Alloc((void**)&m->adam_m, pb) ||
      Alloc((void**)&m->adam_v, pb) || Alloc((void**)&m->grad, pb) ||
      Alloc((void**)&m->loss, sizeof(float)) ||
      Alloc((void**)&m->tiles, kNTiles * kTileShorts * sizeof(short)) ||
      Alloc((void**)&m->bias, (2 * kH + 1) * sizeof(float))) {
    delete m;
    return FVS_HIP_ERROR;
  }
  // He-uniform initialisation of the ReLU layers (Glorot for the linear
  // output) from SplitMix64, zero biases and padding.
  std::vector<float> p(kNParams, 0.0f);
  uint64_t st = seed;
  auto uni = [&](float a) {
    return static_cast<float>((HostSplitMix(st) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * a;
  };
  const float a1 = std::sqrt(6.0f / kIn), a2 = std::sqrt(6.0f / kH),
              a3 = std::sqrt(6.0f / (kH + 1));
  for (int o = 0; o < kH; ++o)
    for (int k = 0; k < kIn; ++k) p[kOffW1 + o * kInPad + k] = uni(a1);
  for (int k = 0; k < kH * kH; ++k) p[kOffW2 + k] = uni(a2);
  for (int k = 0; k < kH; ++k) p[kOffW3 + k] = uni(a3);
  FVS_HIP_CHECK(hipMemcpyAsync(m->params, p.data(), pb, hipMemcpyHostToDevice, m->stream));
  FVS_HIP_CHECK(hipMemsetAsync(m->adam_m, 0, pb, m->stream));
  FVS_HIP_CHECK(hipMemsetAsync(m->adam_v, 0, pb, m->stream));
  if (int rc = Pack(m)) return rc;
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  *out = m;
  return FVS_OK;
}

int fvs_mlp_destroy(fvs_mlp* m) {
  if (!m) return FVS_OK;
  (void)hipStreamSynchronize(m->stream);
  FreeBatch(m);
  void* ps[] = {m->params, m->adam_m, m->adam_v, m->grad, m->loss, m->tiles, m->bias,
                m->prim, m->x, m->target, m->flag, m->iters};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete m;
  return FVS_OK;
}

int fvs_mlp_num_params(void) { return kNParams; }

int fvs_mlp_get_params(fvs_mlp* m, float* host) {
  if (!m || !host) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  FVS_HIP_CHECK(hipMemcpyAsync(host, m->params, kNParams * sizeof(float), hipMemcpyDeviceToHost,
                               m->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  return FVS_OK;
}

int fvs_mlp_set_params(fvs_mlp* m, const float* host) {
  if (!m || !host) {
    SetError("null argument");
    return FVS_BAD_ARGUMENT;
  }
  FVS_HIP_CHECK(hipMemcpyAsync(m->params, host, kNParams * sizeof(float), hipMemcpyHostToDevice,
                               m->stream));
  if (int rc = Pack(m)) return rc;
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  return FVS_OK;
}

int fvs_mlp_generate(fvs_mlp* m, int64_t n, uint64_t seed, double gamma, int64_t counts[3]) {
  if (!m || n <= 0) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  void* ps[] = {m->prim, m->x, m->target, m->flag, m->iters};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  m->prim = nullptr;
  m->x = m->target = nullptr;
  m->flag = m->iters = nullptr;
  if (Alloc((void**)&m->prim, n * 6 * sizeof(double)) ||
      Alloc((void**)&m->x, n * kIn * sizeof(float)) ||
      Alloc((void**)&m->target, n * sizeof(float)) || Alloc((void**)&m->flag, n * sizeof(int)) ||
      Alloc((void**)&m->iters, n * sizeof(int)))
    return FVS_HIP_ERROR;
  m->n = n;
  hipLaunchKernelGGL(MlpSampleKernel, dim3(unsigned((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     m->stream, n, seed, gamma, 5000, 1e-7, m->prim, m->x, m->target, m->flag,
                     m->iters);
  FVS_LAUNCH_CHECK();
  if (counts) {
    std::vector<int> f(n);
    FVS_HIP_CHECK(hipMemcpyAsync(f.data(), m->flag, n * sizeof(int), hipMemcpyDeviceToHost,
                                 m->stream));
    FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
    counts[0] = counts[1] = counts[2] = 0;
    for (int64_t i = 0; i < n; ++i) counts[f[i] < 3 ? f[i] : 2] += 1;
  }
  return FVS_OK;
}

// Copies the generated dataset into caller device buffers (each nullable;
// all null just reports n).
int fvs_mlp_dataset(fvs_mlp* m, int64_t* n, double* prim, float* x, float* target, int* flag,
                    int* iters) {
  if (!m) {
    SetError("null mlp");
    return FVS_BAD_ARGUMENT;
  }
  if (n) *n = m->n;
  if ((prim || x || target || flag || iters) && !m->prim) {
    SetError("no dataset generated");
    return FVS_BAD_ARGUMENT;
  }
  const int64_t k = m->n;
  const hipMemcpyKind d2d = hipMemcpyDeviceToDevice;
  if (prim) FVS_HIP_CHECK(hipMemcpyAsync(prim, m->prim, k * 6 * sizeof(double), d2d, m->stream));
  if (x) FVS_HIP_CHECK(hipMemcpyAsync(x, m->x, k * kIn * sizeof(float), d2d, m->stream));
  if (target) FVS_HIP_CHECK(hipMemcpyAsync(target, m->target, k * sizeof(float), d2d, m->stream));
  if (flag) FVS_HIP_CHECK(hipMemcpyAsync(flag, m->flag, k * sizeof(int), d2d, m->stream));
  if (iters) FVS_HIP_CHECK(hipMemcpyAsync(iters, m->iters, k * sizeof(int), d2d, m->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  return FVS_OK;
}

// FP32 forward of n samples (x: 6 x n device floats) -> y (device floats).
int fvs_mlp_forward(fvs_mlp* m, const float* x, int64_t n, float* y) {
  if (!m || !x || !y || n <= 0) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  hipLaunchKernelGGL(MlpForwardKernel, dim3(unsigned((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     m->stream, m->params, x, n, nullptr, nullptr, y);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

// Loss (mean over samples with flag 0 of e^2, or of the Huber form when
// huber > 0) and its gradient on one batch (device arrays x: 6 x n,
// target: n, flag: n); grad_host gets the kNParams gradient.
int fvs_mlp_gradients(fvs_mlp* m, const float* x, const float* target, const int* flag,
                      int64_t n, float huber, float* grad_host, float* loss_host) {
  if (!m || !x || !target || !flag || n <= 0) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = EnsureBatch(m, n)) return rc;
  FVS_HIP_CHECK(hipMemcpyAsync(m->xb, x, n * kIn * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
  FVS_HIP_CHECK(hipMemcpyAsync(m->tb, target, n * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
  FVS_HIP_CHECK(hipMemcpyAsync(m->fb, flag, n * sizeof(int), hipMemcpyDeviceToDevice, m->stream));
  std::vector<int> f(n);
  FVS_HIP_CHECK(hipMemcpyAsync(f.data(), flag, n * sizeof(int), hipMemcpyDeviceToHost, m->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  int64_t valid = 0;
  for (int64_t i = 0; i < n; ++i) valid += (f[i] == 0);
  if (int rc = Forward(m, n)) return rc;
  if (int rc = Backward(m, n, valid > 0 ? 1.0f / float(valid) : 0.0f, huber)) return rc;
  if (grad_host)
    FVS_HIP_CHECK(hipMemcpyAsync(grad_host, m->grad, kNParams * sizeof(float),
                                 hipMemcpyDeviceToHost, m->stream));
  if (loss_host)
    FVS_HIP_CHECK(hipMemcpyAsync(loss_host, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  return FVS_OK;
}

// Adam training on the generated dataset: `epochs` passes in minibatches of
// `batch` samples (deterministic shuffles from `seed`), learning rate
// lr * (1 + cos(pi t / T)) / 2 over the T steps of the call, Huber loss
// with threshold `huber` (0: squared error).  loss_history (host, nullable)
// receives a smoothed minibatch loss per epoch.
int fvs_mlp_train(fvs_mlp* m, int epochs, int64_t batch, float lr, float huber, uint64_t seed,
                  float* loss_history) {
  if (!m || !m->x || epochs < 0 || batch <= 0) {
    SetError("bad argument (dataset generated?)");
    return FVS_BAD_ARGUMENT;
  }
  if (int rc = EnsureBatch(m, batch)) return rc;
  const int64_t n = m->n;
  std::vector<int> flags(n);
  FVS_HIP_CHECK(hipMemcpyAsync(flags.data(), m->flag, n * sizeof(int), hipMemcpyDeviceToHost,
                               m->stream));
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  std::vector<int64_t> valid;
  valid.reserve(n);
  for (int64_t i = 0; i < n; ++i)
    if (flags[i] == 0) valid.push_back(i);
  const int64_t nv = static_cast<int64_t>(valid.size());
  if (nv < batch) {
    SetError("dataset has %lld valid samples, fewer than one batch", (long long)nv);
    return FVS_BAD_ARGUMENT;
  }
  int64_t* perm = nullptr;
  if (Alloc((void**)&perm, nv * sizeof(int64_t))) return FVS_HIP_ERROR;
  uint64_t st = seed;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const int64_t nbatches = nv / batch;
  const double total = static_cast<double>(epochs) * static_cast<double>(nbatches);
  int64_t step = 0;
  int rc = FVS_OK;
  for (int e = 0; e < epochs && rc == FVS_OK; ++e) {
    for (int64_t i = nv - 1; i > 0; --i) {
      const int64_t j = static_cast<int64_t>(HostSplitMix(st) % static_cast<uint64_t>(i + 1));
      std::swap(valid[i], valid[j]);
    }
    // one upload of the epoch's permutation, complete before the host
    // reshuffles (the previous epoch's gathers are done after the sync)
    if (hipStreamSynchronize(m->stream) != hipSuccess ||
        hipMemcpyAsync(perm, valid.data(), nv * sizeof(int64_t), hipMemcpyHostToDevice,
                       m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess) {
      SetError("permutation upload failed");
      rc = FVS_HIP_ERROR;
      break;
    }
    double loss_sum = 0.0;
    for (int64_t b = 0; b < nbatches; ++b) {
      hipLaunchKernelGGL(MlpGatherKernel, dim3(unsigned((batch + kBlock - 1) / kBlock)),
                         dim3(kBlock), 0, m->stream, m->x, m->target, m->flag, n,
                         perm + b * batch, batch, m->xb, m->tb, m->fb);
      if (hipGetLastError() != hipSuccess) {
        SetError("gather launch failed");
        rc = FVS_HIP_ERROR;
        break;
      }
      if ((rc = Forward(m, batch)) != FVS_OK) break;
      if ((rc = Backward(m, batch, 1.0f / float(batch), huber)) != FVS_OK) break;
      m->adam_t += 1;
      step += 1;
      const float bc1 = 1.0f - std::pow(b1, float(m->adam_t));
      const float bc2 = 1.0f - std::pow(b2, float(m->adam_t));
      const float lr_t =
          static_cast<float>(lr * 0.5 * (1.0 + std::cos(3.141592653589793 * step / total)));
      hipLaunchKernelGGL(MlpAdamKernel, dim3((kNParams + 255) / 256), dim3(256), 0, m->stream,
                         m->params, m->grad, m->adam_m, m->adam_v, lr_t, b1, b2, eps, bc1, bc2);
      if (hipGetLastError() != hipSuccess) {
        SetError("adam launch failed");
        rc = FVS_HIP_ERROR;
        break;
      }
      if (loss_history && (b % 64 == 0 || b == nbatches - 1)) {
        float l = 0.0f;
        if (hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream) !=
                hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess) {
          rc = FVS_HIP_ERROR;
          break;
        }
        loss_sum = (b == 0) ? l : 0.9 * loss_sum + 0.1 * l;
      }
    }
    if (loss_history) loss_history[e] = static_cast<float>(loss_sum);
  }
  (void)hipStreamSynchronize(m->stream);
  (void)hipFree(perm);
  if (rc != FVS_OK) return rc;
  if (int rc = Pack(m)) return rc;
  FVS_HIP_CHECK(hipStreamSynchronize(m->stream));
  return FVS_OK;
}

// Matrix-core (BF16 MFMA) inference on complete state pairs.
int fvs_mlp_infer_pairs(fvs_mlp* m, int dim, int dir, const double* left, const double* right,
                        int64_t n, float* r_out, double* guess) {
  if (!m || !left || !right || n <= 0 || (dim != 2 && dim != 3) || dir < 0 || dir >= dim) {
    SetError("bad argument");
    return FVS_BAD_ARGUMENT;
  }
  const unsigned blocks = unsigned((n + kBlock - 1) / kBlock);
  if (dim == 2)
    hipLaunchKernelGGL(MlpInferPairs<2>, dim3(blocks), dim3(kBlock), 0, m->stream, left, right, n,
                       dir, m->tiles, m->bias, r_out, guess);
  else
    hipLaunchKernelGGL(MlpInferPairs<3>, dim3(blocks), dim3(kBlock), 0, m->stream, left, right, n,
                       dir, m->tiles, m->bias, r_out, guess);
  FVS_LAUNCH_CHECK();
  return FVS_OK;
}

int fvs_mlp_face_guess(fvs_mlp* m, int dim, double gamma, const double* const* cons,
                       const int64_t n[3], int dir, double* guess) {
  if (!m) {
    SetError("null mlp");
    return FVS_BAD_ARGUMENT;
  }
  return MlpFaceGuess(m, m->stream, dim, gamma, cons, n, dir, guess);
}

} // extern "C"
