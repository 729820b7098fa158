// Matrix-core inference of the learned Newton guess, shared by the network's
// own kernels and the exact-Godunov kernels that call it inline.
//
// A 6 -> 64 -> 64 -> 1 ReLU MLP maps a state pair to the log of the star
// pressure relative to sqrt(pL pR).  BF16 operands, FP32 accumulation,
// v_mfma_f32_16x16x16_bf16; activations are kept transposed (hidden units x
// faces) so that a layer's accumulator tile is the next layer's B operand.
// The biases seed the accumulators, so a hidden activation costs one max
// and half a packed bf16 conversion per lane.
#pragma once

#include "fvs_physics.hpp"

#include <cstring>

namespace fvs {
namespace mlp {

// FP32 master parameter layout: W1 [64][8] (inputs padded to 8, columns 6
// and 7 stay zero), b1 [64], W2 [64][64], b2 [64], W3 [64], b3.
constexpr int kH = 64;
constexpr int kIn = 6;
constexpr int kInPad = 8;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kInPad;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kNParams = kOffB3 + 1;  // 4801

// Packed inference operands: 24 A tiles of 64 lanes x 4 bf16
// (layer 1: tiles 0..3 [m]; layer 2: 4 + 4 m + kb; layer 3: 20 + kb).
constexpr int kNTiles = 24;
constexpr int kTileShorts = 64 * 4;

constexpr float kLogScale = 0.14476482730108395f;  // 1 / ln(1000)

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ short ToBf16(float f) {
  unsigned int u;
  std::memcpy(&u, &f, 4);
  const unsigned int lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;  // round to nearest even (finite values)
  return static_cast<short>(u >> 16);
}

// Hidden activation: max(x, 0), then the hardware round-to-nearest-even
// bf16 conversion.
__device__ __forceinline__ float Relu(float x) { return __builtin_fmaxf(x, 0.0f); }
__device__ __forceinline__ short Bf16Dev(float f) {
  return __builtin_bit_cast(short, static_cast<__bf16>(f));
}

__device__ __forceinline__ floatx4_t Mfma16(bf16x4_t a, bf16x4_t b, floatx4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

template <int D>
// Network inputs; the logarithms come from the hardware log2 estimate (the
// features are rounded to bf16 before use).
__device__ __forceinline__ void Features(const double* L, const double* R, int dir, float* x) {
  constexpr float kLog2Scale = 0.6931471805599453f * kLogScale;  // ln 2 / ln 1000
  x[0] = __builtin_amdgcn_logf(static_cast<float>(L[0])) * kLog2Scale;
  x[1] = static_cast<float>(L[1 + dir] / L[0]) * 0.2f;
  x[2] = __builtin_amdgcn_logf(static_cast<float>(L[D + 2])) * kLog2Scale;
  x[3] = __builtin_amdgcn_logf(static_cast<float>(R[0])) * kLog2Scale;
  x[4] = static_cast<float>(R[1 + dir] / R[0]) * 0.2f;
  x[5] = __builtin_amdgcn_logf(static_cast<float>(R[D + 2])) * kLog2Scale;
}

// p0 = sqrt(pL pR) exp(r): the geometric mean in FP64, the network's factor
// from the hardware exp2 estimate (a guess needs no more accuracy).
__device__ __forceinline__ double GuessFromOutput(double pL, double pR, float r) {
  return __builtin_sqrt(pL * pR) *
         static_cast<double>(__builtin_amdgcn_exp2f(r * 1.4426950408889634f));
}

// Matrix-core inference for the 64 faces of one wave (lane = face).
// Operand layouts of v_mfma_f32_16x16x16_bf16 per 16x16x16 block:
//   A[i][k] : lane i + 16 (k / 4), element k % 4
//   B[k][j] : lane j + 16 (k / 4), element k % 4
//   D[i][j] : lane j + 16 (i / 4), element i % 4
// The feature transpose into B and the output gather use cross-lane
// shuffles only (no LDS, no barrier), so a wave may skip the call alone.
__device__ __forceinline__ float MlpWave(const float* xf, const short* __restrict__ tiles,
                                         const float* __restrict__ bias) {
  const int lane = threadIdx.x & 63;
  const int kq = lane >> 4;
  const int r16 = lane & 15;
  float result = 0.0f;
  // One tile of 16 faces at a time through all three layers (few live
  // registers, so the calling kernels keep their occupancy).
#pragma unroll 1
  for (int t = 0; t < 4; ++t) {
    // B[k][j] = feature k of face t * 16 + j (k < 6; rows 6..15 are zero)
    const int src = t * 16 + r16;
    bf16x4_t bx;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float lo = __shfl(xf[e], src, 64);
      const float hi = (e < 2) ? __shfl(xf[4 + e], src, 64) : 0.0f;
      const float v = (kq == 0) ? lo : ((kq == 1) ? hi : 0.0f);
      bx[e] = Bf16Dev(v);
    }
    bf16x4_t h1[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const bf16x4_t aw = *reinterpret_cast<const bf16x4_t*>(tiles + (m * 64 + lane) * 4);
      const float* b1 = bias + m * 16 + kq * 4;
      floatx4_t acc = {b1[0], b1[1], b1[2], b1[3]};
      acc = Mfma16(aw, bx, acc);
      bf16x4_t hb;
#pragma unroll
      for (int e = 0; e < 4; ++e) hb[e] = Bf16Dev(Relu(acc[e]));
      h1[m] = hb;
    }
    bf16x4_t h2[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const float* b2 = bias + kH + m * 16 + kq * 4;
      floatx4_t acc = {b2[0], b2[1], b2[2], b2[3]};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        const bf16x4_t aw =
            *reinterpret_cast<const bf16x4_t*>(tiles + ((4 + m * 4 + kb) * 64 + lane) * 4);
        acc = Mfma16(aw, h1[kb], acc);
      }
      bf16x4_t hb;
#pragma unroll
      for (int e = 0; e < 4; ++e) hb[e] = Bf16Dev(Relu(acc[e]));
      h2[m] = hb;
    }
    floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      const bf16x4_t aw = *reinterpret_cast<const bf16x4_t*>(tiles + ((20 + kb) * 64 + lane) * 4);
      acc = Mfma16(aw, h2[kb], acc);
    }
    // output of face t * 16 + j sits in element 0 of lane j
    const float out = __shfl(acc[0], r16, 64) + bias[2 * kH];
    if (kq == t) result = out;
  }
  return result;
}

// Pairs for which the mean pressure already lies close to the star pressure
// (acoustic estimates |pL - pR| and |uR - uL| (rhoL + rhoR)(aL + aR) / 4 both
// below kWeakPair (pL + pR)): Newton from 0.5 (pL + pR) converges in two or
// three steps there, which a BF16 network cannot beat, so they keep the
// reference guess.
constexpr double kWeakPair = 0.1;

__device__ __forceinline__ bool WeakPair(double rhoL, double uL, double pL, double aL,
                                         double rhoR, double uR, double pR, double aR) {
  const double sp = pL + pR;
  const double ep = fabs(pL - pR);
  const double eu = fabs(uR - uL) * (rhoL + rhoR) * (aL + aR) * 0.25;
  return ep < kWeakPair * sp && eu < kWeakPair * sp;
}


// Waves run the network only when at least this many of their 64 faces are
// non-weak.  A wave's Newton loop runs until its slowest face converges, so
// a few strong faces among weak ones gain little from the guess while the
// whole wave pays for the inference.
constexpr int kWaveStrongLanes = 48;

// The learned initial guess of one face per lane; every lane of the wave
// must call it (the matrix-core step is wave-wide).  Weak pairs, and every
// lane of a wave with fewer than kWaveStrongLanes non-weak faces, keep the
// mean pressure.
template <int D, int DIR>
__device__ __forceinline__ double LearnedGuess(const double* L, const double* R,
                                               const short* __restrict__ tiles,
                                               const float* __restrict__ bias) {
  const double rhoL = L[0], rhoR = R[0];
  const double pL = L[D + 2], pR = R[D + 2];
  const bool weak = WeakPair(rhoL, L[1 + DIR] / rhoL, pL, L[D + 3], rhoR, R[1 + DIR] / rhoR, pR,
                             R[D + 3]);
  double p0 = 0.5 * (pL + pR);
  if (__builtin_popcountll(__ballot(!weak)) >= kWaveStrongLanes) {
    float x[kIn];
    Features<D>(L, R, DIR, x);
    // Opaque copies of the operand pointers: the compiler cannot hoist the
    // weight loads out of a caller's loop (they would stay live in registers
    // across the whole march); they are re-read from cache per call instead.
    const short* t = tiles;
    const float* b = bias;
    __asm__ volatile("" : "+s"(t), "+s"(b));
    const float out = MlpWave(x, t, b);
    if (!weak) p0 = GuessFromOutput(pL, pR, out);
  }
  return p0;
}

} // namespace mlp
} // namespace fvs
