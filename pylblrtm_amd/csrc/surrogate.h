// Learned surrogate of the Voigt function K(x, y) (no counterpart in the
// reference, whose Voigt is scipy.special.wofz at pylblrtm/forward.py:62).
//
// Network: MLP 2 -> 64 -> 64 -> 64 -> 64 -> 1 with SiLU on the inputs
//   u = 2 asinh(|x|) / asinh(50) - 1,  v = (log10 y + 1) / 3
// (domain |x| <= 50, -4 <= log10 y <= 2).  K is even in x, and asinh spreads
// the Doppler core (|x| < ~3, 3% of a linear scale) over half of u.
// output t ~ log(K / K_ref) with the cheap positive reference
//   K_ref(x, y) = exp(-x^2) / (1 + sqrt(pi) y) + y / (sqrt(pi) (x^2 + y^2 + 1/2)),
// so K = K_ref * exp(t) and |t| stays below ~0.7 over the whole domain.
// The hidden 64x64 layers run on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with
// FP32 accumulation; the 2->64 and 64->1 layers are FP32 VALU.
//
// Parameter vector (FP32 master copy), row-major [out][in]:
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbl {
namespace mlp {

constexpr int H = 64;
constexpr int kW1 = 0;                 // [64][2]
constexpr int kB1 = kW1 + H * 2;       // [64]
constexpr int kW2 = kB1 + H;           // [64][64]
constexpr int kB2 = kW2 + H * H;
constexpr int kW3 = kB2 + H;
constexpr int kB3 = kW3 + H * H;
constexpr int kW4 = kB3 + H;
constexpr int kB4 = kW4 + H * H;
constexpr int kW5 = kB4 + H;           // [64]
constexpr int kB5 = kW5 + H;           // [1]
constexpr int kNumParams = kB5 + 1;    // 12,737
constexpr int kHiddenW[3] = {kW2, kW3, kW4};
constexpr int kHiddenB[3] = {kB2, kB3, kB4};

constexpr float kInvAsinh50 = 0.21714253f;   // 1 / asinh(50)
constexpr float kXMax = 50.0f;
constexpr float kLyMin = -4.0f, kLyMax = 2.0f;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 with round-to-nearest-even (inputs are finite).
__device__ __forceinline__ unsigned short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float input_u(float x) {
  return 2.0f * asinhf(fabsf(x)) * kInvAsinh50 - 1.0f;
}
__device__ __forceinline__ float input_v(float log10y) { return (log10y + 1.0f) * (1.0f / 3.0f); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

__device__ __forceinline__ float kref(float x, float y) {
  const float sp = 1.7724538509055159f;
  return __expf(-x * x) / (1.0f + sp * y) + y / (sp * (x * x + y * y + 0.5f));
}

// MFMA 16x16x16 BF16 operand layout (lane l, element i = 0..3):
//   A[m][k]: m = l % 16, k = 4 (l / 16) + i
//   B[k][n]: k = 4 (l / 16) + i, n = l % 16
//   D[m][n]: m = 4 (l / 16) + i, n = l % 16
// With activations stored as [feature][point], D of one layer is exactly the
// B operand of the next (feature block = k block), so the hidden stack runs
// in registers.

// Evaluate the network for the 64 points of a wavefront (lane l holds point
// l's inputs u, v) and return t for lane l's point.  `wbf` = BF16 copies of
// the hidden weights, row-major [3][64][64]; `p` = FP32 parameters.
// Every lane of the wavefront must be active.
__device__ inline float wave_eval(float u, float v, const float* __restrict__ p,
                                  const unsigned short* __restrict__ wbf) {
  const int l = threadIdx.x & 63;
  const int lr = l & 15, lq = l >> 4;
  // inputs of the 4 points this lane feeds into B fragments (n = nt*16 + lr)
  float pu[4], pv[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    pu[nt] = __shfl(u, nt * 16 + lr);
    pv[nt] = __shfl(v, nt * 16 + lr);
  }
  bf16x4 B[4][4];  // [k block][n block]
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int f = kk * 16 + 4 * lq + i;
      const float w0 = p[kW1 + 2 * f], w1 = p[kW1 + 2 * f + 1], b = p[kB1 + f];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        B[kk][nt][i] = (short)f2bf(silu(fmaf(w0, pu[nt], fmaf(w1, pv[nt], b))));
    }
  }
#pragma unroll 1
  for (int layer = 0; layer < 3; ++layer) {
    const unsigned short* W = wbf + layer * H * H;
    const float* bias = p + kHiddenB[layer];
    f32x4 acc[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(W + (mt * 16 + lr) * H + kk * 16 + 4 * lq);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, B[kk][nt], acc[mt][nt], 0, 0, 0);
      }
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float b = bias[mt * 16 + 4 * lq + i];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) B[mt][nt][i] = (short)f2bf(silu(acc[mt][nt][i] + b));
      }
  }
  // output layer 64 -> 1: each lane holds 16 features of 4 points
  float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float w = p[kW5 + kk * 16 + 4 * lq + i];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) part[nt] = fmaf(w, bf2f((unsigned short)B[kk][nt][i]), part[nt]);
    }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    part[nt] += __shfl_xor(part[nt], 16);
    part[nt] += __shfl_xor(part[nt], 32);
  }
  float t = part[0];
#pragma unroll
  for (int nt = 1; nt < 4; ++nt) t = (lq == nt) ? part[nt] : t;
  return t + p[kB5];
}

}  // namespace mlp
}  // namespace lbl
