// Training and stand-alone inference of the Voigt surrogate (see surrogate.h).
//
// Training data: n seeded samples x ~ U[-50, 50], log10 y ~ U[-4, 2]
// (splitmix64 counter hash of (seed, index)), labelled with the FP64 device
// Faddeeva function: t = log(K_fp64(x, y) / K_ref(x, y)).
//
// One training step over a batch of B samples (B a multiple of 64):
//   k_mlp_fwd_bwd: one wavefront per 64 samples; forward (layer 1 VALU,
//     hidden layers BF16 MFMA with FP32 accumulation, output VALU on the
//     BF16 activations exactly as at inference), loss 1/B sum (out - t)^2,
//     backward with the same MFMA tiles (activation gradients W^T dZ and
//     weight gradients dZ H^T), per-wavefront gradient written to its own row
//     of a partial buffer (no atomics: the step is deterministic);
//   k_mlp_reduce: sums the rows;
//   k_mlp_adam: Adam on the FP32 master weights, then BF16 copies of the
//     hidden weights for the MFMA operands.
#include <math.h>

#include "faddeeva.h"
#include "internal.h"
#include "surrogate.h"

namespace lbl {
namespace mlp {
namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double unit(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

__global__ void k_make_data(int64_t n, uint64_t seed, float* x, float* ly, float* t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t h1 = splitmix64(seed ^ splitmix64((uint64_t)(2 * i)));
  const uint64_t h2 = splitmix64(seed ^ splitmix64((uint64_t)(2 * i + 1)));
  const double xd = -50.0 + 100.0 * unit(h1);
  const double lyd = -4.0 + 6.0 * unit(h2);
  const float xf = (float)xd, lyf = (float)lyd;
  const double y = exp10((double)lyf);
  const double xx = (double)xf;
  const double K = K_fp64(xx, y);
  const double sp = 1.7724538509055160273;
  const double kr = exp(-xx * xx) / (1.0 + sp * y) + y / (sp * (xx * xx + y * y + 0.5));
  x[i] = xf;
  ly[i] = lyf;
  t[i] = (float)log(K / kr);
}

// Per-wavefront scratch in LDS for the training step (one wavefront per
// workgroup): activations [feature][sample].
struct TrainLds {
  float z[4][H][H];             // pre-activations of the 4 hidden layers
  unsigned short h[4][H][H];    // BF16 activations
  float dz[H][H];               // current layer's pre-activation gradient
  float u[H], v[H], g[H];
};

__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__global__ __launch_bounds__(64) void k_fwd_bwd(const float* __restrict__ p,
                                                const unsigned short* __restrict__ wbf,
                                                const float* __restrict__ dx,
                                                const float* __restrict__ dly,
                                                const float* __restrict__ dt, int64_t n_data,
                                                uint64_t perm_a, uint64_t perm_b, int64_t base,
                                                float inv_batch, float* __restrict__ partial,
                                                float* __restrict__ loss_partial) {
  __shared__ TrainLds S;
  const int l = threadIdx.x;
  const int lr = l & 15, lq = l >> 4;
  // sample index: an affine bijection of the data set (perm_a odd, n_data a power of two)
  const uint64_t j = (uint64_t)(base + (int64_t)blockIdx.x * 64 + l);
  const int64_t idx = (int64_t)((perm_a * j + perm_b) & (uint64_t)(n_data - 1));
  const float u = input_u(dx[idx]);
  const float v = input_v(dly[idx]);
  const float tgt = dt[idx];
  S.u[l] = u;
  S.v[l] = v;
  // layer 1 (VALU): lane = sample
  for (int f = 0; f < H; ++f) {
    const float z = fmaf(p[kW1 + 2 * f], u, fmaf(p[kW1 + 2 * f + 1], v, p[kB1 + f]));
    S.z[0][f][l] = z;
    S.h[0][f][l] = f2bf(silu(z));
  }
  __syncthreads();
  // hidden layers 2..4 (MFMA): D[f][s] = W[f][k] H[k][s]
  for (int layer = 0; layer < 3; ++layer) {
    const unsigned short* W = wbf + layer * H * H;
    f32x4 acc[4][4];
    for (int mt = 0; mt < 4; ++mt)
      for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int mt = 0; mt < 4; ++mt)
      for (int kk = 0; kk < 4; ++kk) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(W + (mt * 16 + lr) * H + kk * 16 + 4 * lq);
        for (int nt = 0; nt < 4; ++nt) {
          bf16x4 b;
          for (int i = 0; i < 4; ++i) b[i] = (short)S.h[layer][kk * 16 + 4 * lq + i][nt * 16 + lr];
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[mt][nt], 0, 0, 0);
        }
      }
    const float* bias = p + kHiddenB[layer];
    for (int mt = 0; mt < 4; ++mt)
      for (int nt = 0; nt < 4; ++nt)
        for (int i = 0; i < 4; ++i) {
          const int f = mt * 16 + 4 * lq + i, s = nt * 16 + lr;
          const float z = acc[mt][nt][i] + bias[f];
          S.z[layer + 1][f][s] = z;
          S.h[layer + 1][f][s] = f2bf(silu(z));
        }
    __syncthreads();
  }
  // output layer on the BF16 activations (as at inference)
  float out = p[kB5];
  for (int f = 0; f < H; ++f) out = fmaf(p[kW5 + f], bf2f(S.h[3][f][l]), out);
  const float err = out - tgt;
  const float g = 2.0f * err * inv_batch;
  S.g[l] = g;
  float* row = partial + (int64_t)blockIdx.x * kNumParams;
  float lsum = err * err;
  for (int off = 32; off >= 1; off >>= 1) lsum += __shfl_xor(lsum, off);
  float gsum = g;
  for (int off = 32; off >= 1; off >>= 1) gsum += __shfl_xor(gsum, off);
  if (l == 0) {
    loss_partial[blockIdx.x] = lsum;
    row[kB5] = gsum;
  }
  __syncthreads();
  // dW5[f] = sum_s g_s h4[f][s]; dZ4[f][s] = W5[f] g_s silu'(z4)
  {
    const int f = l;
    float a = 0.0f;
    for (int s = 0; s < H; ++s) a = fmaf(S.g[s], bf2f(S.h[3][f][s]), a);
    row[kW5 + f] = a;
  }
  for (int f = 0; f < H; ++f) S.dz[f][l] = p[kW5 + f] * g * dsilu(S.z[3][f][l]);
  __syncthreads();
  // hidden layers backward: layer index 2 (W4), 1 (W3), 0 (W2)
  for (int layer = 2; layer >= 0; --layer) {
    float* dW = row + kHiddenW[layer];
    float* db = row + kHiddenB[layer];
    {
      const int f = l;
      float a = 0.0f;
      for (int s = 0; s < H; ++s) a += S.dz[f][s];
      db[f] = a;
    }
    // dW[m][n] = sum_s dZ[m][s] H_prev[n][s]   (A = dZ, B[k=s][n] = H_prev[n][s])
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc[4];
      for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
      for (int kk = 0; kk < 4; ++kk) {
        bf16x4 a;
        for (int i = 0; i < 4; ++i) a[i] = (short)f2bf(S.dz[mt * 16 + lr][kk * 16 + 4 * lq + i]);
        for (int nt = 0; nt < 4; ++nt) {
          bf16x4 b;
          for (int i = 0; i < 4; ++i) b[i] = (short)S.h[layer][nt * 16 + lr][kk * 16 + 4 * lq + i];
          acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
        }
      }
      for (int nt = 0; nt < 4; ++nt)
        for (int i = 0; i < 4; ++i) dW[(mt * 16 + 4 * lq + i) * H + nt * 16 + lr] = acc[nt][i];
    }
    // dH_prev[k][s] = sum_m W[m][k] dZ[m][s]   (A[k][m] = W[m][k], B[m][s] = dZ[m][s])
    const unsigned short* W = wbf + layer * H * H;
    f32x4 acc[4][4];
    for (int mt = 0; mt < 4; ++mt)
      for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int mt = 0; mt < 4; ++mt)
      for (int kk = 0; kk < 4; ++kk) {
        bf16x4 a;
        for (int i = 0; i < 4; ++i) a[i] = (short)W[(kk * 16 + 4 * lq + i) * H + mt * 16 + lr];
        for (int nt = 0; nt < 4; ++nt) {
          bf16x4 b;
          for (int i = 0; i < 4; ++i) b[i] = (short)f2bf(S.dz[kk * 16 + 4 * lq + i][nt * 16 + lr]);
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[mt][nt], 0, 0, 0);
        }
      }
    __syncthreads();
    for (int mt = 0; mt < 4; ++mt)
      for (int nt = 0; nt < 4; ++nt)
        for (int i = 0; i < 4; ++i) {
          const int f = mt * 16 + 4 * lq + i, s = nt * 16 + lr;
          S.dz[f][s] = acc[mt][nt][i] * dsilu(S.z[layer][f][s]);
        }
    __syncthreads();
  }
  // layer 1: dW1[f][0] = sum_s dZ1[f][s] u_s, dW1[f][1] = ... v_s, db1[f] = sum_s dZ1[f][s]
  {
    const int f = l;
    float a0 = 0.f, a1 = 0.f, ab = 0.f;
    for (int s = 0; s < H; ++s) {
      const float d = S.dz[f][s];
      a0 = fmaf(d, S.u[s], a0);
      a1 = fmaf(d, S.v[s], a1);
      ab += d;
    }
    row[kW1 + 2 * f] = a0;
    row[kW1 + 2 * f + 1] = a1;
    row[kB1 + f] = ab;
  }
}

__global__ void k_reduce(const float* __restrict__ partial, int n_rows, float* __restrict__ grad,
                         const float* __restrict__ loss_partial, float* __restrict__ loss) {
  const int pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx < kNumParams) {
    float a = 0.0f;
    for (int r = 0; r < n_rows; ++r) a += partial[(int64_t)r * kNumParams + pidx];
    grad[pidx] = a;
  }
  if (pidx == 0) {
    double a = 0.0;
    for (int r = 0; r < n_rows; ++r) a += loss_partial[r];
    loss[0] = (float)a;
  }
}

__global__ void k_pack(const float* __restrict__ p, unsigned short* __restrict__ wbf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * H * H) return;
  const int layer = i / (H * H), r = i - layer * H * H;
  wbf[i] = f2bf(p[kHiddenW[layer] + r]);
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNumParams) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(64) void k_eval(int64_t n, const float* __restrict__ x,
                                             const float* __restrict__ ly,
                                             const float* __restrict__ p,
                                             const unsigned short* __restrict__ wbf,
                                             float* __restrict__ t) {
  const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const bool ok = i < n;
  const float u = ok ? input_u(x[i]) : 0.0f;
  const float v = ok ? input_v(ly[i]) : 0.0f;
  const float r = wave_eval(u, v, p, wbf);   // all 64 lanes active
  if (ok) t[i] = r;
}

}  // namespace
}  // namespace mlp

int mlp_num_params() { return mlp::kNumParams; }

hipError_t launch_mlp_make_data(int64_t n, uint64_t seed, float* x, float* ly, float* t,
                                hipStream_t s) {
  if (n <= 0) return hipSuccess;
  mlp::k_make_data<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, seed, x, ly, t);
  return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* p, unsigned short* wbf, hipStream_t s) {
  mlp::k_pack<<<(3 * mlp::H * mlp::H + 255) / 256, 256, 0, s>>>(p, wbf);
  return hipGetLastError();
}

hipError_t launch_mlp_step(float* p, float* m, float* v, unsigned short* wbf, const float* x,
                           const float* ly, const float* t, int64_t n_data, uint64_t perm_a,
                           uint64_t perm_b, int64_t base, int batch, float lr, int adam_t,
                           float* partial, float* loss_partial, float* grad, float* loss,
                           hipStream_t s) {
  const int rows = batch / 64;
  mlp::k_fwd_bwd<<<rows, 64, 0, s>>>(p, wbf, x, ly, t, n_data, perm_a, perm_b, base,
                                     1.0f / (float)batch, partial, loss_partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mlp::k_reduce<<<(mlp::kNumParams + 255) / 256, 256, 0, s>>>(partial, rows, grad, loss_partial,
                                                              loss);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (lr > 0.0f) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)adam_t), c2 = 1.0f - powf(b2, (float)adam_t);
    mlp::k_adam<<<(mlp::kNumParams + 255) / 256, 256, 0, s>>>(p, m, v, grad, lr, b1, b2, 1e-8f,
                                                              c1, c2);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_mlp_pack(p, wbf, s);
  }
  return hipSuccess;
}

hipError_t launch_mlp_eval(int64_t n, const float* x, const float* ly, const float* p,
                           const unsigned short* wbf, float* t, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  mlp::k_eval<<<(unsigned)((n + 63) / 64), 64, 0, s>>>(n, x, ly, p, wbf, t);
  return hipGetLastError();
}

}  // namespace lbl
