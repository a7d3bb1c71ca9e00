// Per-mode corrector-guess network: training and inference kernels.
//
// The network (no reference counterpart) maps 16 features of one spectral
// mode to 6 outputs: x -> silu(W1^T x + b1) -> silu(W2^T h1 + b2) ->
// W3^T h2 + b3, hidden width H (32 or 64).  Parameters are kept in FP32
// (packed: W1 (16,H), b1 (H), W2 (H,H), b2 (H), W3 (H,6), b3 (6)); inference
// feeds the matrix cores bf16 weights and bf16 layer inputs with FP32
// accumulation (v_mfma_f32_16x16x16_bf16).
//
// Inference is computed transposed, Y^T = W^T X^T: weights are the A operand,
// activations the B operand, 16 modes per wavefront.  The 16x16 FP32
// accumulator of one layer (lane l: rows 4(l/16)+j, column l%16) has exactly
// the B-operand layout the next layer needs for that K chunk, so layers chain
// in registers with no exchange.
//
// Training runs one sample per thread in FP32 (forward with the same bf16
// rounding as inference, straight-through gradients), keeps the per-sample
// layer inputs and pre-activation gradients in LDS, and accumulates weight
// gradients per block in a fixed order; a second kernel sums the block
// partials in order (deterministic), and Adam updates the FP32 parameters.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "context.hpp"
#include "dft.hpp"

namespace dnsmi {

constexpr int kDin = 16, kDout = 6;

__host__ __device__ constexpr int mlp_nparams(int H) {
  return kDin * H + H + H * H + H + H * kDout + kDout;
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
// Inference form: the hardware reciprocal (1 ulp) instead of an IEEE divide;
// its error is far below the bf16 rounding of the activations that follows.
__device__ __forceinline__ float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ uint16_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + (((u >> 16) & 1u) + 0x7FFFu)) >> 16);
}
__device__ __forceinline__ float bf16_round(float f) {
  return __uint_as_float(((uint32_t)bf16_bits(f)) << 16);
}

// ------------------------------------------------------------------ training
// One block = S samples (one per thread).  Inputs X (B, 16), targets T (B, 6).
// Row `block` of grad_part (P + 1 floats) receives the block's sum of dL/dp
// in its first P entries and its sum of squared errors in entry P.
// L = (1/B) sum (y - t)^2.
template <int H, int S>
__global__ void __launch_bounds__(S)
    k_mlp_grad_block(const float* __restrict__ X, const float* __restrict__ T, int B,
                     const float* __restrict__ prm, int bf16, float* __restrict__ grad_part) {
  constexpr int P = mlp_nparams(H);
  __shared__ float w[P];
  __shared__ float sx[S][kDin], sh1[S][H], sh2[S][H], sdz1[S][H], sdz2[S][H], sdy[S][kDout];
  for (int i = threadIdx.x; i < P; i += S) w[i] = bf16 && (i < kDin * H || (i >= kDin * H + H && i < kDin * H + H + H * H) || (i >= kDin * H + 2 * H + H * H && i < kDin * H + 2 * H + H * H + H * kDout)) ? bf16_round(prm[i]) : prm[i];
  __syncthreads();
  const float* W1 = w;
  const float* b1 = W1 + kDin * H;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* W3 = b2 + H;
  const float* b3 = W3 + H * kDout;
  const int t = threadIdx.x;
  const int smp = blockIdx.x * S + t;
  const bool live = smp < B;
  float x[kDin], z1[H], z2[H], y[kDout];
#pragma unroll
  for (int k = 0; k < kDin; ++k) {
    const float v = live ? X[(size_t)smp * kDin + k] : 0.f;
    x[k] = bf16 ? bf16_round(v) : v;
    sx[t][k] = x[k];
  }
#pragma unroll 4
  for (int h = 0; h < H; ++h) {
    float a = b1[h];
#pragma unroll
    for (int k = 0; k < kDin; ++k) a += x[k] * W1[k * H + h];
    z1[h] = a;
    const float s = silu(a);
    sh1[t][h] = bf16 ? bf16_round(s) : s;
  }
#pragma unroll 4
  for (int o = 0; o < H; ++o) {
    float a = b2[o];
    for (int h = 0; h < H; ++h) a += sh1[t][h] * W2[h * H + o];
    z2[o] = a;
    const float s = silu(a);
    sh2[t][o] = bf16 ? bf16_round(s) : s;
  }
  float sq = 0.f;
#pragma unroll
  for (int o = 0; o < kDout; ++o) {
    float a = b3[o];
    for (int h = 0; h < H; ++h) a += sh2[t][h] * W3[h * kDout + o];
    y[o] = a;
    const float d = live ? a - T[(size_t)smp * kDout + o] : 0.f;
    sq += d * d;
    sdy[t][o] = live ? 2.0f * d / (float)B : 0.f;
  }
  // backward of the activations
#pragma unroll 4
  for (int h = 0; h < H; ++h) {
    float a = 0.f;
#pragma unroll
    for (int o = 0; o < kDout; ++o) a += W3[h * kDout + o] * sdy[t][o];
    sdz2[t][h] = a * dsilu(z2[h]);
  }
#pragma unroll 4
  for (int h = 0; h < H; ++h) {
    float a = 0.f;
    for (int o = 0; o < H; ++o) a += W2[h * H + o] * sdz2[t][o];
    sdz1[t][h] = a * dsilu(z1[h]);
  }
  __syncthreads();
  // parameter gradients: each thread owns a strided set, sums samples in order
  float* gp = grad_part + (size_t)blockIdx.x * (P + 1);
  for (int p = t; p < P; p += S) {
    float g = 0.f;
    if (p < kDin * H) {
      const int k = p / H, h = p % H;
      for (int s = 0; s < S; ++s) g += sx[s][k] * sdz1[s][h];
    } else if (p < kDin * H + H) {
      const int h = p - kDin * H;
      for (int s = 0; s < S; ++s) g += sdz1[s][h];
    } else if (p < kDin * H + H + H * H) {
      const int q = p - (kDin * H + H), h = q / H, o = q % H;
      for (int s = 0; s < S; ++s) g += sh1[s][h] * sdz2[s][o];
    } else if (p < kDin * H + 2 * H + H * H) {
      const int o = p - (kDin * H + H + H * H);
      for (int s = 0; s < S; ++s) g += sdz2[s][o];
    } else if (p < kDin * H + 2 * H + H * H + H * kDout) {
      const int q = p - (kDin * H + 2 * H + H * H), h = q / kDout, o = q % kDout;
      for (int s = 0; s < S; ++s) g += sh2[s][h] * sdy[s][o];
    } else {
      const int o = p - (kDin * H + 2 * H + H * H + H * kDout);
      for (int s = 0; s < S; ++s) g += sdy[s][o];
    }
    gp[p] = g;
  }
  // block loss (fixed order)
  __shared__ float red[S];
  red[t] = sq;
  __syncthreads();
  if (t == 0) {
    float s = 0.f;
    for (int i = 0; i < S; ++i) s += red[i];
    gp[P] = s;
  }
}

// Deterministic two-level sum of the per-block rows (P parameters and the
// loss, Q = P + 1 columns): chunk q sums blocks [32 q, 32 q + 32) in order
// into part2[q][p]; the final pass sums the chunks in order.
constexpr int kReduceChunk = 32;

__global__ void k_mlp_reduce(const float* __restrict__ part, int nblocks, int Q,
                             float* __restrict__ part2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Q) return;
  const int b0 = blockIdx.y * kReduceChunk;
  const int b1 = b0 + kReduceChunk < nblocks ? b0 + kReduceChunk : nblocks;
  float g = 0.f;
  for (int b = b0; b < b1; ++b) g += part[(size_t)b * Q + p];
  part2[(size_t)blockIdx.y * Q + p] = g;
}

__global__ void k_mlp_reduce_final(const float* __restrict__ part2, int nchunks, int P,
                                   float* __restrict__ grad, float* __restrict__ loss, int B) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P) return;
  float g = 0.f;
  for (int q = 0; q < nchunks; ++q) g += part2[(size_t)q * (P + 1) + p];
  if (p < P)
    grad[p] = g;
  else if (loss)
    *loss = g / (float)B;
}

// Adam (Kingma & Ba 2015) on the FP32 parameters.
__global__ void k_mlp_adam(float* __restrict__ prm, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ v, int P, float lr,
                           float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float gi = g[p];
  const float mi = b1 * m[p] + (1 - b1) * gi;
  const float vi = b2 * v[p] + (1 - b2) * gi * gi;
  m[p] = mi;
  v[p] = vi;
  prm[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------------------ packing
// bf16 A operands in lane order: for each 16x16 tile, lane l holds
// W[k0 + 4(l/16) + j][o0 + l%16], j = 0..3 (4 consecutive uint16).
// Layout: W1 tiles (H/16), W2 tiles (H/16 out x H/16 in, out major),
// W3 tiles (H/16, 6 outputs padded to 16); then FP32 biases b1 | b2 | b3.
template <int H>
__global__ void k_mlp_pack(const float* __restrict__ prm, uint16_t* __restrict__ wb,
                           float* __restrict__ bias) {
  constexpr int NH = H / 16;
  const float* W1 = prm;
  const float* b1 = W1 + kDin * H;
  const float* W2 = b1 + H;
  const float* b2 = W2 + H * H;
  const float* W3 = b2 + H;
  const float* b3 = W3 + H * kDout;
  const int ntiles = NH + NH * NH + NH;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // one (tile, lane) per thread
  if (e < ntiles * 64) {
    const int tile = e / 64, l = e % 64;
    for (int j = 0; j < 4; ++j) {
      const int kk = 4 * (l / 16) + j, oo = l % 16;
      float val = 0.f;
      if (tile < NH) {
        val = W1[kk * H + tile * 16 + oo];
      } else if (tile < NH + NH * NH) {
        const int t2 = tile - NH, o = t2 / NH, k = t2 % NH;
        val = W2[(k * 16 + kk) * H + o * 16 + oo];
      } else {
        const int k = tile - NH - NH * NH;
        val = oo < kDout ? W3[(k * 16 + kk) * kDout + oo] : 0.f;
      }
      wb[(size_t)e * 4 + j] = bf16_bits(val);
    }
  }
  if (e < H) {
    bias[e] = b1[e];
    bias[H + e] = b2[e];
  }
  if (e < 16) bias[2 * H + e] = e < kDout ? b3[e] : 0.f;
}

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4s pack_b(float a0, float a1, float a2, float a3) {
  v4s r;
  r[0] = (short)bf16_bits(a0);
  r[1] = (short)bf16_bits(a1);
  r[2] = (short)bf16_bits(a2);
  r[3] = (short)bf16_bits(a3);
  return r;
}

// The three layers for the 16 samples of this wave: xb = the wave's B
// operand of layer 1 (lane: features 4(l/16)..+3 of sample l%16).  Returns
// the lane's accumulator of the output layer (outputs 4(l/16)+j, sample l%16).
template <int H>
__device__ __forceinline__ v4f mlp_wave(v4s xb, const v4s* __restrict__ wt,
                                        const float* __restrict__ bias, int lane) {
  constexpr int NH = H / 16;
  const int g = lane / 16;
  v4s h1[NH];
#pragma unroll
  for (int o = 0; o < NH; ++o) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wt[o * 64 + lane], xb, acc, 0, 0, 0);
    float a[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = silu_fast(acc[j] + bias[o * 16 + 4 * g + j]);
    h1[o] = pack_b(a[0], a[1], a[2], a[3]);
  }
  v4s h2[NH];
#pragma unroll
  for (int o = 0; o < NH; ++o) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < NH; ++k)
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wt[(NH + o * NH + k) * 64 + lane], h1[k], acc,
                                                      0, 0, 0);
    float a[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = silu_fast(acc[j] + bias[H + o * 16 + 4 * g + j]);
    h2[o] = pack_b(a[0], a[1], a[2], a[3]);
  }
  v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < NH; ++k)
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wt[(NH + NH * NH + k) * 64 + lane], h2[k], acc,
                                                    0, 0, 0);
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] += bias[2 * H + 4 * g + j];
  return acc;
}

// Plain inference on X (B, 16) -> Y (B, 6): for checking the matrix-core
// path against the oracle.
template <int H>
__global__ void __launch_bounds__(256)
    k_mlp_forward(const float* __restrict__ X, int B, const uint16_t* __restrict__ wb,
                  const float* __restrict__ bias, float* __restrict__ Y) {
  constexpr int NT = 2 * (H / 16) + (H / 16) * (H / 16);
  __shared__ v4s wt[NT * 64];
  __shared__ float bs[2 * H + 16];
  for (int i = threadIdx.x; i < NT * 64; i += blockDim.x) wt[i] = reinterpret_cast<const v4s*>(wb)[i];
  for (int i = threadIdx.x; i < 2 * H + 16; i += blockDim.x) bs[i] = bias[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s = (blockIdx.x * 4 + wave) * 16 + (lane % 16);
  const int g = lane / 16;
  float f[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) f[j] = (s < B) ? X[(size_t)s * kDin + 4 * g + j] : 0.f;
  const v4f y = mlp_wave<H>(pack_b(f[0], f[1], f[2], f[3]), wt, bs, lane);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int o = 4 * g + j;
    if (s < B && o < kDout) Y[(size_t)s * kDout + o] = y[j];
  }
}

// ------------------------------------------------------------------ features
// History: N^{n-1}, N^{n-2} kept in FP32 (interleaved re/im).  With
//   D1 = N^n - N^{n-1},  D2 = N^n - 2 N^{n-1} + N^{n-2},
// the quadratic extrapolation of the nonlinear term over one step is
//   N^{n+1} - N^n ~ D1 + D2 = 2 N^n - 3 N^{n-1} + N^{n-2}.
// Features of periodic mode i (flat index over (ky, kz, kx)):
//   0..5  D1 components (re, im) * s_1
//   6..11 D2 components * s_2
//   12    min(dt |k|^2 / Re, 8) ; 13..15 kx/kmax, ky/kmax, kz/kmax
// Target (training): w(k) (N^{n+1} - 3 N^n + 3 N^{n-1} - N^{n-2}) / s_out,
// the residual of the quadratic extrapolation, weighted by
// w = ildt_2 / dt (in (0, 1]) so that the squared loss weights each mode as
// the corrector's error norm does.
// A wall-bounded point (kz, kx, y) has the same features with y (in [-1, 1])
// in place of ky/kmax and |k|^2 = kx^2 + kz^2, and the same weight formula
// (not zeroed at the mean mode, whose profile the wall scheme advances).
struct FeatArgs {
  const double* kx;
  const double* ky;  // periodic: ky; wall-bounded: the y grid
  const double* kz;
  int Nz1, Nx1, Ny;
  int walls;
  long long n;  // points per component
  double kmax, dt, re, c;
  float s_1, s_2, s_out;
};

// Coordinates of flat index i: a periodic mode over (ky, kz, kx) gives
// (kx, ky, kz); a wall-bounded point over (kz, kx, y) gives (kx, y, kz).
// I = unsigned where every index fits in 32 bits (the launchers check it).
template <typename I>
__device__ __forceinline__ void coords(const FeatArgs& a, I i, double& kx, double& kb, double& kz) {
  if (a.walls) {
    const I t = i / (I)a.Ny, u = t / (I)a.Nx1;
    kx = a.kx[t - u * (I)a.Nx1];
    kb = a.ky[i - t * (I)a.Ny];
    kz = a.kz[u];
  } else {
    const I t = i / (I)a.Nx1, u = t / (I)a.Nz1;
    kx = a.kx[i - t * (I)a.Nx1];
    kb = a.ky[u];
    kz = a.kz[t - u * (I)a.Nz1];
  }
}

// |k|^2 of the features and the weight.
__device__ __forceinline__ double ksq(const FeatArgs& a, double kx, double kb, double kz) {
  return a.walls ? kx * kx + kz * kz : (kx * kx + kb * kb) + kz * kz;
}

// Features 12..15.
__device__ __forceinline__ void k_features(const FeatArgs& a, double kx, double kb, double kz,
                                           float* f) {
  f[0] = (float)fmin(a.dt * ksq(a, kx, kb, kz) / a.re, 8.0);
  f[1] = (float)(kx / a.kmax);
  f[2] = a.walls ? (float)kb : (float)(kb / a.kmax);
  f[3] = (float)(kz / a.kmax);
}

// ildt_2 / dt for mode i (0 for the mean mode, which the scheme holds fixed).
__device__ __forceinline__ double mode_weight(const FeatArgs& a, double lapl, bool mean) {
  return mean ? 0.0 : 1.0 / (1.0 - a.dt * a.c * lapl / a.re);
}

// D1 and D2 of one real component from N^n (FP64) and the FP32 history.
__device__ __forceinline__ void diffs(double cur, float h1, float h2, double& d1, double& d2) {
  const double p1 = (double)h1, p2 = (double)h2;
  d1 = cur - p1;
  d2 = (cur - 2.0 * p1) + p2;
}

// Training samples from successive nonlinear terms (FP64 spectral fields);
// the older ones are rounded to FP32 as the history keeps them in the
// features.  With nm3 the target is the residual of cubic extrapolation,
// N^{n+1} - (4 N^n - 6 N^{n-1} + 4 N^{n-2} - N^{n-3}), else of quadratic.
__global__ void k_mlp_features(FeatArgs a, const cx<double>* __restrict__ nm3,
                               const cx<double>* __restrict__ nm2,
                               const cx<double>* __restrict__ nm1,
                               const cx<double>* __restrict__ ncur,
                               const cx<double>* __restrict__ nnext,
                               const long long* __restrict__ modes, int ns,
                               float* __restrict__ X, float* __restrict__ T) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  const long long i = modes[s];
  for (int q = 0; q < 6; ++q) {
    const long long e = (q / 2) * a.n + i;
    const cx<double> v = ncur[e], p1 = nm1[e], p2 = nm2[e];
    double d1, d2;
    diffs((q & 1) ? v.y : v.x, (float)((q & 1) ? p1.y : p1.x), (float)((q & 1) ? p2.y : p2.x), d1, d2);
    X[(size_t)s * kDin + q] = (float)d1 * a.s_1;
    X[(size_t)s * kDin + 6 + q] = (float)d2 * a.s_2;
  }
  double kx, kb, kz;
  coords(a, i, kx, kb, kz);
  float kf[4];
  k_features(a, kx, kb, kz, kf);
  for (int f = 12; f < kDin; ++f) X[(size_t)s * kDin + f] = kf[f - 12];
  if (T && nnext) {
    const double w = mode_weight(a, -ksq(a, kx, kb, kz), !a.walls && i == 0);
    for (int q = 0; q < 6; ++q) {
      const long long e = (q / 2) * a.n + i;
      const cx<double> p2 = nm2[e], p1 = nm1[e], c = ncur[e], x = nnext[e];
      double r;
      if (nm3) {
        const cx<double> p3 = nm3[e];
        r = (q & 1) ? ((x.y - 4 * c.y) + (6 * p1.y - 4 * p2.y) + p3.y)
                    : ((x.x - 4 * c.x) + (6 * p1.x - 4 * p2.x) + p3.x);
      } else {
        r = (q & 1) ? ((x.y - 3 * c.y) + (3 * p1.y - p2.y)) : ((x.x - 3 * c.x) + (3 * p1.x - p2.x));
      }
      T[(size_t)s * kDout + q] = (float)(w * r / a.s_out);
    }
  }
}

// ------------------------------------------------------------------ guess
// ncur = N^n (FP64, from the RHS pipeline); h1 = N^{n-1}, h2 = N^{n-2} and,
// if given, h3 = N^{n-3} (FP32 history).  Per mode, with
//   mode 0: dN = 0                    (the plain predictor)
//   mode 1: dN = D1                   (linear extrapolation of N)
//   mode 3: dN = D1 + D2              (quadratic extrapolation)
//   mode 4: dN = D1 + D2 + D3         (cubic extrapolation; needs h3)
//   mode 2: dN = base + s_out net(x) / w  (learned residual; base cubic
//           with h3, quadratic without)
// and dN Leray-projected (zero at the mean mode):
//   pred     <- u^1 + c ildt_2 dN, u^1 = (u^n ldt_1 + N^n) ildt_2
//   rhs_prev <- N^n + dN                    (so that the incremental
//               corrector from the guess is the same fixed-point map), or
//               with `carry` B = ildt_2 (ldt_1 u^n + (1 - c) N^n), the
//               carry form's corrector input
//   history  <- N^n into the oldest field: h3 when given (the caller then
//               rotates (h1, h2, h3) <- (h3, h1, h2)), else h2 (the caller
//               swaps h1 and h2)
// One lane per mode: each lane reads its mode's N^n and history once
// (coalesced), forms the features and dN; the network runs on the wave's
// 64 modes as four 16-mode matrix-core batches, with operands and outputs
// exchanged through LDS.
template <int H>
__global__ void __launch_bounds__(256)
    k_guess_periodic(FeatArgs a, const cx<double>* __restrict__ state,
                     const cx<double>* __restrict__ ncur, float* __restrict__ h1,
                     float* __restrict__ h2, float* __restrict__ h3, double* __restrict__ pred,
                     double* __restrict__ rhs_prev, const uint16_t* __restrict__ wb,
                     const float* __restrict__ bias, int mode, int carry) {
  constexpr int NT = 2 * (H / 16) + (H / 16) * (H / 16);
  __shared__ v4s wt[NT * 64];
  __shared__ float bs[2 * H + 16];
  __shared__ float feats[4][64][kDin + 1];
  __shared__ float outs[4][64][kDout + 1];
  const bool use_net = mode == 2;
  if (use_net) {
    for (int i = threadIdx.x; i < NT * 64; i += blockDim.x) wt[i] = reinterpret_cast<const v4s*>(wb)[i];
    for (int i = threadIdx.x; i < 2 * H + 16; i += blockDim.x) bs[i] = bias[i];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < a.n;
  const long long ii = live ? i : 0;
  double kx, ky, kz;
  coords(a, (unsigned)ii, kx, ky, kz);
  cx<double> nc[3], d1[3], d2[3], d3[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long long e = q * a.n + ii;
    nc[q] = live ? ncur[e] : cx<double>{0, 0};
    const float2 p1 = live ? reinterpret_cast<const float2*>(h1)[e] : float2{0.f, 0.f};
    const float2 p2 = live ? reinterpret_cast<const float2*>(h2)[e] : float2{0.f, 0.f};
    diffs(nc[q].x, p1.x, p2.x, d1[q].x, d2[q].x);
    diffs(nc[q].y, p1.y, p2.y, d1[q].y, d2[q].y);
    d3[q] = {0.0, 0.0};
    if (h3) {  // D3 = N^n - 3 N^{n-1} + 3 N^{n-2} - N^{n-3}
      const float2 p3 = live ? reinterpret_cast<const float2*>(h3)[e] : float2{0.f, 0.f};
      d3[q] = {((nc[q].x - 3.0 * (double)p1.x) + 3.0 * (double)p2.x) - (double)p3.x,
               ((nc[q].y - 3.0 * (double)p1.y) + 3.0 * (double)p2.y) - (double)p3.y};
    }
  }
  if (use_net) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      feats[wave][lane][2 * q] = (float)d1[q].x * a.s_1;
      feats[wave][lane][2 * q + 1] = (float)d1[q].y * a.s_1;
      feats[wave][lane][6 + 2 * q] = (float)d2[q].x * a.s_2;
      feats[wave][lane][6 + 2 * q + 1] = (float)d2[q].y * a.s_2;
    }
    k_features(a, kx, ky, kz, &feats[wave][lane][12]);
  }
  __syncthreads();  // the weights are loaded by the whole workgroup
  if (use_net) {
    const int g = lane / 16, ml = lane % 16;
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
      const float* fr = feats[wave][sb * 16 + ml];
      const v4f y = mlp_wave<H>(pack_b(fr[4 * g], fr[4 * g + 1], fr[4 * g + 2], fr[4 * g + 3]), wt, bs, lane);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * g + j < kDout) outs[wave][sb * 16 + ml][4 * g + j] = y[j];
    }
  }
  // outs[wave] is written and read by this wavefront only, whose LDS
  // operations complete in issue order
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (!live) return;
  const double lapl = -((kx * kx + ky * ky) + kz * kz);
  const bool mean = (i == 0);
  const double w = mode_weight(a, lapl, mean);
  const double inv_w = mean ? 0.0 : 1.0 - a.dt * a.c * lapl / a.re;  // 1 / w
  cx<double> dn[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (mode == 0)
      dn[q] = {0.0, 0.0};
    else if (mode == 1)
      dn[q] = d1[q];
    else
      dn[q] = {d1[q].x + d2[q].x, d1[q].y + d2[q].y};
    if (mode == 4 || (mode == 2 && h3)) dn[q] = {dn[q].x + d3[q].x, dn[q].y + d3[q].y};
    if (use_net && !mean) {
      dn[q].x += (double)a.s_out * outs[wave][lane][2 * q] * inv_w;
      dn[q].y += (double)a.s_out * outs[wave][lane][2 * q + 1] * inv_w;
    }
  }
  const double il = lapl < 0 ? 1.0 / lapl : 0.0;
  // projection: dn - k (k . dn) / |k|^2 = dn + k (k . dn) * inv_lapl
  const cx<double> kd = {(kx * dn[0].x + ky * dn[1].x) + kz * dn[2].x,
                         (kx * dn[0].y + ky * dn[1].y) + kz * dn[2].y};
  const double kk[3] = {kx, ky, kz};
  const double cil = a.c * a.dt * w;  // c ildt_2
  // the Euler predictor, as the y-forward epilogue forms it
  // (flows/monochromatic.py:320-333): u^1 = (u^n ldt_1 + N^n) ildt_2
  const double ldt1 = mean ? 0.0 : 1.0 / a.dt + (1.0 - a.c) * lapl / a.re;
  const double ildt2 = mean ? 0.0 : 1.0 / (1.0 / a.dt - a.c * lapl / a.re);
  cx<double>* pd = reinterpret_cast<cx<double>*>(pred);
  cx<double>* rp = reinterpret_cast<cx<double>*>(rhs_prev);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    // the mean mode keeps dN = 0: the guess then changes nothing there
    const cx<double> d = mean ? cx<double>{0.0, 0.0}
                              : cx<double>{dn[q].x + kk[q] * kd.x * il, dn[q].y + kk[q] * kd.y * il};
    const long long e = q * a.n + i;
    const cx<double> u = state[e];
    const cx<double> u1 = {(u.x * ldt1 + nc[q].x) * ildt2, (u.y * ldt1 + nc[q].y) * ildt2};
    pd[e] = mode == 0 ? u1 : cx<double>{u1.x + cil * d.x, u1.y + cil * d.y};
    if (carry)  // B = ildt_2 (ldt_1 u^n + (1 - c) N^n), the carry form's corrector input
      rp[e] = {(u.x * ldt1 + (1.0 - a.c) * nc[q].x) * ildt2, (u.y * ldt1 + (1.0 - a.c) * nc[q].y) * ildt2};
    else
      rp[e] = mode == 0 ? nc[q] : cx<double>{nc[q].x + d.x, nc[q].y + d.y};
    // the newest N replaces the oldest in the history (a ring with h3)
    reinterpret_cast<float2*>(h3 ? h3 : h2)[e] = float2{(float)nc[q].x, (float)nc[q].y};
  }
}

// Wall-bounded guess: the nonlinear term fed to the predictor's solve,
//   mode 0: N^n ; 1: N^n + D1 ; 3: N^n + D1 + D2 ;
//   mode 2: N^n + D1 + D2 + s_out net(x) / w     (learned residual)
// per point (kz, kx, y) of the (3, Nz1, Nx1, Ny) layout; h2 <- N^n.  One
// lane per point; the network runs as in k_guess_periodic.
template <int H>
__global__ void __launch_bounds__(256)
    k_extrap_walls(FeatArgs a, const cx<double>* __restrict__ n0, const float2* __restrict__ h1,
                   float2* __restrict__ h2, cx<double>* __restrict__ next,
                   const uint16_t* __restrict__ wb, const float* __restrict__ bias, int mode) {
  constexpr int NT = 2 * (H / 16) + (H / 16) * (H / 16);
  __shared__ v4s wt[NT * 64];
  __shared__ float bs[2 * H + 16];
  __shared__ float feats[4][64][kDin + 1];
  __shared__ float outs[4][64][kDout + 1];
  const bool use_net = mode == 2;
  if (use_net) {
    for (int i = threadIdx.x; i < NT * 64; i += blockDim.x) wt[i] = reinterpret_cast<const v4s*>(wb)[i];
    for (int i = threadIdx.x; i < 2 * H + 16; i += blockDim.x) bs[i] = bias[i];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < a.n;
  const long long ii = live ? i : 0;
  cx<double> nc[3];
  double d1x[3], d1y[3], d2x[3], d2y[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const long long e = q * a.n + ii;
    nc[q] = live ? n0[e] : cx<double>{0, 0};
    const float2 q1 = live ? h1[e] : float2{0.f, 0.f};
    const float2 q2 = live ? h2[e] : float2{0.f, 0.f};
    diffs(nc[q].x, q1.x, q2.x, d1x[q], d2x[q]);
    diffs(nc[q].y, q1.y, q2.y, d1y[q], d2y[q]);
  }
  double kx = 0, kb = 0, kz = 0;
  if (use_net) {
    coords(a, (unsigned)ii, kx, kb, kz);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      feats[wave][lane][2 * q] = (float)d1x[q] * a.s_1;
      feats[wave][lane][2 * q + 1] = (float)d1y[q] * a.s_1;
      feats[wave][lane][6 + 2 * q] = (float)d2x[q] * a.s_2;
      feats[wave][lane][6 + 2 * q + 1] = (float)d2y[q] * a.s_2;
    }
    k_features(a, kx, kb, kz, &feats[wave][lane][12]);
  }
  __syncthreads();  // the weights are loaded by the whole workgroup
  if (use_net) {
    const int g = lane / 16, ml = lane % 16;
#pragma unroll
    for (int sb = 0; sb < 4; ++sb) {
      const float* fr = feats[wave][sb * 16 + ml];
      const v4f y = mlp_wave<H>(pack_b(fr[4 * g], fr[4 * g + 1], fr[4 * g + 2], fr[4 * g + 3]), wt, bs, lane);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * g + j < kDout) outs[wave][sb * 16 + ml][4 * g + j] = y[j];
    }
  }
  // outs[wave] is written and read by this wavefront only
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (!live) return;
  const double inv_w = 1.0 - a.dt * a.c * (-ksq(a, kx, kb, kz)) / a.re;  // 1 / w
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    cx<double> o = nc[q];
    if (mode == 1) o = {nc[q].x + d1x[q], nc[q].y + d1y[q]};
    else if (mode >= 2) o = {nc[q].x + (d1x[q] + d2x[q]), nc[q].y + (d1y[q] + d2y[q])};
    if (use_net) {
      o.x += (double)a.s_out * outs[wave][lane][2 * q] * inv_w;
      o.y += (double)a.s_out * outs[wave][lane][2 * q + 1] * inv_w;
    }
    const long long e = q * a.n + i;
    next[e] = o;
    h2[e] = float2{(float)nc[q].x, (float)nc[q].y};
  }
}

static FeatArgs feat_args(Ctx* c, float s_1, float s_2, float s_out);

int extrap_walls(Ctx* c, const double* n0, const float* h1, float* h2, double* next,
                 const uint16_t* wb, const float* bias, int H, float s_1, float s_2, float s_out,
                 int mode, hipStream_t s) {
  const FeatArgs a = feat_args(c, s_1, s_2, s_out);
  if (a.n >= (1LL << 31)) return c->fail(DNSMI_EINVAL, "guess: too many points for 32-bit indexing");
  const int nb = (int)((a.n + 255) / 256);
  const cx<double>* n = reinterpret_cast<const cx<double>*>(n0);
  const float2* p1 = reinterpret_cast<const float2*>(h1);
  float2* p2 = reinterpret_cast<float2*>(h2);
  cx<double>* o = reinterpret_cast<cx<double>*>(next);
  if (mode != 2 || H == 32)
    hipLaunchKernelGGL(k_extrap_walls<32>, dim3(nb), dim3(256), 0, s, a, n, p1, p2, o, wb, bias, mode);
  else
    hipLaunchKernelGGL(k_extrap_walls<64>, dim3(nb), dim3(256), 0, s, a, n, p1, p2, o, wb, bias, mode);
  return check_launch(c, "walls_extrap");
}

// ------------------------------------------------------------------ launchers
// Floats of scratch mlp_grad needs for B samples: one row of P + 1 per
// block of 32 samples, then one per chunk of 32 blocks.
long long mlp_grad_scratch(int H, int B) {
  constexpr int S = 32;
  const long long nb = (B + S - 1) / S, nq = (nb + kReduceChunk - 1) / kReduceChunk;
  return (nb + nq) * (mlp_nparams(H) + 1);
}

int mlp_grad(int H, int B, const float* X, const float* T, const float* prm, int bf16, float* part,
             float* grad, float* loss, hipStream_t s) {
  constexpr int S = 32;
  const int nb = (B + S - 1) / S;
  if (H == 32) {
    hipLaunchKernelGGL((k_mlp_grad_block<32, S>), dim3(nb), dim3(S), 0, s, X, T, B, prm, bf16, part);
  } else if (H == 64) {
    hipLaunchKernelGGL((k_mlp_grad_block<64, S>), dim3(nb), dim3(S), 0, s, X, T, B, prm, bf16, part);
  } else {
    return DNSMI_EINVAL;
  }
  const int P = mlp_nparams(H), Q = P + 1;
  const int nq = (nb + kReduceChunk - 1) / kReduceChunk;
  float* part2 = part + (size_t)nb * Q;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((Q + 255) / 256, nq), dim3(256), 0, s, part, nb, Q, part2);
  hipLaunchKernelGGL(k_mlp_reduce_final, dim3((Q + 255) / 256), dim3(256), 0, s, part2, nq, P, grad,
                     loss, B);
  return hipGetLastError() == hipSuccess ? DNSMI_OK : DNSMI_EHIP;
}

// ------------------------------------------------------------------ no-allocation gradient
// mlp_grad with its scratch in a file-scope device array instead of a
// workspace argument, for callers that cannot allocate device memory.  Every
// call on one GPU uses that GPU's copy of the array (the kernels overwrite
// every word they read, so nothing needs clearing); calls that overlap on
// one GPU race.  The array's device address is looked up once per device
// (hipGetSymbolAddress) and only that address reaches the kernels.
constexpr int kStaticMaxSamples = 8192;
constexpr long long kStaticScratchFloats =
    ((kStaticMaxSamples + 31) / 32 + ((kStaticMaxSamples + 31) / 32 + kReduceChunk - 1) / kReduceChunk) *
    (mlp_nparams(64) + 1);
static __device__ __align__(256) float scratch[kStaticScratchFloats];

constexpr int kMaxDevices = 64;
static float* g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

int mlp_grad_static(int H, int B, const float* X, const float* T, const float* prm, int bf16,
                    float* grad, float* loss, hipStream_t s, std::string* why) {
  if (B > kStaticMaxSamples) {
    *why = "mlp_grad_static: at most " + std::to_string(kStaticMaxSamples) + " samples per call";
    return DNSMI_EINVAL;
  }
  if (mlp_grad_scratch(H, B) > kStaticScratchFloats) {
    *why = "mlp_grad_static: scratch too small";
    return DNSMI_EINVAL;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    *why = "mlp_grad_static: no current device";
    return DNSMI_EHIP;
  }
  float* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        *why = "mlp_grad_static: hipGetSymbolAddress failed";
        return DNSMI_EHIP;
      }
      g_scratch_dev[dev] = static_cast<float*>(p);
    }
    ptr = g_scratch_dev[dev];
  }
  return mlp_grad(H, B, X, T, prm, bf16, ptr, grad, loss, s);
}

int mlp_adam(int H, float* prm, const float* g, float* m, float* v, int step, float lr,
             hipStream_t s) {
  const int P = mlp_nparams(H);
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((P + 255) / 256), dim3(256), 0, s, prm, g, m, v, P, lr, b1,
                     b2, 1e-8f, c1, c2);
  return hipGetLastError() == hipSuccess ? DNSMI_OK : DNSMI_EHIP;
}

int mlp_pack(int H, const float* prm, uint16_t* wb, float* bias, hipStream_t s) {
  const int NH = H / 16, nt = 2 * NH + NH * NH;
  const int n = nt * 64 > H ? nt * 64 : H;
  if (H == 32)
    hipLaunchKernelGGL(k_mlp_pack<32>, dim3((n + 255) / 256), dim3(256), 0, s, prm, wb, bias);
  else if (H == 64)
    hipLaunchKernelGGL(k_mlp_pack<64>, dim3((n + 255) / 256), dim3(256), 0, s, prm, wb, bias);
  else
    return DNSMI_EINVAL;
  return hipGetLastError() == hipSuccess ? DNSMI_OK : DNSMI_EHIP;
}

int mlp_forward(int H, const float* X, int B, const uint16_t* wb, const float* bias, float* Y,
                hipStream_t s) {
  const int nb = (B + 63) / 64;
  if (H == 32)
    hipLaunchKernelGGL(k_mlp_forward<32>, dim3(nb), dim3(256), 0, s, X, B, wb, bias, Y);
  else if (H == 64)
    hipLaunchKernelGGL(k_mlp_forward<64>, dim3(nb), dim3(256), 0, s, X, B, wb, bias, Y);
  else
    return DNSMI_EINVAL;
  return hipGetLastError() == hipSuccess ? DNSMI_OK : DNSMI_EHIP;
}

static FeatArgs feat_args(Ctx* c, float s_1, float s_2, float s_out) {
  FeatArgs a{};
  a.kx = c->kx;
  a.ky = c->ky;  // wall-bounded contexts keep the y grid here
  a.kz = c->kz;
  a.Nz1 = c->Nz1;
  a.Nx1 = c->Nx1;
  a.Ny = c->ny;
  a.walls = c->walls ? 1 : 0;
  a.n = (long long)c->Ny1 * c->Nz1 * c->Nx1;  // walls: Ny1 = Ny points
  const double kxm = 2 * kPi / c->cfg.lx * (c->nx / 2), kym = 2 * kPi / c->cfg.ly * (c->ny / 2),
               kzm = 2 * kPi / c->cfg.lz * (c->nz / 2);
  a.kmax = c->walls ? fmax(kxm, kzm) : fmax(kxm, fmax(kym, kzm));
  a.dt = c->cfg.dt;
  a.re = c->cfg.re;
  a.c = c->cfg.implicitness;
  a.s_1 = s_1;
  a.s_2 = s_2;
  a.s_out = s_out;
  return a;
}

int mlp_features(Ctx* c, const double* nm3, const double* nm2, const double* nm1, const double* ncur,
                 const double* nnext, const long long* modes, int ns, float s_1, float s_2,
                 float s_out, float* X, float* T, hipStream_t s) {
  const FeatArgs a = feat_args(c, s_1, s_2, s_out);
  if (ns == 0) return DNSMI_OK;
  hipLaunchKernelGGL(k_mlp_features, dim3((ns + 255) / 256), dim3(256), 0, s, a,
                     reinterpret_cast<const cx<double>*>(nm3), reinterpret_cast<const cx<double>*>(nm2),
                     reinterpret_cast<const cx<double>*>(nm1),
                     reinterpret_cast<const cx<double>*>(ncur),
                     reinterpret_cast<const cx<double>*>(nnext), modes, ns, X, T);
  return check_launch(c, "mlp_features");
}

int periodic_guess_rhs(Ctx* c, const double* state, double* rhs, hipStream_t s);

int guess_periodic(Ctx* c, const double* state, double* pred, double* rhs_prev, float* h1,
                   float* h2, float* h3, const uint16_t* wb, const float* bias, int H, float s_1, float s_2,
                   float s_out, int mode, hipStream_t s, int carry) {
  if (c->walls) return c->fail(DNSMI_EINVAL, "guess: periodic contexts only");
  if (mode < 0 || mode > 4) return c->fail(DNSMI_EINVAL, "guess: mode must be 0 to 4");
  if (mode == 4 && !h3) return c->fail(DNSMI_EINVAL, "guess: cubic extrapolation needs hist3");
  if (mode == 2 && H != 32 && H != 64)
    return c->fail(DNSMI_EINVAL, "guess: hidden width must be 32 or 64");
  // N^n; the first corrector pass from the guess records the split reference
  int rc = periodic_guess_rhs(c, state, c->rhs0, s);
  if (rc) return rc;
  const FeatArgs a = feat_args(c, s_1, s_2, s_out);
  if (a.n >= (1LL << 31)) return c->fail(DNSMI_EINVAL, "guess: too many modes for 32-bit indexing");
  const int nb = (int)((a.n + 255) / 256);
  const cx<double>* n0 = reinterpret_cast<const cx<double>*>(c->rhs0);
  if (mode != 2 || H == 32)
    hipLaunchKernelGGL(k_guess_periodic<32>, dim3(nb), dim3(256), 0, s, a,
                       reinterpret_cast<const cx<double>*>(state), n0, h1, h2, h3, pred, rhs_prev, wb,
                       bias, mode, carry);
  else
    hipLaunchKernelGGL(k_guess_periodic<64>, dim3(nb), dim3(256), 0, s, a,
                       reinterpret_cast<const cx<double>*>(state), n0, h1, h2, h3, pred, rhs_prev, wb,
                       bias, mode, carry);
  return check_launch(c, "guess");
}

}  // namespace dnsmi
