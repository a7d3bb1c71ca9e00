// Per-(kz, kx) profile network for the plane Couette corrector guess:
// training and inference on the matrix cores (bf16 operands, FP32
// accumulation, v_mfma_f32_16x16x16_bf16), no reference counterpart.
//
// One sample is one Fourier mode (kz, kx) at one step.  Its input is the
// wall-normal profile of the first and second differences of the nonlinear
// term, D1 = N^n - N^{n-1} and D2 = N^n - 2 N^{n-1} + N^{n-2} (3 components,
// real and imaginary parts, Ny points each: 12 Ny values, each difference
// scaled by its RMS), then dt k^2 / Re, kx / k_max and kz / k_max, zero
// padded to `din` (a multiple of 32).  Its output is the profile of the
// residual of quadratic extrapolation, N(u*) - (3 N^n - 3 N^{n-1} + N^{n-2})
// (6 Ny values) scaled by w = 1 / (1 + c dt k^2 / Re) and 1 / s_out: the
// corrector's fixed point u* = S(u^n; N(u*)) is the converged state of the
// step, so N(u*) = N^{n+1} and the training pairs come from the solver's own
// history.  Network: din -> H -> H -> dout, SiLU, weights FP32 (Adam) and
// bf16 copies for the matrix cores.
//
// Every layer, forward and backward, is one bf16 GEMM (k_gemm_bf16) with a
// fused epilogue: bias + pre-activation store + SiLU (forward), or the SiLU
// derivative (backward); the weight gradients are GEMMs over the batch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "context.hpp"
#include "dft.hpp"

namespace dnsmi {

typedef short pv4s __attribute__((ext_vector_type(4)));
typedef float pv4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t pn_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + (((u >> 16) & 1u) + 0x7FFFu)) >> 16);
}
__device__ __forceinline__ float pn_float(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ float pn_silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float pn_dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// ------------------------------------------------------------------ GEMM
// C (M x N) = op(A) (M x K) op(B) (K x N), bf16 operands, FP32 accumulation.
// op(A)[m][k] = ta ? A[k lda + m] : A[m lda + k]; op(B)[k][n] = tb ? B[n ldb
// + k] : B[k ldb + n].  Epilogue on v = acc (+ bias[n]):
//   EPI_NONE   C = v, Cb = bf16(v)
//   EPI_SILU   C = v (the pre-activation), Cb = bf16(silu(v))
//   EPI_DSILU  C = acc * silu'(Z[m][n]), Cb = bf16 of that (no bias)
// (C or Cb may be null.)  A 64 x 64 tile per workgroup of four waves, each
// wave 2 x 2 MFMA tiles of 16 x 16; K in chunks of 32 through LDS, both
// operands stored k-fastest so each lane's four-element operand is one
// 8-byte LDS read.
enum { EPI_NONE = 0, EPI_SILU = 1, EPI_DSILU = 2 };

struct GemmArgs {
  int M, N, K;
  const uint16_t* A;
  int lda, ta;
  const uint16_t* B;
  int ldb, tb;
  float* C;
  int ldc;
  uint16_t* Cb;
  int ldcb;
  const float* bias;
  const float* Z;
  int ldz;
  int epi;
};

constexpr int kGT = 64, kGK = 32, kGP = 8;

__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ uint16_t As[kGT][kGK + kGP];  // [m][k]
  __shared__ uint16_t Bs[kGT][kGK + kGP];  // [n][k]
  const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  pv4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = pv4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += kGK) {
    // consecutive threads read consecutive memory words
    for (int e = tid; e < kGT * kGK; e += 256) {
      int r, c;
      if (!g.ta) {
        r = e / kGK;
        c = e % kGK;
      } else {
        c = e / kGT;
        r = e % kGT;
      }
      const int m = m0 + r, k = k0 + c;
      uint16_t v = 0;
      if (m < g.M && k < g.K) v = g.ta ? g.A[(size_t)k * g.lda + m] : g.A[(size_t)m * g.lda + k];
      As[r][c] = v;
    }
    for (int e = tid; e < kGT * kGK; e += 256) {
      int r, c;
      if (g.tb) {
        r = e / kGK;
        c = e % kGK;
      } else {
        c = e / kGT;
        r = e % kGT;
      }
      const int n = n0 + r, k = k0 + c;
      uint16_t v = 0;
      if (n < g.N && k < g.K) v = g.tb ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n];
      Bs[r][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGK; kk += 16) {
      pv4s a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = *reinterpret_cast<const pv4s*>(&As[wm + i * 16 + (lane & 15)][kk + 4 * (lane >> 4)]);
        b[i] = *reinterpret_cast<const pv4s*>(&Bs[wn + i * 16 + (lane & 15)][kk + 4 * (lane >> 4)]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // lane holds D[4 (lane / 16) + q][lane % 16] of each 16 x 16 tile
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + q;
        const int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[i][j][q];
        float o;
        if (g.epi == EPI_DSILU) {
          v *= pn_dsilu(g.Z[(size_t)m * g.ldz + n]);
          o = v;
        } else {
          if (g.bias) v += g.bias[n];
          o = g.epi == EPI_SILU ? pn_silu(v) : v;
        }
        if (g.C) g.C[(size_t)m * g.ldc + n] = v;
        if (g.Cb) g.Cb[(size_t)m * g.ldcb + n] = pn_bf16(o);
      }
}

static int gemm(const GemmArgs& g, hipStream_t s) {
  const dim3 grid((g.N + kGT - 1) / kGT, (g.M + kGT - 1) / kGT);
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, s, g);
  return hipGetLastError() == hipSuccess ? DNSMI_OK : DNSMI_EHIP;
}

// ------------------------------------------------------------------ layout
// Parameters (FP32): W1 (din, H) | b1 (H) | W2 (H, H) | b2 (H) | W3 (H, dout) |
// b3 (dout), row-major.  bf16 copies of W1, W2, W3 in the same layout.
struct PNet {
  int din, H, dout;
  __host__ __device__ long long w1() const { return 0; }
  __host__ __device__ long long b1() const { return (long long)din * H; }
  __host__ __device__ long long w2() const { return b1() + H; }
  __host__ __device__ long long b2() const { return w2() + (long long)H * H; }
  __host__ __device__ long long w3() const { return b2() + H; }
  __host__ __device__ long long b3() const { return w3() + (long long)H * dout; }
  __host__ __device__ long long n() const { return b3() + dout; }
  // work floats for a batch of B: Z1, Z2 (B H each), Y / dY (B dout)
  __host__ __device__ long long work_f(int B) const { return (long long)B * (2 * H + dout); }
  // work bf16 words: H1, H2, dZ (B H each), dY (B dout)
  __host__ __device__ long long work_h(int B) const { return (long long)B * (3 * H + dout); }
};

__global__ void k_pn_cast(const float* __restrict__ p, const PNet net, uint16_t* __restrict__ wb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= net.n()) return;
  wb[i] = pn_bf16(p[i]);
}

// dY = 2 (Y - T) / (B dout) in bf16 (the backward GEMMs' operand), and per
// block sums of squares (fixed order) for the loss.
__global__ void __launch_bounds__(256) k_pn_loss(const float* __restrict__ Y, const float* __restrict__ T,
                                                 long long n, float scale, uint16_t* __restrict__ dY,
                                                 float* __restrict__ part) {
  __shared__ float red[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.f;
  if (i < n) {
    const float d = Y[i] - T[i];
    sq = d * d;
    dY[i] = pn_bf16(2.0f * d * scale);
  }
  red[threadIdx.x] = sq;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Column sums of a (B x n) matrix (bf16 values: the backward operand, so
// the bias gradients match the weight gradients' operand), rows in order.
__global__ void k_pn_colsum(const uint16_t* __restrict__ G, int B, int n, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  float s = 0.f;
  for (int r = 0; r < B; ++r) s += pn_float(G[(size_t)r * n + c]);
  out[c] = s;
}

__global__ void k_pn_sum(const float* __restrict__ part, int n, float scale, float* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.f;
  for (int i = 0; i < n; ++i) s += part[i];
  *out = s * scale;
}

__global__ void k_pn_adam(float* __restrict__ prm, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, long long P, float lr, float c1, float c2) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float gi = g[p];
  const float mi = 0.9f * m[p] + 0.1f * gi;
  const float vi = 0.999f * v[p] + 0.001f * gi * gi;
  m[p] = mi;
  v[p] = vi;
  prm[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

static bool pn_ok(const PNet& n) { return n.din > 0 && n.din % 32 == 0 && n.H > 0 && n.H % 64 == 0 && n.dout > 0; }

// Forward pass of B samples: X (B, din) bf16 -> Y (B, dout) FP32; Z1, Z2
// (pre-activations, FP32) and H1, H2 (bf16) kept in the work buffers.
static int pn_forward(const PNet& net, int B, const uint16_t* X, const uint16_t* wb, const float* prm,
                      float* wf, uint16_t* wh, float* Y, hipStream_t s) {
  float* Z1 = wf;
  float* Z2 = Z1 + (size_t)B * net.H;
  uint16_t* H1 = wh;
  uint16_t* H2 = H1 + (size_t)B * net.H;
  GemmArgs g{};
  g = {B, net.H, net.din, X, net.din, 0, wb + net.w1(), net.H, 0, Z1, net.H, H1, net.H, prm + net.b1(),
       nullptr, 0, EPI_SILU};
  int rc = gemm(g, s);
  if (rc) return rc;
  g = {B, net.H, net.H, H1, net.H, 0, wb + net.w2(), net.H, 0, Z2, net.H, H2, net.H, prm + net.b2(),
       nullptr, 0, EPI_SILU};
  if ((rc = gemm(g, s))) return rc;
  g = {B, net.dout, net.H, H2, net.H, 0, wb + net.w3(), net.dout, 0, Y, net.dout, nullptr, 0,
       prm + net.b3(), nullptr, 0, EPI_NONE};
  return gemm(g, s);
}

int pnet_forward(int din, int H, int dout, int B, const uint16_t* X, const float* prm, uint16_t* wb,
                 float* wf, uint16_t* wh, float* Y, hipStream_t s) {
  const PNet net{din, H, dout};
  if (!pn_ok(net) || B <= 0) return DNSMI_EINVAL;
  hipLaunchKernelGGL(k_pn_cast, dim3((unsigned)((net.n() + 255) / 256)), dim3(256), 0, s, prm, net, wb);
  if (hipGetLastError() != hipSuccess) return DNSMI_EHIP;
  return pn_forward(net, B, X, wb, prm, wf, wh, Y, s);
}

// Gradient of L = (1/(B dout)) sum (Y - T)^2 into grad (FP32), loss into
// *loss (may be null).  part: ceil(B dout / 256) floats.
int pnet_grad(int din, int H, int dout, int B, const uint16_t* X, const float* T, const float* prm,
              uint16_t* wb, float* wf, uint16_t* wh, float* part, float* grad, float* loss,
              hipStream_t s) {
  const PNet net{din, H, dout};
  if (!pn_ok(net) || B <= 0) return DNSMI_EINVAL;
  hipLaunchKernelGGL(k_pn_cast, dim3((unsigned)((net.n() + 255) / 256)), dim3(256), 0, s, prm, net, wb);
  float* Z1 = wf;
  float* Z2 = Z1 + (size_t)B * H;
  float* Y = Z2 + (size_t)B * H;
  uint16_t* H1 = wh;
  uint16_t* H2 = H1 + (size_t)B * H;
  uint16_t* dZ = H2 + (size_t)B * H;
  uint16_t* dY = dZ + (size_t)B * H;
  int rc = pn_forward(net, B, X, wb, prm, wf, wh, Y, s);
  if (rc) return rc;
  const long long ny = (long long)B * dout;
  const int nb = (int)((ny + 255) / 256);
  hipLaunchKernelGGL(k_pn_loss, dim3(nb), dim3(256), 0, s, Y, T, ny, 1.0f / (float)ny, dY, part);
  if (loss) hipLaunchKernelGGL(k_pn_sum, dim3(1), dim3(64), 0, s, part, nb, 1.0f / (float)ny, loss);
  GemmArgs g{};
  // dW3 = H2^T dY (H x dout), db3 = colsum dY
  g = {H, dout, B, H2, H, 1, dY, dout, 0, grad + net.w3(), dout, nullptr, 0, nullptr, nullptr, 0, EPI_NONE};
  if ((rc = gemm(g, s))) return rc;
  hipLaunchKernelGGL(k_pn_colsum, dim3((dout + 255) / 256), dim3(256), 0, s, dY, B, dout, grad + net.b3());
  // dZ2 = (dY W3^T) * silu'(Z2)
  g = {B, H, dout, dY, dout, 0, wb + net.w3(), dout, 1, nullptr, 0, dZ, H, nullptr, Z2, H, EPI_DSILU};
  if ((rc = gemm(g, s))) return rc;
  // dW2 = H1^T dZ2, db2
  g = {H, H, B, H1, H, 1, dZ, H, 0, grad + net.w2(), H, nullptr, 0, nullptr, nullptr, 0, EPI_NONE};
  if ((rc = gemm(g, s))) return rc;
  hipLaunchKernelGGL(k_pn_colsum, dim3((H + 255) / 256), dim3(256), 0, s, dZ, B, H, grad + net.b2());
  // dZ1 = (dZ2 W2^T) * silu'(Z1), into the H2 slot (H2 is no longer needed)
  g = {B, H, H, dZ, H, 0, wb + net.w2(), H, 1, nullptr, 0, H2, H, nullptr, Z1, H, EPI_DSILU};
  if ((rc = gemm(g, s))) return rc;
  // dW1 = X^T dZ1, db1
  g = {din, H, B, X, din, 1, H2, H, 0, grad + net.w1(), H, nullptr, 0, nullptr, nullptr, 0, EPI_NONE};
  if ((rc = gemm(g, s))) return rc;
  hipLaunchKernelGGL(k_pn_colsum, dim3((H + 255) / 256), dim3(256), 0, s, H2, B, H, grad + net.b1());
  return hipGetLastError() == hipSuccess ? DNSMI_OK : DNSMI_EHIP;
}

int pnet_adam(long long P, float* prm, const float* g, float* m, float* v, int step, float lr,
              hipStream_t s) {
  const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
  hipLaunchKernelGGL(k_pn_adam, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, prm, g, m, v, P, lr, c1, c2);
  return hipGetLastError() == hipSuccess ? DNSMI_OK : DNSMI_EHIP;
}

// ------------------------------------------------------------------ features
struct PFeat {
  const double* kx;  // per kx index (Nkx)
  const double* kz;  // per kz index (Nkz)
  int Nkx, Ny, din;
  long long M;  // modes
  double kmax, dt, re, c;
  float s_1, s_2, s_out;
};

__device__ __forceinline__ double pf_k2(const PFeat& a, long long m) {
  const double kx = a.kx[m % a.Nkx], kz = a.kz[m / a.Nkx];
  return kx * kx + kz * kz;
}

// One thread per (sample, component, y): D1, D2 of the profile (the older
// terms rounded to FP32, as the guess's history keeps them) into X, and
// with nnext the weighted third difference into T.  X rows: [c][re,im][y]
// for D1, then for D2, then dt k^2 / Re, kx / k_max, kz / k_max, zeros.
__global__ void k_pn_features(PFeat a, const cx<double>* __restrict__ nm2, const cx<double>* __restrict__ nm1,
                              const cx<double>* __restrict__ ncur, const cx<double>* __restrict__ nnext,
                              const long long* __restrict__ modes, int ns, uint16_t* __restrict__ X,
                              float* __restrict__ T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = 3LL * a.Ny;
  if (i >= (long long)ns * per) return;
  const int smp = (int)(i / per);
  const int r = (int)(i - (long long)smp * per);
  const int c = r / a.Ny, y = r - c * a.Ny;
  const long long m = modes[smp];
  const long long e = ((long long)c * a.M + m) * a.Ny + y;
  const cx<double> cu = ncur[e];
  const float h1x = (float)nm1[e].x, h1y = (float)nm1[e].y;
  const float h2x = (float)nm2[e].x, h2y = (float)nm2[e].y;
  const double d1x = cu.x - (double)h1x, d1y = cu.y - (double)h1y;
  const double d2x = (cu.x - 2.0 * (double)h1x) + (double)h2x, d2y = (cu.y - 2.0 * (double)h1y) + (double)h2y;
  uint16_t* xr = X + (size_t)smp * a.din;
  const int o = 2 * c * a.Ny + y;
  xr[o] = pn_bf16((float)d1x * a.s_1);
  xr[o + a.Ny] = pn_bf16((float)d1y * a.s_1);
  xr[6 * a.Ny + o] = pn_bf16((float)d2x * a.s_2);
  xr[6 * a.Ny + o + a.Ny] = pn_bf16((float)d2y * a.s_2);
  if (r == 0) {
    const double k2 = pf_k2(a, m);
    xr[12 * a.Ny] = pn_bf16((float)fmin(a.dt * k2 / a.re, 8.0));
    xr[12 * a.Ny + 1] = pn_bf16((float)(a.kx[m % a.Nkx] / a.kmax));
    xr[12 * a.Ny + 2] = pn_bf16((float)(a.kz[m / a.Nkx] / a.kmax));
    for (int f = 12 * a.Ny + 3; f < a.din; ++f) xr[f] = 0;
  }
  if (T && nnext) {
    const double w = 1.0 / (1.0 + a.dt * a.c * pf_k2(a, m) / a.re);
    const cx<double> p2 = nm2[e], p1 = nm1[e], x = nnext[e];
    const double d3x = (x.x - 3 * cu.x) + (3 * p1.x - p2.x), d3y = (x.y - 3 * cu.y) + (3 * p1.y - p2.y);
    float* tr = T + (size_t)smp * (6 * a.Ny);
    tr[o] = (float)(w * d3x / a.s_out);
    tr[o + a.Ny] = (float)(w * d3y / a.s_out);
  }
}

// The same input rows for the guess: every mode, from N^n (FP64) and the
// FP32 history h1 = N^{n-1}, h2 = N^{n-2} ((re, im) pairs, walls layout).
__global__ void k_pn_features_hist(PFeat a, const cx<double>* __restrict__ ncur, const float2* __restrict__ h1,
                                   const float2* __restrict__ h2, uint16_t* __restrict__ X) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = 3LL * a.Ny;
  if (i >= a.M * per) return;
  const long long m = i / per;
  const int r = (int)(i - m * per);
  const int c = r / a.Ny, y = r - c * a.Ny;
  const long long e = ((long long)c * a.M + m) * a.Ny + y;
  const cx<double> cu = ncur[e];
  const float2 p1 = h1[e], p2 = h2[e];
  const double d1x = cu.x - (double)p1.x, d1y = cu.y - (double)p1.y;
  const double d2x = (cu.x - 2.0 * (double)p1.x) + (double)p2.x, d2y = (cu.y - 2.0 * (double)p1.y) + (double)p2.y;
  uint16_t* xr = X + (size_t)m * a.din;
  const int o = 2 * c * a.Ny + y;
  xr[o] = pn_bf16((float)d1x * a.s_1);
  xr[o + a.Ny] = pn_bf16((float)d1y * a.s_1);
  xr[6 * a.Ny + o] = pn_bf16((float)d2x * a.s_2);
  xr[6 * a.Ny + o + a.Ny] = pn_bf16((float)d2y * a.s_2);
  if (r == 0) {
    const double k2 = pf_k2(a, m);
    xr[12 * a.Ny] = pn_bf16((float)fmin(a.dt * k2 / a.re, 8.0));
    xr[12 * a.Ny + 1] = pn_bf16((float)(a.kx[m % a.Nkx] / a.kmax));
    xr[12 * a.Ny + 2] = pn_bf16((float)(a.kz[m / a.Nkx] / a.kmax));
    for (int f = 12 * a.Ny + 3; f < a.din; ++f) xr[f] = 0;
  }
}

// Guess: next (= the quadratic extrapolation N_ext, FP64) += s_out Y / w per
// point, Y the network's output for every mode in order.
__global__ void k_pn_apply(PFeat a, const float* __restrict__ Y, cx<double>* __restrict__ next) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = 3LL * a.Ny;
  if (i >= a.M * per) return;
  const long long m = i / per;
  const int r = (int)(i - m * per);
  const int c = r / a.Ny, y = r - c * a.Ny;
  const double inv_w = 1.0 + a.dt * a.c * pf_k2(a, m) / a.re;
  const float* yr = Y + (size_t)m * (6 * a.Ny);
  const int o = 2 * c * a.Ny + y;
  const long long e = ((long long)c * a.M + m) * a.Ny + y;
  cx<double> v = next[e];
  v.x += (double)a.s_out * (double)yr[o] * inv_w;
  v.y += (double)a.s_out * (double)yr[o + a.Ny] * inv_w;
  next[e] = v;
}

static PFeat pfeat(Ctx* c, int din, float s_1, float s_2, float s_out) {
  PFeat a{};
  a.kx = c->kx;
  a.kz = c->kz;
  a.Nkx = c->Nx1;
  a.Ny = c->ny;
  a.din = din;
  a.M = (long long)c->Nz1 * c->Nx1;
  const double kxm = 2 * kPi / c->cfg.lx * (c->nx / 2), kzm = 2 * kPi / c->cfg.lz * (c->nz / 2);
  a.kmax = fmax(kxm, kzm);
  a.dt = c->cfg.dt;
  a.re = c->cfg.re;
  a.c = c->cfg.implicitness;
  a.s_1 = s_1;
  a.s_2 = s_2;
  a.s_out = s_out;
  return a;
}

int pnet_features(Ctx* c, const double* nm2, const double* nm1, const double* ncur, const double* nnext,
                  const long long* modes, int ns, int din, float s_1, float s_2, float s_out, uint16_t* X,
                  float* T, hipStream_t s) {
  if (!c->walls) return c->fail(DNSMI_EINVAL, "profile network: plane Couette contexts only");
  if (din < 12 * c->ny + 3 || din % 32) return c->fail(DNSMI_EINVAL, "profile network: din must be >= 12 ny + 3 and a multiple of 32");
  if (ns == 0) return DNSMI_OK;
  const PFeat a = pfeat(c, din, s_1, s_2, s_out);
  const long long n = (long long)ns * 3 * c->ny;
  hipLaunchKernelGGL(k_pn_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a,
                     reinterpret_cast<const cx<double>*>(nm2), reinterpret_cast<const cx<double>*>(nm1),
                     reinterpret_cast<const cx<double>*>(ncur), reinterpret_cast<const cx<double>*>(nnext),
                     modes, ns, X, T);
  return check_launch(c, "profile features");
}

int walls_guess_rhs(Ctx* c, const double* state, double* rhs, hipStream_t s);
int walls_predict(Ctx* c, const double* state, const double* rhs, double* pred, hipStream_t s);
int extrap_walls(Ctx* c, const double* n0, const float* h1, float* h2, double* next, const uint16_t* wb,
                 const float* bias, int H, float s_1, float s_2, float s_out, int mode, hipStream_t s);

// The plane Couette guess with the profile network: N^n (plain; the first
// corrector pass records the split reference), the network's input from N^n
// and the FP32 history,
// its output, then N_ext = quadratic extrapolation + s_out Y / w into
// rhs_prev (h2 <- N^n) and pred = the predictor's solve fed with N_ext.
// Caller-owned buffers, M = the number of modes: wb (the bf16 weight
// copies, one word per parameter), X (M din bf16 words), wf (2 H M floats),
// wh (2 H M bf16 words), Y (M dout floats).
int guess_walls_profile(Ctx* c, const double* state, double* pred, double* rhs_prev, const float* h1, float* h2,
                        int din, int H, int dout, const float* prm, uint16_t* wb, uint16_t* X, float* wf,
                        uint16_t* wh, float* Y, float s_1, float s_2, float s_out, hipStream_t s) {
  if (!c->walls) return c->fail(DNSMI_EINVAL, "profile network: plane Couette contexts only");
  if (dout != 6 * c->ny) return c->fail(DNSMI_EINVAL, "profile network: dout must be 6 ny");
  const long long M = (long long)c->Nz1 * c->Nx1;
  if (M >= (1LL << 31)) return c->fail(DNSMI_EINVAL, "profile network: too many modes");
  int rc = walls_guess_rhs(c, state, c->rhs0, s);
  if (rc) return rc;
  const PFeat a = pfeat(c, din, s_1, s_2, s_out);
  const unsigned nb = (unsigned)((M * 3 * c->ny + 255) / 256);
  // the input reads h1, h2 before the extrapolation overwrites h2
  hipLaunchKernelGGL(k_pn_features_hist, dim3(nb), dim3(256), 0, s, a, reinterpret_cast<const cx<double>*>(c->rhs0),
                     reinterpret_cast<const float2*>(h1), reinterpret_cast<const float2*>(h2), X);
  if ((rc = check_launch(c, "profile input"))) return rc;
  if ((rc = pnet_forward(din, H, dout, (int)M, X, prm, wb, wf, wh, Y, s)))
    return c->fail(rc, "profile network forward");
  if ((rc = extrap_walls(c, c->rhs0, h1, h2, rhs_prev, nullptr, nullptr, 32, s_1, s_2, s_out, 3, s))) return rc;
  hipLaunchKernelGGL(k_pn_apply, dim3(nb), dim3(256), 0, s, a, Y, reinterpret_cast<cx<double>*>(rhs_prev));
  if ((rc = check_launch(c, "profile apply"))) return rc;
  return walls_predict(c, state, rhs_prev, pred, s);
}

}  // namespace dnsmi
