// Plane Couette (wall-bounded) path, gfx950.
//
// Set-up (host C++, once per context; geometries/cartesian.py:575-801,
// fd.py:21-108): CGL grid, Fornberg D1/D2 stencils, per-mode Lk (Neumann)
// and Hk (Dirichlet) band matrices factored with partial pivoting in the
// LAPACK ab layout (dgbtf2 order of operations), homogeneous solutions
// p_i, v_i, q_i and the 2x2 influence matrix inverse, y-quadrature
// weights.  Per-mode tables are stored mode-fastest ([j][row][mode]) so a
// wavefront whose lanes are consecutive (kz, kx) modes reads them
// coalesced.
//
// Per step:
//   RHS (flows/plane_couette.py:274-289): WT1 FD curl + transpose
//   (3,Nkz,Nkx,Ny) -> 6 fields (Ny,Nkz,Nkx); z-inverse; the shared x pass
//   with base flow U = y; the shared z-forward; WT5 transpose back.
//   IMM (flows/plane_couette.py:376-504): one thread per (kz, kx) mode
//   streams over y: divergence and Lk d, the Lk forward sweep fed on the
//   fly, back substitution, the three Hk right-hand sides fed into one
//   forward sweep, back substitution, wall divergence, alpha = -M^-1 d,
//   assembly, correction and the quadrature-weighted norm partials.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "context.hpp"
#include "dft.hpp"
#include "passes.hpp"

namespace dnsmi {

// Influence-matrix work vectors (complex, [vec][mode][y]): f_hat, the
// pressure, the three R_c / velocity columns, three forward-sweep results.
enum ImmVec { V_F = 0, V_P = 1, V_R = 2, V_T = 5, V_COUNT = 8 };
constexpr int kImmT = 128;  // threads per mode in the parallel phases

struct WallTables {
  int p = 0, Ny = 0, Nkz = 0, Nkx = 0, M = 0;
  int curl_ty = 32;  // y points per curl workgroup (walls_curl_ty)
  // device
  double* d1w = nullptr;  // (Ny, p+1)
  int* d1j0 = nullptr;    // (Ny)
  double* d2w = nullptr;  // (Ny, p+2)
  int* d2j0 = nullptr;
  double* qw = nullptr;   // quadrature weight / Ly, (Ny)
  double* lk_ab = nullptr;  // [j][row][mode]
  int* lk_piv = nullptr;    // [j][mode]
  double* hk_ab = nullptr;
  int* hk_piv = nullptr;
  double* vq = nullptr;   // [map][y][4]: v1, v2, q1, q2 of a k^2 class
  int* umode = nullptr;   // per mode: its k^2 class (map)
  double* minv = nullptr;  // [4][mode]
  double* mkx = nullptr;   // per mode
  double* mkz = nullptr;
  double* scratch = nullptr;  // IMM work vectors, complex [vec][mode][y]
  // chunk maps of the band solves (chunked sweeps), one per distinct k^2,
  // packed (map_offset); sched = the modes, with their maps, in the order
  // the chunked solve takes them (two per wavefront)
  double *lk_fmap = nullptr, *lk_bmap = nullptr, *hk_fmap = nullptr, *hk_bmap = nullptr;
  int2* sched = nullptr;  // (mode, map) per slot
  int U = 0;              // distinct k^2 values
  // host copies: the y grid, and each mode's class (walls_table reads the
  // device tables back in the reference layouts through it)
  std::vector<double> h_ys;
  std::vector<int> h_umode;
};

inline int walls_curl_ty(int Ny, int M);  // the curl kernel's tile height

// ------------------------------------------------------------------ host set-up
// Fornberg (1998), SIAM Rev. 40, 685-691: weights of derivatives 0..m at z
// on nodes x[0..n].
static void fornberg(double z, const double* x, int n1, int m, std::vector<double>& C) {
  const int n = n1 - 1;
  C.assign((size_t)n1 * (m + 1), 0.0);
  auto at = [&](int i, int k) -> double& { return C[(size_t)i * (m + 1) + k]; };
  at(0, 0) = 1.0;
  double c1 = 1.0, c4 = x[0] - z;
  for (int i = 1; i <= n; ++i) {
    const int mn = std::min(i, m);
    double c2 = 1.0;
    const double c5 = c4;
    c4 = x[i] - z;
    for (int j = 0; j < i; ++j) {
      const double c3 = x[i] - x[j];
      c2 *= c3;
      if (j == i - 1) {
        for (int k = mn; k > 0; --k) at(i, k) = c1 * (k * at(i - 1, k - 1) - c5 * at(i - 1, k)) / c2;
        at(i, 0) = -c1 * c5 * at(i - 1, 0) / c2;
      }
      for (int k = mn; k > 0; --k) at(j, k) = (c4 * at(j, k) - k * at(j, k - 1)) / c3;
      at(j, 0) = c4 * at(j, 0) / c3;
    }
    c1 = c2;
  }
}

struct Stencil {
  int width;
  std::vector<double> w;  // (Ny, width)
  std::vector<int> j0;    // (Ny)
};

// fd.py:68-108: D1 with p+1 points, D2 with p+2 points, one-sided at walls.
static void build_stencils(const std::vector<double>& y, int p, Stencil& d1, Stencil& d2) {
  const int Ny = (int)y.size();
  d1.width = p + 1;
  d2.width = p + 2;
  d1.w.assign((size_t)Ny * d1.width, 0.0);
  d2.w.assign((size_t)Ny * d2.width, 0.0);
  d1.j0.assign(Ny, 0);
  d2.j0.assign(Ny, 0);
  std::vector<double> C;
  for (int i = 0; i < Ny; ++i) {
    int j0 = std::max(0, std::min(i - d1.width / 2, Ny - d1.width));
    fornberg(y[i], &y[j0], d1.width, 1, C);
    d1.j0[i] = j0;
    for (int k = 0; k < d1.width; ++k) d1.w[(size_t)i * d1.width + k] = C[(size_t)k * 2 + 1];
    j0 = std::max(0, std::min(i - d2.width / 2, Ny - d2.width));
    fornberg(y[i], &y[j0], d2.width, 2, C);
    d2.j0[i] = j0;
    for (int k = 0; k < d2.width; ++k) d2.w[(size_t)i * d2.width + k] = C[(size_t)k * 3 + 2];
  }
}

// Dense row i of a stencil operator, as a function (j -> value).
static double stencil_at(const Stencil& s, int i, int j) {
  const int k = j - s.j0[i];
  return (k >= 0 && k < s.width) ? s.w[(size_t)i * s.width + k] : 0.0;
}

// LAPACK band storage: A(i, j) = ab[(kl + ku + i - j) * n + j] (row-major
// rows of length n here), 2 kl + ku + 1 rows.
struct Band {
  int n, kl, ku, rows;
  std::vector<double> ab;
  std::vector<int> piv;
  Band(int n_, int kl_, int ku_) : n(n_), kl(kl_), ku(ku_), rows(2 * kl_ + ku_ + 1) {
    ab.assign((size_t)rows * n, 0.0);
    piv.assign(n, 0);
  }
  double& at(int i, int j) { return ab[(size_t)(kl + ku + i - j) * n + j]; }
  double& raw(int r, int j) { return ab[(size_t)r * n + j]; }
};

// Partial-pivoting band LU in the order of LAPACK's unblocked dgbtf2
// (the routine dgbtrf uses when its block size exceeds kl): pivot = first
// largest |a| in the column, reciprocal-scaled multipliers, rank-1 update
// up to the last column touched so far.  Returns 0 or the 1-based index of
// a zero pivot.
static int band_lu(Band& B) {
  const int n = B.n, kl = B.kl, ku = B.ku, kv = kl + ku;
  int ju = 0, info = 0;
  for (int j = 0; j < n; ++j) {
    const int km = std::min(kl, n - 1 - j);
    int jp = 0;
    double best = std::fabs(B.raw(kv, j));
    for (int i = 1; i <= km; ++i) {
      const double a = std::fabs(B.raw(kv + i, j));
      if (a > best) {
        best = a;
        jp = i;
      }
    }
    B.piv[j] = j + jp;
    if (B.raw(kv + jp, j) != 0.0) {
      ju = std::max(ju, std::min(j + ku + jp, n - 1));
      if (jp != 0)
        for (int c = j; c <= ju; ++c) std::swap(B.at(j, c), B.at(j + jp, c));
      if (km > 0) {
        const double r = 1.0 / B.raw(kv, j);
        for (int i = 1; i <= km; ++i) B.raw(kv + i, j) *= r;
        for (int c = j + 1; c <= ju; ++c) {
          const double t = -B.at(j, c);
          if (t != 0.0)
            for (int i = 1; i <= km; ++i) B.at(j + i, c) += B.raw(kv + i, j) * t;
        }
      }
    } else if (info == 0) {
      info = j + 1;
    }
  }
  return info;
}

// dgbtrs('N') on one real right-hand side, as the device sweep does it.
static void band_solve(const Band& B, std::vector<double>& b) {
  const int n = B.n, kl = B.kl, ku = B.ku, kuu = kl + ku;
  auto raw = [&](int r, int j) { return B.ab[(size_t)r * n + j]; };
  for (int j = 0; j < n - 1; ++j) {
    const int pj = B.piv[j];
    std::swap(b[j], b[pj]);
    for (int i = 1; i <= kl && j + i < n; ++i) b[j + i] -= raw(kl + ku + i, j) * b[j];
  }
  for (int j = n - 1; j >= 0; --j) {
    const double x = b[j] / raw(kl + ku, j);
    b[j] = x;
    for (int i = 1; i <= kuu && j - i >= 0; ++i) b[j - i] -= raw(kuu - i, j) * x;
  }
}

// Rows per chunk of the chunked band solve, and the register slots of its
// map prefetch (a chunk's map is loaded kChunkSlots - 1 chunks ahead).
constexpr int kChunk = 16;
constexpr int kChunkSlots = 3;
// two slots where three would spill (fd_order 8, three right-hand sides)
constexpr int chunk_slots(int po, int nr) { return po == 8 && nr == 3 ? 2 : kChunkSlots; }
// window words read from LDS per batch: two where eight would not leave
// two waves per SIMD without spilling (one right-hand side, fd_order >= 6)
constexpr int window_batch(int po, int nr) { return nr == 1 && po >= 6 ? 2 : 8; }
// waves per SIMD the chunked solve asks for: two with one right-hand side
// (registers within 256) up to fd_order 6, else one
constexpr int chunk_waves(int po, int nr) { return nr == 1 && po <= 6 ? 2 : 1; }

// Packed chunk maps.  Entry (t, k) of a map (out_t = sum_k map[t][k] w_k)
// is zero by construction for k > t + lag: lag = kl for a forward map (a
// row swap or elimination at window row t reaches at most kl rows down),
// lag = 0 for a back map (row t draws only on the rows solved before it).
// Column k of a map is therefore stored as one zero word (read by the rows
// t < map_lo(k)) followed by rows map_lo(k) .. w-1, map_len(k) words in
// all; the maps' columns lie side by side: chunk c, column k, map u, row t
// at map_offset (lanes t of one map read consecutive words).
__host__ __device__ constexpr int map_lo(int k, int lag) { return k > lag ? k - lag : 0; }
__host__ __device__ constexpr int map_len(int k, int lag, int w) { return w - map_lo(k, lag) + 1; }
// words of columns 0 .. k-1
__host__ __device__ constexpr int map_pre(int k, int lag, int w) {
  return k * (w + 1) - (k > lag + 1 ? (k - lag - 1) * (k - lag) / 2 : 0);
}
// word of row t (t < w) of column k of map u within one chunk's maps
__host__ __device__ constexpr int map_word_at(int k, int u, int t, int lag, int w, int U) {
  return U * map_pre(k, lag, w) + u * map_len(k, lag, w) + (t >= map_lo(k, lag) ? 1 + t - map_lo(k, lag) : 0);
}
__host__ __device__ inline size_t map_offset(int c, int k, int u, int t, int lag, int w, int U) {
  return (size_t)c * U * map_pre(w, lag, w) + (size_t)map_word_at(k, u, t, lag, w, U);
}

// Chunk maps of one band solve (the operation order of band_solve): forward
// chunk c applies the steps j in [cB, min(cB + B, n - 1)) (row swap,
// elimination) to the window of rows cB + t, t < WF = B + kl; its map F_c is
// those row operations applied to the identity.  Back chunk c, top row
// T = n - 1 - cB, applies the steps j = T, T - 1, ... (at most B, j >= 0:
// divide by the diagonal, eliminate upwards) to the window of rows T - t,
// t < WB = B + kl + ku.  Written packed as map u of U; `raw(r, j)` reads
// the factored band, `piv(j)` the pivot row.  Returns false if an entry
// outside the packed region is not zero (it cannot be, by the argument
// above; the check keeps the packing honest).
template <typename Raw, typename Piv>
static bool chunk_maps(int n, int kl, int ku, int u, int U, Raw raw, Piv piv, double* F, double* G) {
  const int B = kChunk, kuu = kl + ku, WF = B + kl, WB = B + kuu;
  const int nC = (n + B - 1) / B;
  bool ok = true;
  std::vector<double> X((size_t)WB * WB);
  for (int c = 0; c < nC; ++c) {
    const int j0 = c * B;
    std::fill(X.begin(), X.end(), 0.0);
    for (int t = 0; t < WF; ++t) X[(size_t)t * WF + t] = 1.0;
    for (int j = j0; j < std::min(j0 + B, n - 1); ++j) {
      const int tj = j - j0, tp = piv(j) - j0;
      if (tp != tj)
        for (int k = 0; k < WF; ++k) std::swap(X[(size_t)tj * WF + k], X[(size_t)tp * WF + k]);
      for (int i = 1; i <= kl && j + i < n; ++i) {
        const double l = raw(kl + ku + i, j);
        for (int k = 0; k < WF; ++k) X[(size_t)(tj + i) * WF + k] -= l * X[(size_t)tj * WF + k];
      }
    }
    for (int k = 0; k < WF; ++k)
      for (int t = 0; t < WF; ++t) {
        const double x = X[(size_t)t * WF + k];
        if (t >= map_lo(k, kl)) F[map_offset(c, k, u, t, kl, WF, U)] = x;
        else ok = ok && x == 0.0;  // read as the column's zero word
      }
    const int T = n - 1 - j0;
    std::fill(X.begin(), X.end(), 0.0);
    for (int t = 0; t < WB; ++t) X[(size_t)t * WB + t] = 1.0;
    for (int j = T; j > T - B && j >= 0; --j) {
      const int tj = T - j;
      const double dg = raw(kl + ku, j);
      for (int k = 0; k < WB; ++k) X[(size_t)tj * WB + k] /= dg;
      for (int i = 1; i <= kuu && j - i >= 0; ++i) {
        const double u = raw(kuu - i, j);
        for (int k = 0; k < WB; ++k) X[(size_t)(tj + i) * WB + k] -= u * X[(size_t)tj * WB + k];
      }
    }
    for (int k = 0; k < WB; ++k)
      for (int t = 0; t < WB; ++t) {
        const double x = X[(size_t)t * WB + k];
        if (t >= map_lo(k, 0)) G[map_offset(c, k, u, t, 0, WB, U)] = x;
        else ok = ok && x == 0.0;
      }
  }
  return ok;
}

// Linear weights of the y quadrature, divided by Ly.  Odd Ny: the
// non-uniform composite Simpson rule (geometries/cartesian.py:125-140).
// Even Ny (the reference exits there): Simpson panels over the first Ny-4
// intervals plus the exact integral of the cubic through the last four
// points.
static std::vector<double> quadrature_weights(const std::vector<double>& y, double Ly) {
  const int n = (int)y.size();
  std::vector<double> w(n, 0.0);
  const int m = (n % 2 == 1) ? n : n - 3;
  for (int i = 0; i + 2 < m; i += 2) {
    const double h0 = y[i + 1] - y[i], h1 = y[i + 2] - y[i + 1];
    const double hs = h0 + h1, hp = h0 * h1, r = h0 / h1;
    w[i] += hs / 6 * (2 - 1 / r);
    w[i + 1] += hs / 6 * (hs * hs / hp);
    w[i + 2] += hs / 6 * (2 - r);
  }
  if (n % 2 == 0) {
    // solve V w = moments, V[m][i] = (x_i - a)^m, on the last 4 points
    const double a = y[n - 4], b = y[n - 1];
    double V[4][5];
    for (int r = 0; r < 4; ++r) {
      for (int i = 0; i < 4; ++i) V[r][i] = std::pow(y[n - 4 + i] - a, r);
      V[r][4] = std::pow(b - a, r + 1) / (r + 1);
    }
    for (int c = 0; c < 4; ++c) {  // Gaussian elimination with partial pivoting
      int pr = c;
      for (int r = c + 1; r < 4; ++r)
        if (std::fabs(V[r][c]) > std::fabs(V[pr][c])) pr = r;
      for (int k = 0; k < 5; ++k) std::swap(V[c][k], V[pr][k]);
      for (int r = c + 1; r < 4; ++r) {
        const double f = V[r][c] / V[c][c];
        for (int k = c; k < 5; ++k) V[r][k] -= f * V[c][k];
      }
    }
    double x[4];
    for (int r = 3; r >= 0; --r) {
      double s = V[r][4];
      for (int k = r + 1; k < 4; ++k) s -= V[r][k] * x[k];
      x[r] = s / V[r][r];
    }
    for (int i = 0; i < 4; ++i) w[n - 4 + i] += x[i];
  }
  for (double& v : w) v /= Ly;
  return w;
}

template <typename V>
static int upload_vec(Ctx* c, V** dst, const std::vector<V>& v) {
  *dst = static_cast<V*>(c->alloc(v.size() * sizeof(V)));
  if (!*dst) return c->fail(DNSMI_ENOMEM, "device allocation failed");
  const hipError_t e = hipMemcpy(*dst, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice);
  if (e != hipSuccess) return c->fail(DNSMI_EHIP, hipGetErrorString(e));
  return DNSMI_OK;
}

static int walls_autotune(Ctx* c);

// ------------------------------------------------------------------ set-up on the GPU
// The per-class tables of geometries/cartesian.py:575-801 built on the GPU:
// the band LU of L_k and H_k (band_lu), the homogeneous solutions and M^-1
// (the band_solve / D1 sequence of walls_setup's class loop) and the chunk
// maps (chunk_maps), with the host code's operations in its order and every
// product, sum and quotient rounded on its own as the host computes them
// (contraction into FMA is switched off in these kernels), so the tables
// equal the host set-up's bit for bit.  One workgroup of one wavefront per
// k^2 class (and matrix, and chunk); a class's band lives in LDS while it is
// factored and used.
struct GpuSetup {
  int n, kl, ku, rows, U, W1, W2;  // W1, W2: widths of the D1 and D2 stencils
  const double* d1w;
  const int* d1j0;
  const double* d2w;
  const int* d2j0;
  const double* k2;  // [U]
  double dt, cc, nu;
  double* band;  // [2][U][rows][n], raw(r, j) at r * n + j; L_k then H_k
  int* piv;      // [2][U][n]
  int* flag;     // [0]: a zero pivot, [1]: a chunk-map entry left out of the packing is not zero
  double* vq;    // [U][n][4]
  double* minv;  // [U][4]
};

__device__ inline double setup_stencil(const double* w, const int* j0, int W, int i, int j) {
  const int k = j - j0[i];
  return (k >= 0 && k < W) ? w[i * W + k] : 0.0;
}

// d1_apply of walls_setup: s += w_k v_(j0 + k), k ascending, from 0.
__device__ inline double setup_d1(const GpuSetup& g, const double* v, int i) {
#pragma clang fp contract(off)
  double s = 0.0;
  for (int k = 0; k < g.W1; ++k) {
    const double t = g.d1w[i * g.W1 + k] * v[g.d1j0[i] + k];
    s = s + t;
  }
  return s;
}

// Build and factor L_k (blockIdx.y = 0) or H_k (1) of class blockIdx.x.
__global__ void __launch_bounds__(64) k_setup_lu(GpuSetup g) {
#pragma clang fp contract(off)
  extern __shared__ double ab[];  // rows x n
  const int u = blockIdx.x, which = blockIdx.y, lane = threadIdx.x;
  const int n = g.n, kl = g.kl, ku = g.ku, kv = kl + ku, rows = g.rows;
  const double k2 = g.k2[u];
  for (int e = lane; e < rows * n; e += 64) ab[e] = 0.0;
  __syncthreads();
  for (int i = lane; i < n; i += 64) {
    const int jlo = i - kl > 0 ? i - kl : 0, jhi = i + ku < n - 1 ? i + ku : n - 1;
    for (int j = jlo; j <= jhi; ++j) {
      const double lraw = setup_stencil(g.d2w, g.d2j0, g.W2, i, j) - (i == j ? k2 : 0.0);
      double v;
      if (which == 0) {
        v = lraw;
        if (i == 0) v = (k2 == 0.0) ? (j == 0 ? 1.0 : 0.0) : setup_stencil(g.d1w, g.d1j0, g.W1, 0, j);
        if (i == n - 1) v = setup_stencil(g.d1w, g.d1j0, g.W1, n - 1, j);
      } else {
        const double cn = g.cc * g.nu;
        const double t = cn * lraw;
        v = (i == j ? 1.0 / g.dt : 0.0) - t;
        if (i == 0 || i == n - 1) v = (i == j) ? 1.0 : 0.0;
      }
      ab[(kv + i - j) * n + j] = v;
    }
  }
  __syncthreads();
  int ju = 0;
  bool zero = false;
  int* piv = g.piv + (size_t)(which * g.U + u) * n;
  for (int j = 0; j < n; ++j) {
    const int km = kl < n - 1 - j ? kl : n - 1 - j;
    int jp = 0;
    double best = fabs(ab[kv * n + j]);
    for (int i = 1; i <= km; ++i) {
      const double a = fabs(ab[(kv + i) * n + j]);
      if (a > best) {
        best = a;
        jp = i;
      }
    }
    const double pv = ab[(kv + jp) * n + j];
    if (lane == 0) piv[j] = j + jp;
    __syncthreads();
    if (pv == 0.0) {
      zero = true;
      continue;
    }
    const int jr = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
    ju = ju > jr ? ju : jr;
    if (jp != 0) {
      for (int c = j + lane; c <= ju; c += 64) {
        const double a0 = ab[(kv + j - c) * n + c];
        ab[(kv + j - c) * n + c] = ab[(kv + j + jp - c) * n + c];
        ab[(kv + j + jp - c) * n + c] = a0;
      }
      __syncthreads();
    }
    if (km > 0) {
      const double r = 1.0 / ab[kv * n + j];
      if (lane >= 1 && lane <= km) ab[(kv + lane) * n + j] = ab[(kv + lane) * n + j] * r;
      __syncthreads();
      const int nc = ju - j;
      for (int e = lane; e < nc * km; e += 64) {
        const int q = e / km, i = 1 + e - q * km, c = j + 1 + q;
        const double t = -ab[(kv + j - c) * n + c];
        if (t != 0.0) {
          const double prod = ab[(kv + i) * n + j] * t;
          ab[(kv + j + i - c) * n + c] = ab[(kv + j + i - c) * n + c] + prod;
        }
      }
      __syncthreads();
    }
  }
  if (zero && lane == 0) g.flag[0] = 1;
  double* out = g.band + (size_t)(which * g.U + u) * rows * n;
  for (int e = lane; e < rows * n; e += 64) out[e] = ab[e];
}

// band_solve on the vectors v + q n (q < nvec), one per 16 lanes, with the
// factored band ab and pivots pl in LDS.
__device__ inline void setup_solve(const GpuSetup& g, const double* ab, const int* pl, double* v, int nvec) {
#pragma clang fp contract(off)
  const int n = g.n, kl = g.kl, kv = g.kl + g.ku, lane = threadIdx.x;
  const int q = lane >> 4, t = lane & 15;
  const bool act = q < nvec;
  double* b = v + (act ? q : 0) * n;
  for (int j = 0; j < n - 1; ++j) {
    if (act && t == 0) {
      const int pj = pl[j];
      const double x = b[j];
      b[j] = b[pj];
      b[pj] = x;
    }
    __syncthreads();
    if (act && t >= 1 && t <= kl && j + t < n) {
      const double prod = ab[(kv + t) * n + j] * b[j];
      b[j + t] = b[j + t] - prod;
    }
    __syncthreads();
  }
  for (int j = n - 1; j >= 0; --j) {
    if (act && t == 0) b[j] = b[j] / ab[kv * n + j];
    __syncthreads();
    const int i = t + 1;
    if (act && i <= kv && j - i >= 0) {
      const double prod = ab[(kv - i) * n + j] * b[j];
      b[j - i] = b[j - i] - prod;
    }
    __syncthreads();
  }
}

// Homogeneous solutions and M^-1 of class blockIdx.x: p_1,2 = L^-1 e_(0, n-1);
// v_1,2 = H^-1 (-D1 p_1,2 with zero wall rows); q_1,2 = H^-1 (p_1,2 with
// zero wall rows); M = D1 v at the walls.
__global__ void __launch_bounds__(64) k_setup_homog(GpuSetup g) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  const int u = blockIdx.x, lane = threadIdx.x, n = g.n, rows = g.rows;
  double* ab = sm;               // rows x n
  double* v = sm + rows * n;     // four vectors of n
  int* pl = reinterpret_cast<int*>(v + 4 * n);
  for (int which = 0; which < 2; ++which) {
    const double* src = g.band + (size_t)(which * g.U + u) * rows * n;
    const int* ps = g.piv + (size_t)(which * g.U + u) * n;
    for (int e = lane; e < rows * n; e += 64) ab[e] = src[e];
    for (int e = lane; e < n; e += 64) pl[e] = ps[e];
    if (which == 0)
      for (int e = lane; e < 2 * n; e += 64) v[e] = (e == 0 || e == 2 * n - 1) ? 1.0 : 0.0;
    __syncthreads();
    setup_solve(g, ab, pl, v, which == 0 ? 2 : 4);
    if (which == 0) {
      // r_1,2 = -D1 p_1,2 (zero on the walls) into vectors 2, 3; then the
      // walls of p_1,2 zeroed in place: the four H_k right-hand sides
      for (int i = lane; i < n; i += 64) {
        const bool wall = i == 0 || i == n - 1;
        v[2 * n + i] = wall ? 0.0 : -setup_d1(g, v, i);
        v[3 * n + i] = wall ? 0.0 : -setup_d1(g, v + n, i);
      }
      __syncthreads();
      if (lane == 0) v[0] = v[n - 1] = v[n] = v[2 * n - 1] = 0.0;
      __syncthreads();
    }
  }
  if (lane == 0) {
    const double m00 = setup_d1(g, v + 2 * n, 0), m01 = setup_d1(g, v + 3 * n, 0);
    const double m10 = setup_d1(g, v + 2 * n, n - 1), m11 = setup_d1(g, v + 3 * n, n - 1);
    double mi[4];
    if (g.k2[u] == 0.0) {
      mi[0] = mi[1] = mi[2] = 0.0;
      mi[3] = 1.0 / m11;
    } else {
      const double a = m00 * m11, b = m01 * m10;
      const double det = a - b;
      mi[0] = m11 / det;
      mi[1] = -m01 / det;
      mi[2] = -m10 / det;
      mi[3] = m00 / det;
    }
    for (int q = 0; q < 4; ++q) g.minv[u * 4 + q] = mi[q];
  }
  for (int j = lane; j < n; j += 64) {
    double* w = g.vq + ((size_t)u * n + j) * 4;
    w[0] = v[2 * n + j];
    w[1] = v[3 * n + j];
    w[2] = v[j];
    w[3] = v[n + j];
  }
}

// Chunk map of chunk blockIdx.x, class blockIdx.y; blockIdx.z: 0 L_k
// forward, 1 L_k back, 2 H_k forward, 3 H_k back (chunk_maps' steps).
__global__ void __launch_bounds__(64) k_setup_maps(GpuSetup g, double* lf, double* lb, double* hf, double* hb) {
#pragma clang fp contract(off)
  constexpr int WMAX = kChunk + 16;  // B + 2p for p <= 8
  __shared__ double X[WMAX * WMAX];
  const int c = blockIdx.x, u = blockIdx.y, which = blockIdx.z >> 1, back = blockIdx.z & 1;
  const int lane = threadIdx.x, n = g.n, kl = g.kl, kuu = g.kl + g.ku, B = kChunk;
  const double* ab = g.band + (size_t)(which * g.U + u) * g.rows * n;
  const int* pv = g.piv + (size_t)(which * g.U + u) * n;
  const int W = back ? B + kuu : B + kl;
  for (int e = lane; e < W * W; e += 64) X[e] = (e / W == e % W) ? 1.0 : 0.0;
  __syncthreads();
  const int j0 = c * B;
  if (!back) {
    const int jend = j0 + B < n - 1 ? j0 + B : n - 1;
    for (int j = j0; j < jend; ++j) {
      const int tj = j - j0, tp = pv[j] - j0;
      if (tp != tj) {
        for (int k = lane; k < W; k += 64) {
          const double x = X[tj * W + k];
          X[tj * W + k] = X[tp * W + k];
          X[tp * W + k] = x;
        }
        __syncthreads();
      }
      for (int e = lane; e < kl * W; e += 64) {
        const int i = 1 + e / W, k = e % W;
        if (j + i < n) {
          const double prod = ab[(kuu + i) * n + j] * X[tj * W + k];
          X[(tj + i) * W + k] = X[(tj + i) * W + k] - prod;
        }
      }
      __syncthreads();
    }
  } else {
    const int T = n - 1 - j0;
    for (int j = T; j > T - B && j >= 0; --j) {
      const int tj = T - j;
      const double dg = ab[kuu * n + j];
      for (int k = lane; k < W; k += 64) X[tj * W + k] = X[tj * W + k] / dg;
      __syncthreads();
      for (int e = lane; e < kuu * W; e += 64) {
        const int i = 1 + e / W, k = e % W;
        if (j - i >= 0) {
          const double prod = ab[(kuu - i) * n + j] * X[tj * W + k];
          X[(tj + i) * W + k] = X[(tj + i) * W + k] - prod;
        }
      }
      __syncthreads();
    }
  }
  double* F = which == 0 ? (back ? lb : lf) : (back ? hb : hf);
  const int lag = back ? 0 : kl;
  bool bad = false;
  for (int e = lane; e < W * W; e += 64) {
    const int t = e / W, k = e % W;
    const double x = X[t * W + k];
    if (t >= map_lo(k, lag)) F[map_offset(c, k, u, t, lag, W, g.U)] = x;
    else bad = bad || x != 0.0;
  }
  if (bad) g.flag[1] = 1;
}

// Per-mode tables from the per-class ones: band factors [j][row][mode],
// pivots [j][mode], M^-1 [4][mode].
__global__ void __launch_bounds__(256) k_setup_scatter(GpuSetup g, const int* __restrict__ umode, int M,
                                                       double* lk_ab, double* hk_ab, int* lk_piv, int* hk_piv,
                                                       double* minv) {
  const long long nab = (long long)g.n * g.rows * M, npv = (long long)g.n * M;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nab; e += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(e % M);
    const long long jr = e / M;
    const int r = (int)(jr % g.rows), j = (int)(jr / g.rows), u = umode[m];
    lk_ab[e] = g.band[((size_t)u * g.rows + r) * g.n + j];
    hk_ab[e] = g.band[((size_t)(g.U + u) * g.rows + r) * g.n + j];
    if (e < npv) {
      const int mm = (int)(e % M), jj = (int)(e / M), uu = umode[mm];
      lk_piv[e] = g.piv[(size_t)uu * g.n + jj];
      hk_piv[e] = g.piv[(size_t)(g.U + uu) * g.n + jj];
    }
    if (e < 4LL * M) {
      const int mm = (int)(e % M), q = (int)(e / M);
      minv[e] = g.minv[umode[mm] * 4 + q];
    }
  }
}

// LDS bytes of the set-up kernels (the GPU set-up runs where both fit 64 KiB).
static size_t setup_lds_lu(int n, int rows) { return (size_t)rows * n * sizeof(double); }
static size_t setup_lds_homog(int n, int rows) {
  return ((size_t)rows * n + 4ull * n) * sizeof(double) + (size_t)n * sizeof(int);
}

int walls_setup(Ctx* c) {
  const dnsmi_config& g = c->cfg;
  if (g.nx % 2 || g.nz % 2 || g.nx < 4 || g.nz < 4 || g.ny < 5)
    return c->fail(DNSMI_EINVAL, "walls: nx, nz even >= 4 and ny >= 5");
  // the y grid is fixed on [-1, 1] (parameters.py:190-191)
  if (g.ly != 2.0) return c->fail(DNSMI_EINVAL, "walls: ly must be 2 (the channel is y in [-1, 1])");
  if (g.oversampling_factor != 3) return c->fail(DNSMI_EINVAL, "oversampling_factor must be 3");
  const int p = g.fd_order;
  if (p != 2 && p != 4 && p != 6 && p != 8) return c->fail(DNSMI_EINVAL, "fd_order must be 2, 4, 6 or 8");
  if (g.ny < p + 3) return c->fail(DNSMI_EINVAL, "ny too small for fd_order");
  c->nx = g.nx;
  c->ny = g.ny;
  c->nz = g.nz;
  c->Nx1 = g.nx / 2;
  c->Nz1 = g.nz - 1;
  c->Ny1 = g.ny;
  c->Px = 3 * g.nx / 2;
  c->Pz = 3 * g.nz / 2;
  c->Py = g.ny;
  if (!supported_length(c->Px) || !supported_length(c->Pz))
    return c->fail(DNSMI_EINVAL, "padded lengths must be one of 12,16,24,32,48,64,72,96,128,144,192");
  const int Ny = g.ny, Nkz = c->Nz1, Nkx = c->Nx1, M = Nkz * Nkx;
  c->spec_elems = 3LL * M * Ny;
  WallTables* t = new WallTables();
  c->wt = t;
  t->p = p;
  t->Ny = Ny;
  t->Nkz = Nkz;
  t->Nkx = Nkx;
  t->M = M;
  // forced variants (tile_cols != 0): bit 64 selects the 64-row curl tiles
  t->curl_ty = g.tile_cols != 0 ? ((g.variants & 64) ? 64 : 32) : walls_curl_ty(Ny, M);
  c->curl_rows = t->curl_ty;

  // buffers: WT1 out (6,Ny,Nkz,Nkx) | z-inverse out (6,Ny,Pz,Nkx) | x-pass out
  // (3,Ny,Pz,Nkx) | z-forward out (3,Ny,Nkz,Nkx)
  const size_t cxs = 16;  // FP64-sized in the mixed form too: escalated passes run in FP64
  const size_t a1 = 6ull * Ny * Nkz * Nkx, a3 = 3ull * Ny * c->Pz * Nkx;
  const size_t b2 = 6ull * Ny * c->Pz * Nkx, b4 = 3ull * Ny * Nkz * Nkx;
  c->bufA_bytes = std::max(a1, a3) * cxs;
  c->bufB_bytes = std::max(b2, b4) * cxs;
  c->bufA = c->alloc(c->bufA_bytes);
  c->bufB = c->alloc(c->bufB_bytes);
  c->rhs0 = static_cast<double*>(c->alloc(c->spec_elems * 16));
  c->nref = c->rhs0;
  if (c->split) {
    // n_ref apart from rhs0, which the corrector uses as scratch
    c->nref = static_cast<double*>(c->alloc(c->spec_elems * 16));
    c->uref = static_cast<double*>(c->alloc(c->spec_elems * 16));
    c->xref = static_cast<float*>(c->alloc(6ull * Ny * c->Pz * c->Px * sizeof(float)));
    if (!c->nref || !c->uref || !c->xref) return c->fail(DNSMI_ENOMEM, "device allocation failed");
  }
  c->max_partials = M + 1024;  // the IMM writes one norm partial per mode
  c->partials = static_cast<double*>(c->alloc(sizeof(double) * c->max_partials));
  t->scratch = static_cast<double*>(c->alloc((size_t)V_COUNT * Ny * M * 16));
  if (!c->bufA || !c->bufB || !c->rhs0 || !c->partials || !t->scratch)
    return c->fail(DNSMI_ENOMEM, "device allocation failed");
  int rc = setup_twiddles(c);
  if (rc) return rc;

  const std::vector<double> kx = real_wavenumbers(g.nx, g.lx);
  const std::vector<double> kz = complex_wavenumbers(g.nz, g.lz);
  if ((rc = upload_doubles(c, &c->kx, kx))) return rc;
  if ((rc = upload_doubles(c, &c->kz, kz))) return rc;

  // CGL grid and base flow U = y (flows/plane_couette.py:73-108)
  std::vector<double> y(Ny);
  for (int j = 0; j < Ny; ++j) y[j] = -std::cos((double)j * kPi / (double)(Ny - 1));
  t->h_ys = y;
  // the y grid in the ky slot: the learned guess's coordinate of a point
  if ((rc = upload_doubles(c, &c->ky, y))) return rc;
  std::vector<double> base(3 * Ny, 0.0), cbase(3 * Ny, 0.0), nbase(3 * Ny, 0.0);
  for (int j = 0; j < Ny; ++j) {
    base[j] = y[j];
    cbase[2 * Ny + j] = -1.0;
    nbase[1 * Ny + j] = y[j] * 1.0;
  }
  if ((rc = upload_doubles(c, &c->base, base))) return rc;
  if ((rc = upload_doubles(c, &c->cbase, cbase))) return rc;
  if ((rc = upload_doubles(c, &c->nbase, nbase))) return rc;

  Stencil d1, d2;
  build_stencils(y, p, d1, d2);
  if ((rc = upload_vec(c, &t->d1w, d1.w))) return rc;
  if ((rc = upload_vec(c, &t->d1j0, d1.j0))) return rc;
  if ((rc = upload_vec(c, &t->d2w, d2.w))) return rc;
  if ((rc = upload_vec(c, &t->d2j0, d2.j0))) return rc;
  if ((rc = upload_vec(c, &t->qw, quadrature_weights(y, g.ly)))) return rc;

  // per-mode IMM tables (geometries/cartesian.py:704-779)
  const int kl = p, ku = p, rows = 3 * p + 1;
  const double dt = g.dt, cc = g.implicitness, nu = 1.0 / g.re;
  // L_k and H_k depend on the mode only through k^2: modes with the same
  // k^2 (bit for bit) share every per-mode table, computed once per class
  // (at the class's first mode, first[u]) and copied to the others
  std::vector<int> umode(M), first;
  std::vector<double> mkx(M), mkz(M), k2u;
  {
    std::unordered_map<uint64_t, int> key;
    for (int iz = 0; iz < Nkz; ++iz)
      for (int ix = 0; ix < Nkx; ++ix) {
        const int m = iz * Nkx + ix;
        const double k2 = kx[ix] * kx[ix] + kz[iz] * kz[iz];
        mkx[m] = kx[ix];
        mkz[m] = kz[iz];
        uint64_t b;
        std::memcpy(&b, &k2, sizeof b);
        const auto it = key.emplace(b, (int)first.size());
        if (it.second) {
          first.push_back(m);
          k2u.push_back(k2);
        }
        umode[m] = it.first->second;
      }
  }
  const int U = (int)first.size();
  t->U = U;
  t->h_umode = umode;
  if ((rc = upload_vec(c, &t->umode, umode))) return rc;
  // the chunk maps of the L_k and H_k solves, one per class (not when the
  // sequential sweeps are requested)
  const int B = kChunk, WF = B + kl, WB = B + kl + ku, nC = (Ny + B - 1) / B;
  const size_t nf = (size_t)nC * U * map_pre(WF, kl, WF), nb = (size_t)nC * U * map_pre(WB, 0, WB);
  // the chunked solve also stages three right-hand sides per mode pair in
  // LDS (2 * 3 * (Ny + 32) complex values), which must fit in 64 KiB
  const bool maps_fit = nf < (1ull << 28) && nb < (1ull << 28) &&
                        (size_t)2 * 3 * (Ny + 32) * sizeof(cx<double>) <= 65536;
  const bool forced = g.tile_cols != 0;  // the variant bits choose the form
  if (forced && (g.variants & 8) == 0 && !maps_fit)
    return c->fail(DNSMI_EINVAL,
                   "chunk maps exceed 2^28 words or their LDS staging 64 KiB: use the sequential band solve");
  const bool want_maps = maps_fit && !(forced && (g.variants & 8) != 0);
  // on the GPU unless asked for on the host, or a class's band and vectors
  // do not fit the set-up kernels' 64 KiB of LDS (or the classes a grid's
  // y dimension)
  const bool on_gpu = !g.setup_on_host && setup_lds_homog(Ny, rows) <= 65536 && U <= 65535;
  const auto t0 = std::chrono::steady_clock::now();
  if (on_gpu) {
    GpuSetup gs{};
    gs.n = Ny;
    gs.kl = kl;
    gs.ku = ku;
    gs.rows = rows;
    gs.U = U;
    gs.W1 = d1.width;
    gs.W2 = d2.width;
    gs.d1w = t->d1w;
    gs.d1j0 = t->d1j0;
    gs.d2w = t->d2w;
    gs.d2j0 = t->d2j0;
    gs.dt = dt;
    gs.cc = cc;
    gs.nu = nu;
    double* k2d = nullptr;
    if ((rc = upload_vec(c, &k2d, k2u))) return rc;
    gs.k2 = k2d;
    gs.band = static_cast<double*>(c->alloc(2ull * U * rows * Ny * sizeof(double)));
    gs.piv = static_cast<int*>(c->alloc(2ull * U * Ny * sizeof(int)));
    gs.flag = static_cast<int*>(c->alloc(2 * sizeof(int)));
    gs.minv = static_cast<double*>(c->alloc(4ull * U * sizeof(double)));
    t->vq = static_cast<double*>(c->alloc(4ull * U * Ny * sizeof(double)));
    t->lk_ab = static_cast<double*>(c->alloc((size_t)Ny * rows * M * sizeof(double)));
    t->hk_ab = static_cast<double*>(c->alloc((size_t)Ny * rows * M * sizeof(double)));
    t->lk_piv = static_cast<int*>(c->alloc((size_t)Ny * M * sizeof(int)));
    t->hk_piv = static_cast<int*>(c->alloc((size_t)Ny * M * sizeof(int)));
    t->minv = static_cast<double*>(c->alloc(4ull * M * sizeof(double)));
    if (!gs.band || !gs.piv || !gs.flag || !gs.minv || !t->vq || !t->lk_ab || !t->hk_ab || !t->lk_piv ||
        !t->hk_piv || !t->minv)
      return c->fail(DNSMI_ENOMEM, "device allocation failed");
    gs.vq = t->vq;
    if (hipMemset(gs.flag, 0, 2 * sizeof(int)) != hipSuccess) return c->fail(DNSMI_EHIP, "memset");
    hipLaunchKernelGGL(k_setup_lu, dim3(U, 2), dim3(64), setup_lds_lu(Ny, rows), nullptr, gs);
    if ((rc = check_launch(c, "setup_lu"))) return rc;
    hipLaunchKernelGGL(k_setup_homog, dim3(U), dim3(64), setup_lds_homog(Ny, rows), nullptr, gs);
    if ((rc = check_launch(c, "setup_homog"))) return rc;
    if (want_maps) {
      for (double** m : {&t->lk_fmap, &t->hk_fmap})
        if (!(*m = static_cast<double*>(c->alloc(nf * sizeof(double)))))
          return c->fail(DNSMI_ENOMEM, "device allocation failed");
      for (double** m : {&t->lk_bmap, &t->hk_bmap})
        if (!(*m = static_cast<double*>(c->alloc(nb * sizeof(double)))))
          return c->fail(DNSMI_ENOMEM, "device allocation failed");
      // each packed column's leading zero word is read, never written
      if (hipMemset(t->lk_fmap, 0, nf * sizeof(double)) != hipSuccess ||
          hipMemset(t->hk_fmap, 0, nf * sizeof(double)) != hipSuccess ||
          hipMemset(t->lk_bmap, 0, nb * sizeof(double)) != hipSuccess ||
          hipMemset(t->hk_bmap, 0, nb * sizeof(double)) != hipSuccess)
        return c->fail(DNSMI_EHIP, "memset");
      hipLaunchKernelGGL(k_setup_maps, dim3(nC, U, 4), dim3(64), 0, nullptr, gs, t->lk_fmap, t->lk_bmap,
                         t->hk_fmap, t->hk_bmap);
      if ((rc = check_launch(c, "setup_maps"))) return rc;
    }
    const long long nab = (long long)Ny * rows * M;
    const int sb = (int)std::min<long long>((nab + 255) / 256, 65536);
    hipLaunchKernelGGL(k_setup_scatter, dim3(sb), dim3(256), 0, nullptr, gs, t->umode, M, t->lk_ab,
                       t->hk_ab, t->lk_piv, t->hk_piv, t->minv);
    if ((rc = check_launch(c, "setup_scatter"))) return rc;
    int flag[2] = {0, 0};
    if (hipMemcpy(flag, gs.flag, sizeof flag, hipMemcpyDeviceToHost) != hipSuccess)
      return c->fail(DNSMI_EHIP, "set-up: read flags");
    for (void* q : {(void*)k2d, (void*)gs.band, (void*)gs.piv, (void*)gs.flag, (void*)gs.minv}) c->release(q);
    if (flag[0]) return c->fail(DNSMI_ESINGULAR, "banded LU hit a zero pivot");
    if (flag[1]) return c->fail(DNSMI_EINVAL, "chunk maps: nonzero entry outside the packed region");
  } else {
    std::vector<double> lk_ab((size_t)Ny * rows * M), hk_ab((size_t)Ny * rows * M);
    std::vector<int> lk_piv((size_t)Ny * M), hk_piv((size_t)Ny * M);
    std::vector<double> minv(4ull * M), vq((size_t)U * Ny * 4);
    auto copy_mode = [&](int m, int r) {
      for (int j = 0; j < Ny; ++j) {
        for (int q = 0; q < rows; ++q) {
          lk_ab[((size_t)j * rows + q) * M + m] = lk_ab[((size_t)j * rows + q) * M + r];
          hk_ab[((size_t)j * rows + q) * M + m] = hk_ab[((size_t)j * rows + q) * M + r];
        }
        lk_piv[(size_t)j * M + m] = lk_piv[(size_t)j * M + r];
        hk_piv[(size_t)j * M + m] = hk_piv[(size_t)j * M + r];
      }
      for (int q = 0; q < 4; ++q) minv[(size_t)q * M + m] = minv[(size_t)q * M + r];
    };
    for (int m = 0; m < M; ++m) {
      if (first[umode[m]] != m) {
        copy_mode(m, first[umode[m]]);
        continue;
      }
      const double k2 = k2u[umode[m]];
      Band L(Ny, kl, ku), H(Ny, kl, ku);
      for (int i = 0; i < Ny; ++i) {
        for (int j = std::max(0, i - kl); j <= std::min(Ny - 1, i + ku); ++j) {
          const double d2ij = stencil_at(d2, i, j);
          const double lraw = d2ij - (i == j ? k2 : 0.0);
          double lv = lraw;
          if (i == 0) lv = (k2 == 0.0) ? (j == 0 ? 1.0 : 0.0) : stencil_at(d1, 0, j);
          if (i == Ny - 1) lv = stencil_at(d1, Ny - 1, j);
          L.at(i, j) = lv;
          double hv = (i == j ? 1.0 / dt : 0.0) - cc * nu * lraw;
          if (i == 0 || i == Ny - 1) hv = (i == j) ? 1.0 : 0.0;
          H.at(i, j) = hv;
        }
      }
      // D1 rows of the walls are p+1 wide and lie inside the band by
      // construction (one-sided stencils start at the wall).
      if (band_lu(L) != 0 || band_lu(H) != 0)
        return c->fail(DNSMI_ESINGULAR, "banded LU hit a zero pivot");
      std::vector<double> p1(Ny, 0.0), p2(Ny, 0.0);
      p1[0] = 1.0;
      p2[Ny - 1] = 1.0;
      band_solve(L, p1);
      band_solve(L, p2);
      auto d1_apply = [&](const std::vector<double>& v, int i) {
        double s = 0.0;
        for (int k = 0; k < d1.width; ++k) s += d1.w[(size_t)i * d1.width + k] * v[d1.j0[i] + k];
        return s;
      };
      std::vector<double> r1(Ny), r2(Ny);
      for (int i = 0; i < Ny; ++i) {
        r1[i] = -d1_apply(p1, i);
        r2[i] = -d1_apply(p2, i);
      }
      r1[0] = r1[Ny - 1] = r2[0] = r2[Ny - 1] = 0.0;
      band_solve(H, r1);
      band_solve(H, r2);
      std::vector<double> s1 = p1, s2 = p2;
      s1[0] = s1[Ny - 1] = s2[0] = s2[Ny - 1] = 0.0;
      band_solve(H, s1);
      band_solve(H, s2);
      const double M00 = d1_apply(r1, 0), M01 = d1_apply(r2, 0);
      const double M10 = d1_apply(r1, Ny - 1), M11 = d1_apply(r2, Ny - 1);
      double Mi[4];
      if (k2 == 0.0) {
        Mi[0] = Mi[1] = Mi[2] = 0.0;
        Mi[3] = 1.0 / M11;
      } else {
        const double det = M00 * M11 - M01 * M10;
        Mi[0] = M11 / det;
        Mi[1] = -M01 / det;
        Mi[2] = -M10 / det;
        Mi[3] = M00 / det;
      }
      for (int q = 0; q < 4; ++q) minv[(size_t)q * M + m] = Mi[q];
      for (int j = 0; j < Ny; ++j) {
        for (int r = 0; r < rows; ++r) {
          lk_ab[((size_t)j * rows + r) * M + m] = L.raw(r, j);
          hk_ab[((size_t)j * rows + r) * M + m] = H.raw(r, j);
        }
        lk_piv[(size_t)j * M + m] = L.piv[j];
        hk_piv[(size_t)j * M + m] = H.piv[j];
        double* w = &vq[((size_t)umode[m] * Ny + j) * 4];
        w[0] = r1[j];
        w[1] = r2[j];
        w[2] = s1[j];
        w[3] = s2[j];
      }
    }
    if ((rc = upload_vec(c, &t->lk_ab, lk_ab))) return rc;
    if ((rc = upload_vec(c, &t->hk_ab, hk_ab))) return rc;
    if (want_maps) {  // built in parallel on the host threads
      std::vector<double> lf(nf), lb(nb), hf(nf), hb(nb);
      const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
      std::vector<char> ok(nth, 1);
      auto work = [&](int w) {
        for (int u = w; u < U; u += nth) {
          const int m = first[u];
          auto lraw = [&](int r, int j) { return lk_ab[((size_t)j * rows + r) * M + m]; };
          auto lpiv = [&](int j) { return lk_piv[(size_t)j * M + m]; };
          auto hraw = [&](int r, int j) { return hk_ab[((size_t)j * rows + r) * M + m]; };
          auto hpiv = [&](int j) { return hk_piv[(size_t)j * M + m]; };
          const bool a = chunk_maps(Ny, kl, ku, u, U, lraw, lpiv, lf.data(), lb.data());
          const bool b = chunk_maps(Ny, kl, ku, u, U, hraw, hpiv, hf.data(), hb.data());
          if (!a || !b) ok[w] = 0;
        }
      };
      std::vector<std::thread> pool;
      for (int w = 1; w < nth; ++w) pool.emplace_back(work, w);
      work(0);
      for (auto& th : pool) th.join();
      for (char o : ok)
        if (!o) return c->fail(DNSMI_EINVAL, "chunk maps: nonzero entry outside the packed region");
      if ((rc = upload_vec(c, &t->lk_fmap, lf))) return rc;
      if ((rc = upload_vec(c, &t->lk_bmap, lb))) return rc;
      if ((rc = upload_vec(c, &t->hk_fmap, hf))) return rc;
      if ((rc = upload_vec(c, &t->hk_bmap, hb))) return rc;
    }
    if ((rc = upload_vec(c, &t->lk_piv, lk_piv))) return rc;
    if ((rc = upload_vec(c, &t->hk_piv, hk_piv))) return rc;
    if ((rc = upload_vec(c, &t->vq, vq))) return rc;
    if ((rc = upload_vec(c, &t->minv, minv))) return rc;
  }
  if (hipDeviceSynchronize() != hipSuccess) return c->fail(DNSMI_EHIP, "set-up: synchronise");
  const double setup_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  char note[96];
  std::snprintf(note, sizeof note, "imm_setup=%s(%.1f ms) curl_rows=%d ", on_gpu ? "gpu" : "host", setup_ms,
                t->curl_ty);
  if (want_maps) {
    // the chunked solve's mode order: by map, so that the two modes of a
    // wavefront share theirs wherever two modes have the same k^2
    std::vector<int> order(M);
    for (int m = 0; m < M; ++m) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return umode[a] < umode[b]; });
    order.resize((size_t)2 * ((M + 1) / 2), order[M - 1]);  // a spare slot repeats the last mode
    std::vector<int2> sched(order.size());
    for (size_t i = 0; i < order.size(); ++i) sched[i] = make_int2(order[i], umode[order[i]]);
    if ((rc = upload_vec(c, &t->sched, sched))) return rc;
  }
  if ((rc = upload_vec(c, &t->mkx, mkx))) return rc;
  if ((rc = upload_vec(c, &t->mkz, mkz))) return rc;
  if (!forced && !maps_fit) {  // the autotuner has one form to choose
    c->sweep_seq = true;
    c->tuning = std::string(note) + "band_solve=sequential(chunk maps exceed 2^28 words or 64 KiB LDS) ";
    return DNSMI_OK;
  }
  if (!forced) {
    rc = walls_autotune(c);
    c->tuning = std::string(note) + c->tuning;
    return rc;
  }
  c->tuning = note;
  return DNSMI_OK;
}

void walls_release(Ctx* c) {
  delete c->wt;
  c->wt = nullptr;
}

// ------------------------------------------------------------------ RHS kernels
// WT1: FD curl + transpose.  In: u (3, M, Ny) with M = Nkz*Nkx modes, FP64.
// Out: V (6, Ny, Nkz, Nkx) in T: u_x, u_y, u_z, w_x = D1 u_z - i kz u_y,
// i kz u_x, -D1 u_x (the i kx parts of w_y, w_z are added by the x pass).
// A block handles TY y-points x kTM modes; loads run along y, stores
// along the mode index.  The curl's tiles are taller than the transposes'
// (kTY): each re-reads a halo of S1 rows on either side, so taller tiles
// re-read less but make fewer workgroups (walls_curl_ty chooses).
constexpr int kTY = 16, kTM = 16;

// The curl's tile height: 64 rows where that still gives at least four
// workgroups per CU (128 x 256 x 128: 2032; the RHS 356 against 386 us
// FP64, 211 against 240 us mixed), else 32 (64 x 128 x 64: 252 workgroups
// of 64 rows would leave CUs idle; 64 against 73 us).
inline int walls_curl_ty(int Ny, int M) {
  const long long blocks64 = (long long)((Ny + 63) / 64) * ((M + kTM - 1) / kTM);
  return blocks64 >= 4 * 256 ? 64 : 32;
}

// EM (EvalMode): EV_RECORD also copies u to ucopy (u_ref); EV_INC takes the
// curl of d = u - uref, formed in FP64.
template <typename T, int S1, int EM, int TY>
__global__ void __launch_bounds__(TY* kTM)
    k_walls_curl_t(const cx<double>* __restrict__ u, cx<T>* __restrict__ v,
                   const double* __restrict__ d1w, const int* __restrict__ d1j0,
                   const double* __restrict__ mkz, int Ny, int M, const int* __restrict__ act,
                   const cx<double>* __restrict__ uref, cx<double>* __restrict__ ucopy) {
  if (act && *act == 0) return;
  constexpr int HALO = S1;
  __shared__ cx<double> tile[3][kTM][TY + 2 * HALO + 1];
  const int ntile_y = (Ny + TY - 1) / TY;
  const int y0 = (blockIdx.x % ntile_y) * TY;
  const int m0 = (blockIdx.x / ntile_y) * kTM;
  const int ylo = max(0, y0 - HALO);
  const int yhi = min(Ny, y0 + TY + HALO);
  const int span = yhi - ylo;
  const size_t comp = (size_t)M * Ny;
  // load: lanes along y
  for (int e = threadIdx.x; e < 3 * kTM * span; e += blockDim.x) {
    const int c = e / (kTM * span);
    const int r = e - c * kTM * span;
    const int mi = r / span, yi = r - mi * span;
    const int m = m0 + mi;
    cx<double> val{0, 0};
    if (m < M) {
      const size_t e = c * comp + (size_t)m * Ny + ylo + yi;
      val = u[e];
      if constexpr (EM == EV_RECORD) {
        const int yy = ylo + yi;  // the tile's own rows (not its halo)
        if (yy >= y0 && yy < y0 + TY) ucopy[e] = val;
      }
      if constexpr (EM == EV_INC) {
        const cx<double> r = uref[e];
        val = {val.x - r.x, val.y - r.y};
      }
    }
    tile[c][mi][yi] = val;
  }
  __syncthreads();
  const int mi = threadIdx.x % kTM, yi = threadIdx.x / kTM;
  const int m = m0 + mi, y = y0 + yi;
  if (m >= M || y >= Ny) return;
  const cx<double> ux = tile[0][mi][y - ylo], uy = tile[1][mi][y - ylo], uz = tile[2][mi][y - ylo];
  cx<double> dux{0, 0}, duz{0, 0};
  const int j0 = d1j0[y];
#pragma unroll
  for (int k = 0; k < S1; ++k) {
    const double w = d1w[y * S1 + k];
    const cx<double> a = tile[0][mi][j0 + k - ylo], b = tile[2][mi][j0 + k - ylo];
    dux.x += w * a.x;
    dux.y += w * a.y;
    duz.x += w * b.x;
    duz.y += w * b.y;
  }
  const double k = mkz[m];
  // flows/plane_couette.py:257-271
  const cx<double> wx = {duz.x + k * uy.y, duz.y - k * uy.x};  // dy w - i kz v
  const cx<double> wyp = {-k * ux.y, k * ux.x};                // i kz u
  const cx<double> wzp = {-dux.x, -dux.y};                     // -dy u
  const size_t f = (size_t)Ny * M;
  const size_t o = (size_t)y * M + m;
  v[0 * f + o] = {static_cast<T>(ux.x), static_cast<T>(ux.y)};
  v[1 * f + o] = {static_cast<T>(uy.x), static_cast<T>(uy.y)};
  v[2 * f + o] = {static_cast<T>(uz.x), static_cast<T>(uz.y)};
  v[3 * f + o] = {static_cast<T>(wx.x), static_cast<T>(wx.y)};
  v[4 * f + o] = {static_cast<T>(wyp.x), static_cast<T>(wyp.y)};
  v[5 * f + o] = {static_cast<T>(wzp.x), static_cast<T>(wzp.y)};
}

// z-inverse of NF independent fields: (NF, Ny, Nkz, Nkx) -> (NF, Ny, Pz, Nkx).
template <typename T, int P, int TC, int NF>
__global__ void __launch_bounds__(TC * ColPlan<P>::threads)
    k_z_inverse_plain(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                      const cx<T>* __restrict__ tw, int nz, int Ny, int Nx1,
                      const int* __restrict__ act) {
  if (act && *act == 0) return;
  constexpr int N1 = ColPlan<P>::N1, N2 = ColPlan<P>::N2;
  __shared__ cx<T> lds[TC * P];
  __shared__ cx<T> tws[P];
  for (int i = threadIdx.x; i < P; i += blockDim.x) tws[i] = tw[i];
  __syncthreads();
  const int tid = threadIdx.x, col = tid % TC, j_ = tid / TC;
  const int ntile = (Nx1 + TC - 1) / TC;
  const int y = blockIdx.x / ntile;
  const int kx = (blockIdx.x - y * ntile) * TC + col;
  const bool ok = kx < Nx1;
  const int Nz1 = nz - 1;
  const size_t in_f = (size_t)Ny * Nz1 * Nx1;
  const size_t in_off = (size_t)y * Nz1 * Nx1 + kx;
  const size_t out_f = (size_t)Ny * P * Nx1;
  const size_t out_off = (size_t)y * P * Nx1 + kx;
#pragma unroll 1
  for (int f = blockIdx.y; f < NF; f += gridDim.y) {  // one field per workgroup row
    const int j = opaque(j_);
    if (j < N2) {
      cx<T> x[N1];
#pragma unroll
      for (int a = 0; a < N1; ++a) {
        const int s = Pad<P>::stored_in(N2 * a, N2, j);
        x[a] = (ok && s >= 0) ? in[f * in_f + in_off + (size_t)s * Nx1] : cx<T>{0, 0};
      }
      step_a_store<P, N1, N2, +1, TC>(x, j, col, lds, tws);
    }
    __syncthreads();
    if (j < N1 && ok) {
      cx<T> o[N2];
      step_b_load<P, N1, N2, +1, TC>(o, j, col, lds);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) out[f * out_f + out_off + (size_t)(j + N1 * k2) * Nx1] = o[k2];
    }
    __syncthreads();
  }
}

// WT5: (3, Ny, Nkz, Nkx) T -> rhs (3, M, Ny) FP64, tiled transpose; with
// nref (split evaluations) rhs = nref + the transformed increment; with
// nref_out (a recording evaluation) rhs is also copied there.
template <typename T>
__global__ void __launch_bounds__(kTY* kTM)
    k_walls_to_spec(const cx<T>* __restrict__ w, cx<double>* __restrict__ rhs, int Ny, int M,
                    const int* __restrict__ act, const cx<double>* __restrict__ nref,
                    cx<double>* __restrict__ nref_out) {
  if (act && *act == 0) return;
  __shared__ cx<double> tile[kTY][kTM + 1];
  const int ntile_y = (Ny + kTY - 1) / kTY;
  const int c = blockIdx.y;
  const int y0 = (blockIdx.x % ntile_y) * kTY;
  const int m0 = (blockIdx.x / ntile_y) * kTM;
  const size_t f = (size_t)Ny * M;
  {
    const int mi = threadIdx.x % kTM, yi = threadIdx.x / kTM;
    const int m = m0 + mi, y = y0 + yi;
    if (m < M && y < Ny) {
      const cx<T> v = w[c * f + (size_t)y * M + m];
      tile[yi][mi] = {static_cast<double>(v.x), static_cast<double>(v.y)};
    }
  }
  __syncthreads();
  const int yi = threadIdx.x % kTY, mi = threadIdx.x / kTY;
  const int m = m0 + mi, y = y0 + yi;
  if (m < M && y < Ny) {
    const size_t e = c * f + (size_t)m * Ny + y;
    cx<double> r = tile[yi][mi];
    if (nref) r = {nref[e].x + r.x, nref[e].y + r.y};
    rhs[e] = r;
    if (nref_out) nref_out[e] = r;
  }
}

// ------------------------------------------------------------------ IMM kernel
struct ImmArgs {
  const cx<double>* un;   // velocity_n (3, M, Ny)
  const cx<double>* uj;   // velocity_j
  const cx<double>* nn;   // nonlin_n
  const cx<double>* nj;   // nonlin_j
  cx<double>* out;        // velocity_new (may alias uj)
  cx<double>* corr;       // correction (may be null)
  double* partials;       // per-block sum of quadrature-weighted k_metric |delta|^2
  const double *d1w, *d2w, *qw;
  const int *d1j0, *d2j0;
  const double *lk_ab, *hk_ab;
  const int *lk_piv, *hk_piv;
  const double* vq;    // [map][y][4]: v1, v2, q1, q2
  const int* umode;    // per mode: its map
  const double *minv, *mkx, *mkz;
  cx<double>* scratch;  // work vectors [vec][mode][y]
  const int* act;       // if set and 0: the launch is skipped
  int Ny, M, Nkx;
  double dt, c, nu;
};

// The influence-matrix step in five phases.  The right-hand sides of the
// banded systems are generated in parallel over (mode, y) (one workgroup
// per mode, lanes along y: the fields' rows are contiguous in y), the
// sequential band sweeps run one thread per (mode, right-hand side) on
// precomputed rows with their loads issued a batch of rows ahead of the
// recurrence, and the assembly is parallel again.  Work vectors are
// [vec][mode][y]; per mode the arithmetic is the one-thread-per-mode
// formulation's, operation for operation.

// Work vector v, mode m: element y at [y * VS] from the returned pointer.
// The rows are contiguous in y (VS = 1): the parallel phases' lanes run
// along y; a sweep lane walks its own row.
constexpr int VS = 1;
__device__ __forceinline__ cx<double>* svec(const ImmArgs& a, int v, int m) {
  return a.scratch + ((size_t)v * a.M + m) * a.Ny;
}
__device__ __forceinline__ cx<double> ikm(double k, cx<double> z) { return {-k * z.y, k * z.x}; }

// D1 / D2 stencils at row y of a row r (global or LDS) whose element y is
// at r[y * stride]
template <int S>
__device__ __forceinline__ cx<double> stencil(const double* w, const int* j0t, const cx<double>* r,
                                              int y, size_t stride = 1) {
  cx<double> s{0, 0};
  const int j0 = j0t[y];
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const double wk = w[y * S + k];
    const cx<double> v = r[(j0 + k) * stride];
    s.x += wk * v.x;
    s.y += wk * v.y;
  }
  return s;
}

// Per-row arithmetic of the parallel phases, shared by the five-launch
// form and the fused chunked form (so both compute the same values, bit for
// bit).  `row` = m Ny, `comp` = M Ny; the rows the stencils read are in LDS.

// d_hat^n at row y (phase 1): i kx u_x + D1 u_y + i kz u_z; un1 = u^n_y's row.
template <int PO>
__device__ __forceinline__ cx<double> imm_dhat_at(const ImmArgs& a, const cx<double>* un1, size_t row,
                                                  size_t comp, double kx, double kz, int y) {
  const cx<double> dv = stencil<PO + 1>(a.d1w, a.d1j0, un1, y);
  const cx<double> t0 = ikm(kx, a.un[row + y]);
  const cx<double> t2 = ikm(kz, a.un[2 * comp + row + y]);
  return {(t0.x + dv.x) + t2.x, (t0.y + dv.y) + t2.y};
}

// f_hat at row y (phase 1), zero on the walls; dh = d_hat^n's row, nj1 and
// nn1 the y rows of N_j and N_n.
template <int PO>
__device__ __forceinline__ cx<double> imm_fhat_at(const ImmArgs& a, const cx<double>* dh, const cx<double>* nj1,
                                                  const cx<double>* nn1, size_t row, size_t comp, double kx,
                                                  double kz, double k2, int y) {
  if (y <= 0 || y >= a.Ny - 1) return {0, 0};
  constexpr int S1 = PO + 1, S2 = PO + 2;
  const double dt = a.dt, cc = a.c, nu = a.nu;
  const cx<double> d = dh[y];
  const cx<double> dvj = stencil<S1>(a.d1w, a.d1j0, nj1, y);
  const cx<double> dvn = stencil<S1>(a.d1w, a.d1j0, nn1, y);
  const cx<double> aj = ikm(kx, a.nj[row + y]), cj = ikm(kz, a.nj[2 * comp + row + y]);
  const cx<double> an = ikm(kx, a.nn[row + y]), cn = ikm(kz, a.nn[2 * comp + row + y]);
  const cx<double> divj = {(aj.x + dvj.x) + cj.x, (aj.y + dvj.y) + cj.y};
  const cx<double> divn = {(an.x + dvn.x) + cn.x, (an.y + dvn.y) + cn.y};
  const cx<double> d2d = stencil<S2>(a.d2w, a.d2j0, dh, y);
  const cx<double> lkd = {d2d.x - k2 * d.x, d2d.y - k2 * d.y};
  const double g = (1 - cc) * nu;
  return {((d.x / dt + cc * divj.x) + (1 - cc) * divn.x) + g * lkd.x,
          ((d.y / dt + cc * divj.y) + (1 - cc) * divn.y) + g * lkd.y};
}

// R_c at row y for the three components (phase 3), zero on the walls; pp =
// the p_P row, ur = the rows of u^n ([c][Ny]).
template <int PO>
__device__ __forceinline__ void imm_r_at(const ImmArgs& a, const cx<double>* pp, const cx<double>* ur, size_t row,
                                         size_t comp, double kx, double kz, double k2, int y,
                                         cx<double> (&r)[3]) {
  const int Ny = a.Ny;
  if (y <= 0 || y >= Ny - 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] = {0, 0};
    return;
  }
  constexpr int S1 = PO + 1, S2 = PO + 2;
  const double dt = a.dt, cc = a.c, nu = a.nu;
  const cx<double> p = pp[y];
  const cx<double> dpp[3] = {ikm(kx, p), stencil<S1>(a.d1w, a.d1j0, pp, y), ikm(kz, p)};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const cx<double>* unc = ur + c * Ny;
    const cx<double> u = unc[y];
    const cx<double> du = stencil<S2>(a.d2w, a.d2j0, unc, y);
    const cx<double> l = {du.x - k2 * u.x, du.y - k2 * u.y};
    const double g = (1.0 - cc) * nu;
    const cx<double> hm = {(1.0 / dt) * u.x + g * l.x, (1.0 / dt) * u.y + g * l.y};
    const cx<double> Nj = a.nj[c * comp + row + y], Nn = a.nn[c * comp + row + y];
    r[c] = {((hm.x - dpp[c].x) + cc * Nj.x) + (1 - cc) * Nn.x, ((hm.y - dpp[c].y) + cc * Nj.y) + (1 - cc) * Nn.y};
  }
}

// alpha = -M^-1 (wall divergence of v_arb) with the k^2 = 0 gauge (phase 5);
// va = v_arb's row (element y at va[y * vs]).
template <int PO>
__device__ __forceinline__ void imm_alpha(const ImmArgs& a, int m, const cx<double>* va, size_t vs,
                                          cx<double>& al1, cx<double>& al2) {
  constexpr int S1 = PO + 1;
  const int Ny = a.Ny;
  const size_t M = a.M;
  const double kx = a.mkx[m], kz = a.mkz[m];
  cx<double> dw0 = stencil<S1>(a.d1w, a.d1j0, va, 0, vs), dw1 = stencil<S1>(a.d1w, a.d1j0, va, Ny - 1, vs);
  if (kx * kx + kz * kz == 0.0) dw0 = {0, 0};
  const double m00 = a.minv[0 * M + m], m01 = a.minv[1 * M + m];
  const double m10 = a.minv[2 * M + m], m11 = a.minv[3 * M + m];
  al1 = {-(m00 * dw0.x + m01 * dw1.x), -(m00 * dw0.y + m01 * dw1.y)};
  al2 = {-(m10 * dw0.x + m11 * dw1.x), -(m10 * dw0.y + m11 * dw1.y)};
}

// Row y of the assembly (phase 5): the corrected velocity into out (and the
// correction into corr) when `store`; returns k_metric |delta|^2 of the row
// (before its quadrature weight).  ua, va, wa: v_arb's rows (stride vs).
__device__ __forceinline__ double imm_assemble_at(const ImmArgs& a, int m, size_t row, size_t comp, double kx,
                                                  double kz, const cx<double>* ua, const cx<double>* va,
                                                  const cx<double>* wa, size_t vs, cx<double> al1,
                                                  cx<double> al2, int y, bool store) {
  const double km = (m % a.Nkx == 0) ? 1.0 : 2.0;
  const double4 w = reinterpret_cast<const double4*>(a.vq)[(size_t)a.umode[m] * a.Ny + y];
  const double v1 = w.x, v2 = w.y, q1 = w.z, q2 = w.w;
  const cx<double> vv = va[(size_t)y * vs];
  const cx<double> vn = {(vv.x + al1.x * v1) + al2.x * v2, (vv.y + al1.y * v1) + al2.y * v2};
  const cx<double> q = {al1.x * q1 + al2.x * q2, al1.y * q1 + al2.y * q2};
  const cx<double> uu = ua[(size_t)y * vs], ww = wa[(size_t)y * vs];
  const cx<double> iq_x = ikm(kx, q), iq_z = ikm(kz, q);
  const cx<double> nw[3] = {{uu.x - iq_x.x, uu.y - iq_x.y}, vn, {ww.x - iq_z.x, ww.y - iq_z.y}};
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t idx = c * comp + row + y;
    const cx<double> old = a.uj[idx];
    const cx<double> d = {nw[c].x - old.x, nw[c].y - old.y};
    if (store) {
      a.out[idx] = nw[c];
      if (a.corr) a.corr[idx] = d;
    }
    s += km * (d.x * d.x) + km * (d.y * d.y);
  }
  return s;
}

// Phase 1 (flows/plane_couette.py:427-445): d_hat^n into LDS, then f_hat
// (zero on the walls) -> V_F.
template <int PO>
__global__ void __launch_bounds__(kImmT) k_imm_fhat(ImmArgs a) {
  if (a.act && *a.act == 0) return;
  // LDS: d_hat, then the y rows of u^n_y, N_j,y and N_n,y (the D1 stencils'
  // operands, read once from HBM)
  extern __shared__ cx<double> dh[];
  const int m = blockIdx.x, Ny = a.Ny;
  const size_t comp = (size_t)a.M * Ny, row = (size_t)m * Ny;
  const double kx = a.mkx[m], kz = a.mkz[m];
  const double k2 = kx * kx + kz * kz;
  cx<double>* un1 = dh + Ny;
  cx<double>* nj1 = dh + 2 * Ny;
  cx<double>* nn1 = dh + 3 * Ny;
  for (int y = threadIdx.x; y < Ny; y += blockDim.x) {
    un1[y] = a.un[comp + row + y];
    nj1[y] = a.nj[comp + row + y];
    nn1[y] = a.nn[comp + row + y];
  }
  __syncthreads();
  for (int y = threadIdx.x; y < Ny; y += blockDim.x) dh[y] = imm_dhat_at<PO>(a, un1, row, comp, kx, kz, y);
  __syncthreads();
  cx<double>* f = svec(a, V_F, m);
  for (int y = threadIdx.x; y < Ny; y += blockDim.x)
    f[y * VS] = imm_fhat_at<PO>(a, dh, nj1, nn1, row, comp, kx, kz, k2, y);
}

// Phases 2 and 4: solve ab x = rhs per (mode, right-hand side r =
// blockIdx.y) with the LU factors and pivots in LAPACK band layout
// (mode-fastest).  Forward sweep rhs vector vin + r -> vtmp + r, back
// substitution -> vout + r.  Loads of U rows are issued before their
// recurrence steps.
template <int PO>
struct SweepBatch {
  static constexpr int value = PO <= 2 ? 32 : (PO <= 4 ? 16 : 8);
};

// Lanes per sweep workgroup (one wave; 16-lane workgroups over 4x the CUs
// measured slower: a sweep is bound by its chain of batch latencies).
constexpr int kSweepLanes = 64;

template <int PO>
__global__ void __launch_bounds__(kSweepLanes)
    k_band_sweeps(ImmArgs a, const double* __restrict__ ab, const int* __restrict__ piv, int vin,
                  int vtmp, int vout) {
  if (a.act && *a.act == 0) return;
  constexpr int KL = PO, KU = PO, KUU = 2 * PO, ROWS = 3 * PO + 1, U = SweepBatch<PO>::value;
  const int m = blockIdx.x * kSweepLanes + threadIdx.x;
  if (m >= a.M) return;
  const int r = blockIdx.y, Ny = a.Ny, M = a.M;
  // rin and out may be the same vector (the H sweeps solve in place): the
  // forward sweep only reads rin and writes tmp, the back substitution only
  // reads tmp and writes out, so each sweep's pointers are restrict within
  // its own scope (a batch's loads need not wait for the previous stores).
  {
    const cx<double>* __restrict__ rin = svec(a, vin + r, m);
    cx<double>* __restrict__ tmp = svec(a, vtmp + r, m);
    cx<double> w[KL + 1];
#pragma unroll
    for (int i = 0; i <= KL; ++i) w[i] = (i < Ny) ? rin[(size_t)i * VS] : cx<double>{0, 0};
    for (int j0 = 0; j0 < Ny - 1; j0 += U) {
      int dd[U];
      double L[U][KL];
      cx<double> nf[U];
      // Loads from clamped (always valid) rows with no branches, so that the
      // whole batch is in flight at once; the guards apply at use.
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u;
        const int jj = j < Ny - 2 ? j : Ny - 2;
        dd[u] = piv[(size_t)jj * M + m] - j;
        const double* col = ab + (size_t)jj * ROWS * M + m;
#pragma unroll
        for (int i = 1; i <= KL; ++i) L[u][i - 1] = col[(size_t)(KL + KU + i) * M];
        const int jf = j + KL + 1 < Ny ? j + KL + 1 : Ny - 1;
        nf[u] = rin[(size_t)jf * VS];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + u;
        if (j < Ny - 1) {
          const cx<double> w0 = w[0];
#pragma unroll
          for (int i = 1; i <= KL; ++i)
            if (dd[u] == i) {
              w[0] = w[i];
              w[i] = w0;
            }
#pragma unroll
          for (int i = 1; i <= KL; ++i) {
            if (j + i < Ny) {
              const double l = L[u][i - 1];
              w[i].x -= l * w[0].x;
              w[i].y -= l * w[0].y;
            }
          }
          tmp[(size_t)j * VS] = w[0];
#pragma unroll
          for (int i = 0; i < KL; ++i) w[i] = w[i + 1];
          w[KL] = (j + KL + 1 < Ny) ? nf[u] : cx<double>{0, 0};
        }
      }
    }
    tmp[(size_t)(Ny - 1) * VS] = w[0];
  }
  {
    const cx<double>* __restrict__ tmp = svec(a, vtmp + r, m);
    cx<double>* __restrict__ out = svec(a, vout + r, m);
    cx<double> z[KUU + 1];
#pragma unroll
    for (int i = 0; i <= KUU; ++i) z[KUU - i] = (Ny - 1 - i >= 0) ? tmp[(size_t)(Ny - 1 - i) * VS] : cx<double>{0, 0};
    for (int j0 = Ny - 1; j0 >= 0; j0 -= U) {
      double dg[U], Uv[U][KUU];
      cx<double> nz[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 - u;
        const double* col = ab + (size_t)(j > 0 ? j : 0) * ROWS * M + m;
        dg[u] = col[(size_t)(KL + KU) * M];
#pragma unroll
        for (int i = 1; i <= KUU; ++i) Uv[u][i - 1] = col[(size_t)(KUU - i) * M];
        nz[u] = tmp[(size_t)(j - KUU - 1 > 0 ? j - KUU - 1 : 0) * VS];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 - u;
        if (j >= 0) {
          const cx<double> x = {z[KUU].x / dg[u], z[KUU].y / dg[u]};
          out[(size_t)j * VS] = x;
#pragma unroll
          for (int i = 1; i <= KUU; ++i) {
            if (j - i >= 0) {
              const double uu = Uv[u][i - 1];
              z[KUU - i].x -= uu * x.x;
              z[KUU - i].y -= uu * x.y;
            }
          }
#pragma unroll
          for (int i = KUU; i > 0; --i) z[i] = z[i - 1];
          z[0] = (j - KUU - 1 >= 0) ? nz[u] : cx<double>{0, 0};
        }
      }
    }
  }
}

// Phases 2 and 4, chunked form: the same solves through the precomputed
// chunk maps (chunk_maps).  32 lanes per mode, two modes per wavefront;
// lane t forms output t of each chunk's map from the window, which the
// mode's lanes share through LDS, so the dependent chain is nC chunks
// instead of Ny rows, and every load is coalesced and independent of the
// recurrence.  The right-hand sides are staged in LDS at the start; a
// chunk's map is loaded S - 1 chunks ahead through S register slots (the
// chunk loop is unrolled by S so the slots are static).  The forward
// results stay in LDS for the back substitution.  NR right-hand sides per
// mode share the maps; the two modes of a wavefront mostly share one map
// (the modes come in the order `sched`, grouped by k^2), whose words are
// then fetched once for both.  Lanes in a packed column's structural zeros
// read its zero word, so the sums are those of the full maps.

// Byte offsets of lane t's words in one chunk's maps (the same for every
// chunk): rows below a column's first stored row read its zero word; lanes
// past the window read row w-1 and are never stored.  They fit 32 bits
// (the maps are at most 2^28 words).
template <int W, int LAG>
__device__ __forceinline__ void map_offsets(unsigned (&o)[W], int U, int u, int t) {
  const int tt = t < W ? t : W - 1;
#pragma unroll
  for (int k = 0; k < W; ++k) o[k] = 8u * (unsigned)map_word_at(k, u, tt, LAG, W, U);
}

// Lane t's row of chunk c's map (chunks past the last repeat it): buffer
// loads from a resource over the chunk's maps (base and size in scalar
// registers) at the lane's 32-bit offsets.  Flat 64-bit addresses would need
// a pair of address registers per load, and the compiler, reusing those of
// loads still in flight, waited for every outstanding load at each chunk
// step, which undid the prefetch; a buffer load needs none.  Offsets past
// the chunk's maps (none, by construction) would read zero.
template <int W, int LAG>
__device__ __forceinline__ void map_row(double (&f)[W], const double* __restrict__ map,
                                        const unsigned (&o)[W], int U, int c, int nC) {
  const int words = U * map_pre(W, LAG, W);
  const double* base = map + (size_t)(c < nC ? c : nC - 1) * words;
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(base), 0, words * (int)sizeof(double), 0x00020000);
#pragma unroll
  for (int k = 0; k < W; ++k)
    f[k] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)o[k], 0, 0));
}

// acc[r] += sum_k map[t][k] win[r][k] in order k = 0, 1, ...; the window
// words are read from LDS in batches of KB before their products, so that
// a batch's reads are in flight together.
template <int W, int NR, int KB>
__device__ __forceinline__ void window_dot(cx<double> (&acc)[NR], const double (&f)[W],
                                           const cx<double>* win) {
#pragma unroll
  for (int k0 = 0; k0 < W; k0 += KB) {
    cx<double> w[KB][NR];
#pragma unroll
    for (int kk = 0; kk < KB; ++kk)
#pragma unroll
      for (int r = 0; r < NR; ++r)
        if (k0 + kk < W) w[kk][r] = win[r * 32 + k0 + kk];
#pragma unroll
    for (int kk = 0; kk < KB; ++kk) {
      if (k0 + kk >= W) break;
      const double fk = f[k0 + kk];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        acc[r].x += fk * w[kk][r].x;
        acc[r].y += fk * w[kk][r].y;
      }
    }
  }
}

// The window is shared by the mode's lanes, all in this wavefront, whose
// LDS operations complete in issue order.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One forward chunk: window position t <-> row cB + t.  vec holds the
// right-hand sides (row stride 1, Ny rows each); the chunk overwrites rows
// below (c + 1) B with results and reads rows from (c + 1) B + KL on.  pf
// receives the map of chunk c + S - 1.
template <int PO, int NR, int S>
__device__ __forceinline__ void fwd_chunk(const double (&f)[kChunk + PO], double (&pf)[kChunk + PO],
                                          const double* __restrict__ fmap,
                                          const unsigned (&o)[kChunk + PO], int U, cx<double>* vec, cx<double>* win, int t, int c, int nC,
                                          int Ny) {
  constexpr int KL = PO, B = kChunk, WF = B + KL;
  const int frow = (c + 1) * B + t;
  const bool fr = t >= KL && t < WF && frow < Ny;
  cx<double> fresh[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) fresh[r] = fr ? vec[r * Ny + frow] : cx<double>{0, 0};
  map_row<WF, KL>(pf, fmap, o, U, c + S - 1, nC);
  cx<double> acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = cx<double>{0, 0};
  window_dot<WF, NR, window_batch(PO, NR)>(acc, f, win);
  wave_sync();
  const int row = c * B + t;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (t < B && row < Ny) vec[r * Ny + row] = acc[r];
    if (t >= B && t < WF) win[r * 32 + t - B] = acc[r];
    if (t >= KL && t < WF) win[r * 32 + t] = fresh[r];
  }
  wave_sync();
}

// One back chunk: window position t <-> row T - t, T = Ny - 1 - cB; pg
// receives the map of chunk c + S - 1.
template <int PO, int NR, int S>
__device__ __forceinline__ void back_chunk(const double (&g)[kChunk + 2 * PO],
                                           double (&pg)[kChunk + 2 * PO],
                                           const double* __restrict__ bmap,
                                           const unsigned (&o)[kChunk + 2 * PO], int U,
                                           cx<double>* __restrict__ out, size_t vs, bool live,
                                           cx<double>* vec, cx<double>* win, int t, int c, int nC,
                                           int Ny) {
  constexpr int KUU = 2 * PO, B = kChunk, WB = B + KUU;
  map_row<WB, 0>(pg, bmap, o, U, c + S - 1, nC);
  const int T = Ny - 1 - c * B;
  const int frow = T - B - t;
  const bool fr = t >= KUU && t < WB && frow >= 0;
  cx<double> fresh[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) fresh[r] = fr ? vec[r * Ny + frow] : cx<double>{0, 0};
  cx<double> acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = cx<double>{0, 0};
  window_dot<WB, NR, window_batch(PO, NR)>(acc, g, win);
  wave_sync();
  const int row = T - t;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    if (t < B && row >= 0 && live) out[r * vs + row] = acc[r];
    if (t >= B && t < WB) win[r * 32 + t - B] = acc[r];
    if (t >= KUU && t < WB) win[r * 32 + t] = fresh[r];
  }
  wave_sync();
}

// A launch's workgroups go to the 8 XCDs (each with its own L2) in turn:
// renumbered so that each XCD takes one contiguous range of slots, where
// neighbouring modes (grouped by k^2) share chunk-map cache lines.
__device__ __forceinline__ int xcd_block(int b, int G) {
  constexpr int X = 8;
  const int q = G / X, r = G % X, x = b % X, i = b / X;
  return x * q + (x < r ? x : r) + i;
}

template <int PO, int NR, int S>
__global__ void __launch_bounds__(64, chunk_waves(PO, NR))
    k_band_chunks(ImmArgs a, const double* __restrict__ fmap, const double* __restrict__ bmap,
                  const int2* __restrict__ sched, int U, int vin, int vout) {
  if (a.act && *a.act == 0) return;
  constexpr int KL = PO, KUU = 2 * PO, B = kChunk, WF = B + KL, WB = B + KUU;
  static_assert(WB <= 32, "a window fits the 32 lanes of a mode");
  static_assert(S >= 2, "a map slot in use and at least one being filled");
  extern __shared__ cx<double> sh[];  // [2][NR][Ny] vectors, then [2][NR][32] windows
  const int Ny = a.Ny, M = a.M, nC = (Ny + B - 1) / B;
  const int lane = threadIdx.x, half = lane >> 5, t = lane & 31;
  const int slot = xcd_block(blockIdx.x, gridDim.x) * 2 + half;
  const bool live = slot < M;  // a spare half repeats the last mode and stores nothing
  const int2 mu = sched[slot];
  const int m = mu.x, u = mu.y;
  cx<double>* vec = sh + (size_t)half * NR * Ny;
  cx<double>* win = sh + (size_t)2 * NR * Ny + (size_t)half * NR * 32;
  const size_t vs = (size_t)M * Ny;  // between work vectors (VS = 1: rows contiguous)
  const cx<double>* __restrict__ rin = a.scratch + ((size_t)vin * M + m) * Ny;
  cx<double>* __restrict__ out = a.scratch + ((size_t)vout * M + m) * Ny;
  // the first maps' loads, then the right-hand sides' to LDS (no global
  // load on the chunk chain), eight rows per lane in flight at a time, all
  // under one wait; lanes past the last row load and store that row again
  // (the same value), so nothing branches
  double f[S][WF];
  unsigned of[WF];
  map_offsets<WF, KL>(of, U, u, t);
#pragma unroll
  for (int q = 0; q < S - 1; ++q) map_row<WF, KL>(f[q], fmap, of, U, q, nC);
  for (int i0 = 0; i0 < Ny; i0 += 8 * 32) {
    cx<double> v[8][NR];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int i = min(i0 + q * 32 + t, Ny - 1);
#pragma unroll
      for (int r = 0; r < NR; ++r) v[q][r] = rin[r * vs + i];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int i = min(i0 + q * 32 + t, Ny - 1);
#pragma unroll
      for (int r = 0; r < NR; ++r) vec[r * Ny + i] = v[q][r];
    }
  }
  {
    wave_sync();
    if (t < WF) {
#pragma unroll
      for (int r = 0; r < NR; ++r) win[r * 32 + t] = t < Ny ? vec[r * Ny + t] : cx<double>{0, 0};
    }
    wave_sync();
    for (int c = 0; c < nC; c += S) {
#pragma unroll
      for (int q = 0; q < S; ++q) {
        if (c + q >= nC) break;
        fwd_chunk<PO, NR, S>(f[q], f[(q + S - 1) % S], fmap, of, U, vec, win, t, c + q, nC, Ny);
      }
    }
  }
  {
    double g[S][WB];
    unsigned og[WB];
    map_offsets<WB, 0>(og, U, u, t);
#pragma unroll
    for (int q = 0; q < S - 1; ++q) map_row<WB, 0>(g[q], bmap, og, U, q, nC);
    if (t < WB) {
#pragma unroll
      for (int r = 0; r < NR; ++r)
        win[r * 32 + t] = (Ny - 1 - t >= 0) ? vec[r * Ny + Ny - 1 - t] : cx<double>{0, 0};
    }
    wave_sync();
    for (int c = 0; c < nC; c += S) {
#pragma unroll
      for (int q = 0; q < S; ++q) {
        if (c + q >= nC) break;
        back_chunk<PO, NR, S>(g[q], g[(q + S - 1) % S], bmap, og, U, out, vs, live, vec, win, t,
                              c + q, nC, Ny);
      }
    }
  }
}

// Fused chunked form: two launches per influence-matrix step instead of
// five.  Phases 1 and 2 (k_imm_lsolve): each mode's f_hat is formed by its
// 32 lanes straight into the LDS row the chunked L_k solve starts from; the
// solution p_P goes to V_P.  Phases 3 to 5 (k_imm_hsolve): R_c for the three
// components straight into LDS, the H_k solve with its results kept in
// LDS, then alpha and the assembly.  The per-row arithmetic and the chunk
// chains are those of the five-launch form (same values, bit for bit); only
// the work vectors' round trips through memory and three launches go.  The
// forward and back chunk loops are those of k_band_chunks.
template <int PO, int NR, int S>
__device__ __forceinline__ void chunk_solve(const double* __restrict__ fmap, const double* __restrict__ bmap,
                                            int U, int u, int t, int nC, int Ny, cx<double>* vec, cx<double>* win,
                                            double (&f)[S][kChunk + PO], const unsigned (&of)[kChunk + PO],
                                            cx<double>* __restrict__ out, size_t vs, bool live) {
  constexpr int KL = PO, KUU = 2 * PO, B = kChunk, WF = B + KL, WB = B + KUU;
  wave_sync();
  if (t < WF) {
#pragma unroll
    for (int r = 0; r < NR; ++r) win[r * 32 + t] = t < Ny ? vec[r * Ny + t] : cx<double>{0, 0};
  }
  wave_sync();
  for (int c = 0; c < nC; c += S) {
#pragma unroll
    for (int q = 0; q < S; ++q) {
      if (c + q >= nC) break;
      fwd_chunk<PO, NR, S>(f[q], f[(q + S - 1) % S], fmap, of, U, vec, win, t, c + q, nC, Ny);
    }
  }
  double g[S][WB];
  unsigned og[WB];
  map_offsets<WB, 0>(og, U, u, t);
#pragma unroll
  for (int q = 0; q < S - 1; ++q) map_row<WB, 0>(g[q], bmap, og, U, q, nC);
  if (t < WB) {
#pragma unroll
    for (int r = 0; r < NR; ++r) win[r * 32 + t] = (Ny - 1 - t >= 0) ? vec[r * Ny + Ny - 1 - t] : cx<double>{0, 0};
  }
  wave_sync();
  for (int c = 0; c < nC; c += S) {
#pragma unroll
    for (int q = 0; q < S; ++q) {
      if (c + q >= nC) break;
      back_chunk<PO, NR, S>(g[q], g[(q + S - 1) % S], bmap, og, U, out, vs, live, vec, win, t, c + q, nC, Ny);
    }
  }
}

// LDS words (complex) per workgroup of the fused kernels
__host__ __device__ constexpr size_t imm_lsolve_lds(int Ny) { return 2 * ((size_t)5 * Ny + 32); }
__host__ __device__ constexpr size_t imm_hsolve_lds(int Ny) { return 2 * ((size_t)7 * Ny + 96); }

template <int PO, int S>
__global__ void __launch_bounds__(64, chunk_waves(PO, 1))
    k_imm_lsolve(ImmArgs a, const double* __restrict__ fmap, const double* __restrict__ bmap,
                 const int2* __restrict__ sched, int U) {
  if (a.act && *a.act == 0) return;
  constexpr int KL = PO, WF = kChunk + KL;
  // per half: the solve's vector [Ny], its window [32], then d_hat and the
  // y rows of u^n_y, N_j,y, N_n,y [4][Ny]
  extern __shared__ cx<double> sh[];
  const int Ny = a.Ny, M = a.M, nC = (Ny + kChunk - 1) / kChunk;
  const int lane = threadIdx.x, half = lane >> 5, t = lane & 31;
  const int slot = xcd_block(blockIdx.x, gridDim.x) * 2 + half;
  const bool live = slot < M;  // a spare half repeats the last mode and stores nothing
  const int2 mu = sched[slot];
  const int m = mu.x, u = mu.y;
  cx<double>* vec = sh + (size_t)half * Ny;
  cx<double>* win = sh + (size_t)2 * Ny + (size_t)half * 32;
  cx<double>* dh = sh + (size_t)2 * Ny + 64 + (size_t)half * 4 * Ny;
  cx<double>* un1 = dh + Ny;
  cx<double>* nj1 = dh + 2 * Ny;
  cx<double>* nn1 = dh + 3 * Ny;
  // the first maps' loads, then the right-hand side
  double f[S][WF];
  unsigned of[WF];
  map_offsets<WF, KL>(of, U, u, t);
#pragma unroll
  for (int q = 0; q < S - 1; ++q) map_row<WF, KL>(f[q], fmap, of, U, q, nC);
  const size_t comp = (size_t)M * Ny, row = (size_t)m * Ny;
  const double kx = a.mkx[m], kz = a.mkz[m];
  const double k2 = kx * kx + kz * kz;
  for (int y = t; y < Ny; y += 32) {
    un1[y] = a.un[comp + row + y];
    nj1[y] = a.nj[comp + row + y];
    nn1[y] = a.nn[comp + row + y];
  }
  wave_sync();
  for (int y = t; y < Ny; y += 32) dh[y] = imm_dhat_at<PO>(a, un1, row, comp, kx, kz, y);
  wave_sync();
  for (int y = t; y < Ny; y += 32) vec[y] = imm_fhat_at<PO>(a, dh, nj1, nn1, row, comp, kx, kz, k2, y);
  chunk_solve<PO, 1, S>(fmap, bmap, U, u, t, nC, Ny, vec, win, f, of, a.scratch + ((size_t)V_P * M + m) * Ny,
                        (size_t)M * Ny, live);
}

template <int PO, int S>
__global__ void __launch_bounds__(64, chunk_waves(PO, 3))
    k_imm_hsolve(ImmArgs a, const double* __restrict__ fmap, const double* __restrict__ bmap,
                 const int2* __restrict__ sched, int U) {
  if (a.act && *a.act == 0) return;
  constexpr int KL = PO, WF = kChunk + KL;
  // per half: the solve's vectors [3][Ny], windows [3][32], then the p_P
  // row and the rows of u^n [4][Ny], whose place the solutions take
  extern __shared__ cx<double> sh[];
  const int Ny = a.Ny, M = a.M, nC = (Ny + kChunk - 1) / kChunk;
  const int lane = threadIdx.x, half = lane >> 5, t = lane & 31;
  const int slot = xcd_block(blockIdx.x, gridDim.x) * 2 + half;
  const bool live = slot < M;
  const int2 mu = sched[slot];
  const int m = mu.x, u = mu.y;
  cx<double>* vec = sh + (size_t)half * 3 * Ny;
  cx<double>* win = sh + (size_t)6 * Ny + (size_t)half * 96;
  cx<double>* pp = sh + (size_t)6 * Ny + 192 + (size_t)half * 4 * Ny;
  cx<double>* ur = pp + Ny;
  double f[S][WF];
  unsigned of[WF];
  map_offsets<WF, KL>(of, U, u, t);
#pragma unroll
  for (int q = 0; q < S - 1; ++q) map_row<WF, KL>(f[q], fmap, of, U, q, nC);
  const size_t comp = (size_t)M * Ny, row = (size_t)m * Ny;
  const double kx = a.mkx[m], kz = a.mkz[m];
  const double k2 = kx * kx + kz * kz;
  const cx<double>* P = a.scratch + ((size_t)V_P * M + m) * Ny;
  for (int y = t; y < Ny; y += 32) {
    pp[y] = P[y];
#pragma unroll
    for (int c = 0; c < 3; ++c) ur[c * Ny + y] = a.un[c * comp + row + y];
  }
  wave_sync();
  for (int y = t; y < Ny; y += 32) {
    cx<double> r[3];
    imm_r_at<PO>(a, pp, ur, row, comp, kx, kz, k2, y, r);
#pragma unroll
    for (int c = 0; c < 3; ++c) vec[c * Ny + y] = r[c];
  }
  // the solutions overwrite the p_P row and the u^n rows, read above
  // before the solve's first wave_sync
  cx<double>* sol = pp;
  chunk_solve<PO, 3, S>(fmap, bmap, U, u, t, nC, Ny, vec, win, f, of, sol, (size_t)Ny, true);
  wave_sync();
  cx<double> al1, al2;
  imm_alpha<PO>(a, m, sol + Ny, 1, al1, al2);
  double acc = 0.0;
  for (int y = t; y < Ny; y += 32)
    acc += a.qw[y] * imm_assemble_at(a, m, row, comp, kx, kz, sol, sol + Ny, sol + 2 * Ny, 1, al1, al2, y, live);
  // the mode's 32 lanes: a fixed pairwise order
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 32);
  if (t == 0 && live && a.partials) a.partials[m] = acc;
}

// Phase 3 (flows/plane_couette.py:449-472): R_c for the three components
// (zero on the walls) -> V_R + c, with the pressure row in LDS.
template <int PO>
__global__ void __launch_bounds__(kImmT) k_imm_rc(ImmArgs a) {
  if (a.act && *a.act == 0) return;
  // LDS: the p_P row, then the rows of u^n (the D2 stencils' operands)
  extern __shared__ cx<double> pp[];
  const int m = blockIdx.x, Ny = a.Ny;
  const size_t comp = (size_t)a.M * Ny, row = (size_t)m * Ny;
  const double kx = a.mkx[m], kz = a.mkz[m];
  const double k2 = kx * kx + kz * kz;
  const cx<double>* P = svec(a, V_P, m);
  cx<double>* ur = pp + Ny;  // [c][y]
  for (int y = threadIdx.x; y < Ny; y += blockDim.x) {
    pp[y] = P[y * VS];
#pragma unroll
    for (int c = 0; c < 3; ++c) ur[c * Ny + y] = a.un[c * comp + row + y];
  }
  __syncthreads();
  for (int y = threadIdx.x; y < Ny; y += blockDim.x) {
    cx<double> r[3];
    imm_r_at<PO>(a, pp, ur, row, comp, kx, kz, k2, y, r);
#pragma unroll
    for (int c = 0; c < 3; ++c) svec(a, V_R + c, m)[y * VS] = r[c];
  }
}

// Phase 5 (flows/plane_couette.py:476-504): alpha = -M^-1 (wall divergence
// of v_arb), the corrected velocity, the correction, and this mode's
// quadrature-weighted k_metric |delta|^2 into partials[m].
template <int PO>
__global__ void __launch_bounds__(kImmT) k_imm_final(ImmArgs a) {
  if (a.act && *a.act == 0) return;
  __shared__ double red[kImmT / 64];
  const int m = blockIdx.x, Ny = a.Ny, M = a.M;
  const size_t comp = (size_t)M * Ny, row = (size_t)m * Ny;
  const double kx = a.mkx[m], kz = a.mkz[m];
  const cx<double>* ua = svec(a, V_R + 0, m);
  const cx<double>* va = svec(a, V_R + 1, m);
  const cx<double>* wa = svec(a, V_R + 2, m);
  cx<double> al1, al2;
  imm_alpha<PO>(a, m, va, VS, al1, al2);
  double acc = 0.0;
  for (int y = threadIdx.x; y < Ny; y += blockDim.x)
    acc += a.qw[y] * imm_assemble_at(a, m, row, comp, kx, kz, ua, va, wa, VS, al1, al2, y, true);
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0 && a.partials) a.partials[m] = s;
}

// Quadrature-weighted norm of a field in (3, M, Ny) layout.
__global__ void __launch_bounds__(256)
    k_walls_norm(const cx<double>* __restrict__ f, const double* __restrict__ qw, int Ny, int M,
                 int Nkx, double* __restrict__ partials) {
  __shared__ double red[4];
  const long long n = (long long)M * Ny;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (i < n) {
    const int m = (int)(i / Ny), y = (int)(i - (long long)m * Ny);
    const double km = (m % Nkx == 0) ? 1.0 : 2.0;
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
      const cx<double> d = f[c * n + i];
      s += km * (d.x * d.x) + km * (d.y * d.y);
    }
    acc = qw[y] * s;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ------------------------------------------------------------------ launchers
template <typename T, int EM, int TY>
static void launch_walls_curl_ty(Ctx* c, const cx<double>* u, cx<T>* A, hipStream_t s) {
  WallTables* t = c->wt;
  const cx<double>* uref = reinterpret_cast<const cx<double>*>(c->uref);
  cx<double>* ucopy = reinterpret_cast<cx<double>*>(c->uref);
  const dim3 grid(((t->Ny + TY - 1) / TY) * ((t->M + kTM - 1) / kTM));
  switch (t->p) {
#define CURL(PO)                                                                                \
  case PO:                                                                                      \
    hipLaunchKernelGGL((k_walls_curl_t<T, PO + 1, EM, TY>), grid, dim3(TY * kTM), 0, s, u, A,   \
                       t->d1w, t->d1j0, t->mkz, t->Ny, t->M, c->act, uref, ucopy);              \
    break;
    CURL(2) CURL(4) CURL(6) CURL(8)
#undef CURL
    default: break;
  }
}
template <typename T, int EM>
static void launch_walls_curl(Ctx* c, const cx<double>* u, cx<T>* A, hipStream_t s) {
  if (c->wt->curl_ty == 64)
    launch_walls_curl_ty<T, EM, 64>(c, u, A, s);
  else
    launch_walls_curl_ty<T, EM, 32>(c, u, A, s);
}

// em: EvalMode; EV_RECORD / EV_INC only in FP32 with the reference buffers.
template <typename T>
static int walls_rhs_t(Ctx* c, const double* state, double* rhs, hipStream_t s, int em) {
  WallTables* t = c->wt;
  const int Ny = t->Ny, M = t->M;
  cx<T>* A = static_cast<cx<T>*>(c->bufA);
  cx<T>* B = static_cast<cx<T>*>(c->bufB);
  const cx<T>* twz = static_cast<const cx<T>*>(c->fp32 ? c->tw32[2] : c->tw64[2]);
  const int ntile_y = (Ny + kTY - 1) / kTY, ntile_m = (M + kTM - 1) / kTM;
  {
    ProfScope ps(c, KID_W_Z_INVERSE, s);
    const cx<double>* u = reinterpret_cast<const cx<double>*>(state);
    if (t->p != 2 && t->p != 4 && t->p != 6 && t->p != 8) return c->fail(DNSMI_EINVAL, "fd_order");
    if (em == EV_PLAIN)
      launch_walls_curl<T, EV_PLAIN>(c, u, A, s);
    else if constexpr (std::is_same<T, float>::value) {
      if (em == EV_RECORD)
        launch_walls_curl<T, EV_RECORD>(c, u, A, s);
      else
        launch_walls_curl<T, EV_INC>(c, u, A, s);
    }
    int rc = check_launch(c, "walls_curl");
    if (rc) return rc;
    const int ntile = (c->Nx1 + 16 - 1) / 16;
    switch (c->Pz) {
#define X(P)                                                                                    \
  case P:                                                                                       \
    hipLaunchKernelGGL((k_z_inverse_plain<T, P, 16, 6>), dim3(Ny * ntile, 6),                   \
                       dim3(16 * ColPlan<P>::threads), 0, s, A, B, twz, c->nz, Ny, c->Nx1,     \
                       c->act);                                                                 \
    break;
      DNSMI_FOR_EACH_P(X)
#undef X
      default: return c->fail(DNSMI_EINVAL, "unsupported padded z length");
    }
    rc = check_launch(c, "walls_z_inverse");
    if (rc) return rc;
  }
  int rc = launch_x_pass<T>(c, B, A, c->base, c->cbase, c->nbase, Ny, KID_W_X_NONLINEAR, s, em);
  if (rc) return rc;
  rc = launch_z_forward_buf<T>(c, A, B, Ny, KID_W_Z_FORWARD, s);
  if (rc) return rc;
  {
    ProfScope ps(c, KID_W_Z_FORWARD, s);
    hipLaunchKernelGGL((k_walls_to_spec<T>), dim3(ntile_y * ntile_m, 3), dim3(kTY * kTM), 0, s, B,
                       reinterpret_cast<cx<double>*>(rhs), Ny, M, c->act,
                       em == EV_INC ? reinterpret_cast<const cx<double>*>(c->nref) : nullptr,
                       em == EV_RECORD ? reinterpret_cast<cx<double>*>(c->nref) : nullptr);
  }
  return check_launch(c, "walls_to_spec");
}

// One evaluation in the context's current precision (see EvalMode; EV_INC
// without a recorded reference records one, every FP64 evaluation is
// EV_PLAIN).
static int walls_rhs_mode(Ctx* c, const double* state, double* rhs, hipStream_t s, int em) {
  if (em == EV_INC && !c->ref_valid) em = EV_RECORD;
  if (!c->fp32 || !c->split) em = EV_PLAIN;
  if (!c->fp32) return walls_rhs_t<double>(c, state, rhs, s, EV_PLAIN);
  const int rc = walls_rhs_t<float>(c, state, rhs, s, em);
  if (rc == DNSMI_OK && em == EV_RECORD) c->ref_valid = true;
  return rc;
}

int walls_get_rhs(Ctx* c, const double* state, double* rhs, hipStream_t s) {
  return walls_rhs_mode(c, state, rhs, s, EV_PLAIN);
}

// N(state), the first evaluation of a guessed step: plain, after which the
// step's first corrector evaluation records the split reference.
int walls_guess_rhs(Ctx* c, const double* state, double* rhs, hipStream_t s) {
  c->ref_valid = false;
  return walls_rhs_mode(c, state, rhs, s, EV_PLAIN);
}

static int walls_imm(Ctx* c, const double* un, const double* uj, const double* nn,
                     const double* nj, double* out, double* corr, double* err, hipStream_t s) {
  WallTables* t = c->wt;
  const int nb = t->M;  // one norm partial per mode
  if (nb > c->max_partials) return c->fail(DNSMI_EINVAL, "imm: too many blocks");
  ImmArgs a{};
  a.un = reinterpret_cast<const cx<double>*>(un);
  a.uj = reinterpret_cast<const cx<double>*>(uj);
  a.nn = reinterpret_cast<const cx<double>*>(nn);
  a.nj = reinterpret_cast<const cx<double>*>(nj);
  a.out = reinterpret_cast<cx<double>*>(out);
  a.corr = reinterpret_cast<cx<double>*>(corr);
  a.partials = c->partials;
  a.d1w = t->d1w;
  a.d2w = t->d2w;
  a.qw = t->qw;
  a.d1j0 = t->d1j0;
  a.d2j0 = t->d2j0;
  a.lk_ab = t->lk_ab;
  a.hk_ab = t->hk_ab;
  a.lk_piv = t->lk_piv;
  a.hk_piv = t->hk_piv;
  a.vq = t->vq;
  a.umode = t->umode;
  a.minv = t->minv;
  a.mkx = t->mkx;
  a.mkz = t->mkz;
  a.scratch = reinterpret_cast<cx<double>*>(t->scratch);
  a.act = c->act;
  a.Ny = t->Ny;
  a.M = t->M;
  a.Nkx = t->Nkx;
  a.dt = c->cfg.dt;
  a.c = c->cfg.implicitness;
  a.nu = 1.0 / c->cfg.re;
  {
    if (!c->sweep_seq && !t->lk_fmap) return c->fail(DNSMI_EINVAL, "imm: chunk maps not built");
    ProfScope ps(c, KID_W_IMM, s);
    const size_t lds = (size_t)t->Ny * sizeof(cx<double>);
    const dim3 gm(t->M), gs((t->M + kSweepLanes - 1) / kSweepLanes, 1),
        gs3((t->M + kSweepLanes - 1) / kSweepLanes, 3), gc((t->M + 1) / 2);
    // chunked solves: per workgroup two modes' vectors and windows
    const size_t lds_c1 = (size_t)2 * 1 * (t->Ny + 32) * sizeof(cx<double>);
    const size_t lds_c3 = (size_t)2 * 3 * (t->Ny + 32) * sizeof(cx<double>);
    if (4 * lds > 65536 || (!c->sweep_seq && lds_c3 > 65536))
      return c->fail(DNSMI_EINVAL, "imm: ny too large for the per-mode LDS rows");
    // the fused form reads u^n while other modes already store the new
    // velocity: it needs them apart
    if (c->imm_fused && !c->sweep_seq && out != un) {
      const size_t lds_l = imm_lsolve_lds(t->Ny) * sizeof(cx<double>);
      const size_t lds_h = imm_hsolve_lds(t->Ny) * sizeof(cx<double>);
      if (lds_h > 65536) return c->fail(DNSMI_EINVAL, "imm: ny too large for the fused form's LDS rows");
      switch (t->p) {
#define IMMF(PO)                                                                                   \
  case PO:                                                                                         \
    hipLaunchKernelGGL((k_imm_lsolve<PO, chunk_slots(PO, 1)>), gc, dim3(64), lds_l, s, a, t->lk_fmap, \
                       t->lk_bmap, t->sched, t->U);                                                \
    hipLaunchKernelGGL((k_imm_hsolve<PO, chunk_slots(PO, 3)>), gc, dim3(64), lds_h, s, a, t->hk_fmap, \
                       t->hk_bmap, t->sched, t->U);                                                \
    break;
        IMMF(2) IMMF(4) IMMF(6) IMMF(8)
#undef IMMF
        default: return c->fail(DNSMI_EINVAL, "fd_order");
      }
    } else
    switch (t->p) {
#define IMM(PO)                                                                                   \
  case PO:                                                                                        \
    hipLaunchKernelGGL(k_imm_fhat<PO>, gm, dim3(kImmT), 4 * lds, s, a);                           \
    if (c->sweep_seq)                                                                             \
      hipLaunchKernelGGL(k_band_sweeps<PO>, gs, dim3(kSweepLanes), 0, s, a, t->lk_ab, t->lk_piv,   \
                         (int)V_F, (int)V_T, (int)V_P);                                           \
    else                                                                                          \
      hipLaunchKernelGGL((k_band_chunks<PO, 1, chunk_slots(PO, 1)>), gc, dim3(64), lds_c1, s, a, t->lk_fmap,          \
                         t->lk_bmap, t->sched, t->U, (int)V_F, (int)V_P);                         \
    hipLaunchKernelGGL(k_imm_rc<PO>, gm, dim3(kImmT), 4 * lds, s, a);                             \
    if (c->sweep_seq)                                                                             \
      hipLaunchKernelGGL(k_band_sweeps<PO>, gs3, dim3(kSweepLanes), 0, s, a, t->hk_ab, t->hk_piv, \
                         (int)V_R, (int)V_T, (int)V_R);                                           \
    else                                                                                          \
      hipLaunchKernelGGL((k_band_chunks<PO, 3, chunk_slots(PO, 3)>), gc, dim3(64), lds_c3, s, a, t->hk_fmap,          \
                         t->hk_bmap, t->sched, t->U, (int)V_R, (int)V_R);                         \
    hipLaunchKernelGGL(k_imm_final<PO>, gm, dim3(kImmT), 0, s, a);                                \
    break;
      IMM(2) IMM(4) IMM(6) IMM(8)
#undef IMM
      default: return c->fail(DNSMI_EINVAL, "fd_order");
    }
  }
  int rc = check_launch(c, "imm");
  if (rc || !err) return rc;
  return finalize_norm(c, nb, err, s);
}

// Chooses the band-solve form by timing one influence-matrix step of each
// on zeroed buffers (both forms are exact solves of the same systems).
static int walls_autotune(Ctx* c) {
  if (hipMemset(c->rhs0, 0, c->spec_elems * 16) != hipSuccess ||
      hipMemset(c->bufA, 0, c->spec_elems * 16) != hipSuccess)
    return c->fail(DNSMI_EHIP, "autotune: memset");
  // zero fields in, the new velocity into a buffer of its own (the fused
  // form needs it apart from u^n); bufA is transform scratch, unused here
  double* z = c->rhs0;
  double* o = static_cast<double*>(c->bufA);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return c->fail(DNSMI_EHIP, "autotune: events");
  // a form that cannot run here (nonzero return) never wins
  auto time_imm = [&](bool seq, bool fused) {
    c->sweep_seq = seq;
    c->imm_fused = fused;
    if (walls_imm(c, z, z, z, z, o, nullptr, nullptr, nullptr) != DNSMI_OK)  // warm-up
      return (double)INFINITY;
    (void)hipEventRecord(e0, nullptr);
    for (int i = 0; i < 3; ++i)
      if (walls_imm(c, z, z, z, z, o, nullptr, nullptr, nullptr) != DNSMI_OK) return (double)INFINITY;
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms / 3.0;
  };
  // three alternating rounds, each form's best kept: a single round is
  // swayed by clock ramps and a profiler's first-launch costs, which fall
  // on whichever form runs first
  double tc = INFINITY, tf = INFINITY, ts = INFINITY;
  for (int round = 0; round < 3; ++round) {
    tc = std::min(tc, time_imm(false, false));
    tf = std::min(tf, time_imm(false, true));
    ts = std::min(ts, time_imm(true, false));
  }
  if (!(ts < INFINITY)) return c->fail(DNSMI_EINVAL, "autotune: the sequential band solve failed: " + c->err);
  c->err.clear();
  c->sweep_seq = ts < tc && ts < tf;
  c->imm_fused = !c->sweep_seq && tf < tc;
  if (c->sweep_seq) {  // the chunk maps are not needed any more
    WallTables* t = c->wt;
    for (double** p : {&t->lk_fmap, &t->lk_bmap, &t->hk_fmap, &t->hk_bmap}) {
      c->release(*p);
      *p = nullptr;
    }
    c->release(t->sched);
    t->sched = nullptr;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  char buf[160];
  snprintf(buf, sizeof buf, "band_solve=%s(chunked %.1f us, fused %.1f us, sequential %.1f us) ",
           c->sweep_seq ? "sequential" : (c->imm_fused ? "fused" : "chunked"), 1e3 * tc, 1e3 * tf, 1e3 * ts);
  c->tuning = buf;
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? DNSMI_OK : c->fail(DNSMI_EHIP, std::string("autotune: ") + hipGetErrorString(e));
}

int walls_predict(Ctx* c, const double* state, const double* rhs, double* pred, hipStream_t s) {
  return walls_imm(c, state, state, rhs, rhs, pred, nullptr, nullptr, s);
}

int walls_correct(Ctx* c, const double* state_prev, const double* pred_in,
                  const double* rhs_prev, const double* rhs_next, double* pred_out,
                  double* corr, hipStream_t s) {
  return walls_imm(c, state_prev, pred_in, rhs_prev, rhs_next, pred_out, corr, nullptr, s);
}

int walls_norm(Ctx* c, const double* corr, double* err, hipStream_t s) {
  WallTables* t = c->wt;
  const long long n = (long long)t->M * t->Ny;
  const int nb = (int)((n + 255) / 256);
  if (nb > c->max_partials) return c->fail(DNSMI_EINVAL, "norm: too many blocks");
  {
    ProfScope ps(c, KID_W_NORM, s);
    hipLaunchKernelGGL(k_walls_norm, dim3(nb), dim3(256), 0, s,
                       reinterpret_cast<const cx<double>*>(corr), t->qw, t->Ny, t->M, t->Nkx,
                       c->partials);
  }
  int rc = check_launch(c, "walls_norm");
  if (rc) return rc;
  return finalize_norm(c, nb, err, s);
}

int extrap_walls(Ctx* c, const double* n0, const float* h1, float* h2, double* next,
                 const uint16_t* wb, const float* bias, int H, float s_1, float s_2, float s_out,
                 int mode, hipStream_t s);

// Wall-bounded corrector guess (no reference counterpart).  The walls
// corrector's fixed point is u* = S(u^n; N(u*)) with S the influence-matrix
// solve (the explicit term is the previous evaluation,
// flows/plane_couette.py:534-539), so the guess is the predictor's solve
// (flows/plane_couette.py:507-519) fed with an extrapolated nonlinear term
// N_ext = N^n + dN (extrap_walls: dN = 0, D1, D1 + D2, or D1 + D2 plus the
// network's residual, D1, D2 the differences of N over the FP32 history
// h1 = N^{n-1}, h2 = N^{n-2}; h2 <- N^n).  rhs_prev <- N_ext, and the
// reference's iterate_correction continues from (pred, rhs_prev).
int guess_walls(Ctx* c, const double* state, double* pred, double* rhs_prev, const float* h1,
                float* h2, const uint16_t* wb, const float* bias, int H, float s_1, float s_2,
                float s_out, int mode, hipStream_t s) {
  if (mode < 0 || mode > 3) return c->fail(DNSMI_EINVAL, "guess: mode must be 0, 1, 2 or 3");
  if (mode == 2 && H != 32 && H != 64)
    return c->fail(DNSMI_EINVAL, "guess: hidden width must be 32 or 64");
  // N^n; the first corrector pass from the guess records the split reference
  int rc = walls_guess_rhs(c, state, c->rhs0, s);
  if (rc) return rc;
  if ((rc = extrap_walls(c, c->rhs0, h1, h2, rhs_prev, wb, bias, H, s_1, s_2, s_out, mode, s)))
    return rc;
  return walls_predict(c, state, rhs_prev, pred, s);
}

// The predictor's evaluation N(u^n) is plain; the first corrector
// evaluation N(u^1) records the split reference (mixed form), so every later
// increment is of the corrector's own size.
int walls_predict_and_correct(Ctx* c, const double* state, double* pred, double* rhs_next,
                              double* err, hipStream_t s) {
  double* n0 = c->rhs0;
  int rc = walls_rhs_mode(c, state, n0, s, EV_PLAIN);
  if (rc) return rc;
  if ((rc = walls_imm(c, state, state, n0, n0, pred, nullptr, nullptr, s))) return rc;
  c->ref_valid = false;
  if ((rc = walls_rhs_mode(c, pred, rhs_next, s, EV_RECORD))) return rc;
  return walls_imm(c, state, pred, n0, rhs_next, pred, nullptr, err, s);
}

int walls_iterate_correction(Ctx* c, const double* state_prev, double* pred,
                             const double* rhs_prev, double* rhs_next, double* err,
                             hipStream_t s) {
  // N_n := rhs_prev, the previous evaluation (flows/plane_couette.py:534-539)
  if ((const void*)rhs_next == (const void*)rhs_prev) {
    int rc = walls_rhs_mode(c, pred, c->rhs0, s, EV_INC);
    if (rc) return rc;
    if ((rc = walls_imm(c, state_prev, pred, rhs_prev, c->rhs0, pred, nullptr, err, s))) return rc;
    const hipError_t e = hipMemcpyAsync(rhs_next, c->rhs0, c->spec_elems * 16,
                                        hipMemcpyDeviceToDevice, s);
    return e == hipSuccess ? DNSMI_OK : c->fail(DNSMI_EHIP, hipGetErrorString(e));
  }
  int rc = walls_rhs_mode(c, pred, rhs_next, s, EV_INC);
  if (rc) return rc;
  return walls_imm(c, state_prev, pred, rhs_prev, rhs_next, pred, nullptr, err, s);
}

// The set-up tables in the reference layouts (for tests), read back from the
// device: which = 0 Lk_ab, 1 Hk_ab (mode, row, y), 2 v1, 3 q1 (mode, y),
// 4 M_inv (mode, 2, 2), 5 ys; 6 Lk_piv, 7 Hk_piv (mode, y; ints written as
// doubles); 8 .. 11 the packed chunk maps as stored (L_k forward, L_k back,
// H_k forward, H_k back; empty without them); 12 the per-class homogeneous
// solutions as stored ([class][y][4]); 13 the word counts of 8 .. 11 and
// the number of k^2 classes.
int walls_table(Ctx* c, int which, double* out, long long n) {
  WallTables* t = c->wt;
  if (!t) return c->fail(DNSMI_EINVAL, "not a wall-bounded context");
  const int Ny = t->Ny, M = t->M, rows = 3 * t->p + 1, U = t->U;
  auto fetch = [&](const void* src, size_t bytes, void* dst) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
  };
  auto want = [&](long long k) { return n == k ? DNSMI_OK : c->fail(DNSMI_EINVAL, "table size mismatch"); };
  int rc;
  switch (which) {
    case 0:
    case 1: {
      if ((rc = want((long long)M * rows * Ny))) return rc;
      std::vector<double> d((size_t)Ny * rows * M);
      if (!fetch(which == 0 ? t->lk_ab : t->hk_ab, d.size() * sizeof(double), d.data()))
        return c->fail(DNSMI_EHIP, "read table");
      for (int m = 0; m < M; ++m)
        for (int r = 0; r < rows; ++r)
          for (int j = 0; j < Ny; ++j) out[((size_t)m * rows + r) * Ny + j] = d[((size_t)j * rows + r) * M + m];
      return DNSMI_OK;
    }
    case 2:
    case 3:
    case 12: {
      std::vector<double> d(4ull * U * Ny);
      if (!fetch(t->vq, d.size() * sizeof(double), d.data())) return c->fail(DNSMI_EHIP, "read table");
      if (which == 12) {
        if ((rc = want((long long)d.size()))) return rc;
        std::copy(d.begin(), d.end(), out);
        return DNSMI_OK;
      }
      if ((rc = want((long long)M * Ny))) return rc;
      const int q = which == 2 ? 0 : 2;  // v1 = H^-1(-D1 p1), q1 = H^-1(p1)
      for (int m = 0; m < M; ++m)
        for (int j = 0; j < Ny; ++j) out[(size_t)m * Ny + j] = d[((size_t)t->h_umode[m] * Ny + j) * 4 + q];
      return DNSMI_OK;
    }
    case 4: {
      if ((rc = want(4LL * M))) return rc;
      std::vector<double> d(4ull * M);
      if (!fetch(t->minv, d.size() * sizeof(double), d.data())) return c->fail(DNSMI_EHIP, "read table");
      for (int m = 0; m < M; ++m)
        for (int q = 0; q < 4; ++q) out[(size_t)m * 4 + q] = d[(size_t)q * M + m];
      return DNSMI_OK;
    }
    case 5:
      if ((rc = want((long long)t->h_ys.size()))) return rc;
      std::copy(t->h_ys.begin(), t->h_ys.end(), out);
      return DNSMI_OK;
    case 6:
    case 7: {
      if ((rc = want((long long)M * Ny))) return rc;
      std::vector<int> d((size_t)Ny * M);
      if (!fetch(which == 6 ? t->lk_piv : t->hk_piv, d.size() * sizeof(int), d.data()))
        return c->fail(DNSMI_EHIP, "read table");
      for (int m = 0; m < M; ++m)
        for (int j = 0; j < Ny; ++j) out[(size_t)m * Ny + j] = d[(size_t)j * M + m];
      return DNSMI_OK;
    }
    case 8:
    case 9:
    case 10:
    case 11: {
      const int kl = t->p, B = kChunk, WF = B + kl, WB = B + 2 * kl, nC = (Ny + B - 1) / B;
      const bool fwd = which == 8 || which == 10;
      const size_t words = (size_t)nC * U * (fwd ? map_pre(WF, kl, WF) : map_pre(WB, 0, WB));
      const double* src = which == 8 ? t->lk_fmap : which == 9 ? t->lk_bmap : which == 10 ? t->hk_fmap : t->hk_bmap;
      if (!src) return want(0);
      if ((rc = want((long long)words))) return rc;
      if (!fetch(src, words * sizeof(double), out)) return c->fail(DNSMI_EHIP, "read table");
      return DNSMI_OK;
    }
    case 13: {  // the words of tables 8 .. 11 (0 where absent), then the class count
      if ((rc = want(5))) return rc;
      out[4] = U;
      const int kl = t->p, B = kChunk, WF = B + kl, WB = B + 2 * kl, nC = (Ny + B - 1) / B;
      const double* maps[4] = {t->lk_fmap, t->lk_bmap, t->hk_fmap, t->hk_bmap};
      for (int i = 0; i < 4; ++i)
        out[i] = maps[i] ? (double)((size_t)nC * U * (i % 2 == 0 ? map_pre(WF, kl, WF) : map_pre(WB, 0, WB))) : 0.0;
      return DNSMI_OK;
    }
    default: return c->fail(DNSMI_EINVAL, "unknown table");
  }
}

}  // namespace dnsmi
