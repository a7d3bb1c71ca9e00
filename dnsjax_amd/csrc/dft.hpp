// Register-resident small DFTs and the two-step (register / LDS / register)
// large transform used by every dealiased FFT pass, written for CDNA4
// (64-lane wavefronts, LDS exchange between the two steps).
//
// A length-P transform (P = N1 * N2, the 3/2-padded grid sizes 12..256 of
// form 2^a 3^b) is split Cooley-Tukey style: with input index
// n = N2 * a + b and output index k = k1 + N1 * k2,
//   step A: thread b (b < N2) holds x[N2 a + b], a = 0..N1-1, in registers,
//           does a DFT-N1 over a, multiplies by w_P^{b k1}, stores to LDS;
//   step B: thread k1 (k1 < N1) reads the N2 values of its k1 from LDS and
//           does a DFT-N2 over b, giving X[k1 + N1 k2], k2 = 0..N2-1.
// Small DFTs are fully unrolled with compile-time twiddles (constexpr
// tables folded into immediates), so each pass costs one LDS round trip.
#pragma once

#include <hip/hip_runtime.h>

namespace dnsmi {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

// A store marked non-temporal: the line is written to memory without being
// kept in the caches, which then keep what the writing kernel still reads.
template <typename T>
__device__ __forceinline__ void st_stream(cx<T>* p, cx<T> v) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
}
// A plain or a non-temporal store, as a wave-uniform flag selects.
template <typename T>
__device__ __forceinline__ void st_out(cx<T>* p, cx<T> v, bool nt) {
  if (nt)
    st_stream(p, v);
  else
    *p = v;
}

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
// i * a
template <typename T>
__device__ __forceinline__ cx<T> cmuli(cx<T> a) { return {-a.y, a.x}; }
// -i * a
template <typename T>
__device__ __forceinline__ cx<T> cmulmi(cx<T> a) { return {a.y, -a.x}; }
// exp(DIR * i * pi / 2) * a, DIR = +1 (inverse) or -1 (forward)
template <int DIR, typename T>
__device__ __forceinline__ cx<T> cmul_dir_i(cx<T> a) { return DIR > 0 ? cmuli(a) : cmulmi(a); }

// ------------------------------------------------------------ constexpr trig
constexpr double kPi = 3.14159265358979323846;

constexpr double ct_sin_reduced(double r) {  // |r| <= pi/4, Taylor to r^25
  double r2 = r * r, t = r, s = r;
  for (int k = 1; k <= 12; ++k) {
    t = -t * r2 / double((2 * k) * (2 * k + 1));
    s += t;
  }
  return s;
}
constexpr double ct_cos_reduced(double r) {
  double r2 = r * r, t = 1.0, s = 1.0;
  for (int k = 1; k <= 12; ++k) {
    t = -t * r2 / double((2 * k - 1) * (2 * k));
    s += t;
  }
  return s;
}

struct CosSin {
  double c, s;
};

// cos and sin of 2 pi m / N, reduced by octant so that the quarter-turn
// multiples are exact.
constexpr CosSin ct_unit(long m, long N) {
  m %= N;
  if (m < 0) m += N;
  const long o = (8 * m) / N;
  const long a = 8 * m - o * N;
  long q = 0;
  double r = 0.0;
  bool neg = false;
  if ((o & 1) == 0) {
    q = o / 2;
    r = (kPi / 4) * double(a) / double(N);
  } else {
    q = (o + 1) / 2;
    r = (kPi / 4) * double(N - a) / double(N);
    neg = true;
  }
  const double C = ct_cos_reduced(r);
  const double S = neg ? -ct_sin_reduced(r) : ct_sin_reduced(r);
  switch (q & 3) {
    case 0: return {C, S};
    case 1: return {-S, C};
    case 2: return {-C, -S};
    default: return {S, -C};
  }
}

template <int N>
struct TwiddleTable {
  double c[N], s[N];
  constexpr TwiddleTable() : c{}, s{} {
    for (int m = 0; m < N; ++m) {
      const CosSin v = ct_unit(m, N);
      c[m] = v.c;
      s[m] = v.s;
    }
  }
};

// a * exp(DIR * 2 pi i e / N) with e a compile-time constant after unrolling.
template <int N, int DIR, typename T>
__device__ __forceinline__ cx<T> twiddle_const(cx<T> a, int e) {
  if (e == 0) return a;
  if (4 * e == N) return cmul_dir_i<DIR>(a);
  if (2 * e == N) return {-a.x, -a.y};
  if (4 * e == 3 * N) return cmul_dir_i<-DIR>(a);
  constexpr TwiddleTable<N> tab{};
  const T c = static_cast<T>(tab.c[e]);
  const T s = static_cast<T>(DIR * tab.s[e]);
  return {a.x * c - a.y * s, a.x * s + a.y * c};
}

// ------------------------------------------------------------ register DFTs
template <int N, int DIR, typename T>
struct Dft;

template <int DIR, typename T>
struct Dft<1, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>*) {}
};

template <int DIR, typename T>
struct Dft<2, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    const cx<T> t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <int DIR, typename T>
struct Dft<3, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    constexpr T h = static_cast<T>(0.86602540378443864676);  // sqrt(3)/2
    const cx<T> s = cadd(a[1], a[2]);
    const cx<T> d = csub(a[1], a[2]);
    const cx<T> m = {a[0].x - static_cast<T>(0.5) * s.x, a[0].y - static_cast<T>(0.5) * s.y};
    // DIR * i * h * d
    const cx<T> r = cmul_dir_i<DIR>(cx<T>{h * d.x, h * d.y});
    a[0] = cadd(a[0], s);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
  }
};

template <int DIR, typename T>
struct Dft<4, DIR, T> {
  __device__ __forceinline__ static void run(cx<T>* a) {
    const cx<T> s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
    const cx<T> s13 = cadd(a[1], a[3]), d13 = cmul_dir_i<DIR>(csub(a[1], a[3]));
    a[0] = cadd(s02, s13);
    a[2] = csub(s02, s13);
    a[1] = cadd(d02, d13);
    a[3] = csub(d02, d13);
  }
};

// Generic mixed radix: N = R * M, decimation in time, all in registers.
template <int N, int DIR, typename T>
struct Dft {
  static constexpr int R = (N % 4 == 0) ? 4 : ((N % 2 == 0) ? 2 : 3);
  static constexpr int M = N / R;
  static_assert(R * M == N, "length must be of form 2^a 3^b");
  __device__ __forceinline__ static void run(cx<T>* a) {
    cx<T> sub[R][M];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int m = 0; m < M; ++m) sub[r][m] = a[R * m + r];
#pragma unroll
    for (int r = 0; r < R; ++r) Dft<M, DIR, T>::run(sub[r]);
#pragma unroll
    for (int k = 0; k < M; ++k) {
      cx<T> t[R];
#pragma unroll
      for (int r = 0; r < R; ++r) t[r] = twiddle_const<N, DIR>(sub[r][k], r * k);
      Dft<R, DIR, T>::run(t);
#pragma unroll
      for (int q = 0; q < R; ++q) a[k + M * q] = t[q];
    }
  }
};

// ------------------------------------------------------------ plans
// (N1, N2) split of each supported padded length; N1 >= N2.
template <int P>
struct Plan;
#define DNSMI_PLAN(P_, N1_, N2_)            \
  template <>                               \
  struct Plan<P_> {                         \
    static constexpr int N1 = N1_;          \
    static constexpr int N2 = N2_;          \
    static_assert(N1_ * N2_ == P_, "plan"); \
  };
DNSMI_PLAN(12, 4, 3)
DNSMI_PLAN(16, 4, 4)
DNSMI_PLAN(24, 6, 4)
DNSMI_PLAN(32, 8, 4)
DNSMI_PLAN(48, 8, 6)
DNSMI_PLAN(64, 8, 8)
DNSMI_PLAN(72, 12, 6)
DNSMI_PLAN(96, 12, 8)
DNSMI_PLAN(128, 16, 8)
DNSMI_PLAN(144, 12, 12)
DNSMI_PLAN(192, 16, 12)
#undef DNSMI_PLAN

// Column passes (y and z) hold two step-A inputs at once, so they take the
// smaller factor first: (N1, N2) with N1 <= N2.
template <int P>
struct ColPlan {
  static constexpr int N1 = Plan<P>::N2;
  static constexpr int N2 = Plan<P>::N1;
  static constexpr int threads = N2;  // per column: max(N1, N2)
};

// The list of instantiated lengths, for host-side dispatch.
#define DNSMI_FOR_EACH_P(X) X(12) X(16) X(24) X(32) X(48) X(64) X(72) X(96) X(128) X(144) X(192)

// A copy of v that the compiler cannot treat as loop-invariant.  Multi-pass
// kernels take their thread index through it at the top of each pass, so
// that the indices, table values and LDS twiddles derived from it are
// recomputed per pass instead of being hoisted out of the pass loop and
// kept live in registers across all passes (which cost occupancy).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Ordering point for an LDS exchange among the lanes of one row.  When every
// row lies inside one wavefront (WAVE), the exchange needs no hardware
// barrier: LDS operations of a single wavefront complete in issue order, so a
// wavefront-scope fence (no instructions; it only stops the compiler moving
// LDS accesses across it) suffices, and the waves of a workgroup run
// unsynchronised.  Rows that straddle wavefronts need the workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void row_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// Padded-axis index maps with the mode count NM known at compile time: NM
// modes stored as NM - 1 (Nyquist dropped) at padded indices [0, NM/2) and
// (P - NM/2, P).  The 3/2 rule gives NM = 2P/3 (Pad<P>); an axis without
// oversampling has NM = P.  A step-A input (m = lo + lane over lanes
// [lo, lo + W)) or step-B output whose whole lane range lies in the zero band
// is known at compile time once the loops are unrolled, so its loads, stores
// and the arithmetic that only feeds them drop out.
template <int P, int NM>
struct PadN {
  static constexpr int n = NM, half = NM / 2;
  // stored index of padded index m, or -1 in the zero band
  static __device__ __forceinline__ int stored(int m) {
    return m < half ? m : (m > P - half ? m - (P - n + 1) : -1);
  }
  static constexpr bool zero(int lo, int hi) { return lo >= half && hi <= P - half; }
  static __device__ __forceinline__ int stored_in(int lo, int W, int lane) {
    return zero(lo, lo + W - 1) ? -1 : stored(lo + lane);
  }
};
template <int P>
using Pad = PadN<P, 2 * P / 3>;

// ------------------------------------------------------------ two-step helpers
// LDS layout for a tile of TC independent transforms (columns): element
// (k1, b) of column col at ((k1 * N2 + b) * TC + col), col fastest, so a
// wavefront whose lanes run over col touches consecutive 16-byte slots.
//
// tw[m] = exp(+2 pi i m / P) in T precision; DIR < 0 uses its conjugate.
template <int P, int N1, int N2, int DIR, int TC, typename T>
__device__ __forceinline__ void step_a_store(cx<T>* x, int b, int col, cx<T>* lds,
                                             const cx<T>* tw) {
  Dft<N1, DIR, T>::run(x);
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
    cx<T> v = x[k1];
    if (k1 != 0) {
      cx<T> w = tw[b * k1];
      if (DIR < 0) w.y = -w.y;
      v = cmul(v, w);
    }
    lds[(k1 * N2 + b) * TC + col] = v;
  }
}

template <int P, int N1, int N2, int DIR, int TC, typename T>
__device__ __forceinline__ void step_b_load(cx<T>* y, int k1, int col, const cx<T>* lds) {
#pragma unroll
  for (int b = 0; b < N2; ++b) y[b] = lds[(k1 * N2 + b) * TC + col];
  Dft<N2, DIR, T>::run(y);
}

// Row-major LDS variant for the x pass (one transform per row, threads of a
// row adjacent in the wavefront): element (k1, b) of row r at
// r * (P + PAD) + k1 * N2 + b.
template <int P>
struct RowPad {
  static constexpr int value = 1;
};

template <int P, int N1, int N2, int DIR, typename T>
__device__ __forceinline__ void row_step_a_store(cx<T>* x, int b, cx<T>* row_lds,
                                                 const cx<T>* tw) {
  Dft<N1, DIR, T>::run(x);
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
    cx<T> v = x[k1];
    if (k1 != 0) {
      cx<T> w = tw[b * k1];
      if (DIR < 0) w.y = -w.y;
      v = cmul(v, w);
    }
    row_lds[k1 * (N2 + 1) + b] = v;
  }
}

// Same, with the twiddles of this plan laid out per k1: tab[k1 * N2 + b] =
// w^(b k1).  Lanes with consecutive b read consecutive entries, so the
// reads do not collide in LDS banks (the flat table's tw[b k1] strides by k1).
template <int P, int N1, int N2, int DIR, typename T>
__device__ __forceinline__ void row_step_a_store_tab(cx<T>* x, int b, cx<T>* row_lds,
                                                     const cx<T>* tab) {
  Dft<N1, DIR, T>::run(x);
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
    cx<T> v = x[k1];
    if (k1 != 0) {
      cx<T> w = tab[k1 * N2 + b];
      if (DIR < 0) w.y = -w.y;
      v = cmul(v, w);
    }
    row_lds[k1 * (N2 + 1) + b] = v;
  }
}

// Fill tab[k1 * N2 + b] = tw[b k1] (k1 < N1, b < N2) cooperatively.
template <int N1, int N2, typename T>
__device__ __forceinline__ void fill_plan_twiddles(cx<T>* tab, const cx<T>* __restrict__ tw) {
  for (int i = threadIdx.x; i < N1 * N2; i += blockDim.x) {
    const int k1 = i / N2, b = i - k1 * N2;
    tab[i] = tw[b * k1];
  }
}

template <int P, int N1, int N2, int DIR, typename T>
__device__ __forceinline__ void row_step_b_load(cx<T>* y, int k1, const cx<T>* row_lds) {
#pragma unroll
  for (int b = 0; b < N2; ++b) y[b] = row_lds[k1 * (N2 + 1) + b];
  Dft<N2, DIR, T>::run(y);
}

// LDS elements a row needs with the padded (N2 + 1) stride.
template <int N1, int N2>
struct RowLds {
  static constexpr int value = N1 * (N2 + 1);
};

}  // namespace dnsmi
