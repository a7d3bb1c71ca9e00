// Library-internal context shared by the periodic and wall-bounded paths.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/dnsmi.h"

namespace dnsmi {

struct WallTables;  // wall-bounded set-up data (walls.hip)

// Kernel ids for the per-kernel event timing (dnsmi_profile_*).
enum KernelId {
  KID_Y_INVERSE = 0,
  KID_Z_INVERSE = 1,
  KID_X_NONLINEAR = 2,
  KID_Z_FORWARD = 3,
  KID_Y_FORWARD = 4,
  KID_POINTWISE = 5,
  KID_FINALIZE = 6,
  KID_W_Z_INVERSE = 7,
  KID_W_X_NONLINEAR = 8,
  KID_W_Z_FORWARD = 9,
  KID_W_IMM = 10,
  KID_W_NORM = 11,
  KID_COUNT = 12
};

struct ProfRec {
  int kid;
  hipEvent_t a, b;
};

// The corrector rule of a device step: the reference's tolerance and
// iteration limit, and the mixed form's escalation to FP64 passes (a pass
// whose error, at or below `cap`, is not below `ratio` times the error two
// passes earlier has reached the FP32 evaluations' resolution; `ratio` 0
// disables it) after at most `primary` passes in the context's precision.
struct CtlRule {
  double tol;
  int max_iter;
  double ratio, cap;
  int primary;
  // split evaluations: a step whose predict-and-correct error exceeds `far`
  // records a new reference at its first pass that starts from an error of
  // at most `near` (far = infinity: never)
  double far, near;
};

// How one evaluation of the nonlinear term treats the mixed form's
// reference state (split-precision evaluations, periodic.hip / walls.hip).
enum EvalMode {
  EV_PLAIN = 0,   // N(u) evaluated in full in the context's precision
  EV_RECORD = 1,  // the same, and u, its physical velocity and vorticity and
                  // N(u) become the reference (u_ref, x_ref, n_ref)
  EV_INC = 2      // N(u) = n_ref + P[dN]: only the increment
                  // dN = (U + u_ref) x w_d + d x (W_ref + w_d), d = u - u_ref,
                  // goes through the FP32 transforms
};

// Launches recorded instead of issued (batched ensembles): the batched twin
// of each kernel (one member per grid.z, its argument struct read from a
// device table) with the member's grid, block and argument struct.
struct LaunchRec {
  const void* kb;
  dim3 grid, block;
  std::vector<unsigned char> args;
};
struct Recorder {
  std::vector<LaunchRec> list;
};

struct Ctx {
  dnsmi_config cfg{};
  bool walls = false;
  int device = 0;
  int nx = 0, ny = 0, nz = 0;
  int Nx1 = 0, Ny1 = 0, Nz1 = 0;  // stored modes: nx/2, ny-1, nz-1 (walls: Ny1 = ny)
  int Px = 0, Py = 0, Pz = 0;     // padded grid (walls: Py = ny)
  long long spec_elems = 0;       // complex elements of one vector field
  bool fp32 = false;              // FFT/nonlinear precision

  // FFT intermediates (T-typed complex), aliased across passes.
  void* bufA = nullptr;
  void* bufB = nullptr;
  size_t bufA_bytes = 0, bufB_bytes = 0;

  // exp(+2 pi i m / P) tables per axis (x, y, z) in both precisions.
  void* tw64[3] = {nullptr, nullptr, nullptr};
  void* tw32[3] = {nullptr, nullptr, nullptr};

  // Stored-mode wavenumbers and base-flow profiles.
  double* kx = nullptr;
  double* ky = nullptr;
  double* kz = nullptr;
  double* base = nullptr;   // (3, Py)
  double* cbase = nullptr;  // (3, Py)
  double* nbase = nullptr;  // (3, Py)

  double* rhs0 = nullptr;      // N^0 of the predictor (spectral vector field)

  // Split-precision evaluations of the mixed form (DNSMI_MIXED, unless
  // cfg.plain_mixed): the step's first corrector evaluation records u_ref
  // (its state), and every later one computes only the increment over it in
  // FP32, so the FP32 rounding scales with the increment, not with N.
  bool split = false;
  bool ref_valid = false;         // u_ref / x_ref / n_ref hold one evaluation
  double* uref = nullptr;         // u_ref, spectral FP64 (spec_elems complex)
  float* xref = nullptr;          // (6, rows, Px): U + u_ref (3), W + w_ref (3)
  double* nref = nullptr;         // N(u_ref), spectral FP64
  bool fp32_primary = false;      // the context's own precision (cfg)
  double* partials = nullptr;  // per-block partial sums of the norm
  int max_partials = 0;

  WallTables* wt = nullptr;

  // Device-side corrector control (graph-capturable steps): when `act` is
  // set, every step kernel first reads *act and skips its launch if 0.
  const int* act = nullptr;
  int* ctl = nullptr;  // device control block (StepCtl, ctl.hpp)
  // While set (the corrector slots of a device step), the final norm
  // reduction of each pass also applies `rule` to the control block.
  bool fuse_ctl = false;
  CtlRule rule{};

  int tile_cols = 16;
  // kernel variants (all exact; the autotuner keeps the fastest)
  bool xpass_split = false;
  bool yinv_seq = false;
  bool zinv_seq = false;
  bool zinv_nt = false;  // z-inverse output through non-temporal stores
  int curl_rows = 32;    // walls: y points per FD-curl workgroup (32 or 64)
  bool sweep_seq = false;  // wall-bounded: sequential band sweeps instead of chunk maps
  bool imm_fused = false;  // wall-bounded, chunk maps: the two-launch fused IMM
  std::string tuning;

  // while set, launches are recorded there instead of issued
  Recorder* rec = nullptr;

  // per-kernel timing (off unless dnsmi_profile_begin was called)
  bool prof = false;
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  hipEvent_t ev_get();

  std::string err;
  std::vector<void*> allocations;

  int fail(int code, const std::string& msg) {
    err = msg;
    return code;
  }
  void* alloc(size_t bytes);
  void release(void* p);  // frees one allocation made by alloc (null: no-op)
};

extern thread_local std::string g_create_error;

int check_launch(Ctx* c, const char* what);

// Records an event pair around the launches in its scope when profiling.
struct ProfScope {
  Ctx* c;
  int kid;
  hipStream_t s;
  hipEvent_t a = nullptr;
  ProfScope(Ctx* c_, int kid_, hipStream_t s_) : c(c_), kid(kid_), s(s_) {
    if (c->prof) {
      a = c->ev_get();
      (void)hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (c->prof && a) {
      hipEvent_t b = c->ev_get();
      (void)hipEventRecord(b, s);
      c->prof_recs.push_back({kid, a, b});
    }
  }
};

// One launch of kernel k with argument struct a, or, while the context
// records, an entry naming its batched twin kb.
template <typename A>
inline void launch(Ctx* c, void (*k)(A), void (*kb)(const A*), dim3 grid, dim3 block, hipStream_t s,
                   const A& a) {
  if (c->rec) {
    LaunchRec r{reinterpret_cast<const void*>(kb), grid, block, std::vector<unsigned char>(sizeof(A))};
    std::memcpy(r.args.data(), &a, sizeof(A));
    c->rec->list.push_back(std::move(r));
    return;
  }
  hipLaunchKernelGGL(k, grid, block, 0, s, a);
}

// Shared small kernels (periodic.hip).
int finalize_norm(Ctx* c, int nblocks, double* err_dev, hipStream_t s);

}  // namespace dnsmi

struct dnsmi_ctx {
  dnsmi::Ctx impl;
};
