// One whole time step with the corrector loop decided on the device, so that
// a step is a fixed sequence of launches a caller can enqueue without host
// synchronisation or capture in a hipGraph.
//
// The host loop of the reference (__main__.py:161-191) reads the corrector
// error after every pass.  Here a device control block plays that role:
// after predict_and_correct, `active` = (err > tol) and at least one extra
// iteration is allowed; each of `slots` corrector slots runs only if
// `active` (every kernel of the slot reads it and returns at once when 0),
// and the pass's final norm reduction then updates the count and `active`
// exactly as the host rule would (ctl_after_iteration).  In the mixed form,
// optional FP64 slots follow, run only by a step whose FP32 passes stopped
// converging (the escalation of ctl_after_iteration), so every step still
// ends at error <= tol.  The step ends with correct_velocity and records the
// step's iteration count.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "context.hpp"
#include "ctl.hpp"

namespace dnsmi {

int periodic_predict_and_correct(Ctx* c, const double* state, double* pred, double* rhs_next,
                                 double* err, hipStream_t s);
int periodic_iterate_correction(Ctx* c, double* pred, const double* rhs_prev, double* rhs_next,
                                double* err, hipStream_t s);
int periodic_predict_and_correct_carry(Ctx* c, const double* state, double* pred, double* carry,
                                       double* err, hipStream_t s);
int periodic_iterate_correction_carry(Ctx* c, double* pred, const double* carry, double* err, hipStream_t s);
int periodic_correct_velocity(Ctx* c, double* state, double* err, hipStream_t s);
int walls_predict_and_correct(Ctx* c, const double* state, double* pred, double* rhs_next,
                              double* err, hipStream_t s);
int walls_iterate_correction(Ctx* c, const double* state_prev, double* pred,
                             const double* rhs_prev, double* rhs_next, double* err,
                             hipStream_t s);

struct CtlBeginArgs {
  StepCtl* ctl;
  double tol;
  int max_iter;
  double far;
};
struct CtlEndArgs {
  StepCtl* ctl;
  int* counts;
  int counts_cap;
  double tol;
};

__device__ inline void ctl_begin_body(StepCtl* ctl, double tol, int max_iter, double far) {
  if (threadIdx.x != 0) return;
  const double e = ctl->err;
  ctl->c = 0;
  ctl->esc = 0;
  ctl->prev = e;
  ctl->prev2 = INFINITY;
  ctl->active = (e > tol) && (max_iter > 0);
  ctl->run_p = ctl->active;
  ctl->run_e = 0;
  ctl->rec = e > far ? 1 : 0;
  ctl_clear_slots(ctl);
  ctl->rp[1] = ctl->active;
}
__global__ void k_ctl_begin(CtlBeginArgs a) { ctl_begin_body(a.ctl, a.tol, a.max_iter, a.far); }
__global__ void k_ctl_begin_b(const CtlBeginArgs* __restrict__ t) {
  const CtlBeginArgs a = t[blockIdx.z];
  ctl_begin_body(a.ctl, a.tol, a.max_iter, a.far);
}

// The re-centring threshold of a step (split evaluations only).
static double recenter_far(const Ctx* c, double tol) {
  return (c->split && c->fp32_primary && c->cfg.recenter_far > 0) ? c->cfg.recenter_far * tol
                                                                  : (double)INFINITY;
}

__device__ inline void ctl_end_body(StepCtl* ctl, int* counts, int counts_cap, double tol) {
  if (threadIdx.x != 0) return;
  const double e = ctl->err;
  if (counts && ctl->steps < counts_cap) counts[ctl->steps] = ctl->c;
  ctl->total_c += ctl->c;
  if (!(e <= tol)) ctl->unconverged += 1;
  if (e != e) ctl->nan += 1;
  ctl->steps += 1;
}
__global__ void k_ctl_end(CtlEndArgs a) { ctl_end_body(a.ctl, a.counts, a.counts_cap, a.tol); }
__global__ void k_ctl_end_b(const CtlEndArgs* __restrict__ t) {
  const CtlEndArgs a = t[blockIdx.z];
  ctl_end_body(a.ctl, a.counts, a.counts_cap, a.tol);
}

// Before the FP64 escalation slots: the passes alternate between the two
// RHS arrays (pass i reads rhs_a when i is odd), so after an odd number of
// passes the latest evaluation is in rhs_b; it is copied to rhs_a, where the
// escalation slots start.  Skipped unless the step escalated.
__global__ void __launch_bounds__(256) k_ctl_settle(const StepCtl* ctl, double* __restrict__ rhs_a,
                                                    const double* __restrict__ rhs_b, long long n) {
  if (!ctl->run_e || (ctl->c & 1) == 0) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rhs_a[i] = rhs_b[i];
}

// One corrector pass of a device step.  Periodic steps use the carry form:
// rhs_a holds B for every pass (rhs_b is unused); wall-bounded passes
// alternate between the two RHS arrays.
static int step_pass(Ctx* c, const double* state, double* pred, const double* rp, double* rn,
                     const double* carry, double* err, hipStream_t s) {
  return c->walls ? walls_iterate_correction(c, state, pred, rp, rn, err, s)
                  : periodic_iterate_correction_carry(c, pred, carry, err, s);
}

// The corrector slots of a device step: `slots` passes in the context's
// precision, then (mixed form) `esc_slots` FP64 passes that run only once
// the step has escalated.
static int step_slots(Ctx* c, const double* state, double* pred, double* rhs_a, double* rhs_b,
                      int slots, int esc_slots, hipStream_t s) {
  StepCtl* ctl = reinterpret_cast<StepCtl*>(c->ctl);
  double* err = &ctl->err;
  int rc = DNSMI_OK;
  c->fuse_ctl = true;
  // with re-centring, each slot after the first also holds the pass as a
  // re-recording evaluation, run instead of the increment when the control
  // block says so (never in slot 1, whose evaluation is an increment after
  // predict-and-correct and the recording one after a guess)
  const bool recenter = c->split && c->fp32_primary && c->rule.far < INFINITY;
  for (int i = 1; i <= slots && rc == DNSMI_OK; ++i) {
    const double* rp = (i % 2) ? rhs_a : rhs_b;
    double* rn = (i % 2) ? rhs_b : rhs_a;
    if (recenter && i > 1) {
      c->act = &ctl->rr[i];
      c->ref_valid = false;  // the evaluation records
      rc = step_pass(c, state, pred, rp, rn, rhs_a, err, s);
      if (rc) break;  // the recording launch left ref_valid set: the next is an increment
    }
    c->act = &ctl->rp[i];
    rc = step_pass(c, state, pred, rp, rn, rhs_a, err, s);
  }
  if (rc == DNSMI_OK && esc_slots > 0) {
    if (c->walls) {
      const long long n = 2 * c->spec_elems;
      hipLaunchKernelGGL(k_ctl_settle, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ctl, rhs_a,
                         rhs_b, n);
      rc = check_launch(c, "ctl_settle");
    }
    c->act = &ctl->run_e;
    c->fp32 = false;
    for (int i = 1; i <= esc_slots && rc == DNSMI_OK; ++i) {
      const double* rp = (i % 2) ? rhs_a : rhs_b;
      double* rn = (i % 2) ? rhs_b : rhs_a;
      rc = step_pass(c, state, pred, rp, rn, rhs_a, err, s);
    }
    c->fp32 = c->fp32_primary;
  }
  c->fuse_ctl = false;
  c->act = nullptr;
  return rc;
}

int step_device(Ctx* c, const double* state, double* pred, double* rhs_a, double* rhs_b, int slots,
                int esc_slots, int* counts, int counts_cap, double tol, int max_iter,
                double esc_ratio, double esc_cap, hipStream_t s) {
  if (!c->ctl) return c->fail(DNSMI_EINVAL, "control block missing");
  if (esc_slots > 0 && !c->fp32_primary)
    return c->fail(DNSMI_EINVAL, "escalation slots need the mixed form");
  StepCtl* ctl = reinterpret_cast<StepCtl*>(c->ctl);
  double* err = &ctl->err;
  if (slots > max_iter) slots = max_iter;
  if (esc_slots > max_iter) esc_slots = max_iter;
  if (slots > kMaxSlots) return c->fail(DNSMI_EINVAL, "step_device: more than 64 corrector slots");
  c->act = nullptr;
  c->fp32 = c->fp32_primary;
  int rc = c->walls ? walls_predict_and_correct(c, state, pred, rhs_a, err, s)
                    : periodic_predict_and_correct_carry(c, state, pred, rhs_a, err, s);
  if (rc) return rc;
  c->rule = CtlRule{tol, max_iter, esc_slots > 0 ? esc_ratio : 0.0, esc_cap,
                    esc_slots > 0 ? slots : max_iter, recenter_far(c, tol), c->cfg.recenter_near * tol};
  launch(c, &k_ctl_begin, &k_ctl_begin_b, dim3(1), dim3(64), s, CtlBeginArgs{ctl, tol, max_iter, c->rule.far});
  if ((rc = check_launch(c, "ctl_begin"))) return rc;
  if ((rc = step_slots(c, state, pred, rhs_a, rhs_b, slots, esc_slots, s))) return rc;
  if (!c->walls && (rc = periodic_correct_velocity(c, pred, nullptr, s))) return rc;
  launch(c, &k_ctl_end, &k_ctl_end_b, dim3(1), dim3(64), s, CtlEndArgs{ctl, counts, counts_cap, tol});
  return check_launch(c, "ctl_end");
}

// The guessed step starts with no error: the first pass always runs.
__global__ void k_ctl_begin_guess(StepCtl* ctl, int max_pass) {
  if (threadIdx.x != 0) return;
  ctl->c = 0;
  ctl->esc = 0;
  ctl->err = INFINITY;
  ctl->prev = INFINITY;
  ctl->prev2 = INFINITY;
  ctl->active = max_pass > 0;
  ctl->run_p = ctl->active;
  ctl->run_e = 0;
  ctl->rec = 3;  // the guess's reference is recorded by its first pass
  ctl_clear_slots(ctl);
  ctl->rp[1] = ctl->active;
}

int guess_periodic(Ctx* c, const double* state, double* pred, double* rhs_prev, float* h1,
                   float* h2, float* h3, const uint16_t* wb, const float* bias, int H, float s_1, float s_2,
                   float s_out, int mode, hipStream_t s, int carry);
int guess_walls(Ctx* c, const double* state, double* pred, double* rhs_prev, const float* h1,
                float* h2, const uint16_t* wb, const float* bias, int H, float s_1, float s_2,
                float s_out, int mode, hipStream_t s);

// A step started from the corrector guess: the guessed predictor, then up
// to `slots` passes (and the escalation slots), each one RHS, with the same
// device rule as step_device (max_pass passes at most; the count recorded is
// the number of passes).
int step_device_guess(Ctx* c, const double* state, double* pred, double* rhs_a, double* rhs_b,
                      float* h1, float* h2, float* h3, const uint16_t* wb, const float* bias, int H,
                      float s_1, float s_2, float s_out, int mode, int slots, int esc_slots,
                      int* counts, int counts_cap, double tol, int max_pass, double esc_ratio,
                      double esc_cap, hipStream_t s) {
  if (!c->ctl) return c->fail(DNSMI_EINVAL, "control block missing");
  if (esc_slots > 0 && !c->fp32_primary)
    return c->fail(DNSMI_EINVAL, "escalation slots need the mixed form");
  StepCtl* ctl = reinterpret_cast<StepCtl*>(c->ctl);
  if (slots > max_pass) slots = max_pass;
  if (esc_slots > max_pass) esc_slots = max_pass;
  if (slots > kMaxSlots) return c->fail(DNSMI_EINVAL, "step_device_guess: more than 64 corrector slots");
  c->act = nullptr;
  c->fp32 = c->fp32_primary;
  int rc = c->walls ? guess_walls(c, state, pred, rhs_a, h1, h2, wb, bias, H, s_1, s_2, s_out, mode, s)
                    : guess_periodic(c, state, pred, rhs_a, h1, h2, h3, wb, bias, H, s_1, s_2, s_out,
                                     mode, s, 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ctl_begin_guess, dim3(1), dim3(64), 0, s, ctl, max_pass);
  if ((rc = check_launch(c, "ctl_begin_guess"))) return rc;
  c->rule = CtlRule{tol, max_pass, esc_slots > 0 ? esc_ratio : 0.0, esc_cap,
                    esc_slots > 0 ? slots : max_pass, recenter_far(c, tol), c->cfg.recenter_near * tol};
  if ((rc = step_slots(c, state, pred, rhs_a, rhs_b, slots, esc_slots, s))) return rc;
  if (!c->walls && (rc = periodic_correct_velocity(c, pred, nullptr, s))) return rc;
  launch(c, &k_ctl_end, &k_ctl_end_b, dim3(1), dim3(64), s, CtlEndArgs{ctl, counts, counts_cap, tol});
  return check_launch(c, "ctl_end");
}

int step_alloc_ctl(Ctx* c) {
  c->ctl = static_cast<int*>(c->alloc(sizeof(StepCtl)));
  if (!c->ctl) return c->fail(DNSMI_ENOMEM, "control block");
  if (hipMemset(c->ctl, 0, sizeof(StepCtl)) != hipSuccess) return c->fail(DNSMI_EHIP, "memset");
  return DNSMI_OK;
}

int step_read_ctl(Ctx* c, double* out7) {
  double* out6 = out7;
  StepCtl h{};
  if (c->ctl) {
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&h, c->ctl, sizeof h, hipMemcpyDeviceToHost) != hipSuccess)
      return c->fail(DNSMI_EHIP, "read control block");
  }
  out6[0] = h.steps;
  out6[1] = h.total_c;
  out6[2] = h.unconverged;
  out6[3] = h.nan;
  out6[4] = h.err;
  out6[5] = h.escalated;
  out7[6] = h.recorded;
  return DNSMI_OK;
}

int step_reset_ctl(Ctx* c) {
  if (c->ctl && hipMemset(c->ctl, 0, sizeof(StepCtl)) != hipSuccess)
    return c->fail(DNSMI_EHIP, "reset control block");
  return DNSMI_OK;
}

}  // namespace dnsmi
