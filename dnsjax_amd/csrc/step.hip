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
// exactly as the host rule would (ctl_after_iteration, including the
// optional noise-floor stops of the mixed form).  The step ends with
// correct_velocity and records the step's iteration count.
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
int periodic_correct_velocity(Ctx* c, double* state, double* err, hipStream_t s);
int walls_predict_and_correct(Ctx* c, const double* state, double* pred, double* rhs_next,
                              double* err, hipStream_t s);
int walls_iterate_correction(Ctx* c, const double* state_prev, double* pred,
                             const double* rhs_prev, double* rhs_next, double* err,
                             hipStream_t s);

__global__ void k_ctl_begin(StepCtl* ctl, double tol, int max_iter) {
  if (threadIdx.x != 0) return;
  const double e = ctl->err;
  ctl->c = 0;
  ctl->floor = 0;
  ctl->prev = e;
  ctl->active = (e > tol) && (max_iter > 0);
}

__global__ void k_ctl_end(StepCtl* ctl, int* counts, int counts_cap, double tol) {
  if (threadIdx.x != 0) return;
  const double e = ctl->err;
  if (counts && ctl->steps < counts_cap) counts[ctl->steps] = ctl->c;
  ctl->total_c += ctl->c;
  if (!ctl->floor && e > tol) ctl->unconverged += 1;
  if (e != e) ctl->nan += 1;
  ctl->steps += 1;
}

int step_device(Ctx* c, const double* state, double* pred, double* rhs_a, double* rhs_b, int slots,
                int* counts, int counts_cap, double tol, int max_iter, double floor_ratio,
                double floor_cap, double floor_margin, hipStream_t s) {
  if (!c->ctl) return c->fail(DNSMI_EINVAL, "control block missing");
  StepCtl* ctl = reinterpret_cast<StepCtl*>(c->ctl);
  double* err = &ctl->err;
  if (slots > max_iter) slots = max_iter;
  c->act = nullptr;
  int rc = c->walls ? walls_predict_and_correct(c, state, pred, rhs_a, err, s)
                    : periodic_predict_and_correct(c, state, pred, rhs_a, err, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ctl_begin, dim3(1), dim3(64), 0, s, ctl, tol, max_iter);
  if ((rc = check_launch(c, "ctl_begin"))) return rc;
  c->act = &ctl->active;
  c->rule = CtlRule{tol, max_iter, floor_ratio, floor_cap, floor_margin};
  c->fuse_ctl = true;
  for (int i = 1; i <= slots && rc == DNSMI_OK; ++i) {
    const double* rp = (i % 2) ? rhs_a : rhs_b;
    double* rn = (i % 2) ? rhs_b : rhs_a;
    rc = c->walls ? walls_iterate_correction(c, state, pred, rp, rn, err, s)
                  : periodic_iterate_correction(c, pred, rp, rn, err, s);
  }
  c->fuse_ctl = false;
  c->act = nullptr;
  if (rc) return rc;
  if (!c->walls && (rc = periodic_correct_velocity(c, pred, nullptr, s))) return rc;
  hipLaunchKernelGGL(k_ctl_end, dim3(1), dim3(64), 0, s, ctl, counts, counts_cap, tol);
  return check_launch(c, "ctl_end");
}

// The guessed step starts with no error: the first pass always runs.
__global__ void k_ctl_begin_guess(StepCtl* ctl, int max_pass) {
  if (threadIdx.x != 0) return;
  ctl->c = 0;
  ctl->floor = 0;
  ctl->err = INFINITY;
  ctl->prev = INFINITY;
  ctl->active = max_pass > 0;
}

int guess_periodic(Ctx* c, const double* state, double* pred, double* rhs_prev, const float* h1,
                   float* h2, const uint16_t* wb, const float* bias, int H, float s_1, float s_2,
                   float s_out, int mode, hipStream_t s);
int guess_walls(Ctx* c, const double* state, double* pred, double* rhs_prev, const float* h1,
                float* h2, const uint16_t* wb, const float* bias, int H, float s_1, float s_2,
                float s_out, int mode, hipStream_t s);

// A step started from the corrector guess (periodic): the guessed predictor,
// then up to `slots` passes, each one RHS, with the same device rule as
// step_device (max_pass passes at most; the count recorded is the number of
// passes).
int step_device_guess(Ctx* c, const double* state, double* pred, double* rhs_a, double* rhs_b,
                      const float* h1, float* h2, const uint16_t* wb, const float* bias, int H,
                      float s_1, float s_2, float s_out, int mode, int slots, int* counts,
                      int counts_cap, double tol, int max_pass, double floor_ratio,
                      double floor_cap, double floor_margin, hipStream_t s) {
  if (!c->ctl) return c->fail(DNSMI_EINVAL, "control block missing");
  StepCtl* ctl = reinterpret_cast<StepCtl*>(c->ctl);
  double* err = &ctl->err;
  if (slots > max_pass) slots = max_pass;
  c->act = nullptr;
  int rc = c->walls ? guess_walls(c, state, pred, rhs_a, h1, h2, wb, bias, H, s_1, s_2, s_out, mode, s)
                    : guess_periodic(c, state, pred, rhs_a, h1, h2, wb, bias, H, s_1, s_2, s_out,
                                     mode, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ctl_begin_guess, dim3(1), dim3(64), 0, s, ctl, max_pass);
  if ((rc = check_launch(c, "ctl_begin_guess"))) return rc;
  c->act = &ctl->active;
  c->rule = CtlRule{tol, max_pass, floor_ratio, floor_cap, floor_margin};
  c->fuse_ctl = true;
  for (int i = 1; i <= slots && rc == DNSMI_OK; ++i) {
    const double* rp = (i % 2) ? rhs_a : rhs_b;
    double* rn = (i % 2) ? rhs_b : rhs_a;
    rc = c->walls ? walls_iterate_correction(c, state, pred, rp, rn, err, s)
                  : periodic_iterate_correction(c, pred, rp, rn, err, s);
  }
  c->fuse_ctl = false;
  c->act = nullptr;
  if (rc) return rc;
  if (!c->walls && (rc = periodic_correct_velocity(c, pred, nullptr, s))) return rc;
  hipLaunchKernelGGL(k_ctl_end, dim3(1), dim3(64), 0, s, ctl, counts, counts_cap, tol);
  return check_launch(c, "ctl_end");
}

int step_alloc_ctl(Ctx* c) {
  c->ctl = static_cast<int*>(c->alloc(sizeof(StepCtl)));
  if (!c->ctl) return c->fail(DNSMI_ENOMEM, "control block");
  if (hipMemset(c->ctl, 0, sizeof(StepCtl)) != hipSuccess) return c->fail(DNSMI_EHIP, "memset");
  return DNSMI_OK;
}

int step_read_ctl(Ctx* c, double* out6) {
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
  out6[5] = h.floor;
  return DNSMI_OK;
}

int step_reset_ctl(Ctx* c) {
  if (c->ctl && hipMemset(c->ctl, 0, sizeof(StepCtl)) != hipSuccess)
    return c->fail(DNSMI_EHIP, "reset control block");
  return DNSMI_OK;
}

}  // namespace dnsmi
