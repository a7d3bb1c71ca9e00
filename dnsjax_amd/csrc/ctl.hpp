// The device control block of a time step whose corrector loop is decided
// on the device (step.hip), and the rule applied after each pass.
#pragma once

#include <hip/hip_runtime.h>

#include "context.hpp"

namespace dnsmi {

struct StepCtl {
  double err;
  double prev;
  int active;
  int c;
  int floor;
  int steps;
  int total_c;
  int unconverged;
  int nan;
  int pad;
  double floor_est;  // noise-floor estimate: the error of the last stall stop
};

// After a pass with error e (the host rule of __main__.py:161-191): count
// the iteration and keep going while e > tol and the limit allows.  The
// mixed form's noise-floor stops, the same rule as the host loops: for an
// error above `tol` and below `cap`, a pass that failed to shrink the error
// by `ratio` is at the floor (a stall stop; its error becomes the floor
// estimate), and with `margin` > 0 a pass whose error is within `margin`
// times the estimate is taken as at the floor without the confirming pass
// (the estimate can then only decrease).  One thread, active passes only.
__device__ inline void ctl_after_iteration(StepCtl* ctl, const CtlRule& r, double e) {
  ctl->c += 1;
  if (r.ratio > 0 && e > r.tol && e <= r.cap && e > r.ratio * ctl->prev) {
    ctl->floor = 1;
    ctl->active = 0;
    ctl->floor_est = e;
  } else if (r.ratio > 0 && r.margin > 0 && ctl->floor_est > 0 && e > r.tol && e <= r.cap &&
             e <= r.margin * ctl->floor_est) {
    ctl->floor = 1;
    ctl->active = 0;
    ctl->floor_est = fmin(ctl->floor_est, e);
  } else {
    ctl->active = (e > r.tol) && (ctl->c < r.max_iter);
  }
  ctl->prev = e;
}

}  // namespace dnsmi
