// The device control block of a time step whose corrector loop is decided
// on the device (step.hip), and the rule applied after each pass.
#pragma once

#include <hip/hip_runtime.h>

#include "context.hpp"

namespace dnsmi {

// Corrector slots of one device step at most (slot i is pass i).
constexpr int kMaxSlots = 64;

struct StepCtl {
  double err;
  double prev;   // error of the previous pass
  double prev2;  // error two passes back
  int active;    // the corrector continues (the reference's rule)
  int c;         // passes of this step so far
  int esc;       // this step's remaining passes run in FP64 (mixed form)
  int run_p;     // active && !esc: the next slot in the context's precision runs
  int run_e;     // active && esc: the next FP64 escalation slot runs
  int steps;
  int total_c;
  int unconverged;
  int nan;
  int escalated;  // steps that escalated
  int rec;        // split reference: 0 as recorded, 1 to re-record (far), 2 re-recorded,
                  // 3 recorded by the first pass (a guessed step): decided after it
  int recorded;   // steps that re-recorded their reference
  // run flags of primary slot i (set by pass i - 1): rp the slot's
  // increment evaluation, rr its re-recording evaluation (at most one of
  // the two runs)
  int rp[kMaxSlots + 1];
  int rr[kMaxSlots + 1];
};

// Start of a step's corrector: no slot runs unless set below.
__device__ inline void ctl_clear_slots(StepCtl* ctl) {
  for (int i = 0; i <= kMaxSlots; ++i) {
    ctl->rp[i] = 0;
    ctl->rr[i] = 0;
  }
}

// After a pass with error e: count it and continue while e > tol and the
// limit allows (__main__.py:161-191).  Mixed form: a pass that did not
// bring the error below `ratio` times the error two passes back, at an error
// of at most `cap`, has reached the resolution of its FP32 evaluations, and
// so has a step whose `primary` passes in the context's precision all ran
// without converging; the step's remaining passes then run in FP64, so it
// still ends at error <= tol.  Split evaluations whose reference lies far
// from the fixed point (the evaluation that recorded it ended with an error
// above `far`) are re-centred: the first pass after the error reached
// `near` records a new reference.  One thread, active passes only.
__device__ inline void ctl_after_iteration(StepCtl* ctl, const CtlRule& r, double e) {
  ctl->c += 1;
  const bool more = (e > r.tol) && (ctl->c < r.max_iter);
  if (more && !ctl->esc && r.ratio > 0 &&
      ((e <= r.cap && e > r.ratio * ctl->prev2) || ctl->c >= r.primary)) {
    ctl->esc = 1;
    ctl->escalated += 1;
  }
  // split evaluations far from the step's fixed point: once the passes have
  // closed in (e <= near), the next pass records a new reference there
  if (ctl->rec == 3) ctl->rec = e > r.far ? 1 : 0;
  const bool rec = more && !ctl->esc && ctl->rec == 1 && e <= r.near;
  if (rec) {
    ctl->rec = 2;
    ctl->recorded += 1;
  }
  ctl->active = more;
  ctl->prev2 = ctl->prev;
  ctl->prev = e;
  ctl->run_p = more && !ctl->esc;
  ctl->run_e = more && ctl->esc;
  const int next = ctl->c + 1;
  if (next <= kMaxSlots) {
    ctl->rp[next] = more && !ctl->esc && !rec;
    ctl->rr[next] = rec;
  }
}

}  // namespace dnsmi
