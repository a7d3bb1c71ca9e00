// Batched device steps for ensembles of equal-size periodic trajectories.
//
// dnsmi_step_device issues a fixed sequence of launches per member (the
// corrector decided on the device).  For a class of members with the same
// sizes, precision and kernel variants that sequence is the same kernel
// with the same grid at every position; only the argument structs differ
// (buffers, control block, Re, dt, tolerance).  A plan records each
// member's step (Ctx::rec: the launches are captured, not issued), checks
// that the sequences agree, stores every position's member arguments side
// by side in one device table, and then issues one launch per position of
// the kernel's batched twin over grid.z = member.  Each member's workgroups
// read its own arguments and run exactly the code of its own launch, so
// its results equal its own dnsmi_step_device's bit for bit; the class
// advances with 1/n of the launches and fills the GPU where one small
// member's grid would not.
//
// The reference runs one trajectory at a time (__main__.py:125-203); this
// has no counterpart there.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"

namespace dnsmi {

int step_device(Ctx* c, const double* state, double* pred, double* rhs_a, double* rhs_b, int slots,
                int esc_slots, int* counts, int counts_cap, double tol, int max_iter, double esc_ratio,
                double esc_cap, hipStream_t s);

struct BatchEntry {
  const void* kb;
  dim3 grid, block;
  void* table;
};

}  // namespace dnsmi

struct dnsmi_batch {
  int device = 0;
  void* mem = nullptr;
  std::vector<dnsmi::BatchEntry> entries;
};

using dnsmi::Ctx;

extern "C" {

int dnsmi_batch_plan(dnsmi_ctx* const* ctxs, int n, const double* const* states, double* const* nexts,
                     double* const* rhs_a, double* const* rhs_b, int slots, int esc_slots,
                     int* const* counts, int counts_cap, double tol, int max_iter, double esc_ratio,
                     double esc_cap, dnsmi_batch** out) {
  if (!ctxs || n < 1 || n > 65535 || !states || !nexts || !rhs_a || !rhs_b || !out || !ctxs[0]) {
    dnsmi::g_create_error = "batch_plan: bad arguments";
    return DNSMI_EINVAL;
  }
  *out = nullptr;
  Ctx* c0 = &ctxs[0]->impl;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return c0->fail(DNSMI_EINVAL, "batch_plan: null member context");
    Ctx* c = &ctxs[i]->impl;
    if (c->walls) return c0->fail(DNSMI_EINVAL, "batch_plan: periodic contexts only");
    if (c->prof) return c0->fail(DNSMI_EINVAL, "batch_plan: a member context is profiling");
    if (c->device != c0->device) return c0->fail(DNSMI_EINVAL, "batch_plan: members on different devices");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return c0->fail(DNSMI_EINVAL, "batch_plan: a context appears twice");
    if (!states[i] || !nexts[i] || !rhs_a[i] || !rhs_b[i])
      return c0->fail(DNSMI_EINVAL, "batch_plan: null member buffer");
  }
  // record every member's step (nothing is issued while a context records)
  std::vector<dnsmi::Recorder> recs(n);
  for (int i = 0; i < n; ++i) {
    Ctx* c = &ctxs[i]->impl;
    c->rec = &recs[i];
    const int rc = dnsmi::step_device(c, states[i], nexts[i], rhs_a[i], rhs_b[i], slots, esc_slots,
                                      counts ? counts[i] : nullptr, counts ? counts_cap : 0, tol, max_iter,
                                      esc_ratio, esc_cap, nullptr);
    c->rec = nullptr;
    if (rc) return c0->fail(rc, "batch_plan: member " + std::to_string(i) + ": " + c->err);
  }
  const size_t L = recs[0].list.size();
  std::vector<size_t> off(L);
  size_t total = 0;
  for (size_t k = 0; k < L; ++k) {
    const dnsmi::LaunchRec& r0 = recs[0].list[k];
    for (int i = 1; i < n; ++i) {
      const std::vector<dnsmi::LaunchRec>& li = recs[i].list;
      const bool same = li.size() == L && li[k].kb == r0.kb && li[k].grid.x == r0.grid.x &&
                        li[k].grid.y == r0.grid.y && li[k].grid.z == r0.grid.z &&
                        li[k].block.x == r0.block.x && li[k].block.y == r0.block.y &&
                        li[k].block.z == r0.block.z && li[k].args.size() == r0.args.size();
      if (!same)
        return c0->fail(DNSMI_EINVAL, "batch_plan: member " + std::to_string(i) +
                                          " issues a different launch sequence (members need the same sizes, "
                                          "precision, kernel variants and slots)");
    }
    if (r0.grid.z != 1) return c0->fail(DNSMI_EINVAL, "batch_plan: a recorded launch uses grid.z");
    total = (total + 255) & ~size_t(255);
    off[k] = total;
    total += (size_t)n * r0.args.size();
  }
  std::vector<unsigned char> host(total ? total : 1);
  for (size_t k = 0; k < L; ++k)
    for (int i = 0; i < n; ++i) {
      const dnsmi::LaunchRec& r = recs[i].list[k];
      std::memcpy(host.data() + off[k] + (size_t)i * r.args.size(), r.args.data(), r.args.size());
    }
  dnsmi_batch* b = new (std::nothrow) dnsmi_batch();
  if (!b) return c0->fail(DNSMI_ENOMEM, "batch_plan: out of host memory");
  b->device = c0->device;
  if (hipMalloc(&b->mem, host.size()) != hipSuccess ||
      hipMemcpy(b->mem, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(b->mem);
    delete b;
    return c0->fail(DNSMI_ENOMEM, "batch_plan: argument tables");
  }
  for (size_t k = 0; k < L; ++k) {
    const dnsmi::LaunchRec& r0 = recs[0].list[k];
    dim3 g = r0.grid;
    g.z = (unsigned)n;
    b->entries.push_back({r0.kb, g, r0.block, static_cast<unsigned char*>(b->mem) + off[k]});
  }
  *out = b;
  return DNSMI_OK;
}

int dnsmi_batch_enqueue(dnsmi_batch* b, void* stream) {
  if (!b) {
    dnsmi::g_create_error = "batch_enqueue: null plan";
    return DNSMI_EINVAL;
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  for (const dnsmi::BatchEntry& e : b->entries) {
    void* table = e.table;
    void* args[] = {&table};
    if (hipLaunchKernel(e.kb, e.grid, e.block, args, 0, s) != hipSuccess) {
      dnsmi::g_create_error = "batch_enqueue: launch failed";
      return DNSMI_EHIP;
    }
  }
  return DNSMI_OK;
}

int dnsmi_batch_launches(const dnsmi_batch* b) { return b ? (int)b->entries.size() : 0; }

void dnsmi_batch_destroy(dnsmi_batch* b) {
  if (!b) return;
  (void)hipFree(b->mem);
  delete b;
}

}  // extern "C"
