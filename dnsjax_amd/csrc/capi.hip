// extern "C" entry points: argument checks, dispatch to the periodic or the
// wall-bounded implementation, error reporting.  No exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <cstdint>
#include <string>

#include "context.hpp"

namespace dnsmi {

thread_local std::string g_create_error;

void* Ctx::alloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 256;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  allocations.push_back(p);
  return p;
}

void Ctx::release(void* p) {
  if (!p) return;
  for (size_t i = 0; i < allocations.size(); ++i) {
    if (allocations[i] == p) {
      (void)hipFree(p);
      allocations.erase(allocations.begin() + (long)i);
      return;
    }
  }
}

hipEvent_t Ctx::ev_get() {
  if (ev_used == ev_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ev_pool.push_back(e);
  }
  return ev_pool[ev_used++];
}

int check_launch(Ctx* c, const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return c->fail(DNSMI_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return DNSMI_OK;
}

// periodic.hip
int periodic_setup(Ctx* c);
int periodic_get_rhs(Ctx* c, const double* state, double* rhs, hipStream_t s);
int periodic_predict(Ctx* c, const double* state, const double* rhs, double* pred, hipStream_t s);
int periodic_correct(Ctx* c, const double* pred_in, const double* rhs_prev,
                     const double* rhs_next, double* pred_out, double* corr, hipStream_t s);
int periodic_norm(Ctx* c, const double* corr, double* err, hipStream_t s);
int periodic_predict_and_correct(Ctx* c, const double* state, double* pred, double* rhs_next,
                                 double* err, hipStream_t s);
int periodic_predict_and_correct_carry(Ctx* c, const double* state, double* pred, double* carry,
                                       double* err, hipStream_t s);
int periodic_iterate_correction_carry(Ctx* c, double* pred, const double* carry, double* err, hipStream_t s);
int periodic_iterate_correction(Ctx* c, double* pred, const double* rhs_prev, double* rhs_next,
                                double* err, hipStream_t s);
int periodic_correct_velocity(Ctx* c, double* state, double* err, hipStream_t s);

// walls.hip
int walls_setup(Ctx* c);
void walls_release(Ctx* c);
int walls_get_rhs(Ctx* c, const double* state, double* rhs, hipStream_t s);
int walls_predict(Ctx* c, const double* state, const double* rhs, double* pred, hipStream_t s);
int walls_correct(Ctx* c, const double* state_prev, const double* pred_in,
                  const double* rhs_prev, const double* rhs_next, double* pred_out,
                  double* corr, hipStream_t s);
int walls_norm(Ctx* c, const double* corr, double* err, hipStream_t s);
int walls_predict_and_correct(Ctx* c, const double* state, double* pred, double* rhs_next,
                              double* err, hipStream_t s);
int walls_iterate_correction(Ctx* c, const double* state_prev, double* pred,
                             const double* rhs_prev, double* rhs_next, double* err,
                             hipStream_t s);
int walls_table(Ctx* c, int which, double* out, long long n);
int get_stats(Ctx* c, const double* state, double* out, hipStream_t s);
int step_device(Ctx* c, const double* state, double* pred, double* rhs_a, double* rhs_b, int slots,
                int esc_slots, int* counts, int counts_cap, double tol, int max_iter,
                double esc_ratio, double esc_cap, hipStream_t s);
int step_read_ctl(Ctx* c, double* out7);
int step_device_guess(Ctx* c, const double* state, double* pred, double* rhs_a, double* rhs_b,
                      float* h1, float* h2, float* h3, const uint16_t* wb, const float* bias, int H,
                      float s_1, float s_2, float s_out, int mode, int slots, int esc_slots,
                      int* counts, int counts_cap, double tol, int max_pass, double esc_ratio,
                      double esc_cap, hipStream_t s);
int step_alloc_ctl(Ctx* c);
int mlp_grad(int H, int B, const float* X, const float* T, const float* prm, int bf16, float* part,
             float* grad, float* loss, hipStream_t s);
int mlp_grad_static(int H, int B, const float* X, const float* T, const float* prm, int bf16,
                    float* grad, float* loss, hipStream_t s, std::string* why);
long long mlp_grad_scratch(int H, int B);
int mlp_adam(int H, float* prm, const float* g, float* m, float* v, int step, float lr,
             hipStream_t s);
int pnet_forward(int din, int H, int dout, int B, const uint16_t* X, const float* prm, uint16_t* wb,
                 float* wf, uint16_t* wh, float* Y, hipStream_t s);
int pnet_grad(int din, int H, int dout, int B, const uint16_t* X, const float* T, const float* prm,
              uint16_t* wb, float* wf, uint16_t* wh, float* part, float* grad, float* loss,
              hipStream_t s);
int pnet_adam(long long P, float* prm, const float* g, float* m, float* v, int step, float lr,
              hipStream_t s);
int pnet_features(Ctx* c, const double* nm2, const double* nm1, const double* ncur, const double* nnext,
                  const long long* modes, int ns, int din, float s_1, float s_2, float s_out, uint16_t* X,
                  float* T, hipStream_t s);
int guess_walls_profile(Ctx* c, const double* state, double* pred, double* rhs_prev, const float* h1, float* h2,
                        int din, int H, int dout, const float* prm, uint16_t* wb, uint16_t* X, float* wf,
                        uint16_t* wh, float* Y, float s_1, float s_2, float s_out, hipStream_t s);
int mlp_pack(int H, const float* prm, uint16_t* wb, float* bias, hipStream_t s);
int mlp_forward(int H, const float* X, int B, const uint16_t* wb, const float* bias, float* Y,
                hipStream_t s);
int mlp_features(Ctx* c, const double* nm3, const double* nm2, const double* nm1, const double* ncur,
                 const double* nnext, const long long* modes, int ns, float s_1, float s_2,
                 float s_out, float* X, float* T, hipStream_t s);
int guess_periodic(Ctx* c, const double* state, double* pred, double* rhs_prev, float* h1,
                   float* h2, float* h3, const uint16_t* wb, const float* bias, int H, float s_1, float s_2,
                   float s_out, int mode, hipStream_t s, int carry);
int guess_walls(Ctx* c, const double* state, double* pred, double* rhs_prev, const float* h1,
                float* h2, const uint16_t* wb, const float* bias, int H, float s_1, float s_2,
                float s_out, int mode, hipStream_t s);
int step_reset_ctl(Ctx* c);

}  // namespace dnsmi

using dnsmi::Ctx;

static hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

#define DNSMI_CHECK_CTX(ctx)                 \
  if (!(ctx)) {                              \
    dnsmi::g_create_error = "null context";  \
    return DNSMI_EINVAL;                     \
  }                                          \
  Ctx* c = &(ctx)->impl;

#define DNSMI_REQUIRE(cond, msg) \
  if (!(cond)) return c->fail(DNSMI_EINVAL, msg);

extern "C" {

int dnsmi_create(const dnsmi_config* cfg, dnsmi_ctx** out) {
  if (!cfg || !out) {
    dnsmi::g_create_error = "null argument";
    return DNSMI_EINVAL;
  }
  *out = nullptr;
  dnsmi_ctx* h = new (std::nothrow) dnsmi_ctx();
  if (!h) {
    dnsmi::g_create_error = "host allocation failed";
    return DNSMI_ENOMEM;
  }
  Ctx* c = &h->impl;
  c->cfg = *cfg;
  int rc = DNSMI_OK;
  if (hipGetDevice(&c->device) != hipSuccess) {
    rc = c->fail(DNSMI_EHIP, "no HIP device");
  } else if (cfg->precision != DNSMI_FP64 && cfg->precision != DNSMI_MIXED) {
    rc = c->fail(DNSMI_EINVAL, "precision must be DNSMI_FP64 or DNSMI_MIXED");
  } else if (cfg->tile_cols != 0 && cfg->tile_cols != 8 && cfg->tile_cols != 16) {
    rc = c->fail(DNSMI_EINVAL, "tile_cols must be 0 (autotune), 8 or 16");
  } else if (!(cfg->re > 0) || !(cfg->dt > 0) || cfg->implicitness < 0 || cfg->implicitness > 1) {
    rc = c->fail(DNSMI_EINVAL, "need re > 0, dt > 0, 0 <= implicitness <= 1");
  } else if (cfg->recenter_far < 0 || cfg->recenter_near < 0 ||
             (cfg->recenter_far > 0 && !(cfg->recenter_near < cfg->recenter_far))) {
    rc = c->fail(DNSMI_EINVAL, "need 0 <= recenter_near < recenter_far (or recenter_far = 0)");
  } else {
    c->fp32 = cfg->precision == DNSMI_MIXED;
    c->fp32_primary = c->fp32;
    c->split = c->fp32 && !cfg->plain_mixed;
    if (cfg->tile_cols > 0) {
      c->tile_cols = cfg->tile_cols;
      c->yinv_seq = (cfg->variants & 1) != 0;
      c->zinv_seq = (cfg->variants & 2) != 0;
      c->xpass_split = (cfg->variants & 4) != 0;
      c->sweep_seq = (cfg->variants & 8) != 0;
      c->imm_fused = (cfg->variants & 16) != 0 && !c->sweep_seq;
      c->zinv_nt = (cfg->variants & 32) != 0;
    }
    if (cfg->system == DNSMI_PLANE_COUETTE) {
      c->walls = true;
      rc = dnsmi::walls_setup(c);
    } else {
      rc = dnsmi::periodic_setup(c);
    }
    if (rc == DNSMI_OK) rc = dnsmi::step_alloc_ctl(c);
  }
  if (rc != DNSMI_OK) {
    dnsmi::g_create_error = c->err;
    dnsmi_destroy(h);
    return rc;
  }
  *out = h;
  return DNSMI_OK;
}

void dnsmi_destroy(dnsmi_ctx* ctx) {
  if (!ctx) return;
  Ctx* c = &ctx->impl;
  if (c->walls) dnsmi::walls_release(c);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (void* p : c->allocations) (void)hipFree(p);
  delete ctx;
}

const char* dnsmi_last_error(const dnsmi_ctx* ctx) {
  return ctx ? ctx->impl.err.c_str() : dnsmi::g_create_error.c_str();
}

long long dnsmi_spec_elems(const dnsmi_ctx* ctx) { return ctx ? ctx->impl.spec_elems : -1; }

int dnsmi_get_rhs(dnsmi_ctx* ctx, const double* state, double* rhs, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state && rhs && (const void*)state != (const void*)rhs, "get_rhs: bad pointers");
  return c->walls ? dnsmi::walls_get_rhs(c, state, rhs, as_stream(stream))
                  : dnsmi::periodic_get_rhs(c, state, rhs, as_stream(stream));
}

int dnsmi_predict(dnsmi_ctx* ctx, const double* state, const double* rhs, double* pred,
                  void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state && rhs && pred, "predict: null pointer");
  return c->walls ? dnsmi::walls_predict(c, state, rhs, pred, as_stream(stream))
                  : dnsmi::periodic_predict(c, state, rhs, pred, as_stream(stream));
}

int dnsmi_correct(dnsmi_ctx* ctx, const double* state_prev, const double* pred_in,
                  const double* rhs_prev, const double* rhs_next, double* pred_out,
                  double* correction_out, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state_prev && pred_in && rhs_prev && rhs_next && pred_out,
                "correct: null pointer");
  return c->walls ? dnsmi::walls_correct(c, state_prev, pred_in, rhs_prev, rhs_next, pred_out,
                                         correction_out, as_stream(stream))
                  : dnsmi::periodic_correct(c, pred_in, rhs_prev, rhs_next, pred_out,
                                            correction_out, as_stream(stream));
}

int dnsmi_norm(dnsmi_ctx* ctx, const double* correction, double* err_dev, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(correction && err_dev, "norm: null pointer");
  return c->walls ? dnsmi::walls_norm(c, correction, err_dev, as_stream(stream))
                  : dnsmi::periodic_norm(c, correction, err_dev, as_stream(stream));
}

int dnsmi_predict_and_correct(dnsmi_ctx* ctx, const double* state, double* pred_out,
                              double* rhs_next_out, double* err_dev, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state && pred_out && rhs_next_out && err_dev, "predict_and_correct: null pointer");
  DNSMI_REQUIRE((const void*)state != (const void*)pred_out &&
                    (const void*)state != (const void*)rhs_next_out &&
                    (const void*)pred_out != (const void*)rhs_next_out,
                "predict_and_correct: outputs must not alias the state or each other");
  return c->walls ? dnsmi::walls_predict_and_correct(c, state, pred_out, rhs_next_out, err_dev,
                                                     as_stream(stream))
                  : dnsmi::periodic_predict_and_correct(c, state, pred_out, rhs_next_out,
                                                        err_dev, as_stream(stream));
}

int dnsmi_iterate_correction(dnsmi_ctx* ctx, const double* state_prev, double* pred,
                             const double* rhs_prev, double* rhs_next_out, double* err_dev,
                             void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state_prev && pred && rhs_prev && rhs_next_out && err_dev,
                "iterate_correction: null pointer");
  DNSMI_REQUIRE((const void*)pred != (const void*)state_prev &&
                    (const void*)pred != (const void*)rhs_prev &&
                    (const void*)pred != (const void*)rhs_next_out,
                "iterate_correction: pred must not alias the other arrays");
  return c->walls ? dnsmi::walls_iterate_correction(c, state_prev, pred, rhs_prev, rhs_next_out,
                                                    err_dev, as_stream(stream))
                  : dnsmi::periodic_iterate_correction(c, pred, rhs_prev, rhs_next_out, err_dev,
                                                       as_stream(stream));
}

int dnsmi_predict_and_correct_carry(dnsmi_ctx* ctx, const double* state, double* pred_out,
                                    double* carry_out, double* err_dev, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(!c->walls, "predict_and_correct_carry: periodic contexts only");
  DNSMI_REQUIRE(state && pred_out && carry_out && err_dev, "predict_and_correct_carry: null pointer");
  DNSMI_REQUIRE((const void*)state != pred_out && (const void*)state != carry_out && pred_out != carry_out,
                "predict_and_correct_carry: buffers must be distinct");
  return dnsmi::periodic_predict_and_correct_carry(c, state, pred_out, carry_out, err_dev, as_stream(stream));
}

int dnsmi_iterate_correction_carry(dnsmi_ctx* ctx, double* pred, const double* carry, double* err_dev,
                                   void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(!c->walls, "iterate_correction_carry: periodic contexts only");
  DNSMI_REQUIRE(pred && carry && err_dev && (const void*)pred != (const void*)carry,
                "iterate_correction_carry: null or aliased pointer");
  return dnsmi::periodic_iterate_correction_carry(c, pred, carry, err_dev, as_stream(stream));
}

int dnsmi_correct_velocity(dnsmi_ctx* ctx, double* state, double* err_dev, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state, "correct_velocity: null pointer");
  if (c->walls) {
    DNSMI_REQUIRE(!err_dev, "correct_velocity: walls measure no correction");
    return DNSMI_OK;
  }
  return dnsmi::periodic_correct_velocity(c, state, err_dev, as_stream(stream));
}

int dnsmi_get_stats(dnsmi_ctx* ctx, const double* state, double* stats_dev, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state && stats_dev, "get_stats: null pointer");
  return dnsmi::get_stats(c, state, stats_dev, as_stream(stream));
}

int dnsmi_step_device(dnsmi_ctx* ctx, const double* state, double* next, double* rhs_a,
                      double* rhs_b, int slots, int esc_slots, int* counts, int counts_cap,
                      double tol, int max_iter, double esc_ratio, double esc_cap, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state && next && rhs_a && rhs_b && slots >= 0 && esc_slots >= 0 && max_iter >= 0,
                "step_device: bad arguments");
  DNSMI_REQUIRE((const void*)state != next && rhs_a != rhs_b && (void*)next != rhs_a &&
                    (void*)next != rhs_b && (const void*)state != rhs_a &&
                    (const void*)state != rhs_b,
                "step_device: buffers must be distinct");
  return dnsmi::step_device(c, state, next, rhs_a, rhs_b, slots, esc_slots, counts, counts_cap, tol,
                            max_iter, esc_ratio, esc_cap, as_stream(stream));
}

int dnsmi_split_rerecord(dnsmi_ctx* ctx) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(c->split, "split_rerecord: the context has no split evaluations");
  c->ref_valid = false;
  return DNSMI_OK;
}

int dnsmi_step_stats(dnsmi_ctx* ctx, double* out6) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(out6, "step_stats: null output");
  return dnsmi::step_read_ctl(c, out6);
}

int dnsmi_step_reset(dnsmi_ctx* ctx) {
  DNSMI_CHECK_CTX(ctx)
  return dnsmi::step_reset_ctl(c);
}

int dnsmi_escalate(dnsmi_ctx* ctx, int fp64) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(!fp64 || c->fp32_primary, "escalate: FP64 passes need the mixed form");
  c->fp32 = fp64 ? false : c->fp32_primary;
  return DNSMI_OK;
}

// Context-free entry points report through dnsmi_last_error(NULL).
static int ctxfree(int rc, const char* what) {
  if (rc == DNSMI_EINVAL)
    dnsmi::g_create_error = std::string(what) + ": invalid argument (hidden width must be 32 or 64)";
  else if (rc)
    dnsmi::g_create_error = std::string(what) + ": kernel launch failed";
  return rc;
}

int dnsmi_mlp_sizes(int H, long long* nparams, long long* packed_weight_bytes,
                    long long* bias_floats) {
  if (H != 32 && H != 64) return ctxfree(DNSMI_EINVAL, "mlp_sizes");
  const long long nh = H / 16;
  if (nparams) *nparams = 16LL * H + H + (long long)H * H + H + 6LL * H + 6;
  if (packed_weight_bytes) *packed_weight_bytes = (2 * nh + nh * nh) * 64 * 4 * 2;
  if (bias_floats) *bias_floats = 2LL * H + 16;
  return DNSMI_OK;
}

int dnsmi_mlp_grad(int H, int B, const float* X, const float* T, const float* params, int bf16,
                   float* part, float* grad, float* loss_dev, void* stream) {
  if (!X || !T || !params || !part || !grad || B <= 0) return ctxfree(DNSMI_EINVAL, "mlp_grad");
  return ctxfree(dnsmi::mlp_grad(H, B, X, T, params, bf16, part, grad, loss_dev, as_stream(stream)),
                 "mlp_grad");
}

int dnsmi_mlp_grad_static(int H, int B, const float* X, const float* T, const float* params,
                          int bf16, float* grad, float* loss_dev, void* stream) {
  if (!X || !T || !params || !grad || B <= 0) return ctxfree(DNSMI_EINVAL, "mlp_grad_static");
  if (H != 32 && H != 64) return ctxfree(DNSMI_EINVAL, "mlp_grad_static");
  std::string why;
  const int rc = dnsmi::mlp_grad_static(H, B, X, T, params, bf16, grad, loss_dev, as_stream(stream), &why);
  if (rc && !why.empty()) {
    dnsmi::g_create_error = why;
    return rc;
  }
  return ctxfree(rc, "mlp_grad_static");
}

int dnsmi_mlp_adam(int H, float* params, const float* grad, float* m, float* v, int step, float lr,
                   void* stream) {
  if (!params || !grad || !m || !v || step < 1) return ctxfree(DNSMI_EINVAL, "mlp_adam");
  return ctxfree(dnsmi::mlp_adam(H, params, grad, m, v, step, lr, as_stream(stream)), "mlp_adam");
}

long long dnsmi_pnet_params(int din, int H, int dout) {
  if (din <= 0 || din % 32 || H <= 0 || H % 64 || dout <= 0) return -1;
  return (long long)din * H + H + (long long)H * H + H + (long long)H * dout + dout;
}

int dnsmi_pnet_grad(int din, int H, int dout, int B, const void* X, const float* T, const float* params,
                    void* wb, float* wf, void* wh, float* part, float* grad, float* loss_dev, void* stream) {
  if (!X || !T || !params || !wb || !wf || !wh || !part || !grad || B <= 0 || dnsmi_pnet_params(din, H, dout) < 0)
    return ctxfree(DNSMI_EINVAL, "pnet_grad");
  return ctxfree(dnsmi::pnet_grad(din, H, dout, B, static_cast<const uint16_t*>(X), T, params,
                                  static_cast<uint16_t*>(wb), wf, static_cast<uint16_t*>(wh), part, grad,
                                  loss_dev, as_stream(stream)),
                 "pnet_grad");
}

int dnsmi_pnet_forward(int din, int H, int dout, int B, const void* X, const float* params, void* wb,
                       float* wf, void* wh, float* Y, void* stream) {
  if (!X || !params || !wb || !wf || !wh || !Y || B <= 0 || dnsmi_pnet_params(din, H, dout) < 0)
    return ctxfree(DNSMI_EINVAL, "pnet_forward");
  return ctxfree(dnsmi::pnet_forward(din, H, dout, B, static_cast<const uint16_t*>(X), params,
                                     static_cast<uint16_t*>(wb), wf, static_cast<uint16_t*>(wh), Y,
                                     as_stream(stream)),
                 "pnet_forward");
}

int dnsmi_pnet_adam(long long P, float* params, const float* grad, float* m, float* v, int step, float lr,
                    void* stream) {
  if (!params || !grad || !m || !v || step < 1 || P <= 0) return ctxfree(DNSMI_EINVAL, "pnet_adam");
  return ctxfree(dnsmi::pnet_adam(P, params, grad, m, v, step, lr, as_stream(stream)), "pnet_adam");
}

int dnsmi_mlp_pack(int H, const float* params, void* packed_weights, float* bias, void* stream) {
  if (!params || !packed_weights || !bias) return ctxfree(DNSMI_EINVAL, "mlp_pack");
  return ctxfree(dnsmi::mlp_pack(H, params, static_cast<uint16_t*>(packed_weights), bias,
                                 as_stream(stream)),
                 "mlp_pack");
}

int dnsmi_mlp_forward(int H, const float* X, int B, const void* packed_weights, const float* bias,
                      float* Y, void* stream) {
  if (!X || !packed_weights || !bias || !Y || B < 0) return ctxfree(DNSMI_EINVAL, "mlp_forward");
  return ctxfree(dnsmi::mlp_forward(H, X, B, static_cast<const uint16_t*>(packed_weights), bias, Y,
                                    as_stream(stream)),
                 "mlp_forward");
}

int dnsmi_mlp_features(dnsmi_ctx* ctx, const double* n_m3, const double* n_m2, const double* n_m1,
                       const double* n_cur, const double* n_next, const long long* modes,
                       int nsamples, float s_1, float s_2, float s_out, float* X, float* T,
                       void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(n_m2 && n_m1 && n_cur && modes && X && nsamples >= 0,
                "mlp_features: bad arguments");
  DNSMI_REQUIRE(!T == !n_next, "mlp_features: targets need n_next and T together");
  DNSMI_REQUIRE(!n_m3 || !c->walls, "mlp_features: cubic targets are for periodic contexts");
  return dnsmi::mlp_features(c, n_m3, n_m2, n_m1, n_cur, n_next, modes, nsamples, s_1, s_2, s_out, X, T,
                             as_stream(stream));
}

int dnsmi_pnet_features(dnsmi_ctx* ctx, const double* n_m2, const double* n_m1, const double* n_cur,
                        const double* n_next, const long long* modes, int nsamples, int din, float s_1,
                        float s_2, float s_out, void* X, float* T, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(n_m2 && n_m1 && n_cur && modes && X && nsamples >= 0, "pnet_features: null pointer");
  return dnsmi::pnet_features(c, n_m2, n_m1, n_cur, n_next, modes, nsamples, din, s_1, s_2, s_out,
                              static_cast<uint16_t*>(X), T, as_stream(stream));
}

int dnsmi_guess_profile(dnsmi_ctx* ctx, const double* state, double* pred, double* rhs_prev,
                        const float* hist1, float* hist2, int din, int H, int dout, const float* params,
                        void* wb, void* X, float* wf, void* wh, float* Y, float s_1, float s_2, float s_out,
                        void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state && pred && rhs_prev && hist1 && hist2 && params && wb && X && wf && wh && Y,
                "guess_profile: null pointer");
  DNSMI_REQUIRE(dnsmi_pnet_params(din, H, dout) > 0, "guess_profile: din a multiple of 32, H of 64");
  DNSMI_REQUIRE((const void*)state != pred && (void*)pred != rhs_prev && (const void*)hist1 != hist2,
                "guess_profile: buffers must be distinct");
  return dnsmi::guess_walls_profile(c, state, pred, rhs_prev, hist1, hist2, din, H, dout, params,
                                    static_cast<uint16_t*>(wb), static_cast<uint16_t*>(X), wf,
                                    static_cast<uint16_t*>(wh), Y, s_1, s_2, s_out, as_stream(stream));
}

int dnsmi_guess(dnsmi_ctx* ctx, const double* state, double* pred, double* rhs_prev,
                float* hist1, float* hist2, float* hist3, const void* packed_weights, const float* bias,
                int H, float s_1, float s_2, float s_out, int mode, int carry, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state && pred && rhs_prev && hist1 && hist2, "guess: null pointer");
  DNSMI_REQUIRE(mode != 2 || (packed_weights && bias), "guess: network weights missing");
  DNSMI_REQUIRE((const void*)state != pred && (const void*)state != rhs_prev &&
                    (void*)pred != rhs_prev && hist1 != hist2 && hist1 != hist3 && hist2 != hist3,
                "guess: buffers must be distinct");
  DNSMI_REQUIRE(!hist3 || !c->walls, "guess: hist3 (cubic extrapolation) is for periodic contexts");
  DNSMI_REQUIRE(!carry || !c->walls, "guess: the carry output is for periodic contexts");
  const uint16_t* wb = static_cast<const uint16_t*>(packed_weights);
  if (c->walls)
    return dnsmi::guess_walls(c, state, pred, rhs_prev, hist1, hist2, wb, bias, H, s_1, s_2, s_out,
                              mode, as_stream(stream));
  return dnsmi::guess_periodic(c, state, pred, rhs_prev, hist1, hist2, hist3, wb, bias, H, s_1, s_2,
                               s_out, mode, as_stream(stream), carry);
}

int dnsmi_step_device_guess(dnsmi_ctx* ctx, const double* state, double* next, double* rhs_a,
                            double* rhs_b, float* hist1, float* hist2, float* hist3,
                            const void* packed_weights, const float* bias, int H, float s_1,
                            float s_2, float s_out, int mode, int slots, int esc_slots,
                            int* counts, int counts_cap, double tol, int max_pass,
                            double esc_ratio, double esc_cap, void* stream) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(state && next && rhs_a && rhs_b && hist1 && hist2, "step_device_guess: null pointer");
  DNSMI_REQUIRE(mode != 2 || (packed_weights && bias), "step_device_guess: network weights missing");
  DNSMI_REQUIRE(slots >= 0 && esc_slots >= 0 && max_pass >= 1 && (counts || counts_cap == 0),
                "step_device_guess: bad slots / max_pass / counts");
  DNSMI_REQUIRE((const void*)state != next && (void*)rhs_a != rhs_b && (void*)next != rhs_a &&
                    (void*)next != rhs_b && hist1 != hist2 && hist1 != hist3 && hist2 != hist3,
                "step_device_guess: buffers must be distinct");
  DNSMI_REQUIRE(!hist3 || !c->walls, "step_device_guess: hist3 (cubic extrapolation) is for periodic contexts");
  return dnsmi::step_device_guess(c, state, next, rhs_a, rhs_b, hist1, hist2, hist3,
                                  static_cast<const uint16_t*>(packed_weights), bias, H, s_1, s_2,
                                  s_out, mode, slots, esc_slots, counts, counts_cap, tol, max_pass,
                                  esc_ratio, esc_cap, as_stream(stream));
}

const char* dnsmi_tuning(const dnsmi_ctx* ctx) { return ctx ? ctx->impl.tuning.c_str() : ""; }

int dnsmi_variants(const dnsmi_ctx* ctx, int* tile_cols, int* variants) {
  if (!ctx || !tile_cols || !variants) {
    dnsmi::g_create_error = "variants: bad arguments";
    return DNSMI_EINVAL;
  }
  const Ctx* c = &ctx->impl;
  *tile_cols = c->tile_cols;
  *variants = (c->yinv_seq ? 1 : 0) | (c->zinv_seq ? 2 : 0) | (c->xpass_split ? 4 : 0) | (c->sweep_seq ? 8 : 0) |
              (c->imm_fused ? 16 : 0) | (c->zinv_nt ? 32 : 0) | (c->walls && c->curl_rows == 64 ? 64 : 0);
  return DNSMI_OK;
}

int dnsmi_walls_table(dnsmi_ctx* ctx, int which, double* out, long long n) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(out && n >= 0, "walls_table: bad output");
  return dnsmi::walls_table(c, which, out, n);
}

static const char* const kKernelNames[dnsmi::KID_COUNT] = {
    "y_inverse", "z_inverse", "x_nonlinear", "z_forward", "y_forward", "pointwise", "finalize_norm",
    "walls_z_inverse", "walls_x_nonlinear", "walls_z_forward", "walls_imm", "walls_norm"};

int dnsmi_profile_begin(dnsmi_ctx* ctx) {
  DNSMI_CHECK_CTX(ctx)
  c->prof_recs.clear();
  c->ev_used = 0;
  c->prof = true;
  return DNSMI_OK;
}

int dnsmi_profile_end(dnsmi_ctx* ctx, int n_ids, double* total_ms, long long* launches) {
  DNSMI_CHECK_CTX(ctx)
  DNSMI_REQUIRE(n_ids >= 0 && (n_ids == 0 || (total_ms && launches)), "profile_end: bad arrays");
  c->prof = false;
  for (int i = 0; i < n_ids; ++i) {
    total_ms[i] = 0.0;
    launches[i] = 0;
  }
  for (const dnsmi::ProfRec& r : c->prof_recs) {
    if (hipEventSynchronize(r.b) != hipSuccess) return c->fail(DNSMI_EHIP, "profile_end: event sync");
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess)
      return c->fail(DNSMI_EHIP, "profile_end: elapsed time");
    if (r.kid < n_ids) {
      total_ms[r.kid] += ms;
      launches[r.kid] += 1;
    }
  }
  c->prof_recs.clear();
  c->ev_used = 0;
  return DNSMI_OK;
}

const char* dnsmi_kernel_name(int id) {
  return (id >= 0 && id < dnsmi::KID_COUNT) ? kKernelNames[id] : "";
}

}  // extern "C"
