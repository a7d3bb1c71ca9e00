/*
 * dnsmi: MI355X-native time-step hot path of the dnsjax pseudo-spectral
 * Navier-Stokes solver, as a C ABI over hand-written HIP kernels (gfx950).
 *
 * Every array argument is a device pointer to a caller-owned, contiguous
 * buffer in the reference's layouts:
 *   periodic spectral state / rhs : (3, ny-1, nz-1, nx/2)  [ky, kz, kx]
 *   wall-bounded spectral state   : (3, nz-1, nx/2, ny)    [kz, kx, y]
 * complex values interleaved (re, im) float64.  `err_dev` is a device
 * double.  `stream` is a hipStream_t (NULL = the default stream); all work is
 * stream-ordered and no call synchronises the host unless its comment says
 * so.  Functions return 0 on success and a DNSMI_E* code otherwise; the
 * message is in dnsmi_last_error(ctx) (or dnsmi_last_error(NULL) for a
 * failed create).  A context is used from one host thread at a time.
 *
 * Reference interfaces replaced (paths under the reference's src/dnsjax/):
 *   dnsmi_get_rhs               <- get_rhs_fn: flows/monochromatic.py:366-386,
 *                                  flows/plane_couette.py:274-289
 *   dnsmi_predict               <- predict_fn: flows/monochromatic.py:389-393,
 *                                  flows/plane_couette.py:507-519
 *   dnsmi_correct               <- correct_fn: flows/monochromatic.py:396-405,
 *                                  flows/plane_couette.py:522-540
 *   dnsmi_norm                  <- norm_fn: flows/monochromatic.py:408-410,
 *                                  flows/plane_couette.py:543-549
 *   dnsmi_predict_and_correct   <- timestep.py:68-94 (make_stepper result 1)
 *   dnsmi_iterate_correction    <- timestep.py:96-120 (make_stepper result 2)
 *   dnsmi_correct_velocity      <- flows/monochromatic.py:545-572,
 *                                  flows/plane_couette.py:575-582
 *   dnsmi_create / dnsmi_destroy<- the flow set-up done at module import:
 *                                  flows/monochromatic.py:113-288,
 *                                  flows/plane_couette.py:64-222,
 *                                  geometries/cartesian.py:575-801
 */
#ifndef DNSMI_H
#define DNSMI_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DNSMI_OK = 0,
  DNSMI_EINVAL = 1,  /* bad argument or unsupported configuration */
  DNSMI_EHIP = 2,    /* a HIP runtime call failed */
  DNSMI_ENOMEM = 3,  /* device allocation failed */
  DNSMI_ESINGULAR = 4 /* banded LU factorisation hit a zero pivot */
};

enum {
  DNSMI_KOLMOGOROV = 0,
  DNSMI_WALEFFE = 1,
  DNSMI_DECAYING_BOX = 2,
  DNSMI_PLANE_COUETTE = 3
};

enum {
  DNSMI_FP64 = 0,  /* everything in float64 */
  DNSMI_MIXED = 1  /* FFT passes and nonlinear product in float32; state,
                      rhs, Helmholtz/IMM solves, norms and projections
                      in float64 */
};

typedef struct dnsmi_config {
  int system;              /* DNSMI_KOLMOGOROV ... DNSMI_PLANE_COUETTE */
  int nx, ny, nz;          /* modes before dealiasing (parameters.py:55-63) */
  int oversampling_factor; /* 3 (the 3/2 rule); parameters.py:43 */
  int oversample_y;        /* periodic only; parameters.py:44 */
  double lx, ly, lz;       /* box; ly is 2 for walls (parameters.py:188-191) */
  double re, dt, implicitness; /* parameters.py:36,90-91 */
  double tilt_degree;      /* periodic forcing tilt; parameters.py:52 */
  int fd_order;            /* walls: stencil order p; parameters.py:62 */
  int precision;           /* DNSMI_FP64 or DNSMI_MIXED */
  int correct_divergence;  /* parameters.py:108 */
  int tile_cols;           /* 0: autotune the kernel variants and the column
                              tile of the y/z passes at create; 8 or 16: use
                              that tile and `variants` as given */
  int variants;            /* bit mask when tile_cols != 0: 1 sequential
                              y-inverse, 2 sequential z-inverse, 4 split x pass,
                              8 sequential band sweeps (walls; else chunk maps),
                              16 the chunk maps' fused two-launch influence-
                              matrix step (walls, without 8), 32 non-temporal
                              stores of the z-inverse output (periodic), 64
                              64-row workgroups of the FD curl (walls) */
  int plain_mixed;         /* DNSMI_MIXED only: 0 (default) evaluates every
                              corrector pass of a step as an FP32 increment over
                              the step's first evaluation (split precision);
                              1 evaluates each pass in full in FP32 */
  double recenter_far;     /* split evaluations only, device steps (host
                              loops: dnsmi_split_rerecord): a step whose
                              predict-and-correct error exceeds recenter_far
                              times the tolerance records a new reference at
                              its first pass that starts from an error of at
                              most recenter_near times the tolerance; 0: never */
  double recenter_near;
  int setup_on_host;       /* walls: 1 builds the influence-matrix tables
                              (band LU, homogeneous solutions, M^-1, chunk
                              maps) on the host; 0 (default) on the GPU, with
                              the same operations in the same order (the two
                              agree bit for bit), falling back to the host
                              where a class's band does not fit 64 KiB of LDS
                              (geometries/cartesian.py:575-801) */
} dnsmi_config;

typedef struct dnsmi_ctx dnsmi_ctx;

int dnsmi_create(const dnsmi_config* cfg, dnsmi_ctx** out);
void dnsmi_destroy(dnsmi_ctx* ctx);
const char* dnsmi_last_error(const dnsmi_ctx* ctx);

/* Sizes: number of complex elements of one spectral vector field. */
long long dnsmi_spec_elems(const dnsmi_ctx* ctx);

/* rhs_no_lapl of `state` (rhs may not alias state). */
int dnsmi_get_rhs(dnsmi_ctx* ctx, const double* state, double* rhs, void* stream);

/* Euler predictor (pred may alias state). */
int dnsmi_predict(dnsmi_ctx* ctx, const double* state, const double* rhs, double* pred,
                  void* stream);

/* One corrector application: pred_out = correction applied to pred_in,
 * correction_out = pred_out - pred_in (either output may alias pred_in; the
 * correction may be NULL). */
int dnsmi_correct(dnsmi_ctx* ctx, const double* state_prev, const double* pred_in,
                  const double* rhs_prev, const double* rhs_next, double* pred_out,
                  double* correction_out, void* stream);

/* err_dev = convergence norm of a correction field. */
int dnsmi_norm(dnsmi_ctx* ctx, const double* correction, double* err_dev, void* stream);

/* Fused predict + one correct: state is not modified. */
int dnsmi_predict_and_correct(dnsmi_ctx* ctx, const double* state, double* pred_out,
                              double* rhs_next_out, double* err_dev, void* stream);

/* One more corrector iteration; pred is updated in place (the donation of
 * timestep.py:97).  rhs_next_out may alias rhs_prev. */
int dnsmi_iterate_correction(dnsmi_ctx* ctx, const double* state_prev, double* pred,
                             const double* rhs_prev, double* rhs_next_out, double* err_dev,
                             void* stream);

/* Carry form of the two calls above (periodic contexts only; no
 * counterpart in the reference, whose iteration it reproduces in its full
 * form): carry_out receives B = ildt_2 (ldt_1 u^n + (1 - c) N^n) instead of
 * N, and each corrector pass is pred <- B + c ildt_2 N(pred), reading carry
 * and pred and writing pred only (the incremental form u + c ildt_2 (N(u) -
 * N_prev), which the reference writes, telescopes to it; results agree to
 * rounding).  err_dev as above.  The device steps use this form. */
int dnsmi_predict_and_correct_carry(dnsmi_ctx* ctx, const double* state, double* pred_out,
                                    double* carry_out, double* err_dev, void* stream);
int dnsmi_iterate_correction_carry(dnsmi_ctx* ctx, double* pred, const double* carry, double* err_dev,
                                   void* stream);

/* Divergence correction + mean-mode zeroing in place (periodic); a
 * pass-through for walls.  If err_dev is not NULL it receives the norm of
 * the divergence correction (debug.measure_corrections). */
int dnsmi_correct_velocity(dnsmi_ctx* ctx, double* state, double* err_dev, void* stream);

/* The kernel variants the autotuner chose at create (tile_cols == 0), as
 * "name=base|alt(us base/us alt) ..." (no counterpart in the reference). */
const char* dnsmi_tuning(const dnsmi_ctx* ctx);

/* The kernel variants the context runs, in the config's convention
 * (tile_cols 8 or 16; variants bit mask as for dnsmi_config.variants), so
 * that further contexts can be created with the same choice (no
 * counterpart in the reference). */
int dnsmi_variants(const dnsmi_ctx* ctx, int* tile_cols, int* variants);

/* Host copy (out: n host doubles) of a wall-bounded set-up table in the
 * reference layout, for inspection (no counterpart in the reference, which
 * keeps these as module attributes, flows/plane_couette.py:139-219):
 * which = 0 Lk_ab (Nkz,Nkx,3p+1,Ny), 1 Hk_ab, 2 v1 (Nkz,Nkx,Ny), 3 q1,
 * 4 M_inv (Nkz,Nkx,2,2), 5 y grid (Ny), 6 Lk_piv, 7 Hk_piv (0-based pivots
 * as doubles); as stored on the device: 8 .. 11 the packed chunk maps of the
 * L_k forward, L_k back, H_k forward and H_k back solves, 12 the per-class
 * homogeneous solutions ([class][y][v1, v2, q1, q2]), 13 the word counts of
 * tables 8 .. 11 (0 where a table is absent) and the number of k^2 classes
 * (5 values).  Read back from the device tables; synchronous, host memory
 * only. */
int dnsmi_walls_table(dnsmi_ctx* ctx, int which, double* out, long long n);

/* One full time step with the corrector loop decided on the device
 * (__main__.py:161-191 without host reads): predict_and_correct from
 * `state` into `next`, then up to `slots` iterate_correction passes, each
 * skipped on the device once error <= tol or max_iter extra passes ran, then
 * correct_velocity(next).  Mixed form: `esc_slots` FP64 passes follow, run
 * only by a step that escalated: a pass whose error (tol < error <= esc_cap)
 * is not below esc_ratio times the error two passes earlier has reached the
 * resolution of the FP32 evaluations, and so has a step whose `slots` passes
 * all ran without converging; its remaining passes run in FP64 and it still
 * ends at error <= tol (esc_slots 0: such a step ends unconverged).
 * rhs_a / rhs_b are two RHS work arrays.  No host synchronisation; the
 * launch sequence is fixed, so it can be captured in a hipGraph.  counts
 * (device int[counts_cap], may be NULL) receives each step's
 * extra-iteration count in order.  `state` is not modified. */
int dnsmi_step_device(dnsmi_ctx* ctx, const double* state, double* next, double* rhs_a,
                      double* rhs_b, int slots, int esc_slots, int* counts, int counts_cap,
                      double tol, int max_iter, double esc_ratio, double esc_cap, void* stream);
/* Synchronising read of the device step counters: out7 = {steps, total extra
 * iterations, steps not converged (error > tol at the end), NaN steps, last
 * error, steps that escalated to FP64 passes, steps that re-recorded their
 * split reference}; dnsmi_step_reset zeroes them. */
int dnsmi_step_stats(dnsmi_ctx* ctx, double* out7);
int dnsmi_step_reset(dnsmi_ctx* ctx);

/* Split evaluations only (no counterpart in the reference): the next
 * corrector evaluation (dnsmi_iterate_correction) records a new reference
 * at its iterate, and the later ones are increments over it.  A host loop
 * re-centres as the device steps do with recenter_far / recenter_near. */
int dnsmi_split_rerecord(dnsmi_ctx* ctx);

/* Mixed form only (no counterpart in the reference): while fp64 != 0 the
 * corrector passes (dnsmi_iterate_correction and the passes of the device
 * steps) evaluate in FP64, the host loop's escalation after a pass that no
 * longer converges in FP32 (see dnsmi_step_device); 0 returns to the mixed
 * form. */
int dnsmi_escalate(dnsmi_ctx* ctx, int fp64);

/* Diagnostics of a state into stats_dev[4] = {E, I, D, E'} (device doubles).
 * Replaces get_stats: flows/monochromatic.py:433-510 (E' = |u|^2/2, input,
 * dissipation, energy; the forcing constants use F = nu (2 pi/Ly)^2, equal
 * to the reference's pi^2/(4 Re) at Ly = 4) and flows/plane_couette.py:591-618
 * (E' only; E = E', I = D = 0). */
int dnsmi_get_stats(dnsmi_ctx* ctx, const double* state, double* stats_dev, void* stream);

/* ---- Learned corrector guess (no counterpart in the reference) ----------
 * A per-mode network x (16 features) -> 6 outputs, hidden width H = 32 or
 * 64, SiLU, parameters in FP32 packed as W1 (16,H) | b1 (H) | W2 (H,H) | b2 (H)
 * | W3 (H,6) | b3 (6).  It predicts the second difference of the nonlinear
 * term per mode so that the first corrector pass starts near the implicit
 * solution; the corrector loop and its stopping rule are unchanged
 * (__main__.py:161-191), only its starting point moves.  All buffers are
 * device memory; entry points without a context report through
 * dnsmi_last_error(NULL). */

/* Sizes: parameter count, packed bf16 weight bytes and bias floats. */
int dnsmi_mlp_sizes(int H, long long* nparams, long long* packed_weight_bytes,
                    long long* bias_floats);
/* Gradient of L = (1/B) sum (y - t)^2 over X (B,16), T (B,6), FP32, samples
 * summed in a fixed order (deterministic).  part: (ceil(B/32) +
 * ceil(B/1024)) * (nparams + 1) floats of scratch; grad: nparams floats;
 * loss_dev: 1 float or NULL.
 * bf16 != 0 rounds weights and layer inputs to bfloat16 as inference does
 * (straight-through gradients). */
int dnsmi_mlp_grad(int H, int B, const float* X, const float* T, const float* params, int bf16,
                   float* part, float* grad, float* loss_dev, void* stream);
/* dnsmi_mlp_grad without the `part` workspace and without any device
 * allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under graph capture, serving loops).
 * The scratch is a file-scope device array of 6.0 MB per GPU (held until
 * the process ends), which limits a call to B <= 8192 samples (larger B:
 * DNSMI_EINVAL).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other streams, other threads) race.  Same kernels and sums as
 * dnsmi_mlp_grad, so the results are identical. */
int dnsmi_mlp_grad_static(int H, int B, const float* X, const float* T, const float* params,
                          int bf16, float* grad, float* loss_dev, void* stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8; step counts from 1). */
int dnsmi_mlp_adam(int H, float* params, const float* grad, float* m, float* v, int step, float lr,
                   void* stream);
/* Pack params into matrix-core operand tiles (bf16) and a bias block. */
int dnsmi_mlp_pack(int H, const float* params, void* packed_weights, float* bias, void* stream);
/* Matrix-core inference: Y (B,6) = net(X (B,16)). */
int dnsmi_mlp_forward(int H, const float* X, int B, const void* packed_weights, const float* bias,
                      float* Y, void* stream);
/* Training samples from four successive nonlinear terms N^{n-2}, N^{n-1},
 * N^n, N^{n+1} (spectral layout, complex128) of a periodic context: X (ns,16)
 * for the flat mode indices `modes` (over (ky, kz, kx)) and, when n_next
 * and T are both given, T (ns,6) = w(k) (N^{n+1} - 3 N^n + 3 N^{n-1} -
 * N^{n-2}) / s_out, w = 1 / (1 + c dt |k|^2 / Re): the residual of quadratic
 * extrapolation; with n_m3 = N^{n-3} (periodic only, else NULL) the
 * residual of cubic extrapolation, w (N^{n+1} - 4 N^n + 6 N^{n-1} -
 * 4 N^{n-2} + N^{n-3}) / s_out.  Features: D1 = N^n - N^{n-1} times s_1, D2 = N^n -
 * 2 N^{n-1} + N^{n-2} times s_2 (older terms rounded to FP32 as the step
 * history keeps them), min(dt |k|^2/Re, 8), k / k_max.  Wall-bounded
 * contexts: `modes` index points over (kz, kx, y); y itself replaces
 * ky / k_max, |k|^2 = kx^2 + kz^2, and the weight is not zeroed at the mean
 * mode. */
int dnsmi_mlp_features(dnsmi_ctx* ctx, const double* n_m3, const double* n_m2, const double* n_m1,
                       const double* n_cur, const double* n_next, const long long* modes,
                       int nsamples, float s_1, float s_2, float s_out, float* X, float* T,
                       void* stream);
/* Predictor with a guess.  Periodic contexts: N^n = rhs(state), u^1 =
 * predict(state, N^n) (flows/monochromatic.py:320-333), then per mode
 * dN = 0 (mode 0), D1 (mode 1, linear extrapolation), D1 + D2 (mode 3,
 * quadratic), D1 + D2 + D3 (mode 4, cubic, D3 = N^n - 3 N^{n-1} +
 * 3 N^{n-2} - N^{n-3}) or the quadratic (cubic with hist3) extrapolation
 * plus the network's residual (mode 2), Leray-projected and zero at the
 * mean mode; pred = u^1 + c ildt_2 dN and rhs_prev = N^n + dN, so that
 * iterate_correction(state, pred, rhs_prev) is the reference's fixed-point
 * iteration in its full form u <- ildt_2 (ldt_1 u^n + c N(u) + (1 - c) N^n),
 * started from the guess.  hist1 = N^{n-1}, hist2 = N^{n-2}, hist3 =
 * N^{n-3} or NULL: 3 * spec_elems complex values each as FP32 re/im pairs.
 * N^n overwrites the oldest history: without hist3, hist2, so the caller
 * swaps hist1 and hist2 for the next step; with hist3 (periodic contexts
 * only), hist3, so the caller rotates (hist1, hist2, hist3) <- (hist3,
 * hist1, hist2).  Mode 0 (the plain predictor) reads no
 * history.  carry != 0 (periodic contexts): rhs_prev receives the carry
 * form's B = ildt_2 (ldt_1 u^n + (1 - c) N^n) instead, for
 * dnsmi_iterate_correction_carry.
 * Wall-bounded contexts: pred = the predictor's solve
 * (flows/plane_couette.py:507-519) fed with N_ext = N^n + dN per point
 * (same modes, with the network's residual divided by w = 1 / (1 + c dt
 * (kx^2 + kz^2) / Re)) and rhs_prev = N_ext: the walls corrector's fixed
 * point u* = S(u^n; N(u*)) is then reached from the guess. */
int dnsmi_guess(dnsmi_ctx* ctx, const double* state, double* pred, double* rhs_prev,
                float* hist1, float* hist2, float* hist3, const void* packed_weights, const float* bias,
                int H, float s_1, float s_2, float s_out, int mode, int carry, void* stream);

/* One time step started from the corrector guess, decided on the device
 * (periodic and wall-bounded): dnsmi_guess from `state` into `next` (with rhs_a as
 * rhs_prev), then up to `slots` corrector passes, each skipped on the device
 * once error <= tol or max_pass passes ran, and the escalation slots of
 * dnsmi_step_device; then correct_velocity(next).  counts[i] receives
 * step i's number of passes (one RHS each).  hist1..hist3 as in dnsmi_guess
 * (swap or rotate them, and swap state and next, for the next step).  A
 * fixed launch sequence with no host synchronisation: capturable. */
int dnsmi_step_device_guess(dnsmi_ctx* ctx, const double* state, double* next, double* rhs_a,
                            double* rhs_b, float* hist1, float* hist2, float* hist3,
                            const void* packed_weights, const float* bias, int H, float s_1,
                            float s_2, float s_out, int mode, int slots, int esc_slots,
                            int* counts, int counts_cap, double tol, int max_pass,
                            double esc_ratio, double esc_cap, void* stream);

/* ---- Profile network for the plane Couette guess (no counterpart in the
 * reference) -------------------------------------------------------------
 * One network per context: din -> H -> H -> dout, SiLU, FP32 parameters W1
 * (din,H) | b1 (H) | W2 (H,H) | b2 (H) | W3 (H,dout) | b3 (dout), row-major.
 * Training and inference run every layer as a bf16 matrix-core GEMM with
 * FP32 accumulation (backward included).  One sample is one (kz, kx) mode:
 * input the wall-normal profiles of D1 = N^n - N^{n-1} and D2 = N^n -
 * 2 N^{n-1} + N^{n-2} (3 components, real and imaginary: 12 Ny values
 * scaled by s_1, s_2), dt k^2/Re, kx/k_max, kz/k_max, zero padded to din (a
 * multiple of 32, >= 12 Ny + 3); output (dout = 6 Ny) the profile of the
 * residual of quadratic extrapolation, N^{n+1} - (3 N^n - 3 N^{n-1} +
 * N^{n-2}), times w = 1/(1 + c dt k^2/Re) / s_out.  H a multiple of 64.
 * X and the bf16 buffers are 16-bit words; wb holds one word per parameter. */
long long dnsmi_pnet_params(int din, int H, int dout); /* -1 if unsupported */
/* Gradient of L = (1/(B dout)) sum (Y - T)^2 over X (B,din) bf16, T (B,dout)
 * FP32 into grad (FP32), loss into loss_dev (may be NULL).  Work: wf
 * B (2H + dout) floats, wh B (3H + dout) words, part ceil(B dout/256)
 * floats. */
int dnsmi_pnet_grad(int din, int H, int dout, int B, const void* X, const float* T, const float* params,
                    void* wb, float* wf, void* wh, float* part, float* grad, float* loss_dev, void* stream);
/* Y (B,dout) = net(X); work wf 2 H B floats, wh 2 H B words. */
int dnsmi_pnet_forward(int din, int H, int dout, int B, const void* X, const float* params, void* wb,
                       float* wf, void* wh, float* Y, void* stream);
/* One Adam step over P parameters (beta1 0.9, beta2 0.999, eps 1e-8). */
int dnsmi_pnet_adam(long long P, float* params, const float* grad, float* m, float* v, int step, float lr,
                    void* stream);
/* Training rows X (nsamples, din) and, with n_next, targets T (nsamples,
 * dout) from four successive nonlinear terms of a plane Couette context
 * (walls layout, complex128), for the mode indices `modes` (over (kz, kx));
 * the two older terms are rounded to FP32 as the guess's history keeps
 * them. */
int dnsmi_pnet_features(dnsmi_ctx* ctx, const double* n_m2, const double* n_m1, const double* n_cur,
                        const double* n_next, const long long* modes, int nsamples, int din, float s_1,
                        float s_2, float s_out, void* X, float* T, void* stream);
/* dnsmi_guess (plane Couette) with the profile network: N^n (recorded as
 * the step's reference) into the context, rhs_prev = quadratic
 * extrapolation plus s_out Y / w per point, pred = the predictor's solve fed
 * with rhs_prev, hist2 <- N^n.  Work, M = (nz-1) nx/2 modes: X M din words,
 * wf 2 H M floats, wh 2 H M words, Y M dout floats, wb one word per
 * parameter. */
int dnsmi_guess_profile(dnsmi_ctx* ctx, const double* state, double* pred, double* rhs_prev,
                        const float* hist1, float* hist2, int din, int H, int dout, const float* params,
                        void* wb, void* X, float* wf, void* wh, float* Y, float s_1, float s_2, float s_out,
                        void* stream);

/* Per-kernel timing with HIP events on the launch stream (no counterpart in
 * the reference, whose bench.timer (bench.py:26-67) times whole calls).
 * Between begin and end every kernel launch of the context is bracketed by
 * an event pair; end synchronises those events and returns, per kernel id
 * (< n_ids), the summed milliseconds and the launch count. */
int dnsmi_profile_begin(dnsmi_ctx* ctx);
int dnsmi_profile_end(dnsmi_ctx* ctx, int n_ids, double* total_ms, long long* launches);
const char* dnsmi_kernel_name(int id);

/* Batched device steps of an ensemble (no counterpart in the reference,
 * which runs one trajectory, __main__.py:125-203): for n periodic contexts
 * with the same sizes, precision and kernel variants (each with its own
 * Re, dt, tolerance and buffers), dnsmi_batch_plan records the launches of
 * one dnsmi_step_device per member (member i: states[i] -> nexts[i], work
 * arrays rhs_a[i], rhs_b[i], iteration counts into counts[i] when counts
 * is given) and merges them into one launch per kernel over the members;
 * dnsmi_batch_enqueue issues the whole step of all members on `stream`
 * (capturable).  Each member's results equal its own dnsmi_step_device's
 * bit for bit.  A plan keeps the buffers' addresses (plan each buffer
 * parity once) and its argument tables (a few KB of device memory) until
 * dnsmi_batch_destroy.  Errors of a plan are reported on ctxs[0]. */
typedef struct dnsmi_batch dnsmi_batch;
int dnsmi_batch_plan(dnsmi_ctx* const* ctxs, int n, const double* const* states, double* const* nexts,
                     double* const* rhs_a, double* const* rhs_b, int slots, int esc_slots,
                     int* const* counts, int counts_cap, double tol, int max_iter, double esc_ratio,
                     double esc_cap, dnsmi_batch** out);
int dnsmi_batch_enqueue(dnsmi_batch* b, void* stream);
/* Launches one dnsmi_batch_enqueue issues. */
int dnsmi_batch_launches(const dnsmi_batch* b);
void dnsmi_batch_destroy(dnsmi_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* DNSMI_H */
