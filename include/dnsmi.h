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
  int tile_cols;           /* 0: autotune the kernel variants at create;
                              otherwise use `variants` as given */
  int variants;            /* bit mask when tile_cols != 0: 1 sequential
                              y-inverse, 2 sequential z-inverse, 4 split x pass */
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

/* Divergence correction + mean-mode zeroing in place (periodic); a
 * pass-through for walls.  If err_dev is not NULL it receives the norm of
 * the divergence correction (debug.measure_corrections). */
int dnsmi_correct_velocity(dnsmi_ctx* ctx, double* state, double* err_dev, void* stream);

/* The kernel variants the autotuner chose at create (tile_cols == 0), as
 * "name=base|alt(us base/us alt) ..." (no counterpart in the reference). */
const char* dnsmi_tuning(const dnsmi_ctx* ctx);

/* Host copy (out: n host doubles) of a wall-bounded set-up table in the
 * reference layout, for inspection (no counterpart in the reference, which
 * keeps these as module attributes, flows/plane_couette.py:139-219):
 * which = 0 Lk_ab (Nkz,Nkx,3p+1,Ny), 1 Hk_ab, 2 v1 (Nkz,Nkx,Ny), 3 q1,
 * 4 M_inv (Nkz,Nkx,2,2), 5 y grid (Ny), 6 Lk_piv, 7 Hk_piv (0-based pivots
 * as doubles).  Synchronous, host memory only. */
int dnsmi_walls_table(dnsmi_ctx* ctx, int which, double* out, long long n);

/* Per-kernel timing with HIP events on the launch stream (no counterpart in
 * the reference, whose bench.timer (bench.py:26-67) times whole calls).
 * Between begin and end every kernel launch of the context is bracketed by
 * an event pair; end synchronises those events and returns, per kernel id
 * (< n_ids), the summed milliseconds and the launch count. */
int dnsmi_profile_begin(dnsmi_ctx* ctx);
int dnsmi_profile_end(dnsmi_ctx* ctx, int n_ids, double* total_ms, long long* launches);
const char* dnsmi_kernel_name(int id);

#ifdef __cplusplus
}
#endif

#endif /* DNSMI_H */
