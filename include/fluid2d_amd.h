/*
 * fluid2d_amd.h -- C ABI of libf2dgpu.so, the MI355X (gfx950) implementation
 * of the Fluid2d timestep hot path.
 *
 * The reference binds its Fortran kernels into Python with f2py
 * (build.py:14-39); the Python classes Operators, Gmg, Timescheme, Euler and
 * Boussinesq call them.  This header is the replacement boundary: plain C
 * types, device pointers, an explicit stream, and an int status instead of
 * Fortran `stop` / Python `exit`.
 *
 * Conventions
 *  - Fields are row-major float64 arrays of m rows (ny + 2*nh) by n columns
 *    (nx + 2*nh), x contiguous: NumPy C order.  Fortran x(j,i) of the
 *    reference is element (j-1)*n + (i-1).  Masks are int8 0/1.  nh must be 3.
 *  - Kernel-level entry points take device pointers and a stream
 *    (hipStream_t, NULL = the null stream) and only enqueue work; results
 *    are in place, as with the reference's f2py intent(inplace) arguments.
 *    Scalar results are written to a device buffer.
 *  - Handle-level entry points (f2d_gmg_*, f2d_model_*) own their device
 *    buffers.  Calls on one handle must not overlap; different handles are
 *    independent.
 *  - Every function returns F2D_OK (0) or an F2D_E_* code;
 *    f2d_last_error() describes the last failure of the calling thread.
 */
#ifndef FLUID2D_AMD_H
#define FLUID2D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define F2D_OK 0
#define F2D_E_ARG 1         /* invalid argument: nh != 3, bad shape, bad option */
#define F2D_E_HIP 2         /* HIP runtime error */
#define F2D_E_NOMEM 3       /* device allocation failed */
#define F2D_E_DIVERGED 4    /* multigrid solve diverged (core/gmg/hierarchy.py:182-188 exits) */
#define F2D_E_UNSUPPORTED 5 /* option not built */

typedef void *f2d_stream_t; /* a hipStream_t */

const char *f2d_last_error(void);
int f2d_version(void);
int f2d_device_count(int *count);

/* ---- device memory helpers (for callers without another allocator) ---- */
int f2d_malloc(void **ptr, size_t bytes);
int f2d_free(void *ptr);
int f2d_memcpy_h2d(void *dst, const void *src, size_t bytes);
int f2d_memcpy_d2h(void *dst, const void *src, size_t bytes);
int f2d_memcpy_d2d(void *dst, const void *src, size_t bytes);
int f2d_synchronize(void);

/* ---- advection (core/operators.py:214-236 calls these) ---------------- */

/* core/fortran_advection.f90:2-165 adv_upwind(msk,x,y,u,v,cst,nh,method,order);
 * with xflx/yflx non-NULL: core/fortran_fluxes.f90:2-170 (edge fluxes stored).
 * cst = {dx, dy, logcosh, umax, aparab} (host array, core/operators.py:112-117).
 * method 0 minmax, 1 parabolic, 2 logcosh; order 1, 3 or 5.
 * Writes y on the interior only (no halo fill). */
int f2d_adv_upwind_f64(const int8_t *msk, const double *x, double *y, const double *u,
                       const double *v, double *xflx, double *yflx, const double *cst,
                       int nh, int method, int order, int m, int n, f2d_stream_t stream);

/* core/fortran_advection.f90:169-284 adv_centered (orders 2, 4, 6); with
 * xflx/yflx non-NULL: core/fortran_fluxes.f90:174-279 (orders 2, 4). */
int f2d_adv_centered_f64(const int8_t *msk, const double *x, double *y, const double *u,
                         const double *v, double *xflx, double *yflx, const double *cst,
                         int nh, int order, int m, int n, f2d_stream_t stream);

/* ---- grid operators ---------------------------------------------------- */

/* core/fortran_operators.f90:44-64 celltocorner(xr,xp) */
int f2d_celltocorner_f64(const double *xr, double *xp, int m, int n, f2d_stream_t stream);

/* core/fortran_operators.f90:2-39 computeorthogradient(msk,psi,dx,dy,nh,u,v) */
int f2d_orthogradient_f64(const int8_t *msk, const double *psi, double dx, double dy, int nh,
                          double *u, double *v, int m, int n, f2d_stream_t stream);

/* core/fortran_operators.f90:125-156 add_diffusion(msk,trac,dx,nh,Kdiff,dtrac) */
int f2d_add_diffusion_f64(const int8_t *msk, const double *trac, double dx, int nh, double kdiff,
                          double *dtrac, int m, int n, f2d_stream_t stream);

/* core/fortran_operators.f90:330-381 add_torque(msk,buoy,dx,nh,gravity,domega) */
int f2d_add_torque_f64(const int8_t *msk, const double *buoy, double dx, int nh, double gravity,
                       double *domega, int m, int n, f2d_stream_t stream);

/* ---- multigrid kernels (core/gmg/level.py calls these) ----------------- */

/* core/gmg/fortran_multigrid.f90:365-412 fillhalo(x,nh): periodic halo */
int f2d_fillhalo_f64(double *x, int nh, int m, int n, f2d_stream_t stream);

/* core/gmg/fortran_multigrid.f90:2-127 smoothtwicewithA(msk,A,x,b,omega,yo)
 * followed by fillhalo (core/gmg/level.py:351-365).  A: 5 planes of m*n
 * (A(:,:,1..5) of the reference).  work: m*n scratch (replaces yo). */
int f2d_smooth_f64(const int8_t *msk, const double *A, double *x, const double *b, double omega,
                   double *work, int m, int n, f2d_stream_t stream);

/* core/gmg/fortran_multigrid.f90:320-362 computeresidualwithA + fillhalo
 * (core/gmg/level.py:370-387) */
int f2d_residual_f64(const int8_t *msk, const double *A, const double *x, const double *b,
                     double *r, int m, int n, f2d_stream_t stream);

/* core/gmg/fortran_multigrid.f90:501-546 restrict(msk2,x1,nh,x2) + fillhalo
 * on the coarse grid (core/gmg/level.py:473-496) */
int f2d_restrict_f64(const int8_t *msk2, const double *x1, int nh, double *x2, int m2, int n2,
                     int m1, int n1, f2d_stream_t stream);

/* core/gmg/fortran_multigrid.f90:415-498 interpolate(msk1,msk2,x2,nh,x1)
 * (core/gmg/level.py:447-470).  accumulate=1 adds into x1 instead
 * (the x += r of core/gmg/hierarchy.py:125, fused). */
int f2d_interpolate_f64(const int8_t *msk1, const int8_t *msk2, const double *x2, int nh,
                        double *x1, int m2, int n2, int m1, int n1, int accumulate,
                        f2d_stream_t stream);

/* core/gmg/fortran_multigrid.f90:706-811 coarsenmatrix(Afine,msk1,msk2,nh)
 * + fillhalo of each plane (core/gmg/level.py:321-327).  Af, Ac: 9 planes. */
int f2d_coarsenmatrix_f64(const double *Af, const int8_t *msk1, const int8_t *msk2, int nh,
                          double *Ac, int m1, int n1, int m2, int n2, f2d_stream_t stream);

/* core/gmg/fortran_multigrid.f90:813-839 computenorm: out[0] = sum of x^2
 * over the fluid interior (device scalar; summation order differs from the
 * serial reference by rounding only). */
int f2d_norm2_f64(const int8_t *msk, const double *x, int nh, int m, int n, double *out,
                  f2d_stream_t stream);

/* core/fortran_diag.f90:155-195 computekemaxu: out = {ke, maxu} (device) */
int f2d_kemaxu_f64(const int8_t *msk, const double *u, const double *v, int nh, int m, int n,
                   double *out, f2d_stream_t stream);

/* core/fortran_diag.f90:3-30 computedotprod(msk,x,y,nh) -> z: out[0] = sum of
 * x*y over the interior where msk != 0 (device scalar; tree order, rounding
 * differs from the serial reference by rounding only).  Called by
 * core/euler.py:197-206 (momentum, enstrophy source terms). */
int f2d_dotprod_f64(const int8_t *msk, const double *x, const double *y, int nh, int m, int n,
                    double *out, f2d_stream_t stream);

/* core/fortran_diag.f90:63-90 computesum(msk,x,nh) -> y: out[0] = sum of x
 * over the interior where msk != 0 (device scalar, tree order). */
int f2d_sum_f64(const int8_t *msk, const double *x, int nh, int m, int n, double *out,
                f2d_stream_t stream);

/* core/fortran_diag.f90:93-120 computesumandnorm(msk,x,nh) -> (y, y2):
 * out[0] = sum of x, out[1] = sum of x^2 over the interior where msk == 1
 * (device scalars, tree order).  Called by core/euler.py:194-196 and
 * core/boussinesq.py:132-139. */
int f2d_sumandnorm_f64(const int8_t *msk, const double *x, int nh, int m, int n, double *out,
                       f2d_stream_t stream);

/* core/grid.py:132-140 domain_integration: np.sum over the interior in
 * NumPy's own summation order (bit-identical to NumPy 2.x); out[0] (device). */
int f2d_domain_sum_f64(const double *x, int nh, int m, int n, double *out, f2d_stream_t stream);

/* Allocation-free form of f2d_domain_sum_f64 (same kernels, same result bit
 * for bit).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its chunk partials live in a
 * library-owned device array of 64 KiB per GPU (held until the process
 * ends), so interiors of at most 8192*8192 cells are accepted; larger ones
 * return F2D_E_ARG.  It calls no hipMalloc/hipFree.  Calls on one GPU share
 * that array: calls that overlap on the same GPU (other streams or host
 * threads) race and must be serialised by the caller. */
int f2d_domain_sum_f64_noalloc(const double *x, int nh, int m, int n, double *out,
                               f2d_stream_t stream);

/* ---- multigrid hierarchy (core/gmg/hierarchy.py Gmg, one subdomain) ---- */
typedef struct f2d_gmg f2d_gmg;

/* core/gmg/hierarchy.py:23-84 + core/gmg/level.py:24-335.  mskp: level-0
 * corner mask (host, m*n int8).  npre = npost = 1, ndeepest = 16. */
int f2d_gmg_create(int nx, int ny, int nh, double dx, double dy, double hydroepsilon,
                   double omega, const int8_t *mskp_host, f2d_gmg **out);
int f2d_gmg_destroy(f2d_gmg *g);
/* core/gmg/hierarchy.py:80-84 (param 'qgoperator'): turn the Laplacian into
 * the Helmholtz operator A - 1/Rd^2: on every level, the centre coefficient
 * A5 becomes A5 - 1/Rd^2 wherever it is nonzero (halo included).  Call once,
 * after f2d_gmg_create and before any cycle. */
int f2d_gmg_set_helmholtz(f2d_gmg *g, double Rd);
int f2d_gmg_nlevels(const f2d_gmg *g, int *nlev);
/* which: 0 msk (int8), 1 A (5 planes), 2 x, 3 b, 4 r */
int f2d_gmg_level(const f2d_gmg *g, int lev, int which, void **dev_ptr, int *m, int *n);
/* core/gmg/hierarchy.py:98-126 Vcycle(lev1), :131-151 Fcycle(lev1) */
int f2d_gmg_vcycle(f2d_gmg *g, int lev1);
int f2d_gmg_fcycle(f2d_gmg *g, int lev1);
/* core/gmg/hierarchy.py:207-218 twoVcycle(x,b): x (device, in/out), b (device) */
int f2d_gmg_twovcycle(f2d_gmg *g, double *x, const double *b);
/* core/gmg/hierarchy.py:154-192 solve(x,b,{maxite,tol}) -> (nite,res);
 * returns F2D_E_DIVERGED where the reference exits. */
int f2d_gmg_solve(f2d_gmg *g, double *x, const double *b, int maxite, double tol, int *nite,
                  double *res);

/* ---- models: Euler / Boussinesq with RK3_SSP or RK3 --------------------- */
typedef struct f2d_model f2d_model;

typedef struct {
  int model;            /* 0 euler (core/euler.py), 1 boussinesq (core/boussinesq.py) */
  int nx, ny;           /* interior size; powers of two down to the coarsest level */
  double Lx, Ly;
  int order;            /* 1..6 (core/defaults.json: 5) */
  int flux_splitting;   /* 0 minmax, 1 parabolic (core/operators.py:124-131) */
  double aparab;        /* 0.05 */
  int timestepping;     /* 0 RK3_SSP, 1 RK3 (core/timescheme.py) */
  int perio;            /* geometry == 'perio': zero-mean psi after a full solve */
  int diffusion;        /* core/operators.py:292-301 */
  double kdiff_vort, kdiff_buoy;
  double gravity;
  double hydroepsilon;
} f2d_model_params;

/* Operators.__init__ + model __init__ (core/operators.py:12-155,
 * core/euler.py:19-101, core/boussinesq.py:17-85).  msk: cell mask (host). */
int f2d_model_create(const f2d_model_params *p, const int8_t *msk_host, f2d_model **out);
int f2d_model_destroy(f2d_model *M);
/* state: nvar x m x n float64 on the device (Var.state, core/variables.py:25);
 * Euler nvar=5 [vorticity,psi,u,v,source], Boussinesq 6 [..,buoyancy,banom] */
int f2d_model_state(f2d_model *M, double **state_dev, int *nvar, int *m, int *n);
/* tendency buffers dx0, dx1, dx2 of core/timescheme.py:35-41 */
int f2d_model_tendency(f2d_model *M, int k, double **dev);
int f2d_model_corner_mask(f2d_model *M, int8_t **dev);
int f2d_model_gmg(f2d_model *M, f2d_gmg **g);
/* core/euler.py:182-183 set_psi_from_vorticity (full inversion) */
int f2d_model_set_psi_from_vorticity(f2d_model *M);
/* Operators.invert_vorticity(x, flag, island) on a model-shaped device
 * array (core/operators.py:421-498); flags bit 0: 'full' (else 'fast'),
 * bit 1: island (the islands set by f2d_model_set_islands, if any) */
int f2d_model_invert_vorticity(f2d_model *M, double *x_dev, int flags);
/* No-slip walls (param.noslip): the vorticity source of
 * Operators.rhs_noslip (core/operators.py:245-290) is added after each step
 * (core/euler.py:108-112, core/boussinesq.py:92-94).  Host inputs, m*n
 * each: msknoslip and mskbc (int8) and xr0, yr0, with bcarea, x2bc, y2bc of
 * Operators.set_boundary_msk (core/operators.py:157-185);
 * enforce_momentum: the Operators' flag (param.enforce_momentum). */
int f2d_model_set_noslip(f2d_model *M, const int8_t *msknoslip_host, const int8_t *mskbc_host,
                         const double *xr0_host, const double *yr0_host, double bcarea,
                         double x2bc, double y2bc, int enforce_momentum);
/* Islands (param.isisland, core/island.py:21-43): rhs correction and
 * boundary psi (host, m*n each) used by the state's inversions. */
int f2d_model_set_islands(f2d_model *M, const double *rhsp_host, const double *psi_host);
/* Operators.rhs_adv (core/operators.py:214-236) into dxdt (device) */
int f2d_model_rhs_adv(f2d_model *M, const double *x_dev, double *dxdt_dev);
/* cst[3] (umax) used by the next advection (core/fluid2d.py:397) */
int f2d_model_set_umax(f2d_model *M, double umax);
/* model.step(t, dt) (core/euler.py:103-147, core/boussinesq.py:87-100) */
int f2d_model_step(f2d_model *M, double dt);
/* model.diagnostics raw sums (core/euler.py:185-221, core/boussinesq.py:123-153):
 * out[8] = {maxspeed, ke, sum w, sum w^2, px | pe, py | sum b, sum psi | sum b^2, source | 0}
 * xr, yr: host coordinate arrays (m*n) or NULL for the last ones given. */
int f2d_model_diagnostics(f2d_model *M, const double *xr_host, const double *yr_host, double *out);
/* Fluid2d.loop body with fixed dt (core/fluid2d.py:235-268): per step
 * cst[3] = maxspeed, step, diagnostics.  nite_log (host, nsteps ints) or NULL. */
int f2d_model_run(f2d_model *M, int nsteps, double dt, int *nite_log);
/* Fluxes.diag_fluxes (core/fluxes.py:100-170) of the model-shaped device
 * state x: reversible and irreversible advective edge fluxes over one step
 * dt (one step forward, one step -dt from the time-reversed state), into
 * flx_dev, 4*ntr planes of m*n (ntr = 3 Euler: vorticity uc vc; 4
 * Boussinesq: vorticity buoyancy uc vc) in the order of fullflx_list
 * (core/fluxes.py:45-50): rev then irr; per quantity x then y.  Uses the
 * model's umax (f2d_model_set_umax).  dt != 0. */
int f2d_model_diag_fluxes(f2d_model *M, const double *x_dev, double dt, double *flx_dev);
/* F-cycles and final relative residual of the last full solve */
int f2d_model_last_solve(f2d_model *M, int *nite, double *res);

/* ---- batched engine ------------------------------------------------------
 * Any number of independent Euler/Boussinesq instances (different sizes,
 * masks and physics; one time scheme per batch) advanced together.  Every
 * kernel launch covers all instances taking part in a phase; one step is a
 * fixed launch program without host round trips (the solve stopping test of
 * core/gmg/hierarchy.py:169 and dt live on the device), captured once into
 * a hipGraph and replayed.  Each instance performs exactly the reference's
 * operations, so its FP64 result is independent of the batch it runs in.
 */
typedef struct f2d_batch f2d_batch;
const char *f2d_batch_last_error(void);
/* Launch configuration of a batch, chosen by the autotuner
 * (fluid2d_amd/autotune.py) or left at the defaults (NULL / zeros).
 * coarse_lds_budget: bytes of LDS for the single-workgroup coarse-level
 *   kernel (0 = the device maximum); levels are handed to it from the
 *   coarsest up while they fit.
 * min_tiles: a fine-level phase uses the tallest tile (16, 8 or 4 rows of
 *   64 cells) that still gives at least this many workgroups (0 = always
 *   16 rows).
 * coarse_wave_cells: inside the coarse kernel, levels of at most this many
 *   cells (halo included) run on one wavefront without workgroup barriers;
 *   larger ones use the whole workgroup (0 = none on one wavefront;
 *   negative = the default, 256).
 * persist: > 0 runs each run of consecutive smoother and
 *   residual/restriction phases below the finest level as one persistent
 *   launch of at most `persist` workgroups per compute unit, with a bounded
 *   device-wide barrier between phases (same results bit for bit); 0 = one
 *   launch per phase.  A batch with persistent launches must not share the
 *   device with another such batch at the same time (their workgroups could
 *   not all be resident); a barrier that cannot complete gives up and the
 *   step reports F2D_E_HIP instead of hanging.
 * No reference counterpart (the reference has no launches). */
typedef struct {
  int coarse_lds_budget;
  int min_tiles;
  int coarse_wave_cells;
  int persist;
} f2d_batch_tuning;
/* precision: 0 FP64 throughout (bit-identical to the reference algorithm);
 *   1 mixed: FP64 state, RK combination, full-solve rhs, outer residual and
 *   correction (iterative refinement) and coarse-level solve; FP32
 *   advection, tendencies and fine-level V-cycles.
 * msks[i]: host cell mask of instance i; tuning may be NULL. */
int f2d_batch_create(int ninst, const f2d_model_params *params, const int8_t *const *msks,
                     int precision, const f2d_batch_tuning *tuning, f2d_batch **out);
/* new launch configuration for an existing batch (between steps) */
int f2d_batch_retune(f2d_batch *b, const f2d_batch_tuning *tuning);
int f2d_batch_tuning_info(f2d_batch *b, f2d_batch_tuning *out);
int f2d_batch_destroy(f2d_batch *b);
int f2d_batch_info(f2d_batch *b, int inst, int *nlev, int *nfine, long long *launches_per_step);
int f2d_batch_set_field(f2d_batch *b, int inst, int field, const double *host);
int f2d_batch_get_state(f2d_batch *b, int inst, double *host);
int f2d_batch_state_ptr(f2d_batch *b, int inst, double **dev);
/* fixed dt of instance inst (inst < 0: all) */
int f2d_batch_set_dt(f2d_batch *b, int inst, double dt);
/* Step budget of instance inst (inst < 0: all; nsteps < 0: unlimited).  A
 * gate at the top of every step lets an instance run while its budget
 * lasts; exhausted instances are skipped by every kernel, so instances of
 * one batch can stop at different step counts. */
int f2d_batch_set_steps(f2d_batch *b, int inst, int nsteps);
int f2d_batch_progress(f2d_batch *b, int inst, int *steps_left, long long *steps_done);
/* adaptive dt = cfl*min(dx,dy)/maxspeed capped at dtmax (core/fluid2d.py:376-397) */
int f2d_batch_set_adaptive(f2d_batch *b, int on, double cfl, double dtmax);
/* Fast (per-stage) inversions by gmg.solve(maxite, tol) instead of
 * twoVcycle: the alternative core/operators.py:452-455 keeps commented out
 * (ite, res = self.gmg.solve(psi, self.work, {'maxite': .., 'tol': ..})),
 * then psi *= mskp and (u, v) as on the fast path.  In mixed precision the
 * F-cycles run in FP32 between FP64 residuals and FP64 corrections
 * (iterative refinement to tol).  maxite = 0 restores twoVcycle (the
 * reference).  Rebuilds the step program; call between steps. */
int f2d_batch_set_fast_solve(f2d_batch *b, int maxite, double tol);
/* No-slip walls on instance inst (Operators.set_boundary_msk and
 * rhs_noslip, core/operators.py:157-185, :245-290, run after each time
 * step as core/euler.py:108-112 and core/boussinesq.py:92-94 do): the
 * wall source from psi, minus its mean over the boundary cells mskbc (and,
 * with enforce, its momentum against xr0, yr0 normalised by x2bc, y2bc),
 * filled and subtracted from the vorticity, then the fast inversion of the
 * state (with the instance's islands); Euler stores source/dt in its
 * source field.  Host arrays of (ny+6) x (nx+6); once per instance;
 * rebuilds the programs.  Same arguments as f2d_model_set_noslip. */
int f2d_batch_set_noslip(f2d_batch *b, int inst, const int8_t *msknoslip, const int8_t *mskbc,
                         const double *xr0, const double *yr0, double bcarea, double x2bc,
                         double y2bc, int enforce);
/* Islands on instance inst (core/island.py:21-43, core/fluid2d.py:110-113):
 * the state's inversions (the full solve and the no-slip fast inversion)
 * subtract rhsp from their right-hand side and add psi after the mask
 * (core/operators.py:444-446, :483-489).  Once per instance. */
int f2d_batch_set_islands(f2d_batch *b, int inst, const double *rhsp, const double *psi);
/* the corner mask mskp of instance inst, (ny+6) x (nx+6) bytes (the
 * operator-level f2d_model_corner_mask; Operators.mskp of
 * core/operators.py:62-67) */
int f2d_batch_corner_mask(f2d_batch *b, int inst, int8_t *host);
/* The loop's momentum enforcement (Fluid2d.enforce_zero_momentum,
 * core/fluid2d.py:352-358): set_momentum enables it with the centred
 * coordinates xr0, yr0; enforce_momentum then subtracts xr0 px + yr0 py
 * from the vorticity (px = diags px / x2, py = diags py / y2, from the
 * host), runs the fast inversion of the state and refreshes the
 * diagnostics (synchronous). */
int f2d_batch_set_momentum(f2d_batch *b, int inst, const double *xr0, const double *yr0);
int f2d_batch_enforce_momentum(f2d_batch *b, int inst, double px, double py);
int f2d_batch_use_graph(f2d_batch *b, int on);
/* set_psi_from_vorticity of every instance, then diagnostics (maxspeed) */
int f2d_batch_set_psi_from_vorticity(f2d_batch *b);
/* Continuous batching (no reference counterpart: the reference runs one
 * model per process).  reset_instance: instance `inst` becomes a fresh slot
 * -- every buffer that carries values between steps is zeroed and its
 * device record reset as at creation; operators, masks and the other
 * instances are untouched (synchronises the batch's stream).
 * init_instance: set_psi_from_vorticity (core/euler.py / boussinesq.py,
 * the full solve + diagnostics) of that instance alone, enqueued on the
 * batch's stream.  With set_field, set_dt and set_steps in between, a
 * request admitted into a running batch computes exactly what it computes
 * in a batch of its own. */
int f2d_batch_reset_instance(f2d_batch *b, int inst);
int f2d_batch_init_instance(f2d_batch *b, int inst);
/* diagnostics (and the next step's umax, core/fluid2d.py:397) of every
 * instance's current state, e.g. after a host write of the state
 * (Euler.diagnostics, core/euler.py:185-221) */
int f2d_batch_refresh_diagnostics(f2d_batch *b);
/* nsteps x (RK step + diagnostics), synchronous / enqueue only */
int f2d_batch_step(f2d_batch *b, int nsteps);
int f2d_batch_step_async(f2d_batch *b, int nsteps);
int f2d_batch_synchronize(f2d_batch *b);
/* done = 1 when all work enqueued on the batch's stream has completed */
int f2d_batch_query(f2d_batch *b, int *done);
/* Runs nsteps steps eagerly; every launch of kind (0 smoother, 1
 * residual/restriction, 2 advection, 3 coarse levels, 4 persistent
 * multigrid, 5 fused down-leg, 6 two-level up-leg) at top-down level
 * `level` (-1: any) is issued 8 times back to back between a pair of HIP
 * events (the repeats leave the state as one launch does).  avg_ms: mean
 * duration per launch (elapsed / 8); alg_bytes / impl_bytes: mean
 * algorithmic / implemented bytes per launch. */
int f2d_batch_profile(f2d_batch *b, int kind, int level, int nsteps, double *avg_ms,
                      long long *nlaunch, double *alg_bytes, double *impl_bytes);
/* out8 as f2d_model_diagnostics (raw sums; maxspeed and ke scaled) */
int f2d_batch_diagnostics(f2d_batch *b, int inst, double *out8);
int f2d_batch_solve_info(f2d_batch *b, int inst, int *nite, double *res, int *diverged);

#ifdef __cplusplus
}
#endif
#endif
