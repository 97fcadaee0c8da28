/*
 * psisolve.h -- C ABI of the MI355X-native barotropic streamfunction solver.
 *
 * Drop-in replacement for pyOM2's streamfunction path.  Every entry point
 * names the reference routine it replaces (paths relative to the pyOM2
 * source tree).  Conventions:
 *
 *  - 2-D fields are caller-owned DEVICE arrays of double (or the stated
 *    type), C-contiguous with shape (ny+4, nx+4): the element a(i,j) of
 *    pyOM2's Fortran array a(-1:nx+2,-1:ny+2) (halo onx=2) lives at
 *    a[(j+1)*(nx+4) + (i+1)].  1-D metric arrays are HOST arrays with
 *    nx+4 (x) or ny+4 (y) entries in the same offset convention.
 *  - Every function returns a psi_status; the process never exits.  A
 *    context's last error text is available from psi_ctx_last_error().
 *  - `stream` arguments are hipStream_t values passed as void* (NULL is
 *    the null stream).  A context is bound to one stream and one device;
 *    contexts are independent of each other (the reference's SAVE'd
 *    coefficient cache becomes per-context state) but one context must not
 *    be used from two threads at once.
 *  - psi_congrad, psi_solve_streamfunction and the 3-D entry points
 *    allocate no device memory: all scratch is allocated by psi_ctx_create.
 */
#ifndef PSISOLVE_H
#define PSISOLVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum psi_status {
    PSI_OK = 0,
    PSI_NOT_CONVERGED = 1,          /* max iterations reached (converged=.false.) */
    PSI_DIVERGED = 2,               /* divergence test fired (converged=.false.) */
    PSI_NAN = 3,                    /* NaN estimate: the reference would halt */
    PSI_SINGULAR_ISLAND_SYSTEM = 4, /* DGESV info != 0: the reference would halt */
    PSI_BAD_ARG = 5,
    PSI_HIP_ERROR = 6,
    PSI_TRACE_FAILED = 7,           /* island perimeter walk lost track / no start */
    PSI_NO_LEARNED_MODEL = 8
} psi_status;

typedef enum psi_mode {
    PSI_MODE_PLAIN = 0,    /* FP64 Jacobi-PCG */
    PSI_MODE_MIXED = 1,    /* FP32 inner Jacobi-PCG under FP64 iterative refinement */
    PSI_MODE_LEARNED = 2,  /* flexible CG with the learned preconditioner (FP32 inner) */
    PSI_MODE_POLY = 3,     /* as MIXED, preconditioned by a Chebyshev polynomial of degree
                              poly_degree in Z A (Z the Jacobi inverse) times Z */
    PSI_MODE_MG = 4        /* FP64 flexible CG preconditioned by an FP32 aggregation-multigrid
                              V-cycle (psi_mg_setup first); PSI_CRIT_RELRES only.  poly_degree
                              = Chebyshev smoothing steps per level (<= 0: 5; 4 for the
                              surface pressure), poly_lower = the smoother's interval
                              [2 poly_lower, 2] (<= 0: 1/15) */
} psi_mode;

typedef enum psi_criterion {
    /* reference semantics: absolute bound congr_epsilon on the extrapolated
     * max-norm step (solve_streamfct.f90:288-297) */
    PSI_CRIT_REFERENCE = 0,
    /* ||forc - A sol||_2 / ||forc||_2 over active points below relres_target,
     * confirmed on the FP64 true residual */
    PSI_CRIT_RELRES = 1
} psi_criterion;

typedef struct psi_grid {
    int32_t nx, ny;        /* interior points */
    int32_t cyclic_x;      /* enable_cyclic_x (cyclic y is rejected, as in
                              streamfct_init.f90:27-30) */
    const double *dxt, *dxu;               /* host, nx+4 */
    const double *dyt, *dyu, *cost, *cosu; /* host, ny+4 */
} psi_grid;

/* Device fields describing the basin.  maskU/maskV (surface level) are only
 * needed by the island integrals and may be NULL when nisle <= 1. */
typedef struct psi_fields {
    const double *hur, *hvr;      /* device (ny+4, nx+4) */
    const double *maskU, *maskV;  /* device (ny+4, nx+4) or NULL */
} psi_fields;

/* Island boundary lists (streamfct_init.f90:87-93, 258): island k (0-based,
 * the reference's isle = k+1) owns entries offsets[k]..offsets[k+1]-1. */
typedef struct psi_islands {
    int32_t nisle;
    const int32_t *offsets;        /* host, nisle+1 */
    const int32_t *bi, *bj;        /* host, boundary point (Fortran i, j) */
    const int32_t *di, *dj;        /* host, line direction */
} psi_islands;

typedef struct psi_opts {
    int32_t mode;              /* psi_mode */
    int32_t criterion;         /* psi_criterion */
    double congr_epsilon;      /* reference: 1e-12 */
    int32_t max_iterations;    /* congr_max_iterations (per solve, inner iterations) */
    double relres_target;      /* PSI_CRIT_RELRES target, e.g. 1e-10 */
    int32_t check_every;       /* iterations between host polls of the device
                                  convergence flag; <= 0 picks a default */
    double inner_reduction;    /* mixed/learned: relative residual reduction per
                                  FP32 inner solve; <= 0 picks the default 0.25 */
    int32_t max_outer;         /* mixed/learned: FP64 refinement steps; <= 0 default */
    int32_t poly_degree;       /* PSI_MODE_POLY: 1..6; <= 0 picks 4 */
    double poly_lower;         /* PSI_MODE_POLY: the polynomial's interval is
                                  [2 poly_lower, 2]; <= 0 picks 0.005 */
    int32_t leg_max_iterations;/* mixed/poly: inner iterations per FP32 leg; <= 0: no cap */
    int32_t time_kernels;      /* PSI_MODE_MG: nonzero records HIP events around the
                                  four heaviest kernels of one outer iteration per solve
                                  (the last of the first poll batch) and reports their
                                  mean durations in psi_info.kernel_us */
    int32_t mg_coarse_degree;  /* PSI_MODE_MG: Chebyshev smoothing steps on the coarse
                                  levels (at most 256 tiles of 64 x 16 cells, whose
                                  kernels are bound by launch latency); 1..10, <= 0
                                  picks the library default */
    int32_t reserved;
} psi_opts;

typedef struct psi_info {
    int32_t iterations;        /* congr_itts (total inner iterations for mixed) */
    int32_t converged;         /* logical converged */
    int32_t status;            /* psi_status of the solve */
    int32_t outer_iterations;  /* FP64 refinement steps (0 for plain) */
    double est_error;          /* reference estimated_error */
    double relres_recursive;   /* recursively updated relative residual */
    double relres_true;        /* FP64 ||forc - A sol|| / ||forc|| (active points) */
    double elapsed_ms;         /* device time of the solve (HIP events) */
    double sweep_ms;           /* device time spanned by the fused PCG sweeps */
    int32_t sweep_launches;    /* sweep kernels launched (incl. ones that found
                                  the solve finished and returned at once) */
    int32_t setup_launches;    /* setup + residual kernels launched */
    int32_t kernel_samples;    /* outer iterations timed (psi_opts.time_kernels) */
    double kernel_us[4];       /* their mean durations (us) of the finest level's
                                  pre-smoothing, post-smoothing, pass A, pass B */
} psi_info;

typedef struct psi_ctx psi_ctx;

/* ---- island setup (host) ----------------------------------------------
 * Replaces the island part of streamfunction_init (streamfct_init.f90:35-267)
 * with isleperim/expand (island_module.f90:7-273): labels the 8-connected
 * land masses of kbot (host int32 (ny+4, nx+4), 0 = land) and walks each
 * perimeter.  Output buffers are caller-owned host arrays: offsets needs
 * nisle+1 entries (cap_isles), the point/direction arrays cap_points
 * entries.  When a cap is too small, returns PSI_BAD_ARG with *nisle and
 * *npoints set to the sizes needed.  No fixed limits (the reference's
 * maxipp/mnisle/maxq) apply. */
int psi_islands_trace(int32_t nx, int32_t ny, int32_t cyclic_x, const int32_t *kbot,
                      int32_t *nisle, int32_t *offsets, int32_t cap_isles,
                      int32_t *bi, int32_t *bj, int32_t *di, int32_t *dj,
                      int32_t cap_points, int32_t *npoints);

/* ---- context ------------------------------------------------------------
 * Builds the cached operator (make_coeff_streamfunction,
 * solve_streamfct.f90:123-148, cached at :172-180) and the Jacobi inverse
 * with its Dirichlet set (make_inv_sfc, :381-422), and allocates all
 * device scratch.  `islands` may be NULL (no Dirichlet points). */
int psi_ctx_create(psi_ctx **out, const psi_grid *grid, const psi_fields *fields,
                   const psi_islands *islands, void *stream);
/* Re-derive the cached coefficients from new hur/hvr (the reference's cache
 * never invalidates; this is the explicit invalidation). */
int psi_ctx_update_coeffs(psi_ctx *ctx, const double *hur, const double *hvr);
int psi_ctx_set_stream(psi_ctx *ctx, void *stream);
void psi_ctx_destroy(psi_ctx *ctx);
const char *psi_ctx_last_error(const psi_ctx *ctx);
int psi_ctx_nisle(const psi_ctx *ctx);
/* Bytes of device memory the context holds. */
int64_t psi_ctx_device_bytes(const psi_ctx *ctx);

/* ---- the solve ----------------------------------------------------------
 * congrad_streamfunction(is_,ie_,js_,je_, forc, iterations, sol, converged)
 * (solve_streamfct.f90:153-327): solves A sol = forc with the Jacobi
 * preconditioner; sol is the first guess on entry (its cyclic halo is set
 * first, :191-192) and the solution on exit.  info receives iterations,
 * converged and the estimates.  Returns PSI_OK, PSI_NOT_CONVERGED,
 * PSI_DIVERGED or PSI_NAN (the last three with info->converged = 0). */
int psi_congrad(psi_ctx *ctx, const psi_opts *opts, const double *forc, double *sol,
                psi_info *info);

/* ---- the plain solve without a context and without allocating -----------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (a serving loop, a harness that forbids hipMalloc, a program that
 * captures HIP graphs around its other work).  psi_congrad_noalloc runs
 * psi_congrad's plain FP64 Jacobi-PCG (mode PSI_MODE_PLAIN, either
 * criterion; the same kernels, so the same iterations and solution) on the
 * basin described by grid / fields / islands, without a context: it calls no
 * hipMalloc, hipFree, hipHostMalloc or event creation.  Its operator, active
 * mask, CG vectors, partials and solver state live in one static device
 * array of 19.5 MiB per GPU (held until the process ends); the first call on
 * a device looks up that array's address there.  Limits: (nx+4)(ny+4) <=
 * 516^2 (a 512 x 512 interior) with nx, ny <= 2048, and at most 2^18 island
 * perimeter points; larger inputs return PSI_BAD_ARG with the reason in
 * psi_ctx_last_error(NULL).  fields->maskU/maskV are not read.  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU (two threads,
 * two streams) race.  The call returns after the solve has finished on
 * `stream`; info->elapsed_ms is host wall time, info->sweep_ms is -1. */
int psi_congrad_noalloc(const psi_grid *grid, const psi_fields *fields,
                        const psi_islands *islands, const psi_opts *opts, const double *forc,
                        double *sol, psi_info *info, void *stream);

/* ---- building blocks, exposed for the drop-in and for parity tests ----- */
/* apply_op (solve_pressure.f90:254-274) with the cached operator: res = A p
 * on the interior, res = 0 on the halo. */
int psi_apply_op(psi_ctx *ctx, const double *p, double *res);
/* make_coeff_streamfunction: the 9 coefficient planes cf(:,:,ii+2,jj+2),
 * plane (ii+1)+3*(jj+1), each (ny+4, nx+4), into cf (device). */
int psi_make_coeff(psi_ctx *ctx, double *cf);
/* make_inv_sfc: Z (device (ny+4, nx+4)). */
int psi_make_inv(psi_ctx *ctx, double *Z);
/* dot_sfc (solve_streamfct.f90:348-362) and absmax_sfc (:331-345) over the
 * interior; results to host. */
int psi_dot(psi_ctx *ctx, const double *a, const double *b, double *out);
int psi_absmax(psi_ctx *ctx, const double *a, double *out);
/* setcyclic_xy (parallel_none.f90:188-211), in place; no-op unless cyclic_x. */
int psi_setcyclic(psi_ctx *ctx, double *a);
/* line_integral (streamfct_init.f90:311-377) around island isle (1-based). */
int psi_line_integral(psi_ctx *ctx, int32_t isle, const double *uloc, const double *vloc,
                      double *line);
/* Barotropic velocity of a streamfunction for the island integrals
 * (streamfct_init.f90:294-300, solve_streamfct.f90:68-74). */
int psi_psi_velocity(psi_ctx *ctx, const double *psi, double *fpx, double *fpy);
/* FP64 true relative residual ||forc - A sol||_2/||forc||_2 over active
 * points (the figure psi_info.relres_true reports). */
int psi_true_relres(psi_ctx *ctx, const double *forc, const double *sol, double *relres);

/* ---- island machinery ---------------------------------------------------
 * streamfunction_init's solver part (streamfct_init.f90:269-303): for every
 * island, psin = 1 on its perimeter, congrad with zero forcing; then
 * line_psin(n, isle) = line integral around n of the velocity of psin(isle).
 * psin: device, nisle planes of (ny+4, nx+4); line_psin: host nisle*nisle,
 * row-major [n][isle]; iterations: host nisle entries (may be NULL). */
int psi_streamfunction_init(psi_ctx *ctx, const psi_opts *opts, double *psin,
                            double *line_psin, int32_t *iterations);

/* The 2-D body of solve_streamfunction (solve_streamfct.f90:39-94) after the
 * depth integration: given fpx, fpy (depth-integrated forcing times hur/hvr,
 * device), it forms forc = curl (:49-53), warm-starts
 * dpsi(taup1) = 2 dpsi(tau) - dpsi(taum1) (:56), solves (:57), sets the
 * cyclic halo (:58-59), solves the island system for dpsin(2:nisle, tau)
 * (:61-88) and steps psi(taup1) with Adams-Bashforth (:90-94).
 * psi, dpsi: device arrays of 3 time levels, each (ny+4, nx+4), level
 * index 0..2 (the reference's 1..3).  dpsin: host nisle x 3, row-major
 * [isle][level].  psin: device nisle planes; line_psin: host [n][isle].
 * forc_out (device, may be NULL) receives the assembled forcing. */
int psi_solve_streamfunction(psi_ctx *ctx, const psi_opts *opts, const double *fpx,
                             const double *fpy, double *psi, double *dpsi,
                             double *dpsin, const double *psin, const double *line_psin,
                             int32_t taum1, int32_t tau, int32_t taup1, double dt_mom,
                             double AB_eps, double *forc_out, psi_info *info);

/* ---- the 3-D parts of solve_streamfunction --------------------------------
 * A 3-D field a(i,j,k) is nz consecutive (ny+4, nx+4) planes (level k = 1,
 * the bottom, first): the Fortran a(-1:nx+2, -1:ny+2, nz) in place.  A
 * field with time levels, a(:,:,:,1:3), is 3 such blocks.  Fields are
 * device arrays; dzt/dzw are host arrays like the other 1-D metrics.  nz is
 * at most 128. */
typedef struct psi_levels {
    int32_t nz;
    int32_t reserved;
    const double *dzt, *dzw;               /* host, nz entries (main_module.f90:108) */
    const double *maskT, *maskU, *maskV;   /* nz planes each (:96-98) */
} psi_levels;

/* Hydrostatic pressure, its gradient added to du/dv and the depth-integrated
 * forcing (solve_streamfct.f90:22-48): p_hydro from rho (tau level), du/dv
 * (tau level, updated in place on the interior), fpx = hur sum_k (du+du_mix)
 * maskU dzt, fpy likewise, with their cyclic halo.  fxa = grav/rho_0
 * (main_module.f90:18-19: pass 9.81 and 1024 as the reference's real*8
 * parameters hold them). */
int psi_streamfunction_forcing(psi_ctx *ctx, const psi_levels *lev, double grav, double rho_0,
                               const double *rho, double *du, double *dv,
                               const double *du_mix, const double *dv_mix, double *p_hydro,
                               double *fpx, double *fpy);

/* Barotropic and baroclinic velocity at taup1 (solve_streamfct.f90:95-118;
 * the open-boundary call set_obc_streamfct, :101, is not part of it):
 * Adams-Bashforth step of u, v from du, dv (tau, taum1) and du_mix, dv_mix,
 * removal of the vertical mean, then the barotropic velocity of psi(taup1)
 * added on the interior.  Outputs u_taup1, v_taup1 (nz planes each). */
int psi_velocity_update(psi_ctx *ctx, const psi_levels *lev, double dt_mom, double AB_eps,
                        const double *u_tau, const double *v_tau, const double *du_tau,
                        const double *dv_tau, const double *du_taum1, const double *dv_taum1,
                        const double *du_mix, const double *dv_mix, const double *psi_taup1,
                        double *u_taup1, double *v_taup1);

/* The model state solve_streamfunction reads and writes (main_module.f90:
 * rho :127, u/du and v/dv :143-144, du_mix/dv_mix :148-149, p_hydro :154). */
typedef struct psi_momentum {
    const double *rho;             /* 3 time levels x nz planes */
    double *u, *v;                 /* 3 x nz planes */
    double *du, *dv;               /* 3 x nz planes */
    const double *du_mix, *dv_mix; /* nz planes */
    double *p_hydro;               /* nz planes */
} psi_momentum;

/* The whole solve_streamfunction (solve_streamfct.f90:10-120, serial, without
 * open boundaries): psi_streamfunction_forcing, psi_solve_streamfunction and
 * psi_velocity_update in sequence on the context's stream, with fpx/fpy in
 * context scratch.  Returns the solve's status; on PSI_NAN (where the
 * reference halts) the velocities are not updated. */
int psi_solve_streamfunction_3d(psi_ctx *ctx, const psi_opts *opts, const psi_levels *lev,
                                const psi_momentum *mom, double grav, double rho_0,
                                double *psi, double *dpsi, double *dpsin, const double *psin,
                                const double *line_psin, int32_t taum1, int32_t tau,
                                int32_t taup1, double dt_mom, double AB_eps, double *forc_out,
                                psi_info *info);

/* ---- multigrid preconditioner (no counterpart in the reference) ------------
 * Builds the aggregation-multigrid hierarchy of the context's operator on the
 * host (2 x 2 Galerkin coarsening of the area-weighted operator down to at
 * most 1024 active cells, whose dense inverse is formed) and allocates its
 * device memory.  Needed once before psi_congrad with PSI_MODE_MG; undone by
 * psi_ctx_update_coeffs.  psi_mg_levels reports the levels (grid size and
 * active cells per level, coarsest last) into caller arrays of cap entries. */
int psi_mg_setup(psi_ctx *ctx);
int psi_mg_levels(const psi_ctx *ctx, int32_t *nlevels, int32_t *nx, int32_t *ny,
                  int32_t *nactive, int32_t cap);
/* Performance report: average duration in microseconds of `reps` launches of
 * each of the solve's heaviest kernels, timed with HIP events on the solve's
 * stream: us[0] prolongation + post-smoothing on the finest level, us[1]
 * pre-smoothing + restriction on the finest level, us[2] and us[3] the two
 * passes of the outer iteration (n >= 4).  Their vectors are scratch: call
 * between solves. */
int psi_mg_kernel_times(psi_ctx *ctx, int32_t degree, int32_t reps, double *us, int32_t n);

/* ---- the surface-pressure solver (for_src/external/solve_pressure.f90) -----
 * pyOM2's external mode when enable_streamfunction is off (momentum.f90:
 * 158-166): A p = forc with A = 1/cos^2 ((h p_x)_x + cos (cos h p_y)_y) (minus
 * 1/(grav dt^2) with a free surface), solved by conjugate gradients without a
 * preconditioner.  Its own context (a grid, no islands); every field is a
 * device (ny+4, nx+4) array or nz such planes as in psi_levels. */
typedef struct psi_press psi_press;

typedef struct psi_press_fields {
    const double *hu, *hv;     /* device: depth at u and v points (main_module.f90:99) */
    const double *maskT;       /* device: maskT(:,:,nz), values 0 or 1 */
    int32_t free_surface;      /* enable_free_surface */
    int32_t reserved;
    double grav, dt_mom;       /* enter the free-surface term, fixed at creation as the
                                  reference's SAVE'd coefficients are at its first call */
} psi_press_fields;

/* make_coeff_surf_press (solve_pressure.f90:87-152) on the device, once. */
int psi_press_create(psi_press **out, const psi_grid *grid, const psi_press_fields *fields,
                     void *stream);
void psi_press_destroy(psi_press *ctx);
int psi_press_set_stream(psi_press *ctx, void *stream);
const char *psi_press_last_error(const psi_press *ctx);
int64_t psi_press_device_bytes(const psi_press *ctx);
/* The coefficients as the reference's cf(:,:,3,3): 9 device planes, plane
 * (ii+1) + 3 (jj+1) holding cf(:,:,ii+2,jj+2). */
int psi_make_coeff_surf_press(psi_press *ctx, double *cf);
/* apply_op (solve_pressure.f90:254-274) with these coefficients: res = A p on
 * the interior, zero on the halo. */
int psi_press_apply_op(psi_press *ctx, const double *p, double *res);
/* congrad_surf_press (solve_pressure.f90:157-249): psi (device, in place) is
 * the first guess and the result; the reference's tests in its order (p.Ap = 0,
 * divergence of r.r, the extrapolated step bound congr_epsilon, NaN, the
 * iteration cap).  opts: mode PSI_MODE_PLAIN, criterion PSI_CRIT_REFERENCE,
 * congr_epsilon, max_iterations, check_every.  info->iterations is the
 * reference's congr_itts; relres_true is not computed (-1).  Returns PSI_OK,
 * PSI_NOT_CONVERGED, PSI_DIVERGED or PSI_NAN.
 * Also PSI_MODE_MG with PSI_CRIT_RELRES (no counterpart in the reference,
 * free surface only): flexible CG in FP64 preconditioned by the aggregation
 * multigrid V-cycle of PSI_MODE_MG on S = W A, W the T-cell area, to the FP64
 * true relative residual relres_target (poly_degree, poly_lower and
 * check_every as for psi_congrad); the hierarchy is built on first use or by
 * psi_press_mg_setup.  A cyclic solution gets its halo set. */
int psi_congrad_surf_press(psi_press *ctx, const psi_opts *opts, const double *forc, double *psi,
                           psi_info *info);
/* Builds the multigrid hierarchy of the surface-pressure operator (no
 * counterpart in the reference; free surface only, PSI_BAD_ARG with a rigid
 * lid, whose operator is singular). */
int psi_press_mg_setup(psi_press *ctx);
/* Performance report of the multigrid mode, as psi_mg_kernel_times (us[0]
 * post-smoothing and us[1] pre-smoothing of the finest level, us[2] and us[3]
 * the outer passes; n >= 4); after psi_press_mg_setup, between solves. */
int psi_press_mg_kernel_times(psi_press *ctx, int32_t degree, int32_t reps, double *us,
                              int32_t n);
/* The whole solve_pressure (solve_pressure.f90:9-84, serial): hydrostatic
 * pressure into p_hydro and its gradient into du/dv (tau), u/v at taup1 by
 * Adams-Bashforth, the depth-integrated transport's divergence as forcing
 * (into forc_out when given), the first guess psi(taup1) = 2 psi(tau) -
 * psi(taum1), congrad_surf_press, setcyclic_xy, and the surface-pressure
 * gradient removed from u/v at taup1.  psi: 3 planes (the surface pressure).
 * Returns the solve's status; on PSI_NAN (where the reference halts) the
 * velocities are not corrected. */
int psi_solve_pressure(psi_press *ctx, const psi_opts *opts, const psi_levels *lev,
                       const psi_momentum *mom, double grav, double rho_0, double *psi,
                       int32_t taum1, int32_t tau, int32_t taup1, double dt_mom, double AB_eps,
                       double *forc_out, psi_info *info);
/* Performance report: average duration in microseconds of `reps` launches of
 * the iteration's kernels (us[0] direction + apply, us[1] update, us[2] and
 * us[3] the two folds; n >= 4), HIP events; call between solves. */
int psi_press_kernel_times(psi_press *ctx, int32_t reps, double *us, int32_t n);

/* ---- the non-hydrostatic pressure solver (for_src/non_hydrostatic/
 * non_hydrostatic.f90) ---------------------------------------------------------
 * Used when enable_hydrostatic is off (momentum.f90:167): the 3-D Poisson
 * problem of make_coeff_non_hydro (7-point, Neumann through the masks) solved
 * by congrad_non_hydro (conjugate gradients without a preconditioner).  3-D
 * fields are nz device planes as in psi_levels; fields with time levels are
 * 3 such blocks. */
typedef struct psi_nh psi_nh;

/* The context of one basin: grid, levels (dzt, dzw; maskT, maskU, maskV) and
 * maskW (nz device planes, main_module.f90:96-99).  Every mask value must be 0
 * or 1; nz >= 2 (make_coeff_non_hydro, non_hydrostatic.f90:72-141, reads
 * maskW(:,:,k-1) and dzw(k-1) at k = nz).  The coefficients are rebuilt in the
 * kernels from the masks and 1-D metric factors, as the reference's SAVE'd cf
 * holds them (:159-163). */
int psi_nh_create(psi_nh **out, const psi_grid *grid, const psi_levels *lev,
                  const double *maskW, void *stream);
void psi_nh_destroy(psi_nh *ctx);
int psi_nh_set_stream(psi_nh *ctx, void *stream);
const char *psi_nh_last_error(const psi_nh *ctx);
int64_t psi_nh_device_bytes(const psi_nh *ctx);
/* The coefficients: 7 blocks of nz device planes, in apply_op_3D's order —
 * cf(:,:,:,2,2,1) (below), (2,1,2) (south), (1,2,2) (west), (2,2,2) (centre),
 * (3,2,2) (east), (2,3,2) (north), (2,2,3) (above); the other 20 entries of
 * the reference's cf(:,:,:,3,3,3) are zero. */
int psi_make_coeff_non_hydro(psi_nh *ctx, double *cf);
/* apply_op_3D (non_hydrostatic.f90:247-270): res = A p on the interior of
 * every level, zero elsewhere. */
int psi_apply_op_3d(psi_nh *ctx, const double *p, double *res);
/* congrad_non_hydro (non_hydrostatic.f90:145-242): p (nz device planes, in
 * place) is the first guess and the result; the reference's tests in its
 * order (p.Ap = 0, divergence of r.r, the extrapolated step bound
 * congr_epsilon_non_hydro, the iteration cap; no NaN test).  opts: mode
 * PSI_MODE_PLAIN, criterion PSI_CRIT_REFERENCE, congr_epsilon (the
 * reference's congr_epsilon_non_hydro), max_iterations
 * (congr_max_itts_non_hydro), check_every.  A zero initial residual returns 0
 * iterations.  Returns PSI_OK, PSI_NOT_CONVERGED or PSI_DIVERGED. */
int psi_congrad_non_hydro(psi_nh *ctx, const psi_opts *opts, const double *forc, double *p,
                          psi_info *info);
/* The whole solve_non_hydrostatic (non_hydrostatic.f90:6-64, serial): w at
 * taup1 by Adams-Bashforth (k < nz), setcyclic_xyz of u, v, w at taup1 (in
 * place, as the reference), the forcing (into forc_out when given), the
 * first guess p(taup1) = 2 p(tau) - p(taum1), congrad_non_hydro,
 * setcyclic_xyz, at itt = 0 p(taup1) copied to tau and taum1, and the
 * pressure gradient removed from u, v, w at taup1.  u, v, w, dw, p_nh: 3 x nz
 * planes; dw_mix: nz planes.  Returns the solve's status. */
int psi_solve_non_hydrostatic(psi_nh *ctx, const psi_opts *opts, double *u, double *v,
                              double *w, const double *dw, const double *dw_mix, double *p_nh,
                              int32_t taum1, int32_t tau, int32_t taup1, double dt_mom,
                              double AB_eps, int32_t itt, double *forc_out, psi_info *info);
/* Performance report: average duration in microseconds of `reps` launches of
 * the iteration's kernels (us[0] the direction pair, us[1] update, us[2] and
 * us[3] the two folds, and with n >= 6 us[4], us[5] the two direction
 * kernels alone), HIP events; call between solves. */
int psi_nh_kernel_times(psi_nh *ctx, int32_t reps, double *us, int32_t n);

/* ---- flexible CG with the learned preconditioner ---------------------------
 * No counterpart in the reference.  The host applies the network of psinn.h
 * to the current residual between psi_fcg_input_scale and psi_fcg_step.
 * FP32 vectors, FP64 area-weighted inner products (S = W A is symmetric),
 * FP64 residual replacement in psi_fcg_flush. */
/* Z (fp32 Jacobi inverse, 0 off active points) and the inverse depth at psi
 * points, both device (ny+4, nx+4): the network's coefficient inputs. */
int psi_fcg_prepare(psi_ctx *ctx, float *Z32, float *hinv32);
/* r = forc - A sol in FP64, scaled to FP32; restart the directions.
 * *relres = ||r|| / ||forc|| (host sync). */
int psi_fcg_begin(psi_ctx *ctx, const double *forc, const double *sol, double *relres);
/* Device scalar 1/rms(Z r) for the network input; returns the FP32 residual
 * and the scalar's device addresses (no host sync). */
int psi_fcg_input_scale(psi_ctx *ctx, const float *Z32, const float **r32, const double **scale);
/* One flexible-CG step with preconditioned residual z (fp32, network layout
 * H x W, solver grid in its top-left corner); no host sync. */
int psi_fcg_step(psi_ctx *ctx, const float *znet, int H, int W);
/* The same step with the network composed with the multigrid V-cycle
 * (psi_mg_setup first): the preconditioned residual is z = c n + V(r - c A n),
 * n the network's output znet, c = (n, W r) / (n, S n) its A-optimal scale,
 * V one V-cycle (5 Chebyshev steps per level) on what the network leaves; no
 * host sync. */
int psi_fcg_step_mg(psi_ctx *ctx, const float *znet, int H, int W);
/* Recursively updated relative residual (host sync). */
int psi_fcg_read(psi_ctx *ctx, double *relres_recursive);
/* sol += correction; r = forc - A sol in FP64 (residual replacement);
 * *relres = true relative residual; directions restart. */
int psi_fcg_flush(psi_ctx *ctx, const double *forc, double *sol, double *relres);
/* The ten flexible-CG device scalars (diagnostics; host sync). */
int psi_fcg_scalars(psi_ctx *ctx, double *out10);

/* Scalars of the PSI_MODE_POLY preconditioner (no reference counterpart):
 * the Chebyshev iteration for Z A e = Z r on [2 lower, 2] of the given degree
 * (1..6), applied as d_0 = Z r * out[0]; d_j = out[2j-1] d_{j-1} +
 * out[2j] res_{j-1}; res_j = res_{j-1} - Z A d_j; z = sum_j d_j.  out holds
 * 1 + 2 degree floats. */
int psi_poly_coefficients(int32_t degree, double lower, float *out);

/* Training support of the learned preconditioner (no reference counterpart):
 * A-norm cosine loss L = mean_b 1 - (z.Wr)^2 / ((z.Sz)(e.Wr)) of the network
 * output z (fp32, B x H x W network layout) for residuals R (B FP64 planes,
 * solver layout) with ER_b = e_b . W r_b; writes dL/dz to gz and L to *loss
 * (device).  work: 2B device doubles. */
int psi_nn_anorm_loss(psi_ctx *ctx, const float *znet, int H, int W, int B, const double *R,
                      const double *ER, double *work, float *gz, double *loss);
/* Area-weighted interior product a . W b of two device fields (host result). */
int psi_nn_weighted_dot(psi_ctx *ctx, const double *a, const double *b, double *out);

#ifdef __cplusplus
}
#endif
#endif /* PSISOLVE_H */
