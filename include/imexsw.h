/*
 * imexsw.h -- C-ABI of the MI355X (gfx950) implementation of the IMEX_SfloW2D
 * per-timestep hot path: limited reconstruction + central-upwind (Kurganov-
 * Tadmor) fluxes, IMEX Runge-Kutta stages, the per-cell implicit Newton solve
 * (complex-step Jacobian, line search) for Voellmy friction, and the explicit
 * erosion/deposition and post-step diagnostics epilogues.
 *
 * The boundary replaces the public procedures and public state of the
 * reference Fortran module `solver_2d` (src/solver_2d.f90:12-247), which the
 * unchanged main program calls once per timestep (src/IMEX_SfloW2D.f90:337-531).
 * Every entry point below names the reference routine it stands in for.
 *
 * Conventions
 *  - All arrays crossing the ABI are host arrays in the reference's Fortran
 *    column-major layout: q(var, j, k) with the variable index fastest, then
 *    x (j = 1..nx), then y (k = 1..ny) (src/solver_2d.f90:276-282).  Static
 *    per-cell fields are (j, k); staggered x-fields are (nx+1, ny), y-fields
 *    (nx, ny+1).  Device-side layout is private to the library.
 *  - Everything is FP64 (wp = dp, src/parameters_2d.f90:15).
 *  - Every call returns int: 0 = OK, < 0 = error (message via
 *    imexsw_last_error), > 0 = count of numerical events (NaN / negative
 *    thickness / non-converged Newton cells) that the reference would have
 *    reported by printing and blocking on stdin (src/solver_2d.f90:1379-1471).
 *    Nothing crosses the ABI as an exception and no call ever blocks on stdin.
 *  - One host thread per context; calls on one context are not re-entrant.
 *    Many contexts per process are allowed.
 */
#ifndef IMEXSW_H
#define IMEXSW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMEXSW_MAX_SOLID 6
#define IMEXSW_MAX_VARS (4 + IMEXSW_MAX_SOLID)
#define IMEXSW_MAX_LEVELS 8

/* Side order used by bc_flag / bc_value: west, east, south, north. */
enum { IMEXSW_W = 0, IMEXSW_E = 1, IMEXSW_S = 2, IMEXSW_N = 3 };

/* Source-ring direction bits (geometry_2d sourceE/W/S/N,
 * src/geometry_2d.f90:872-1102). */
enum { IMEXSW_SRC_E = 1, IMEXSW_SRC_W = 2, IMEXSW_SRC_S = 4, IMEXSW_SRC_N = 8 };

/*
 * Plain-old-data mirror of the module variables the hot path reads
 * (parameters_2d, constitutive_2d, geometry_2d).  Filled by the host from the
 * reference's namelists (read_param, src/inpout_2d.f90:957-4886); derived
 * quantities (rho_a_amb, inv_pres, eps_sing, one_by_dx, ...) are computed by
 * the library exactly as read_param / init_grid compute them.
 *
 * Supported model subset: gas carrier (GAS_FLAG=T), no liquid, no additional
 * gas components, no stochastic or pore-pressure variables, no Sutherland
 * viscosity, rheology model 1 (Voellmy-Salm), 6 (turbulent drag), 8
 * (turbulent friction integrated in closed form) or RHEOLOGY_FLAG=F, with or
 * without vertical profiles, solver scheme "KT".  n_vars = 4 + n_solid (src/inpout_2d.f90:1429,
 * 1852).
 */
typedef struct imexsw_params {
  /* grid: geometry_2d comp_cells_x/y, x0, y0, cell_size */
  int32_t nx, ny;
  double x0, y0, cell_size;

  /* layout / scheme: parameters_2d */
  int32_t n_solid;          /* number of solid classes, 0..IMEXSW_MAX_SOLID */
  int32_t n_rk;             /* IMEX RK stages N_RK (1..4) */

  /* flags (LOGICALs as 0/1) */
  int32_t energy_flag;      /* q4 = total energy (T) or h*rho*cp*T (F) */
  int32_t alpha_flag;       /* qp carries alphas (T) or h*alphas (F) */
  int32_t slope_correction_flag;
  int32_t curvature_term_flag;
  int32_t topo_change_flag;
  int32_t radial_source_flag;
  int32_t rheology_flag;
  int32_t rheology_model;   /* 1 = Voellmy-Salm, 6 = turbulent drag
                               rho f |u|^2, 8 = the same drag integrated in
                               closed form */
  int32_t settling_flag;
  int32_t entrainment_flag;
  int32_t erodible_deposit_flag;

  /* numerics: NUMERIC_PARAMETERS */
  double cfl;               /* CFL (-1 => fixed max_dt) */
  double max_dt;
  double reconstr_coeff;
  double theta;             /* generalized minmod parameter */
  int32_t limiter[IMEXSW_MAX_VARS + 2];   /* limiter(1:n_vars+2) */

  /* boundary conditions bc{W,E,S,N}(1:n_vars) */
  int32_t bc_flag[4][IMEXSW_MAX_VARS];
  double bc_value[4][IMEXSW_MAX_VARS];

  /* physics: EXPL_TERMS / GAS_TRANSPORT / SOLID_TRANSPORT / RHEOLOGY */
  double grav;
  double pres;
  double T_ambient;
  double sp_heat_a;
  double sp_gas_const_a;
  double kin_visc_c;        /* carrier kinematic viscosity (= KIN_VISC_A) */
  double rho_s[IMEXSW_MAX_SOLID];
  double diam_s[IMEXSW_MAX_SOLID];
  double sp_heat_s[IMEXSW_MAX_SOLID];
  double mu;
  double xi;

  /* erosion / deposition */
  double erosion_coeff;
  double erodible_porosity;
  double erodible_fract[IMEXSW_MAX_SOLID];
  double T_erodible;
  double alphastot_min;
  double maximum_solid_packing;

  /* radial source (RADIAL_SOURCE_PARAMETERS, h_source already derived from
   * MFR_SOURCE by read_param, src/inpout_2d.f90:4379-4482) */
  double x_source, y_source, r_source;
  double h_source, vel_source, T_source;
  double alphas_source[IMEXSW_MAX_SOLID];
  double time_param[4];

  /* hazard tables (THICKNESS_LEVELS0 / DYN_PRES_LEVELS0); n = 0 disables */
  int32_t n_thickness_levels;
  int32_t n_dyn_pres_levels;
  double thickness_levels[IMEXSW_MAX_LEVELS];
  double dyn_pres_levels[IMEXSW_MAX_LEVELS];

  /* vertical profiles (VERTICAL_PROFILES_FLAG and the namelist
   * VERTICAL_PROFILES_PARAMETERS, src/inpout_2d.f90:367-368, 3695-3796):
   * log velocity / exponential concentration closures of
   * src/constitutive_2d.f90:797-920, 970-1123, 1976-2107, 2532-2986.
   * They require ENERGY_FLAG = F and RHEOLOGY_FLAG = F or
   * RHEOLOGY_MODEL = 8, as read_param does (src/inpout_2d.f90:3699-3718). */
  int32_t vertical_profiles_flag;
  int32_t bottom_conc_flag;   /* deposition scaled by the bottom concentration */
  int32_t n_quad;             /* Gauss-Legendre points, 1..IMEXSW_MAX_QUAD */
  double vonK;                /* von Karman constant */
  double k_s;                 /* roughness length [m] */
  double Sc;                  /* Schmidt number */
  double z_dyn;               /* read but without effect on the state: the
                                 profile dynamic pressure is overwritten by
                                 r_phys_var (src/constitutive_2d.f90:957) */
  /* RHEOLOGY_PARAMETERS friction_factor: rheology models 6 and 8, and the shear
   * velocity of the profiles (src/inpout_2d.f90:357, 3350-3365) */
  double friction_factor;
} imexsw_params;

#define IMEXSW_MAX_QUAD 64

/* Per-context counters (cumulative since create or the last reset). */
typedef struct imexsw_stats {
  int64_t steps;
  int64_t cell_updates_active;  /* sum over steps of solve_cells */
  int64_t cell_updates_grid;    /* nx*ny per step */
  int64_t cell_updates_computed;/* sum over steps of the computed cells (the
                                   active cells minus those that provably
                                   keep their state) */
  int64_t newton_cells;         /* cells that entered the Newton solve */
  int64_t newton_iters;         /* total Newton iterations */
  int64_t newton_nonconverged;  /* cells that hit max_nl_iter */
  int64_t nan_events;           /* NaN trap, src/solver_2d.f90:1379-1394 */
  int64_t neg_h_events;         /* h < -1e-7, src/solver_2d.f90:1405-1433 */
  int64_t neg_solid_events;     /* solids <= -1e-7, src/solver_2d.f90:1439 */
  int64_t sum_solid_events;     /* sum(solids) > h, src/solver_2d.f90:1569 */
  int64_t low_T_events;         /* T < 273 K, src/solver_2d.f90:1538 */
  int64_t line_search_evals;    /* residual evaluations inside lnsrch */
  int64_t newton_iter_hist[8];  /* cells by Newton iterations: 1, 2, 3, 4,
                                   5-8, 9-16, 17-32, 33-100 */
  int64_t newton_iter_max;      /* most iterations any cell took */
  int64_t warm_cells;           /* Newton solves started from the learned guess */
  int64_t warm_fallbacks;       /* ... that fell back to the plain predictor */
  int64_t sparse_steps;         /* steps run in the few-computed-cells form
                                   (row stencil, one pass per row for the
                                   stage update and erosion); same results */
  int32_t solve_cells;          /* active cells of the last check_solve */
  int32_t compute_cells;        /* of those, cells whose update is computed
                                   (the rest provably keep their state) */
} imexsw_stats;

typedef struct imexsw_ctx imexsw_ctx;

/* Library version string. */
const char *imexsw_version(void);

/* Last error message of a context (or of the last failed create when ctx is
 * NULL). */
const char *imexsw_last_error(const imexsw_ctx *ctx);

/* Allocate device state for one grid.  Replaces allocate_solver_variables
 * (src/solver_2d.f90:263-529) together with the derived-constant setup of
 * read_param / init_grid / init_problem_param (src/inpout_2d.f90:1330, 1353,
 * 1852-1886; src/geometry_2d.f90:282-290; src/constitutive_2d.f90:469-518). */
int imexsw_create(const imexsw_params *params, imexsw_ctx **ctx);

/* Release everything.  Replaces deallocate_solver_variables
 * (src/solver_2d.f90:543-657). */
void imexsw_destroy(imexsw_ctx *ctx);

/* Upload the topography B_cent(nx,ny) and B_nodata(nx,ny) (may be NULL) and
 * derive every per-cell static field on the device the way
 * topography_reconstruction does (src/geometry_2d.f90:535-860). */
int imexsw_set_topography(imexsw_ctx *ctx, const double *B_cent,
                          const uint8_t *B_nodata);

/* Alternative to imexsw_set_topography for a host that already holds the
 * geometry_2d fields: B_cent, B_prime_{x,y}_geom, B_second_{xx,xy,yy}_geom,
 * grav_coeff, d_grav_coeff_dx (nx,ny); grav_coeff_stag_x (nx+1,ny);
 * grav_coeff_stag_y (nx,ny+1); B_nodata (may be NULL).  Any double pointer
 * except B_cent may be NULL (zero, or one for the grav_coeff family). */
int imexsw_set_geometry(imexsw_ctx *ctx, const double *B_cent,
                        const double *Bx_geom, const double *By_geom,
                        const double *Bxx_geom, const double *Bxy_geom,
                        const double *Byy_geom, const double *grav_coeff,
                        const double *grav_stag_x, const double *grav_stag_y,
                        const double *dgrav_dx, const uint8_t *B_nodata);

/* Radial source ring: source_cell(nx,ny) in {0,1,2}, dir_bits(nx,ny) of
 * IMEXSW_SRC_*, vect(8,nx,ny) = (E_x,E_y,W_x,W_y,S_x,S_y,N_x,N_y) as produced
 * by init_source (src/geometry_2d.f90:872-1102). */
int imexsw_set_source(imexsw_ctx *ctx, const int32_t *source_cell,
                      const uint8_t *dir_bits, const double *vect);

/* Upload the conservative state q(n_vars,nx,ny) (read_solution /
 * init_2d write the module array q: src/inpout_2d.f90:5013,
 * src/init_2d.f90:36, 89).  The upload starts a new run: the next
 * imexsw_advance initialises dt_old and dt_old_old from its dt0, as the
 * main program does after the initial state is set
 * (src/IMEX_SfloW2D.f90:248-249). */
int imexsw_upload_q(imexsw_ctx *ctx, const double *q);

/* Upload erodible(n_solid,nx,ny) and optionally deposit/erosion
 * (nx,ny,n_solid) accumulators (may be NULL = keep). */
int imexsw_upload_erosion(imexsw_ctx *ctx, const double *erodible,
                          const double *deposit, const double *erosion);

/* Active-cell mask: wet cells, source ring, N_RK-fold 4-neighbour dilation,
 * source interior excluded.  Replaces check_solve(solve_all)
 * (src/solver_2d.f90:666-857).  solve_cells (may be NULL) receives the count
 * (copied back to the host: this call synchronises). */
int imexsw_check_solve(imexsw_ctx *ctx, int solve_all, double t,
                       int32_t *solve_cells);

/* CFL time step over the active cells.  Replaces timestep
 * (src/solver_2d.f90:872-1005).  Synchronises and returns dt on the host. */
int imexsw_timestep(imexsw_ctx *ctx, double *dt);

/* One IMEX Runge-Kutta step of size dt at time t.  Replaces imex_RK_solver
 * (src/solver_2d.f90:1021-1632). */
int imexsw_rk_step(imexsw_ctx *ctx, double t, double dt);

/* Operator-split erosion/deposition/entrainment update.  Replaces
 * update_erosion_deposition_cell (src/solver_2d.f90:2380-2595). */
int imexsw_erosion_deposition(imexsw_ctx *ctx, double dt);

/* Per-step topography update derivatives.  Replaces the per-step call of
 * topography_reconstruction (src/IMEX_SfloW2D.f90:387,
 * src/geometry_2d.f90:535-860). */
int imexsw_topography_reconstruction(imexsw_ctx *ctx);

/* Post-step diagnostics: qc->qp, running maxima of h, dynamic pressure and
 * speed, hazard tables, and the per-step sums printed by the host
 * (src/IMEX_SfloW2D.f90:391-483).  sums (may be NULL) receives
 * {sum q1, sum qp1, count(q1 > 1e-7), sum of solid masses}, each still to be
 * multiplied by dx*dy as the host prints them (synchronises when non-NULL). */
int imexsw_post_step(imexsw_ctx *ctx, double sums[4]);

/* Run nsteps of the reference's main loop entirely on the device, including
 * the dt clipping and growth limiter of src/IMEX_SfloW2D.f90:357-381 (dt_old
 * and dt_old_old are carried in the context, initialised from dt0 by the
 * first call), with no host round trip per step.  t_io = current time
 * (updated), t_end and t_output as in the host loop.  Stops early when t
 * reaches t_output (t_output within 1e-7 of t_end counts as t_end): there
 * the reference's loop writes its output and output_solution raises
 * t_output (src/IMEX_SfloW2D.f90:510-527, src/inpout_2d.f90:6166), so the
 * host does the same and calls again; a call with t >= t_output takes no
 * step.  Steps run as replays of captured 20-step graphs (direct
 * launches for a remainder, and for every step while per-kernel timing is
 * on); each replay runs the form for few or for many computed cells, chosen
 * from the computed-cell count of an earlier step (stats.sparse_steps). */
int imexsw_advance(imexsw_ctx *ctx, int32_t nsteps, double dt0, double *t_io,
                   double t_end, double t_output, int32_t *steps_done);

/* Download state.  Any pointer may be NULL.  q(n_vars,nx,ny),
 * qp(n_vars+2,nx,ny), hmax/pdynmax/mod_vel_max(nx,ny),
 * vuln(n_thk*n_pdyn,nx,ny) as 0/1 bytes. */
int imexsw_download(imexsw_ctx *ctx, double *q, double *qp, double *hmax,
                    double *pdynmax, double *mod_vel_max, uint8_t *vuln);

/* Allocation-free download of q(n_vars,nx,ny) and qp(n_vars+2,nx,ny) (either
 * may be NULL), the same values as imexsw_download gives.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under stream capture, serving loops, hosts that forbid hipMalloc).
 * The rows are converted through a static staging array of 2^21 doubles
 * (16 MiB) per GPU, held until the process ends, in as many chunks as
 * needed; a grid whose single row holds more values (nvar * nx > 2^21) is
 * refused with an error.  Calls on one GPU share that array: calls that
 * overlap in time on one GPU (contexts used from several host threads)
 * race, so serialise them.  Synchronises the context's stream. */
int imexsw_download_noalloc(imexsw_ctx *ctx, double *q, double *qp);

/* Download erosion state: erodible(n_solid,nx,ny), deposit(nx,ny,n_solid),
 * erosion(nx,ny,n_solid), B_cent(nx,ny).  Any pointer may be NULL. */
int imexsw_download_erosion(imexsw_ctx *ctx, double *erodible, double *deposit,
                            double *erosion, double *B_cent);

/* Download solve_mask(nx,ny) as 0/1 bytes (src/solver_2d.f90:159). */
int imexsw_download_mask(imexsw_ctx *ctx, uint8_t *mask);

/* Last device dt (after clipping when produced by imexsw_advance). */
int imexsw_get_dt(imexsw_ctx *ctx, double *dt);

/* Counters.  reset != 0 zeroes them after reading. */
int imexsw_stats_get(imexsw_ctx *ctx, imexsw_stats *stats, int reset);

/* Per-kernel device time, measured with HIP events recorded on the
 * context's stream around every launch.  names: mask, stencil, local,
 * topo, dtpolicy (the kernel between steps that commits the time-step
 * policy), post, erosion.  ms[i] accumulates milliseconds, calls[i]
 * launches.  Returns the number of kernels (<= cap).  Timing is enabled by
 * imexsw_set_timing(ctx, 1); while it is on, imexsw_advance launches every
 * kernel directly (the events slow the step). */
int imexsw_set_timing(imexsw_ctx *ctx, int enable);
int imexsw_kernel_times(imexsw_ctx *ctx, double *ms, int64_t *calls, int cap);
const char *imexsw_kernel_name(int index);

/* Launch shape of the stencil pass (no reference counterpart: the
 * reference's OpenMP loops have no tiling).  Each step the stencil kernel
 * picks the tallest strip of 2^m rows, m >= 0, up to max_strip_rows, that
 * still leaves about target_strips strips for the step's computed cells.
 * max_strip_rows: 1, 2, 4, 8 or 16; target_strips >= 1.  Results do not
 * depend on these settings; defaults 16 and 2048. */
int imexsw_set_tuning(imexsw_ctx *ctx, int max_strip_rows, int target_strips);

/* Wait for all work queued on the context's stream. */
int imexsw_sync(imexsw_ctx *ctx);

/* ---------------------------------------------------------------------------
 * Point conversions with the context's model (vertical profiles included),
 * evaluated on the device.  States are packed one after another:
 * qp(n_vars+2) = h, hu, hv, T, (h*)alphas, u, v; qc(n_vars).  For the
 * host's set-up work (initial states, source parameters), not per step.
 * ------------------------------------------------------------------------- */

/* qp_to_qc of n states (src/constitutive_2d.f90:1709-2140), as init_2d
 * converts the initial physical states (src/init_2d.f90:53, 149-153). */
int imexsw_qp_to_qc(imexsw_ctx *ctx, int64_t n, const double *qp, double *qc);

/* qc_to_qp of n states (src/constitutive_2d.f90:1628-1681). */
int imexsw_qc_to_qp(imexsw_ctx *ctx, int64_t n, const double *qc, double *qp);

/* eval_flux_coeffs of n physical states with slopes Bx[n], By[n]
 * (src/constitutive_2d.f90:2532-2823): shape(n_vars) per state, all 1
 * without vertical profiles or for a dry state.  read_param derives the
 * source thickness from MFR_SOURCE with shape(1) of the source state
 * (src/inpout_2d.f90:4440-4480). */
int imexsw_flux_coeffs(imexsw_ctx *ctx, int64_t n, const double *qp, const double *Bx,
                       const double *By, double *shape);

/* Lambert W of n arguments on branch 0 or -1 as the device evaluates it
 * (lambertw0 / lambertwm1, src/geometry_2d.f90:1830-2038; -1 outside
 * [-1/e, 0) on branch -1 and below -1/e on branch 0, as there). */
int imexsw_lambertw(int32_t branch, int64_t n, const double *x, double *w);

/* Gauss-Legendre points and weights on [-1, 1], ascending, n in
 * 1..IMEXSW_MAX_QUAD (gaulegf, src/geometry_2d.f90:2407-2449); host only. */
int imexsw_gauss_legendre(int32_t n, double *x, double *w);

/* ---------------------------------------------------------------------------
 * .q_2d state files (host only; no context, no device work)
 * ------------------------------------------------------------------------- */

/* Write a state in the .q_2d text format of output_solution
 * (src/inpout_2d.f90:5766-5891, the conservative-variable file): one record
 * per cell, j fastest within k, holding x_comp[j], y_comp[k] and
 * q(1..nv, j, k), each as Fortran E20.12; values with |q| < 1e-20 are
 * written as zero (the reference also zeroes them in its state array); a
 * record with a blank after every row of cells and two at the end.  q is the
 * host's column-major (var, j, k) array.  Returns 0, or -1 when the file
 * cannot be written. */
int imexsw_q2d_write(const char *path, int32_t nv, int32_t nx, int32_t ny,
                     const double *x_comp, const double *y_comp, const double *q);

/* Read a .q_2d restart file as read_solution does
 * (src/inpout_2d.f90:5432-5453): for each row k, nx records of 20-column
 * E fields x, y, q(1..nv) followed by one record that is skipped; a cell
 * with q(1) <= 0 is set to zero, and negative solid masses q(5..4+n_solid)
 * to zero.  Fields beyond the end of a record read as zero.  Returns 0, or
 * -1 when the file is missing, short or has a malformed field. */
int imexsw_q2d_read(const char *path, int32_t nv, int32_t n_solid, int32_t nx, int32_t ny,
                    double *q);

/* ---------------------------------------------------------------------------
 * Ensembles: many independent grids (members) of possibly different sizes
 * and parameters advanced together, one set of batched launches per time
 * step for all members, the steps captured in a HIP graph.  The reference
 * runs an ensemble as a farm of separate processes, one per member
 * (EXAMPLES/EXAMPLE_ENSEMBLE_ETNA/launch_jobs.py); there is no solver_2d
 * counterpart.  Every member advances exactly as it would alone in its own
 * context (same kernels per cell, same bits).
 * Supported: 1..3 solid classes (the same for all members), 2-D grids, no
 * topography change, radial source, slope correction or curvature.
 * ------------------------------------------------------------------------- */
typedef struct imexsw_ensemble imexsw_ensemble;

/* params[n_members]: one imexsw_params per member. */
int imexsw_ensemble_create(int32_t n_members, const imexsw_params *params,
                           imexsw_ensemble **ens);
void imexsw_ensemble_destroy(imexsw_ensemble *ens);
const char *imexsw_ensemble_last_error(const imexsw_ensemble *ens);

/* Member i as an ordinary context for set-up and output (topography, q,
 * check_solve, downloads, stats).  Owned by the ensemble. */
imexsw_ctx *imexsw_ensemble_member(imexsw_ensemble *ens, int32_t i);

/* Stencil strip height for all members (4, 8 or 16 rows) and the number of
 * time steps per captured graph (1..1000).  Results do not depend on these. */
int imexsw_ensemble_set_tuning(imexsw_ensemble *ens, int strip_rows, int graph_steps);

/* Run nsteps batched steps of the main loop for every member (the dt policy
 * of src/IMEX_SfloW2D.f90:357-381 per member, dt history initialised from
 * dt0 on a member's first run).  t[n_members] holds each member's time in
 * and out (NULL: keep the members' device times); members stop at t_end.
 * steps_done = steps member 0 took in this call. */
int imexsw_ensemble_advance(imexsw_ensemble *ens, int32_t nsteps, double dt0, double *t,
                            double t_end, int32_t *steps_done);

/* Autotuner: runs steps_per_trial real steps with each strip height, keeps
 * the fastest (best_strip_rows may be NULL).  The trials advance the
 * members (t as in imexsw_ensemble_advance). */
int imexsw_ensemble_autotune(imexsw_ensemble *ens, int32_t steps_per_trial, double dt0,
                             double *t, double t_end, int *best_strip_rows);

/* Device time of the last imexsw_ensemble_advance, milliseconds (HIP
 * events on the ensemble's stream). */
double imexsw_ensemble_last_ms(const imexsw_ensemble *ens);

/* ---------------------------------------------------------------------------
 * Learned Newton warm start (no reference counterpart: an addition).  A
 * per-cell MLP 9 -> 64 -> 64 -> 64 -> 2 (GELU, BF16 weights, FP32
 * accumulation on the matrix cores) predicts the converged momenta of the
 * stage-2 Newton solve (src/solver_2d.f90:1207-1212) from the state that
 * enters it.  Newton then starts from whichever of the prediction and the
 * plain predictor q_si (:1196) has the smaller residual and still iterates
 * to the reference tolerance, so the plain method stays both fallback and
 * judge.  Features: log h, predictor velocity u, v, T, three solid volume
 * fractions, log dt, |grad B|; targets: q*_x / q_si_x - 1, q*_y / q_si_y - 1.
 * ------------------------------------------------------------------------- */
typedef struct imexsw_mlp imexsw_mlp;

int imexsw_mlp_num_params(void);
/* He-normal initial weights, deterministic in seed. */
int imexsw_mlp_create(imexsw_mlp **mlp, uint64_t seed);
void imexsw_mlp_destroy(imexsw_mlp *mlp);
const char *imexsw_mlp_last_error(const imexsw_mlp *mlp);
/* params: imexsw_mlp_num_params() floats (W1[64][16] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[16][64] b4[16]); norm: 22 floats (input mean[9],
 * input std[9], target mean[2], target std[2]).  Either may be NULL. */
int imexsw_mlp_set_params(imexsw_mlp *mlp, const float *params, const float *norm);
int imexsw_mlp_get_params(imexsw_mlp *mlp, float *params, float *norm);
/* Inference: raw features x[n][9] -> raw targets y[n][2]. */
int imexsw_mlp_forward(imexsw_mlp *mlp, const float *x, int64_t n, float *y);
/* Loss (mean over samples of the squared error summed over the outputs) and
 * its parameter gradient for one minibatch given in the standardised space:
 * x[n][9], t[n][2]; grad: imexsw_mlp_num_params() floats. */
int imexsw_mlp_grad(imexsw_mlp *mlp, const float *x, const float *t, int32_t n, float *grad,
                    float *loss);
/* Train on raw pairs[n][11] (9 features then 2 targets): standardisation
 * from the data, Adam (lr) over seeded shuffled minibatches of `batch`
 * samples; loss_hist[epochs] (may be NULL). */
int imexsw_mlp_train(imexsw_mlp *mlp, const float *pairs, int64_t n, int32_t epochs,
                     int32_t batch, float lr, uint64_t seed, float *loss_hist);
/* D = A . B on the matrix cores for 16 x 16 row-major bf16 bit patterns
 * (operand-layout check). */
int imexsw_mfma_probe(const uint16_t *A, const uint16_t *B, float *D);

/* Collect (features, targets) pairs of every converged stage-2 Newton solve
 * of the following steps into a device buffer of `capacity` pairs (0 stops
 * and frees).  imexsw_harvest_get copies out at most cap pairs and reports
 * how many were produced. */
int imexsw_harvest(imexsw_ctx *ctx, int64_t capacity);
int imexsw_harvest_get(imexsw_ctx *ctx, float *pairs, int64_t cap, int64_t *n);
/* Use `mlp` for the warm start of the following steps (NULL: off).  The MLP
 * must outlive its use by the context. */
int imexsw_set_warm_start(imexsw_ctx *ctx, imexsw_mlp *mlp);

/* Closed-form Newton start (enable != 0) for the following stage-2 solves,
 * no reference counterpart: the momenta of the Voellmy implicit equation
 * m = m_si + dt NH(m) are parallel to m_si, and with the mixture density
 * frozen their magnitude is the root of a quadratic; refreshing the density
 * (which depends on the speed only through the kinetic energy) makes it
 * the solution to rounding.  Newton starts there when that point meets the
 * tolerance or has a smaller residual than the plain predictor q_si
 * (src/solver_2d.f90:1196), and iterates to the reference tolerance as
 * always.  Exclusive with the learned warm start; N_RK = 2 only. */
int imexsw_set_closed_form_start(imexsw_ctx *ctx, int enable);

#ifdef __cplusplus
}
#endif

#endif /* IMEXSW_H */
