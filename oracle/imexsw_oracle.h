/*
 * imexsw_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, OpenMP over cells like the reference) of the
 * IMEX_SfloW2D per-timestep hot path, used as the correctness oracle for the
 * MI355X library and as the CPU baseline ("port") in bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * It keeps the reference's data layout (Fortran column-major (var, j, k),
 * 1-based logic) and its module-level persistent arrays, so that the
 * behaviour of src/solver_2d.f90 -- including stale interface values of
 * cells that leave the active set -- is reproduced.  `fresh` mode (see
 * orc_set_fresh) instead rebuilds every non-source cell's interface values
 * from the current state at every reconstruction; this is the semantics the
 * device implementation follows, and the tests measure the difference
 * between the two modes.
 */
#ifndef IMEXSW_ORACLE_H
#define IMEXSW_ORACLE_H

#include <stdint.h>
#include "../include/imexsw.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

orc_t *orc_create(const imexsw_params *p);
void orc_destroy(orc_t *o);
void orc_set_fresh(orc_t *o, int fresh);
void orc_set_threads(int n);

int orc_n_vars(const orc_t *o);

/* geometry: B_cent(nx,ny) -> topography_reconstruction; nodata may be NULL */
void orc_set_topography(orc_t *o, const double *B_cent, const uint8_t *nodata);
/* radial source ring from params (init_source) */
void orc_init_source(orc_t *o);
void orc_get_source(const orc_t *o, int32_t *source_cell, uint8_t *dir_bits,
                    double *vect);
void orc_get_geometry(const orc_t *o, double *B_cent, double *Bx, double *By,
                      double *Bxx, double *Bxy, double *Byy);

void orc_set_q(orc_t *o, const double *q);
void orc_get_q(const orc_t *o, double *q);
void orc_get_qp(const orc_t *o, double *qp);
void orc_set_erodible(orc_t *o, const double *erodible);
void orc_get_erosion(const orc_t *o, double *erodible, double *deposit,
                     double *erosion, double *B_cent);
void orc_get_post(const orc_t *o, double *hmax, double *pdynmax,
                  double *mod_vel_max, uint8_t *vuln);
void orc_get_mask(const orc_t *o, uint8_t *mask);

/* solver_2d public procedures */
int orc_check_solve(orc_t *o, int solve_all);
double orc_timestep(orc_t *o);
void orc_imex_rk(orc_t *o, double t, double dt);
void orc_erosion_deposition(orc_t *o, double dt);
void orc_topography_reconstruction(orc_t *o);
/* post-step loop + per-step sums (src/IMEX_SfloW2D.f90:391-483) */
void orc_post_step(orc_t *o, double sums[4]);
/* initial diagnostics loop (src/IMEX_SfloW2D.f90:255-311) */
void orc_init_post(orc_t *o);

/* the main loop: nsteps (or until t_end), with the host's dt policy */
int orc_run(orc_t *o, int nsteps, double dt0, double *t_io, double t_end,
            double t_output, double *dt_last);

/* counters: [newton_cells, newton_iters, nonconverged, stale_events,
 *            nan, neg_h, neg_solid, sum_solid, low_T, active_cell_updates] */
void orc_counters(const orc_t *o, int64_t *c10);

/* ---- component entry points for unit tests (one cell / one face) ---- */
void orc_qc_to_qp(const orc_t *o, const double *qc, double *qp, double *p_dyn);
void orc_qp_to_qc(const orc_t *o, const double *qp, double *qc);
void orc_mixt_var(const orc_t *o, const double *qp, double *out6);
double orc_limit(const orc_t *o, const double *v, const double *z, int lim);
double orc_settling_velocity(const orc_t *o, double diam, double rhos,
                             double rhoc, double inv_kin_visc);
void orc_eval_fluxes(const orc_t *o, const double *qc, const double *qp,
                     double grav_coeff, int dir, double *flux);
void orc_local_speeds(const orc_t *o, const double *qp, double grav_coeff,
                      int dir, double *vmin_vmax);
void orc_implicit_terms(const orc_t *o, const double *q, double *nh);
void orc_implicit_terms_cstep(const orc_t *o, const double *q, int col,
                              double h, double *re, double *im);
void orc_semi_impl_terms(const orc_t *o, const double *qc, const double *qp,
                         double grav_coeff, double *nh);
/* one stage-2 (N_RK=2) local update of one cell from q0 and the explicit
 * residual D = divFlux - expl_terms: returns Newton iterations */
int orc_local_update(const orc_t *o, const double *q0, const double *D,
                     double dt, double *q_out);
double orc_det_pow(double x, double y);
double orc_det_cos(double x);
void orc_source_bdry(const orc_t *o, double time, double vx, double vy,
                     double *out);
/* vertical profiles: special functions and closures */
double orc_lambertw0(double x);
double orc_lambertwm1(double x);
void orc_gauss_legendre(int n, double *x, double *w);
double orc_vp_h0(const orc_t *o, double h);
void orc_flux_coeffs(const orc_t *o, const double *qp, double Bx, double By,
                     double *shape);
void orc_dep_coeffs(const orc_t *o, const double *qp, double mod_vel,
                    double *dep);

#ifdef __cplusplus
}
#endif
#endif
