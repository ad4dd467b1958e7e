/*
 * imexsw_oracle.c -- TEST INFRASTRUCTURE ONLY (see imexsw_oracle.h).
 *
 * Plain-C restatement of the reference's per-timestep hot path.  Every
 * routine cites the Fortran it follows (paths relative to the reference
 * tree).  Evaluation order follows the Fortran expressions left to right,
 * intrinsic reductions (SUM, DOT_PRODUCT, MATMUL) are sequential from zero,
 * and it must be built without FMA contraction or fast-math
 * (-ffp-contract=off -fno-fast-math), so results are those of the reference
 * built with gfortran -O2.
 *
 * Complex arithmetic (complex-step Jacobian) follows gfortran's default
 * lowering (-fcx-fortran-rules): products are (ar*br - ai*bi, ar*bi + ai*br),
 * quotients use Smith's range-reduced algorithm, SQRT is glibc's csqrt
 * formula and DOT_PRODUCT conjugates its first argument.
 *
 * Supported subset: gas carrier, no liquid / additional gas / stochastic /
 * pore pressure / bottom or lateral sources / Sutherland viscosity,
 * rheology model 1 (Voellmy-Salm), 6 (turbulent drag), 8 (closed-form
 * turbulent friction) or no rheology, with or without vertical profiles,
 * solver scheme KT.
 */
#include "imexsw_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXV IMEXSW_MAX_VARS
#define MAXS IMEXSW_MAX_SOLID
#define EPS_WP 2.220446049250313e-16 /* EPSILON(1.0_wp) */


/* ------------------------------------------------------------------------ */
/* elementary functions with a fixed operation sequence                     */
/* The reference calls libm pow, cos and (inside csqrt) hypot.  Vendor      */
/* libraries differ in the last bits between CPU and GPU; both the oracle   */
/* and the device implement these with the same IEEE operation sequence so */
/* that device and oracle can agree bit for bit.  Methods: Cody & Waite     */
/* (1980) argument reduction (ln2 and pi/2 split), log(m) = 2 atanh(s),     */
/* s = (m-1)/(m+1), degree-14 Taylor exp (both polynomials by Estrin's     */
/* scheme, for short dependency chains on the GPU), Taylor sin/cos on      */
/* |r| <= pi/4.                                                             */
/* They are within a few ulp of the correctly rounded values over the       */
/* ranges used, i.e. 1e-15 relative to libm.                                */
/* ------------------------------------------------------------------------ */
static inline int64_t d2i(double x) { int64_t i; memcpy(&i, &x, 8); return i; }
static inline double i2d(int64_t i) { double x; memcpy(&x, &i, 8); return x; }

static double det_log(double x) {
  int64_t bits = d2i(x);
  int e = (int)((bits >> 52) & 0x7ff);
  if (e == 0) {
    x = x * 18014398509481984.0;
    bits = d2i(x);
    e = (int)((bits >> 52) & 0x7ff) - 54;
  }
  e -= 1022;
  double m = i2d((bits & 0x000fffffffffffffLL) | 0x3fe0000000000000LL);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  /* 1/3 + s2/5 + ... + s2^10/23 by Estrin's scheme (depth 7, not 20) */
  const double x2 = s2 * s2, x4 = x2 * x2, x8 = x4 * x4;
  const double a0 = 1.0 / 3.0 + 1.0 / 5.0 * s2, a1 = 1.0 / 7.0 + 1.0 / 9.0 * s2;
  const double a2 = 1.0 / 11.0 + 1.0 / 13.0 * s2, a3 = 1.0 / 15.0 + 1.0 / 17.0 * s2;
  const double a4 = 1.0 / 19.0 + 1.0 / 21.0 * s2;
  const double b0 = a0 + a1 * x2, b1 = a2 + a3 * x2, b2 = a4 + 1.0 / 23.0 * x2;
  const double p = (b0 + b1 * x4) + b2 * x8;
  const double logm = 2.0 * s + 2.0 * s * (s2 * p);
  const double de = (double)e;
  return (de * 6.93147180369123816490e-01 + logm) + de * 1.90821492927058770002e-10;
}

static double det_exp(double y) {
  if (y > 709.0) return INFINITY;
  if (y < -745.0) return 0.0;
  const double n = rint(y * 1.44269504088896338700e+00);
  const double r = (y - n * 6.93147180369123816490e-01) - n * 1.90821492927058770002e-10;
  /* sum of r^k / k!, k = 0..14, by Estrin's scheme (depth 8, not 28) */
  const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
  const double a0 = 1.0 + r, a1 = 0.5 + 1.0 / 6.0 * r;
  const double a2 = 1.0 / 24.0 + 1.0 / 120.0 * r, a3 = 1.0 / 720.0 + 1.0 / 5040.0 * r;
  const double a4 = 1.0 / 40320.0 + 1.0 / 362880.0 * r;
  const double a5 = 1.0 / 3628800.0 + 1.0 / 39916800.0 * r;
  const double a6 = 1.0 / 479001600.0 + 1.0 / 6227020800.0 * r;
  const double b0 = a0 + a1 * r2, b1 = a2 + a3 * r2, b2 = a4 + a5 * r2;
  const double b3 = a6 + 1.0 / 87178291200.0 * r2;
  const double p = (b0 + b1 * r4) + (b2 + b3 * r4) * r8;
  const int ni = (int)n;
  const int n1 = ni / 2, n2 = ni - n1;
  const double s1 = i2d((int64_t)(1023 + n1) << 52);
  const double s2 = i2d((int64_t)(1023 + n2) << 52);
  return (p * s1) * s2;
}

static double det_pow(double x, double y) {
  if (x == 0.0) return (y > 0.0) ? 0.0 : (y == 0.0 ? 1.0 : INFINITY);
  if (x == 1.0 || y == 0.0) return 1.0;
  if (x != x || y != y) return x + y;
  return det_exp(y * det_log(x));
}

static double det_cos(double x) {
  const double P1 = 1.57079632673412561417e+00, P2 = 6.07710050650619224932e-11,
               P3 = 2.02226624879595063154e-21;
  const double k = rint(x * 6.36619772367581382433e-01);
  const double r = ((x - k * P1) - k * P2) - k * P3;
  const double r2 = r * r;
  double c = -1.0 / 87178291200.0;
  c = c * r2 + 1.0 / 479001600.0;
  c = c * r2 - 1.0 / 3628800.0;
  c = c * r2 + 1.0 / 40320.0;
  c = c * r2 - 1.0 / 720.0;
  c = c * r2 + 1.0 / 24.0;
  c = c * r2 - 0.5;
  c = c * r2 + 1.0;
  double s = -1.0 / 1307674368000.0;
  s = s * r2 + 1.0 / 6227020800.0;
  s = s * r2 - 1.0 / 39916800.0;
  s = s * r2 + 1.0 / 362880.0;
  s = s * r2 - 1.0 / 5040.0;
  s = s * r2 + 1.0 / 120.0;
  s = s * r2 - 1.0 / 6.0;
  s = r + r * (r2 * s);
  const int q = ((int)k) & 3;
  return q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}

static double det_hypot(double x, double y) {
  double a = fabs(x), b = fabs(y);
  if (a < b) {
    double t = a;
    a = b;
    b = t;
  }
  if (a == 0.0)
    return 0.0;
  const double r = b / a;
  return a * sqrt(1.0 + r * r);
}
double orc_det_pow(double x, double y) { return det_pow(x, y); }
double orc_det_cos(double x) { return det_cos(x); }

/* ------------------------------------------------------------------------ */
/* complex numbers with gfortran semantics                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
  double re, im;
} cx;

static inline cx C(double re, double im) {
  cx z = {re, im};
  return z;
}
static inline cx cadd(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
static inline cx cmul(cx a, cx b) {
  return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* complex * real (real operand promoted with zero imaginary part) */
static inline cx cmulr(cx a, double b) { return C(a.re * b, a.im * b); }
/* complex / real */
static inline cx cdivr(cx a, double b) { return C(a.re / b, a.im / b); }
/* Smith's algorithm (GCC expand_complex_div_wide) */
static inline cx cdiv(cx a, cx b) {
  double ratio, div, tr, ti;
  if (fabs(b.re) < fabs(b.im)) {
    ratio = b.re / b.im;
    div = (b.re * ratio) + b.im;
    tr = (a.re * ratio) + a.im;
    ti = (a.im * ratio) - a.re;
  } else {
    ratio = b.im / b.re;
    div = (b.im * ratio) + b.re;
    tr = (a.im * ratio) + a.re;
    ti = a.im - (a.re * ratio);
  }
  return C(tr / div, ti / div);
}
/* glibc csqrt for finite moderate arguments */
static inline cx csqrt_f(cx x) {
  if (x.im == 0.0) {
    if (x.re < 0.0)
      return C(0.0, copysign(sqrt(-x.re), x.im));
    return C(fabs(sqrt(x.re)), copysign(0.0, x.im));
  }
  if (x.re == 0.0) {
    double r;
    if (fabs(x.im) >= 2.0 * 2.2250738585072014e-308)
      r = sqrt(0.5 * fabs(x.im));
    else
      r = 0.5 * sqrt(2.0 * fabs(x.im));
    return C(r, copysign(r, x.im));
  }
  double d = det_hypot(x.re, x.im), r, s;
  if (x.re > 0.0) {
    r = sqrt(0.5 * (d + x.re));
    s = 0.5 * (x.im / r);
  } else {
    s = sqrt(0.5 * (d - x.re));
    r = fabs(0.5 * (x.im / s));
  }
  return C(r, copysign(s, x.im));
}
static inline cx cconj(cx a) { return C(a.re, -a.im); }

/* ------------------------------------------------------------------------ */
/* context: module variables of parameters_2d / constitutive_2d /           */
/* geometry_2d / solver_2d                                                  */
/* ------------------------------------------------------------------------ */
struct orc {
  imexsw_params p;
  int nx, ny, nix, niy, nv, ns, ne, nrk;
  int idx_u, idx_v; /* 1-based */
  int fresh;
  double dx, dy, dx2, dy2, one_by_dx, one_by_dy;
  double eps_sing, eps_sing4;
  double rho_a_amb, inv_pres, inv_rho_s[MAXS], rho_c_sub, coeff_porosity;
  double pi_g, h_cs, one_by_h;
  double T_erodible;
  /* vertical profiles: Gauss-Legendre points on [-1, 1] (z_quad, w_quad of
   * geometry_2d) and H_crit_rel (src/inpout_2d.f90:3763, 3784-3791) */
  int vp, nq;
  double zq[IMEXSW_MAX_QUAD], wq[IMEXSW_MAX_QUAD], h_crit_rel, sqrt_ff;
  double dt_old, dt_old_old; /* host loop state (src/IMEX_SfloW2D.f90:250) */
  int implicit_flag[MAXV + 1]; /* 1-based */
  int n_nh;
  double a_tilde_ij[4][4], a_dirk_ij[4][4], omega_tilde[4], omega[4];
  double *x_comp, *x_stag, *y_comp, *y_stag;
  /* state */
  double t;
  double *q, *q0, *q_fv, *qp, *q_rk, *qp_rk, *divFlux, *NH, *SI_NH,
      *expl_terms, *residual_term;
  double *qL, *qR, *qB, *qT, *qpL, *qpR, *qpB, *qpT;
  double *axNeg, *axPos, *ayNeg, *ayPos, *Hx, *Hy, *axmax, *aymax;
  uint8_t *solve_mask, *solve_mask_temp, *smx, *smy;
  int *j_cent, *k_cent, *j_stag_x, *k_stag_x, *j_stag_y, *k_stag_y;
  int solve_cells, sif_x, sif_y;
  double *hmax, *pdynmax, *mod_vel_max;
  uint8_t *vuln_table, *thck_table, *pdyn_table;
  double *B_cent, *Bx, *By, *Bxx, *Bxy, *Byy; /* *_geom fields */
  double *grav_coeff, *gsx, *gsy, *dgdx;
  uint8_t *B_nodata;
  int *source_cell;
  uint8_t *sE, *sW, *sS, *sN;
  double *sEx, *sEy, *sWx, *sWy, *sSx, *sSy, *sNx, *sNy;
  double *deposit, *erosion, *erodible;
  /* counters */
  int64_t cnt[10];
};

/* 1-based index helpers (Fortran column-major) */
#define I2(j, k) ((size_t)((j)-1) + (size_t)o->nx * (size_t)((k)-1))
#define IQ(i, j, k) ((size_t)((i)-1) + (size_t)o->nv * I2(j, k))
#define IQP(i, j, k) ((size_t)((i)-1) + (size_t)(o->nv + 2) * I2(j, k))
#define IRK(i, j, k, r, nvar)                                                  \
  ((size_t)((i)-1) +                                                           \
   (size_t)(nvar) * (I2(j, k) + (size_t)o->nx * o->ny * (size_t)((r)-1)))
#define IFX(j, k) ((size_t)((j)-1) + (size_t)o->nix * (size_t)((k)-1))
#define IFY(j, k) ((size_t)((j)-1) + (size_t)o->nx * (size_t)((k)-1))

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* geometry_2d: limiters (src/geometry_2d.f90:1555-1676)                    */
/* ------------------------------------------------------------------------ */
static double minmod(double a, double b) {
  /* src/geometry_2d.f90:1634-1655 */
  if (fmin(fabs(a), fabs(b)) <= 0.0)
    return 0.0;
  double sa = a / fabs(a), sb = b / fabs(b);
  return 0.5 * (sa + sb) * fmin(fabs(a), fabs(b));
}
static double maxmod(double a, double b) {
  /* src/geometry_2d.f90:1657-1676 */
  if (fabs(a * b) <= 0.0)
    return 0.0;
  double sa = a / fabs(a), sb = b / fabs(b);
  return 0.5 * (sa + sb) * fmax(fabs(a), fabs(b));
}
static double limit(const orc_t *o, const double v[3], const double z[3],
                    int lim) {
  /* src/geometry_2d.f90:1555-1620 */
  double a = (v[2] - v[1]) / (z[2] - z[1]);
  double b = (v[1] - v[0]) / (z[1] - z[0]);
  double c = (v[2] - v[0]) / (z[2] - z[0]);
  switch (lim) {
  case 0:
    return 0.0;
  case 1:
    return minmod(a, b);
  case 2: {
    double s1 = minmod(a, 2.0 * b), s2 = minmod(2.0 * a, b);
    return maxmod(s1, s2);
  }
  case 3:
    return minmod(c, o->p.theta * minmod(a, b));
  case 4:
    return minmod(c, 2.0 * minmod(a, b));
  case 5:
    return c;
  case 6:
    return a;
  case 7:
    return b;
  }
  return 0.0;
}
double orc_limit(const orc_t *o, const double *v, const double *z, int lim) {
  return limit(o, v, z, lim);
}

/* ------------------------------------------------------------------------ */
/* constitutive_2d                                                          */
/* ------------------------------------------------------------------------ */

/* vertical profiles (defined after settling_velocity) */
static double vp_velocity(const orc_t *o, double r_h, double q2, double q3,
                          double r_rho_m, double r_rho_c,
                          const double *r_alphas);
static void vp_momentum(const orc_t *o, double r_h, double r_u, double r_v,
                        double r_rho_c, const double *r_alphas, double *qc2,
                        double *qc3);
static void vp_flux_coeffs(const orc_t *o, const double *qpj, double Bx,
                           double By, double r_rho_c, double r_rho_m,
                           double *shape);
static void vp_dep_coeffs(const orc_t *o, const double *qpj, double mod_vel,
                          double r_rho_c, double *dep);

/* r_phys_var (src/constitutive_2d.f90:541-967), with the profile branch.
 * Outputs h, u, v, alphas(1:ns), rho_m, T, red_grav, p_dyn. */
static void r_phys_var(const orc_t *o, const double *qj, double *r_h,
                       double *r_u, double *r_v, double *r_alphas,
                       double *r_rho_m, double *r_T, double *r_red_grav,
                       double *p_dyn) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns;
  double r_xs[MAXS], inv_qj1;
  if (qj[0] > EPS_WP) {
    inv_qj1 = 1.0 / qj[0];
    for (int i = 0; i < ns; i++)
      r_xs[i] = qj[4 + i] * inv_qj1;
    double sq = 0.0;
    for (int i = 0; i < ns; i++)
      sq += qj[4 + i];
    if (sq == qj[0]) {
      double sx = 0.0;
      for (int i = 0; i < ns; i++)
        sx += r_xs[i];
      for (int i = 0; i < ns; i++)
        r_xs[i] = r_xs[i] / sx;
    }
  } else {
    *r_h = 0.0;
    *r_u = 0.0;
    *r_v = 0.0;
    for (int i = 0; i < ns; i++)
      r_alphas[i] = 0.0;
    *r_rho_m = o->rho_a_amb;
    *r_T = P->T_ambient;
    *r_red_grav = 0.0;
    *p_dyn = 0.0;
    return;
  }
  double r_xs_tot = 0.0;
  for (int i = 0; i < ns; i++)
    r_xs_tot += r_xs[i];
  /* gas carrier, no liquid, no additional gas */
  double r_xc = 1.0 - r_xs_tot;
  double r_sp_heat_c_by_xc = (r_xc - 0.0) * P->sp_heat_a + 0.0;
  double dsh = 0.0;
  for (int i = 0; i < ns; i++)
    dsh += r_xs[i] * P->sp_heat_s[i];
  double r_sp_heat_mix = dsh + r_sp_heat_c_by_xc;
  double T;
  if (qj[0] > o->eps_sing) {
    if (P->energy_flag)
      T = (qj[3] - 0.5 * (qj[1] * qj[1] + qj[2] * qj[2]) * inv_qj1) /
          (qj[0] * r_sp_heat_mix);
    else
      T = qj[3] / (qj[0] * r_sp_heat_mix);
    if (T <= 0.0)
      T = P->T_ambient;
  } else {
    T = P->T_ambient;
  }
  double r_sp_gas_const_c;
  if (r_xc > EPS_WP)
    r_sp_gas_const_c = ((r_xc - 0.0) * P->sp_gas_const_a + 0.0) / r_xc;
  else
    r_sp_gas_const_c = P->sp_gas_const_a;
  double r_inv_rho_c = r_sp_gas_const_c * T * o->inv_pres;
  double dinv = 0.0;
  for (int i = 0; i < ns; i++)
    dinv += r_xs[i] * o->inv_rho_s[i];
  double r_inv_rhom = dinv + r_xc * r_inv_rho_c;
  double rho_m = 1.0 / r_inv_rhom;
  for (int i = 0; i < ns; i++)
    r_alphas[i] = r_xs[i] * rho_m * o->inv_rho_s[i];
  *r_h = qj[0] * r_inv_rhom;
  *r_red_grav = (rho_m - o->rho_a_amb) * r_inv_rhom * P->grav;
  double u, v;
  if (o->vp) {
    /* profile branch (:797-942): the mean speed of the profiles */
    double r_rho_c = P->pres / (r_sp_gas_const_c * T);
    double ua = vp_velocity(o, *r_h, qj[1], qj[2], rho_m, r_rho_c, r_alphas);
    u = ua * qj[1] / (sqrt(qj[1] * qj[1] + qj[2] * qj[2]));
    v = ua * qj[2] / (sqrt(qj[1] * qj[1] + qj[2] * qj[2]));
  } else if (qj[0] > o->eps_sing) {
    u = qj[1] * inv_qj1;
    v = qj[2] * inv_qj1;
  } else {
    double q1sq = qj[0] * qj[0];
    u = sqrt(2.0) * qj[0] * qj[1] / sqrt(q1sq * q1sq + o->eps_sing4);
    v = sqrt(2.0) * qj[0] * qj[2] / sqrt(q1sq * q1sq + o->eps_sing4);
  }
  *r_u = u;
  *r_v = v;
  *r_rho_m = rho_m;
  *r_T = T;
  *p_dyn = 0.5 * rho_m * (u * u + v * v);
}

/* qc_to_qp (src/constitutive_2d.f90:1628-1681) */
static void qc_to_qp(const orc_t *o, const double *qc, double *qp,
                     double *p_dyn) {
  double h, u, v, al[MAXS], rho_m, T, rg;
  r_phys_var(o, qc, &h, &u, &v, al, &rho_m, &T, &rg, p_dyn);
  qp[0] = h;
  qp[1] = h * u;
  qp[2] = h * v;
  qp[3] = T;
  for (int i = 0; i < o->ns; i++)
    qp[4 + i] = o->p.alpha_flag ? al[i] : al[i] * h;
  qp[o->idx_u - 1] = u;
  qp[o->idx_v - 1] = v;
}
void orc_qc_to_qp(const orc_t *o, const double *qc, double *qp, double *pd) {
  double d;
  qc_to_qp(o, qc, qp, pd ? pd : &d);
}

/* eval_sp_heat, no-liquid branch (src/constitutive_2d.f90:1504-1598) */
static void eval_sp_heat(const orc_t *o, const double *r_alphas,
                         double r_inv_rhom, double *sp_heat_c,
                         double *sp_heat_mix) {
  const imexsw_params *P = &o->p;
  double r_xs[MAXS], sxs = 0.0;
  for (int i = 0; i < o->ns; i++)
    r_xs[i] = r_alphas[i] * P->rho_s[i] * r_inv_rhom;
  for (int i = 0; i < o->ns; i++)
    sxs += r_xs[i];
  double r_xc = 1.0 - sxs;
  double c;
  if (r_xc > EPS_WP)
    c = ((r_xc - 0.0) * P->sp_heat_a + 0.0) / r_xc;
  else
    c = P->sp_heat_a;
  double d = 0.0;
  for (int i = 0; i < o->ns; i++)
    d += r_xs[i] * P->sp_heat_s[i];
  *sp_heat_c = c;
  *sp_heat_mix = d + r_xc * c;
}

/* mixt_var (src/constitutive_2d.f90:1333-1483) */
static void mixt_var(const orc_t *o, const double *qpj, double *r_Ri,
                     double *r_rho_m, double *r_rho_c, double *r_red_grav,
                     int sp_heat_flag, double *r_sp_heat_c,
                     double *r_sp_heat_mix) {
  const imexsw_params *P = &o->p;
  double r_h = qpj[0];
  if (qpj[0] <= EPS_WP) {
    *r_red_grav = 0.0;
    *r_rho_m = o->rho_a_amb;
    *r_rho_c = o->rho_a_amb;
    *r_sp_heat_c = P->sp_heat_a;
    *r_sp_heat_mix = P->sp_heat_a;
    *r_Ri = 0.0;
    return;
  }
  double r_u = qpj[o->idx_u - 1], r_v = qpj[o->idx_v - 1], r_T = qpj[3];
  double r_alphas[MAXS];
  for (int i = 0; i < o->ns; i++)
    r_alphas[i] = P->alpha_flag ? qpj[4 + i] : qpj[4 + i] / qpj[0];
  double alphas_tot = 0.0;
  for (int i = 0; i < o->ns; i++)
    alphas_tot += r_alphas[i];
  double r_alphal = 0.0;
  double r_rho_a = P->pres / (P->sp_gas_const_a * r_T);
  double rho_c = ((1.0 - r_alphal - alphas_tot - 0.0) * r_rho_a + 0.0) /
                 (1.0 - r_alphal - alphas_tot);
  double d = 0.0;
  for (int i = 0; i < o->ns; i++)
    d += r_alphas[i] * P->rho_s[i];
  double rho_m = (1.0 - alphas_tot) * rho_c + d;
  double r_inv_rhom = 1.0 / rho_m;
  double rg = (rho_m - o->rho_a_amb) / rho_m * P->grav;
  double ri;
  if ((r_u * r_u + r_v * r_v) > 0.0)
    ri = fmin(1.0e15, rg * r_h / (r_u * r_u + r_v * r_v));
  else
    ri = 0.0;
  *r_Ri = ri;
  *r_rho_m = rho_m;
  *r_rho_c = rho_c;
  *r_red_grav = rg;
  if (sp_heat_flag)
    eval_sp_heat(o, r_alphas, r_inv_rhom, r_sp_heat_c, r_sp_heat_mix);
}
void orc_mixt_var(const orc_t *o, const double *qp, double *out6) {
  mixt_var(o, qp, &out6[0], &out6[1], &out6[2], &out6[3], 1, &out6[4],
           &out6[5]);
}

/* qp_to_qc, no-liquid branch (src/constitutive_2d.f90:1709-2140) */
static void qp_to_qc(const orc_t *o, const double *qp, double *qc) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns;
  double r_h = qp[0], r_hu, r_hv, r_u, r_v;
  if (r_h > EPS_WP) {
    r_hu = qp[1];
    r_hv = qp[2];
    r_u = qp[o->idx_u - 1];
    r_v = qp[o->idx_v - 1];
  } else {
    for (int i = 0; i < o->nv; i++)
      qc[i] = 0.0;
    return;
  }
  double r_T = qp[3];
  double r_alphas[MAXS], r_alphas_rhos[MAXS], r_xs[MAXS];
  for (int i = 0; i < ns; i++)
    r_alphas[i] = P->alpha_flag ? qp[4 + i] : qp[4 + i] / qp[0];
  double alphas_tot = 0.0;
  for (int i = 0; i < ns; i++)
    alphas_tot += r_alphas[i];
  for (int i = 0; i < ns; i++)
    r_alphas_rhos[i] = r_alphas[i] * P->rho_s[i];
  double r_alphal = 0.0;
  double r_rho_a = P->pres / (P->sp_gas_const_a * r_T);
  double r_rho_c = ((1.0 - r_alphal - alphas_tot - 0.0) * r_rho_a + 0.0) /
                   (1.0 - r_alphal - alphas_tot);
  /* the alphas corrections of :1911-1919 reach only the profile branch */
  double r_alphac = 1.0 - alphas_tot;
  double sar = 0.0;
  for (int i = 0; i < ns; i++)
    sar += r_alphas_rhos[i];
  double r_rho_m = r_alphac * r_rho_c + sar;
  double r_inv_rhom = 1.0 / r_rho_m;
  for (int i = 0; i < ns; i++)
    r_xs[i] = r_alphas_rhos[i] * r_inv_rhom;
  double r_xc = r_alphac * r_rho_c * r_inv_rhom;
  double r_sp_heat_c;
  if (r_xc > EPS_WP)
    r_sp_heat_c = ((r_xc - 0.0) * P->sp_heat_a + 0.0) / r_xc;
  else
    r_sp_heat_c = P->sp_heat_a;
  double dsh = 0.0;
  for (int i = 0; i < ns; i++)
    dsh += r_xs[i] * P->sp_heat_s[i];
  double r_sp_heat_mix = dsh + r_xc * r_sp_heat_c;
  qc[0] = r_rho_m * r_h;
  if (o->vp) {
    /* the corrections of :1911-1919, which only the profiles read */
    double ra[MAXS];
    for (int i = 0; i < ns; i++)
      ra[i] = (alphas_tot > 1.0)   ? r_alphas[i] / alphas_tot
              : (alphas_tot < 0.0) ? 0.0
                                   : r_alphas[i];
    vp_momentum(o, r_h, r_u, r_v, r_rho_c, ra, &qc[1], &qc[2]);
  } else {
    qc[1] = r_rho_m * r_hu;
    qc[2] = r_rho_m * r_hv;
  }
  if (P->energy_flag) {
    if (r_h > 0.0)
      qc[3] = r_h * r_rho_m * (r_sp_heat_mix * r_T + 0.5 * (r_u * r_u + r_v * r_v));
    else
      qc[3] = 0.0;
  } else {
    qc[3] = r_h * r_rho_m * r_sp_heat_mix * r_T;
  }
  for (int i = 0; i < ns; i++)
    qc[4 + i] = r_xs[i] * qc[0];
}
void orc_qp_to_qc(const orc_t *o, const double *qp, double *qc) {
  qp_to_qc(o, qp, qc);
}

/* qp_to_qp2 (src/constitutive_2d.f90:2155-2179) */
static void qp_to_qp2(const double *qpj, double Bj, double *qp2) {
  qp2[0] = qpj[0] + Bj;
  if (qpj[0] <= 0.0) {
    qp2[1] = 0.0;
    qp2[2] = 0.0;
  } else {
    qp2[1] = qpj[1] / qpj[0];
    qp2[2] = qpj[2] / qpj[0];
  }
}

/* eval_local_speeds_x / _y (src/constitutive_2d.f90:2194-2298) */
static void local_speeds(const orc_t *o, const double *qpj, double gc,
                         int dir, double *vmin, double *vmax) {
  double ri, rho_m, rho_c, rg, c1, c2;
  mixt_var(o, qpj, &ri, &rho_m, &rho_c, &rg, 0, &c1, &c2);
  double r_h = qpj[0];
  double vel = (dir == 1) ? qpj[o->idx_u - 1] : qpj[o->idx_v - 1];
  if (rg * r_h < 0.0) {
    *vmin = vel;
    *vmax = vel;
  } else {
    double cel = (dir == 1) ? sqrt(rg * r_h * gc) : sqrt(gc * rg * r_h);
    *vmin = vel - cel;
    *vmax = vel + cel;
  }
}
void orc_local_speeds(const orc_t *o, const double *qp, double gc, int dir,
                      double *mm) {
  local_speeds(o, qp, gc, dir, &mm[0], &mm[1]);
}

/* eval_fluxes (src/constitutive_2d.f90:2316-2509); Bx, By: the slope of
 * the cell the face state comes from (read by the profile coefficients) */
static void eval_fluxes(const orc_t *o, const double *qcj, const double *qpj,
                        double Bx, double By, double gc, int dir,
                        double *flux) {
  const int ns = o->ns;
  if (qpj[0] > EPS_WP) {
    double r_h = qpj[0], r_u = qpj[o->idx_u - 1], r_v = qpj[o->idx_v - 1];
    double ri, rho_m, rho_c, rg, c1, c2;
    mixt_var(o, qpj, &ri, &rho_m, &rho_c, &rg, 0, &c1, &c2);
    double sc[MAXV];
    for (int i = 0; i < o->ne; i++)
      sc[i] = 1.0;
    if (o->vp)
      vp_flux_coeffs(o, qpj, Bx, By, rho_c, rho_m, sc);
    double vel = (dir == 1) ? r_u : r_v;
    double pres_term = 0.5 * rho_m * gc * rg * (r_h * r_h);
    flux[0] = vel * qcj[0] * sc[0];
    if (dir == 1) {
      flux[1] = r_u * qcj[1] * sc[1] + pres_term;
      flux[2] = r_u * qcj[2] * sc[2];
    } else {
      flux[1] = r_v * qcj[1] * sc[1];
      flux[2] = r_v * qcj[2] * sc[2] + pres_term;
    }
    if (o->p.energy_flag)
      flux[3] = vel * (qcj[3] * sc[3] + pres_term);
    else
      flux[3] = vel * qcj[3] * sc[3];
    for (int i = 0; i < ns; i++)
      flux[4 + i] = vel * qcj[4 + i] * sc[4 + i];
    double sf = 0.0;
    for (int i = 0; i < ns; i++)
      sf += flux[4 + i];
    if ((flux[0] > 0.0) && (sf / flux[0] > 1.0)) {
      for (int i = 0; i < ns; i++)
        flux[4 + i] = flux[4 + i] / sf * flux[0];
    }
  } else {
    for (int i = 0; i < o->ne; i++)
      flux[i] = 0.0;
  }
}
void orc_eval_fluxes(const orc_t *o, const double *qc, const double *qp,
                     double gc, int dir, double *flux) {
  eval_fluxes(o, qc, qp, 0.0, 0.0, gc, dir, flux);
}

/* eval_expl_terms, no bottom source (src/constitutive_2d.f90:3015-3245) */
static void eval_expl_terms(const orc_t *o, double Bx, double By, double Bxx,
                            double Bxy, double Byy, double gc, double dgdx,
                            double dgdy, const double *qpj, double *expl) {
  const imexsw_params *P = &o->p;
  for (int i = 0; i < o->ne; i++)
    expl[i] = 0.0;
  double cell_fract = 0.0;
  if ((qpj[0] <= EPS_WP) && (cell_fract == 0.0))
    return;
  double r_h = qpj[0], r_u = qpj[o->idx_u - 1], r_v = qpj[o->idx_v - 1];
  double ri, rho_m, rho_c, rg, shc, shm;
  mixt_var(o, qpj, &ri, &rho_m, &rho_c, &rg, 1, &shc, &shm);
  double cft;
  if (P->curvature_term_flag)
    cft = Bxx * (r_u * r_u) + 2.0 * Bxy * r_u * r_v + Byy * (r_v * r_v);
  else
    cft = 0.0;
  double tg = rg + cft;
  expl[1] = -(gc * rho_m * tg * r_h * Bx) +
            0.5 * rho_m * rg * (r_h * r_h) * dgdx;
  expl[2] = -(gc * rho_m * tg * r_h * By) +
            0.5 * rho_m * rg * (r_h * r_h) * dgdy;
  if (P->energy_flag)
    expl[3] = expl[1] * r_u + expl[2] * r_v;
  else
    expl[3] = 0.0;
}

/* c_phys_var reduced to the outputs the Voellmy term reads (rho_m, u, v)
 * (src/constitutive_2d.f90:1145-1313) */
static void c_phys_var(const orc_t *o, const cx *cq, cx *rho_m, cx *u, cx *v) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns;
  cx xs[MAXS], inv_cqj1;
  if (cq[0].re > EPS_WP) {
    inv_cqj1 = cdiv(C(1.0, 0.0), cq[0]);
    for (int i = 0; i < ns; i++)
      xs[i] = cmul(cq[4 + i], inv_cqj1);
  } else {
    *u = C(0.0, 0.0);
    *v = C(0.0, 0.0);
    *rho_m = C(o->rho_a_amb, 0.0);
    return;
  }
  cx xs_tot = C(0.0, 0.0);
  for (int i = 0; i < ns; i++)
    xs_tot = cadd(xs_tot, xs[i]);
  cx xc = csub(C(1.0, 0.0), xs_tot);
  cx sp_heat_c = cdiv(cadd(cmulr(csub(xc, C(0.0, 0.0)), P->sp_heat_a),
                           C(0.0, 0.0)),
                      xc);
  cx dsh = C(0.0, 0.0);
  for (int i = 0; i < ns; i++)
    dsh = cadd(dsh, cmul(cconj(xs[i]), C(P->sp_heat_s[i], 0.0)));
  cx sp_heat_mix = cadd(dsh, cmul(xc, sp_heat_c));
  cx T;
  if (cq[0].re > o->eps_sing) {
    if (P->energy_flag) {
      /* 0.5_wp * (q2**2 + q3**2) * inv_cqj1: real*complex, then complex */
      cx ke = cmul(cmulr(cadd(cmul(cq[1], cq[1]), cmul(cq[2], cq[2])), 0.5),
                   inv_cqj1);
      T = cdiv(csub(cq[3], ke), cmul(cq[0], sp_heat_mix));
    } else {
      T = cdiv(cq[3], cmul(cq[0], sp_heat_mix));
    }
    if (T.re <= 0.0)
      T = C(P->T_ambient, 0.0);
  } else {
    T = C(P->T_ambient, 0.0);
  }
  cx sp_gas_const_c = cdiv(cadd(cmulr(csub(xc, C(0.0, 0.0)),
                                      P->sp_gas_const_a),
                                C(0.0, 0.0)),
                           xc);
  cx inv_rho_c = cmulr(cmul(sp_gas_const_c, T), o->inv_pres);
  cx dinv = C(0.0, 0.0);
  for (int i = 0; i < ns; i++)
    dinv = cadd(dinv, cmul(cconj(xs[i]), C(o->inv_rho_s[i], 0.0)));
  cx inv_rhom = cadd(dinv, cmul(xc, inv_rho_c));
  *rho_m = cdiv(C(1.0, 0.0), inv_rhom);
  if (cq[0].re > o->eps_sing) {
    *u = cmul(cq[1], inv_cqj1);
    *v = cmul(cq[2], inv_cqj1);
  } else {
    cx q1_2 = cmul(cq[0], cq[0]);
    cx q1_4 = cmul(q1_2, q1_2);
    cx den = csqrt_f(cadd(q1_4, C(o->eps_sing4, 0.0)));
    *u = cdiv(cmul(cmulr(cq[0], sqrt(2.0)), cq[1]), den);
    *v = cdiv(cmul(cmulr(cq[0], sqrt(2.0)), cq[2]), den);
  }
}

/* eval_implicit_terms, Voellmy (model 1, :3474-3497) and turbulent-drag
 * (model 6, :3639-3650) branches (src/constitutive_2d.f90:3327-3823);
 * complex in, complex out.  The real-argument call
 * of the reference goes through the same complex arithmetic with zero
 * imaginary parts (:3431-3441). */
static void implicit_terms_cs(const orc_t *o, const cx *cq, double Bx,
                              double By, cx *nh) {
  const imexsw_params *P = &o->p;
  for (int i = 0; i < o->ne; i++)
    nh[i] = C(0.0, 0.0);
  if (!P->rheology_flag)
    return;
  cx rho_m, u, v;
  c_phys_var(o, cq, &rho_m, &u, &v);
  cx red_grav = cmulr(cdiv(csub(rho_m, C(o->rho_a_amb, 0.0)), rho_m), P->grav);
  cx w = P->slope_correction_flag ? cadd(cmulr(u, Bx), cmulr(v, By))
                                   : C(0.0, 0.0);
  cx mod_vel2 = cadd(cadd(cmul(u, u), cmul(v, v)), cmul(w, w));
  cx mod_vel = csqrt_f(mod_vel2);
  if (P->rheology_model == 1) {
    if (mod_vel.re != 0.0) {
      cx turb = cmul(cdivr(cmul(rho_m, red_grav), P->xi), mod_vel2);
      nh[1] = csub(nh[1], cmul(turb, cdiv(u, mod_vel)));
      nh[2] = csub(nh[2], cmul(turb, cdiv(v, mod_vel)));
    }
  } else if (P->rheology_model == 6) {
    /* turbulent drag rho_m f |u|^2 along the horizontal velocity
     * (:3639-3650); the real factor multiplies both components */
    cx mod_vel_hor = csqrt_f(cadd(cmul(u, u), cmul(v, v)));
    if (mod_vel.re != 0.0) {
      nh[1] = csub(nh[1], cmul(cmulr(cmul(rho_m, cdiv(u, mod_vel_hor)),
                                     P->friction_factor),
                               mod_vel2));
      nh[2] = csub(nh[2], cmul(cmulr(cmul(rho_m, cdiv(v, mod_vel_hor)),
                                     P->friction_factor),
                               mod_vel2));
    }
  }
}

static void implicit_terms_r(const orc_t *o, const double *q, double Bx,
                             double By, double *nh) {
  cx cq[MAXV], cn[MAXV];
  for (int i = 0; i < o->nv; i++)
    cq[i] = C(q[i], 0.0);
  implicit_terms_cs(o, cq, Bx, By, cn);
  for (int i = 0; i < o->ne; i++)
    nh[i] = cn[i].re;
}
void orc_implicit_terms(const orc_t *o, const double *q, double *nh) {
  implicit_terms_r(o, q, 0.0, 0.0, nh);
}
void orc_implicit_terms_cstep(const orc_t *o, const double *q, int col,
                              double h, double *re, double *im) {
  cx cq[MAXV], cn[MAXV];
  for (int i = 0; i < o->nv; i++)
    cq[i] = C(q[i], 0.0);
  cq[col] = C(q[col], h);
  implicit_terms_cs(o, cq, 0.0, 0.0, cn);
  for (int i = 0; i < o->ne; i++) {
    re[i] = cn[i].re;
    im[i] = cn[i].im;
  }
}

/* eval_nh_semi_impl_terms, Voellmy branch (src/constitutive_2d.f90:3841-4397,
 * model 1 at :3979-4028) */
static void semi_impl_terms(const orc_t *o, double Bx, double By, double Bxx,
                            double Bxy, double Byy, double gc,
                            const double *qcj, const double *qpj,
                            double *out) {
  const imexsw_params *P = &o->p;
  (void)qcj;
  for (int i = 0; i < o->ne; i++)
    out[i] = 0.0;
  if (!P->rheology_flag)
    return;
  double r_h = qpj[0], r_u = qpj[o->idx_u - 1], r_v = qpj[o->idx_v - 1];
  double ri, rho_m, rho_c, rg, c1, c2;
  mixt_var(o, qpj, &ri, &rho_m, &rho_c, &rg, 0, &c1, &c2);
  double r_w = P->slope_correction_flag ? r_u * Bx + r_v * By : 0.0;
  double mod_vel = sqrt(r_u * r_u + r_v * r_v + r_w * r_w);
  if (P->rheology_model == 1) {
    if (mod_vel > 0.0) {
      double cft;
      if (P->curvature_term_flag)
        cft = Bxx * (r_u * r_u) + 2.0 * Bxy * r_u * r_v + Byy * (r_v * r_v);
      else
        cft = 0.0;
      double tt = P->mu * rho_m * (rg + cft) * r_h * sqrt(gc);
      tt = fmax(0.0, tt);
      tt = fmax(0.0, tt);
      out[1] = out[1] - tt * r_u / mod_vel;
      out[2] = out[2] - tt * r_v / mod_vel;
    }
  }
}
void orc_semi_impl_terms(const orc_t *o, const double *qc, const double *qp,
                         double gc, double *nh) {
  semi_impl_terms(o, 0.0, 0.0, 0.0, 0.0, 0.0, gc, qc, qp, nh);
}

/* settling_velocity (src/constitutive_2d.f90:5148-5214) */
static double settling_velocity(const orc_t *o, double diam, double rhos,
                                double rhoc, double inv_kin_visc) {
  double inv_sqrt_C_D = 1.0;
  if (rhos <= rhoc)
    return 0.0;
  double const_part = sqrt(0.75 * (rhos / rhoc - 1.0) * diam * o->p.grav);
  double sv = const_part * inv_sqrt_C_D;
  double Rey = diam * sv * inv_kin_visc;
  if (Rey <= 1000.0) {
    for (int i = 1; i <= 20; i++) {
      double old = sv;
      inv_sqrt_C_D = sqrt(Rey / (24.0 * (1.0 + 0.15 * det_pow(Rey, 0.687))));
      sv = const_part * inv_sqrt_C_D;
      if (fabs(old - sv) / old < 1.0e-6)
        return sv;
      Rey = diam * sv * inv_kin_visc;
    }
  }
  return sv;
}
double orc_settling_velocity(const orc_t *o, double diam, double rhos,
                             double rhoc, double ikv) {
  return settling_velocity(o, diam, rhos, rhoc, ikv);
}

/* ------------------------------------------------------------------------ */
/* vertical profiles (VERTICAL_PROFILES_FLAG)                               */
/* Velocity u(z) ~ log(30 z / k_s + 1) up to a boundary-layer top h0 and    */
/* constant above it; solid concentration ~ exp(a z) with a = -w_s /        */
/* epsilon_s (Rouse).  Depth averages use n_quad Gauss-Legendre points on   */
/* [0, h0] plus the constant part on [h0, h].                               */
/* ------------------------------------------------------------------------ */

/* Series of W about the branch point -1/e in p = +-sqrt(2 (e x + 1)):
 * W = sum mu_k p^k (Corless, Gonnet, Hare, Jeffrey & Knuth 1996, "On the
 * Lambert W function", eqs. 4.22-4.24; the rationals follow from their
 * recurrence).  Used for |p| < 0.1, where the k = 13 term is below 1e-16. */
static const double kWmu[13] = {
    -1.0, 1.0, -1.0 / 3.0, 11.0 / 72.0, -43.0 / 540.0, 769.0 / 17280.0,
    -221.0 / 8505.0, 680863.0 / 43545600.0, -1963.0 / 204120.0,
    226287557.0 / 37623398400.0, -5776369.0 / 1515591000.0,
    169709463197.0 / 69528040243200.0, -1118511313.0 / 709296588000.0};
#define W_E 2.718281828459045 /* e rounded to double */

static double lw_series(double p) {
  double s = kWmu[12];
  for (int k = 11; k >= 0; k--)
    s = s * p + kWmu[k];
  return s;
}

/* Halley's iteration for w e^w = x (Corless et al. 1996, eq. 5.9) from a
 * starting point on the wanted branch; stops when the step is below
 * 1e-15 |w| (the iteration is cubic, so the last step already was). */
static double lw_halley(double x, double w) {
  for (int it = 0; it < 20; it++) {
    double ew = det_exp(w);
    double f = w * ew - x;
    double wp1 = w + 1.0;
    double dw = f / (ew * wp1 - (w + 2.0) * f / (2.0 * wp1));
    w = w - dw;
    if (!(fabs(dw) > 1.0e-15 * fabs(w)))
      break;
  }
  return w;
}

/* W_0 (principal branch) for x >= -1/e.  Outside the domain (x < -1/e)
 * the value is -1, as lambertw0 returns it (src/geometry_2d.f90:1900-1915). */
double orc_lambertw0(double x) {
  double p2 = 2.0 * (W_E * x + 1.0);
  if (!(p2 > 0.0))
    return -1.0;
  if (p2 < 0.01)
    return lw_series(sqrt(p2));
  double w;
  if (x < -0.25) {
    double p = sqrt(p2);
    w = -1.0 + p * (1.0 + p * (-1.0 / 3.0 + p * (11.0 / 72.0)));
  } else {
    /* Winitzki (2003): W ~ L (1 - log(1 + L) / (2 + L)), L = log(1 + x) */
    double L = det_log(1.0 + x);
    w = (x == 0.0) ? 0.0 : L * (1.0 - det_log(1.0 + L) / (2.0 + L));
  }
  return lw_halley(x, w);
}

/* W_{-1} for -1/e <= x < 0.  Outside the domain, and for the arguments
 * above -64 e^-64 that lambertwm1 declares too small, the value is -1 as
 * there (src/geometry_2d.f90:1985-2010). */
double orc_lambertwm1(double x) {
  double p2 = 2.0 * (W_E * x + 1.0);
  if (!(p2 > 0.0) || !(x < 0.0) || x > -1.0288423856755186e-26)
    return -1.0;
  if (p2 < 0.01)
    return lw_series(-sqrt(p2));
  double w;
  if (x < -0.25) {
    double p = -sqrt(p2);
    w = -1.0 + p * (1.0 + p * (-1.0 / 3.0 + p * (11.0 / 72.0)));
  } else {
    /* asymptotic start W ~ L1 - L2 + L2 / L1, L1 = log(-x), L2 = log(-L1) */
    double L1 = det_log(-x), L2 = det_log(-L1);
    w = L1 - L2 + L2 / L1;
  }
  return lw_halley(x, w);
}

/* Gauss-Legendre points and weights on [-1, 1] in ascending order (the
 * layout gaulegf returns, src/geometry_2d.f90:2407-2449): Newton's method
 * on P_n from the starting points cos(pi (i - 1/4) / (n + 1/2)), P_n and
 * P_n' by the three-term recurrence, weights 2 / ((1 - x^2) P_n'(x)^2). */
void orc_gauss_legendre(int n, double *x, double *w) {
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(pi * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; it++) {
      double pm = 1.0, pc = z; /* P_0, P_1 */
      for (int m = 2; m <= n; m++) {
        double pn = ((2.0 * m - 1.0) * z * pc - (m - 1.0) * pm) / m;
        pm = pc;
        pc = pn;
      }
      if (n == 1) {
        pc = z;
        pm = 1.0;
      }
      dp = n * (z * pc - pm) / (z * z - 1.0);
      double dz = pc / dp;
      z = z - dz;
      if (!(fabs(dz) > 2.220446049250313e-15))
        break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    w[n - 1 - i] = w[i];
  }
}

/* The boundary-layer top h0 of a layer of thickness r_h: for
 * h/k_s > H_crit_rel the smaller root of b log(c z + 1) - z = a
 * (z = -b W(-e^d / (b c)) - 1/c on both real branches), else h
 * (src/constitutive_2d.f90:821-854, 2011-2044, 2697-2730, 2929-2958). */
static double vp_h0(const orc_t *o, double r_h) {
  const imexsw_params *P = &o->p;
  double h_rel = r_h / P->k_s, h0_rel;
  if (h_rel > o->h_crit_rel) {
    double a = h_rel * P->vonK / o->sqrt_ff;
    double b = 1.0 / 30.0 + h_rel, c = 30.0;
    double d = a / b - 1.0 / (b * c);
    double arg = -det_exp(d) / (b * c);
    double h1 = -(b * orc_lambertw0(arg)) - 1.0 / c;
    double h2 = -(b * orc_lambertwm1(arg)) - 1.0 / c;
    h0_rel = fmin(h1, h2);
  } else {
    h0_rel = h_rel;
  }
  return h0_rel * P->k_s;
}
double orc_vp_h0(const orc_t *o, double h) { return vp_h0(o, h); }

/* the quadrature of [0, h0] and log(b z + 1) at its points */
static void vp_quad(const orc_t *o, double h0, double b, double *z,
                    double *w, double *lz) {
  for (int i = 0; i < o->nq; i++) {
    z[i] = 0.5 * h0 * (o->zq[i] + 1.0);
    w[i] = 0.5 * h0 * o->wq[i];
    lz[i] = det_log(b * z[i] + 1.0);
  }
}

/* depth average of the log profile: (int_0^h0 log(bz+1) + log(b h0+1)(h-h0))/h */
static double vp_u_log_avg(const orc_t *o, const double *w, const double *lz,
                           double lh0, double h, double h0) {
  double s = 0.0;
  for (int i = 0; i < o->nq; i++)
    s += w[i] * lz[i];
  return (s + lh0 * (h - h0)) / h;
}

/* Rouse numbers w_s / (vonK u*) with the settling velocity in a carrier of
 * density rho_c (zero without shear) */
static void vp_rouse(const orc_t *o, double shear_vel, double rho_c,
                     double *rouse) {
  const imexsw_params *P = &o->p;
  double inv_kin_visc = 1.0 / P->kin_visc_c;
  for (int i = 0; i < o->ns; i++) {
    double sv = settling_velocity(o, P->diam_s[i], P->rho_s[i], rho_c,
                                  inv_kin_visc);
    rouse[i] = (shear_vel > 0.0) ? sv / (P->vonK * shear_vel) : 0.0;
  }
}

/* avg_profiles_mix (src/constitutive_2d.f90:970-1123): the depth-averaged
 * mixture momentum per unit depth of a flow with mean velocity u_guess and
 * depth-averaged solid partial densities rho_alphas_avg.  Its dynamic
 * pressure output is not returned: r_phys_var overwrites it (:957). */
static double vp_avg_mix(const orc_t *o, double h, const double *sv,
                         const double *rho_alphas_avg, double u_guess,
                         double h0, double b, double u_rel0, double rho_c) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns, nq = o->nq;
  double shear_vel = u_guess * o->sqrt_ff;
  double rouse[MAXS];
  for (int i = 0; i < ns; i++)
    rouse[i] = (shear_vel > 0.0) ? sv[i] / (P->vonK * shear_vel) : 0.0;
  double z[IMEXSW_MAX_QUAD], w[IMEXSW_MAX_QUAD], lz[IMEXSW_MAX_QUAD];
  vp_quad(o, h0, b, z, w, lz);
  double eps_s = P->Sc * shear_vel * P->vonK * ((h0 / 6.0) + (P->k_s / 60.0));
  double a_coeff = -(P->vonK * shear_vel / eps_s);
  double lh0 = det_log(b * h0 + 1.0);
  double ncu = 1.0 / vp_u_log_avg(o, w, lz, lh0, h, h0);
  double rua[MAXS];
  for (int i = 0; i < ns; i++) {
    double a = a_coeff * rouse[i];
    double se = 0.0, sq = 0.0;
    for (int m = 0; m < nq; m++)
      se += w[m] * det_exp(a * z[m]);
    double eh0 = det_exp(a * h0);
    double nca = 1.0 / ((se + eh0 * (h - h0)) / h);
    for (int m = 0; m < nq; m++)
      sq += w[m] * (det_exp(a * z[m]) * lz[m]);
    double d1 = ncu * nca * sq;
    double ar0 = nca * eh0;
    double d2 = (h - h0) * u_rel0 * ar0;
    rua[i] = rho_alphas_avg[i] * u_guess * (d1 + d2) / h;
  }
  double s = 0.0;
  for (int i = 0; i < ns; i++)
    s += (P->rho_s[i] - rho_c) / P->rho_s[i] * rua[i];
  return u_guess * rho_c + s;
}

/* The mean speed of r_phys_var's profile branch (src/constitutive_2d.f90:
 * 797-942): u_avg such that the profile-averaged momentum of a flow with
 * mean speed u_avg is |(q2, q3)| / h, found by the fixed-point map
 * u <- u |q| / (h M(u)) with Aitken's delta-squared acceleration, at most
 * ten rounds, tolerances 1e-8 (single-precision literals in the reference). */
static double vp_velocity(const orc_t *o, double r_h, double q2, double q3,
                          double r_rho_m, double r_rho_c,
                          const double *r_alphas) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns;
  double sv[MAXS], ra[MAXS];
  double inv_kin_visc = 1.0 / P->kin_visc_c;
  for (int i = 0; i < ns; i++) {
    ra[i] = r_alphas[i] * P->rho_s[i];
    sv[i] = settling_velocity(o, P->diam_s[i], P->rho_s[i], r_rho_c,
                              inv_kin_visc);
  }
  double h0 = vp_h0(o, r_h);
  double b = 30.0 / P->k_s;
  double z[IMEXSW_MAX_QUAD], w[IMEXSW_MAX_QUAD], lz[IMEXSW_MAX_QUAD];
  vp_quad(o, h0, b, z, w, lz);
  double lh0 = det_log(b * h0 + 1.0);
  double ncu = 1.0 / vp_u_log_avg(o, w, lz, lh0, r_h, h0);
  double u_rel0 = ncu * lh0;
  double uRho_avg = sqrt(q2 * q2 + q3 * q3) / r_h;
  double u_guess = uRho_avg / r_rho_m, u_new = u_guess;
  const double rel_tol = (double)1.e-8f, abs_tol = (double)1.e-8f;
  for (int it = 0; it < 10; it++) {
    double x0 = u_guess;
    double m = vp_avg_mix(o, r_h, sv, ra, u_guess, h0, b, u_rel0, r_rho_c);
    u_new = u_guess * uRho_avg / m;
    double x1 = u_new;
    m = vp_avg_mix(o, r_h, sv, ra, u_new, h0, b, u_rel0, r_rho_c);
    u_new = u_new * uRho_avg / m;
    double x2 = u_new;
    double den = (x2 - x1) - (x1 - x0);
    if (fabs(den) < (double)0.1f * abs_tol)
      break;
    u_new = x2 - ((x2 - x1) * (x2 - x1)) / den;
    if ((fabs(u_guess - u_new) / u_guess < rel_tol) ||
        (fabs(u_guess - u_new) < abs_tol))
      break;
    u_guess = u_new;
  }
  return u_new;
}

/* qp_to_qc's profile branch (src/constitutive_2d.f90:1976-2107): the
 * momentum of a flow of mean velocity (r_u, r_v) with the log velocity and
 * exponential concentration profiles */
static void vp_momentum(const orc_t *o, double r_h, double r_u, double r_v,
                        double r_rho_c, const double *r_alphas, double *qc2,
                        double *qc3) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns, nq = o->nq;
  double mod_vel = sqrt(r_u * r_u + r_v * r_v + 0.0 * 0.0);
  double shear_vel = o->sqrt_ff * mod_vel;
  double rouse[MAXS];
  vp_rouse(o, shear_vel, r_rho_c, rouse);
  double h0 = vp_h0(o, r_h);
  double b = 30.0 / P->k_s;
  double eps_s = P->Sc * shear_vel * P->vonK * ((h0 / 6.0) + (P->k_s / 60.0));
  double a_coeff = -(P->vonK * shear_vel / eps_s);
  double z[IMEXSW_MAX_QUAD], w[IMEXSW_MAX_QUAD], lz[IMEXSW_MAX_QUAD];
  vp_quad(o, h0, b, z, w, lz);
  double lh0 = det_log(b * h0 + 1.0);
  double ncu = 1.0 / vp_u_log_avg(o, w, lz, lh0, r_h, h0);
  double u_rel0 = ncu * lh0;
  double rua[MAXS];
  for (int i = 0; i < ns; i++) {
    double a = a_coeff * rouse[i];
    double se = 0.0, sq = 0.0;
    for (int m = 0; m < nq; m++)
      se += w[m] * det_exp(a * z[m]);
    double eh0 = det_exp(a * h0);
    double nca = 1.0 / ((se + eh0 * (r_h - h0)) / r_h);
    for (int m = 0; m < nq; m++)
      sq += w[m] * (det_exp(a * z[m]) * lz[m]);
    double d1 = (mod_vel * ncu) * (r_alphas[i] * nca) * sq;
    double ar0 = nca * eh0;
    double d2 = (r_h - h0) * (mod_vel * u_rel0) * (ar0 * r_alphas[i]);
    rua[i] = P->rho_s[i] * (d1 + d2) / r_h;
  }
  double s = 0.0;
  for (int i = 0; i < ns; i++)
    s += (P->rho_s[i] - r_rho_c) / P->rho_s[i] * rua[i];
  double uRho_avg = mod_vel * r_rho_c + s;
  *qc2 = r_h * uRho_avg * r_u / mod_vel;
  *qc3 = r_h * uRho_avg * r_v / mod_vel;
}

/* eval_flux_coeffs (src/constitutive_2d.f90:2532-2823): the ratios of the
 * profile-integrated advective fluxes to those of uniform profiles, per
 * equation (mass, momentum, temperature, solids) */
static void vp_flux_coeffs(const orc_t *o, const double *qpj, double Bx,
                           double By, double r_rho_c, double r_rho_m,
                           double *shape) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns, nq = o->nq;
  for (int i = 0; i < o->ne; i++)
    shape[i] = 1.0;
  double r_h = qpj[0], r_hu = qpj[1], r_hv = qpj[2];
  double r_u = qpj[o->idx_u - 1], r_v = qpj[o->idx_v - 1];
  double r_alphas[MAXS];
  for (int i = 0; i < ns; i++)
    r_alphas[i] = P->alpha_flag ? qpj[4 + i] : qpj[4 + i] / qpj[0];
  double r_hw = 0.0, r_w = 0.0;
  if (P->slope_correction_flag) {
    r_hw = r_hu * Bx + r_hv * By;
    r_w = r_u * Bx + r_v * By;
  }
  double mod_hvel = sqrt(r_hu * r_hu + r_hv * r_hv + r_hw * r_hw);
  double mod_vel = sqrt(r_u * r_u + r_v * r_v + r_w * r_w);
  double shear_vel = o->sqrt_ff * mod_vel;
  double rouse[MAXS];
  vp_rouse(o, shear_vel, r_rho_c, rouse);
  double h0 = vp_h0(o, r_h);
  double b = 30.0 / P->k_s;
  double z[IMEXSW_MAX_QUAD], w[IMEXSW_MAX_QUAD], lz[IMEXSW_MAX_QUAD];
  vp_quad(o, h0, b, z, w, lz);
  double lh0 = det_log(b * h0 + 1.0);
  double log_term_h0 = lh0 * lh0;
  double ncu = 1.0 / vp_u_log_avg(o, w, lz, lh0, r_h, h0);
  double u_rel0 = ncu * lh0;
  double eps_s = P->Sc * shear_vel * P->vonK * ((h0 / 6.0) + (P->k_s / 60.0));
  double a_coeff = -(P->vonK * shear_vel / eps_s);
  double irhv = 0.0, rua[MAXS];
  for (int i = 0; i < ns; i++) {
    double a = a_coeff * rouse[i];
    double se = 0.0, sq = 0.0, sl = 0.0;
    for (int m = 0; m < nq; m++)
      se += w[m] * det_exp(a * z[m]);
    double eh0 = det_exp(a * h0);
    double nca = 1.0 / ((se + eh0 * (r_h - h0)) / r_h);
    for (int m = 0; m < nq; m++)
      sq += w[m] * (det_exp(a * z[m]) * lz[m]);
    double d1 = (mod_vel * ncu) * (r_alphas[i] * nca) * sq;
    double ar0 = nca * eh0;
    double d2 = (r_h - h0) * (mod_vel * u_rel0) * (r_alphas[i] * ar0);
    rua[i] = P->rho_s[i] * (d1 + d2) / r_h;
    for (int m = 0; m < nq; m++)
      sl += det_exp(a * z[m]) * (w[m] * (lz[m] * lz[m]));
    d2 = (r_h - h0) * eh0 * log_term_h0;
    irhv = irhv + (P->rho_s[i] - r_rho_c) * (nca * r_alphas[i]) *
                      (mod_hvel * mod_vel * (ncu * ncu)) * (sl + d2);
  }
  double s = 0.0;
  for (int i = 0; i < ns; i++)
    s += (P->rho_s[i] - r_rho_c) / P->rho_s[i] * rua[i];
  double uRho_avg = mod_vel * r_rho_c + s;
  double s2 = 0.0;
  for (int m = 0; m < nq; m++)
    s2 += w[m] * (lz[m] * lz[m]);
  double ihv = mod_hvel * mod_vel * (ncu * ncu) * (s2 + (r_h - h0) * log_term_h0);
  double rhom_vel_vel = (irhv + r_rho_c * ihv) / (r_h * r_h);
  shape[0] = uRho_avg / (mod_vel * r_rho_m);
  shape[1] = rhom_vel_vel / (uRho_avg * mod_vel);
  shape[2] = shape[1];
  shape[3] = shape[0];
  for (int i = 0; i < ns; i++)
    shape[4 + i] = rua[i] / (P->rho_s[i] * r_alphas[i] * mod_vel);
}
void orc_flux_coeffs(const orc_t *o, const double *qp, double Bx, double By,
                     double *shape) {
  double ri, rho_m, rho_c, rg, c1, c2;
  mixt_var(o, qp, &ri, &rho_m, &rho_c, &rg, 0, &c1, &c2);
  vp_flux_coeffs(o, qp, Bx, By, rho_c, rho_m, shape);
}

/* eval_dep_coeffs (src/constitutive_2d.f90:2844-2986): ratio of the bottom
 * to the depth-averaged concentration of each class.  (The reference fills
 * dep_coeff(1:n_eqns) of an n_solid array first; only n_solid entries are
 * written here, and the loop overwrites all of them.) */
static void vp_dep_coeffs(const orc_t *o, const double *qpj, double mod_vel,
                          double r_rho_c, double *dep) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns, nq = o->nq;
  double r_h = qpj[0];
  double shear_vel = o->sqrt_ff * mod_vel;
  double rouse[MAXS];
  vp_rouse(o, shear_vel, r_rho_c, rouse);
  double h0 = vp_h0(o, r_h);
  double eps_s = P->Sc * shear_vel * P->vonK * ((h0 / 6.0) + (P->k_s / 60.0));
  double a_coeff = -(P->vonK * shear_vel / eps_s);
  double z[IMEXSW_MAX_QUAD], w[IMEXSW_MAX_QUAD];
  for (int m = 0; m < nq; m++) {
    z[m] = 0.5 * h0 * (o->zq[m] + 1.0);
    w[m] = 0.5 * h0 * o->wq[m];
  }
  for (int i = 0; i < ns; i++) {
    double a = a_coeff * rouse[i];
    double se = 0.0;
    for (int m = 0; m < nq; m++)
      se += w[m] * det_exp(a * z[m]);
    dep[i] = 1.0 / ((se + det_exp(a * h0) * (r_h - h0)) / r_h);
  }
}
void orc_dep_coeffs(const orc_t *o, const double *qp, double mod_vel,
                    double *dep) {
  double ri, rho_m, rho_c, rg, c1, c2;
  mixt_var(o, qp, &ri, &rho_m, &rho_c, &rg, 0, &c1, &c2);
  vp_dep_coeffs(o, qp, mod_vel, rho_c, dep);
}

/* integrate_friction_term (src/constitutive_2d.f90:3252-3306), rheology
 * model 8: the speed after dt of du/dt = -f u^2 / h in closed form, the
 * direction kept, through qc_to_qp and back through qp_to_qc */
static void integrate_friction(const orc_t *o, double *qj, double dt) {
  if (qj[0] == 0.0)
    return;
  double qp[MAXV + 2], pd;
  qc_to_qp(o, qj, qp, &pd);
  double r_h = qp[0], r_u = qp[o->idx_u - 1], r_v = qp[o->idx_v - 1];
  double mvh = sqrt(r_u * r_u + r_v * r_v);
  double mv = 1.0 / (1.0 / mvh + o->p.friction_factor / r_h * dt);
  r_u = r_u * (mv / mvh);
  r_v = r_v * (mv / mvh);
  qp[1] = r_h * r_u;
  qp[2] = r_h * r_v;
  qp[o->idx_u - 1] = r_u;
  qp[o->idx_v - 1] = r_v;
  qp_to_qc(o, qp, qj);
}

/* eval_mass_exchange_terms, supported branches
 * (src/constitutive_2d.f90:4424-5003) */
static void mass_exchange_terms(orc_t *o, const double *qpj, double Bx,
                                double By, const double *erodible, double dt,
                                double *erosion_term, double *deposition_term,
                                double *eqns_term, double *topo_term) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns;
  for (int i = 0; i < ns; i++) {
    erosion_term[i] = 0.0;
    deposition_term[i] = 0.0;
  }
  double cont_erosion = 0.0, cont_loss = 0.0;
  for (int i = 0; i < o->ne; i++)
    eqns_term[i] = 0.0;
  *topo_term = 0.0;
  if (qpj[0] <= EPS_WP)
    return;
  const double alpha_max = 0.6, hind_exp = 4.65;
  double r_h = qpj[0], r_u = qpj[o->idx_u - 1], r_v = qpj[o->idx_v - 1];
  double r_alphas[MAXS];
  for (int i = 0; i < ns; i++)
    r_alphas[i] = P->alpha_flag ? qpj[4 + i] : qpj[4 + i] / qpj[0];
  double alphas_tot = 0.0;
  for (int i = 0; i < ns; i++)
    alphas_tot += r_alphas[i];
  double r_w = P->slope_correction_flag ? r_u * Bx + r_v * By : 0.0;
  double mod_vel2 = r_u * r_u + r_v * r_v + r_w * r_w;
  double mod_vel = sqrt(mod_vel2);
  double tot_erosion, tot_solid_erosion;
  if (P->erosion_coeff > 0.0) {
    tot_erosion = P->erosion_coeff * mod_vel * r_h * (1.0 - alphas_tot);
    tot_solid_erosion = tot_erosion * (1.0 - P->erodible_porosity);
    for (int i = 0; i < ns; i++)
      erosion_term[i] = P->erodible_fract[i] * tot_solid_erosion;
  } else {
    tot_solid_erosion = 0.0;
    for (int i = 0; i < ns; i++)
      erosion_term[i] = 0.0;
  }
  for (int i = 0; i < ns; i++)
    erosion_term[i] = fmax(0.0, fmin(erosion_term[i], erodible[i] / dt));
  tot_solid_erosion = 0.0;
  for (int i = 0; i < ns; i++)
    tot_solid_erosion += erosion_term[i];
  tot_erosion = tot_solid_erosion / (1.0 - P->erodible_porosity);
  cont_erosion = tot_erosion * P->erodible_porosity;
  if (alphas_tot <= P->alphastot_min)
    return;
  double r_T = qpj[3];
  double ri, rho_m, rho_c, rg, shc, shm;
  mixt_var(o, qpj, &ri, &rho_m, &rho_c, &rg, 1, &shc, &shm);
  double inv_kin_visc = 1.0 / P->kin_visc_c;
  double rhoc = rho_c;
  double dep_coeff[MAXS];
  if (o->vp && P->bottom_conc_flag)
    vp_dep_coeffs(o, qpj, mod_vel, rho_c, dep_coeff);
  else
    for (int i = 0; i < ns; i++)
      dep_coeff[i] = 1.0;
  for (int i = 0; i < ns; i++) {
    if ((r_alphas[i] > 0.0) && P->settling_flag) {
      double sv = settling_velocity(o, P->diam_s[i], P->rho_s[i], rhoc,
                                    inv_kin_visc);
      deposition_term[i] = r_alphas[i] * sv * dep_coeff[i];
      deposition_term[i] =
          deposition_term[i] *
          det_pow(1.0 - fmin(1.0, alphas_tot / alpha_max), hind_exp);
      deposition_term[i] = fmin(deposition_term[i], r_h * r_alphas[i] / dt);
    }
  }
  cont_loss = 0.0;
  double sdep = 0.0;
  for (int i = 0; i < ns; i++)
    sdep += deposition_term[i];
  cont_loss = cont_loss + o->coeff_porosity * sdep;
  double sal = 0.0;
  for (int i = 0; i < ns; i++)
    sal += r_alphas[i];
  cont_loss = fmin(cont_loss,
                   r_h * fmax(0.0, P->maximum_solid_packing - sal) / dt);
  cont_loss = fmin(cont_loss, (r_h * (1.0 - sal) -
                               o->coeff_porosity * (r_h * sal - dt * sdep)) /
                                  dt);
  double T_erodible = o->T_erodible;
  if (P->erodible_deposit_flag)
    T_erodible = r_T;
  double air_entr;
  if (P->entrainment_flag && (r_h > 0.0) && (ri > 0.0)) {
    double entr_coeff = 0.075 / sqrt(1.0 + 718.0 * det_pow(ri, 2.4));
    air_entr = entr_coeff * sqrt(mod_vel2);
  } else {
    air_entr = 0.0;
  }
  for (int i = 0; i < o->ne; i++)
    eqns_term[i] = 0.0;
  double dep_tot = 0.0, ers_tot = 0.0, rho_dep_tot = 0.0, rho_ers_tot = 0.0;
  for (int i = 0; i < ns; i++)
    dep_tot += deposition_term[i];
  for (int i = 0; i < ns; i++)
    ers_tot += erosion_term[i];
  for (int i = 0; i < ns; i++)
    rho_dep_tot += P->rho_s[i] * deposition_term[i];
  for (int i = 0; i < ns; i++)
    rho_ers_tot += P->rho_s[i] * erosion_term[i];
  eqns_term[0] = o->rho_a_amb * air_entr + rho_ers_tot - rho_dep_tot +
                 o->rho_c_sub * cont_erosion - rho_c * cont_loss;
  eqns_term[1] = -(r_u * (rho_dep_tot + rho_c * cont_loss));
  eqns_term[2] = -(r_v * (rho_dep_tot + rho_c * cont_loss));
  double sdepc = 0.0, sersc = 0.0;
  for (int i = 0; i < ns; i++)
    sdepc += P->rho_s[i] * P->sp_heat_s[i] * deposition_term[i];
  for (int i = 0; i < ns; i++)
    sersc += P->rho_s[i] * P->sp_heat_s[i] * erosion_term[i];
  if (P->energy_flag)
    eqns_term[3] = -(r_T * (sdepc + rho_c * shc * cont_loss)) -
                   0.5 * mod_vel2 * (rho_dep_tot + rho_c * cont_loss) +
                   T_erodible * (sersc + o->rho_c_sub * shc * cont_erosion) +
                   P->T_ambient * P->sp_heat_a * o->rho_a_amb * air_entr;
  else
    eqns_term[3] = -(r_T * (sdepc + rho_c * shc * cont_loss)) +
                   T_erodible * (sersc + o->rho_c_sub * shc * cont_erosion) +
                   P->T_ambient * P->sp_heat_a * o->rho_a_amb * air_entr;
  for (int i = 0; i < ns; i++)
    eqns_term[4 + i] = P->rho_s[i] * (erosion_term[i] - deposition_term[i]);
  *topo_term = (dep_tot - ers_tot) / (1.0 - P->erodible_porosity);
}

/* eval_source_bdry (src/constitutive_2d.f90:5024-5130) */
static void source_bdry(const orc_t *o, double time, double vx, double vy,
                        double *sb) {
  const imexsw_params *P = &o->p;
  const int ns = o->ns;
  const double *tp = P->time_param;
  if (time >= tp[3]) {
    sb[0] = 0.0;
    sb[1] = 0.0;
    sb[2] = 0.0;
    sb[3] = P->T_source;
    for (int i = 0; i < ns; i++)
      sb[4 + i] = 0.0;
    sb[o->idx_u - 1] = 0.0;
    sb[o->idx_v - 1] = 0.0;
    return;
  }
  double t_rem = fmod(time, tp[0]);
  double t_coeff = 0.0;
  if (tp[2] == 0.0) {
    if (t_rem <= tp[1])
      t_coeff = 1.0;
  } else {
    if (t_rem < tp[2])
      t_coeff = 0.5 * (1.0 - det_cos(o->pi_g * t_rem / tp[2]));
    else if (t_rem <= (tp[1] - tp[2]))
      t_coeff = 1.0;
    else if (t_rem <= tp[1])
      t_coeff = 0.5 * (1.0 + det_cos(o->pi_g * ((t_rem - tp[1]) / tp[2] + 1.0)));
  }
  sb[0] = t_coeff * P->h_source;
  sb[1] = det_pow(t_coeff, 1.5) * P->h_source * P->vel_source * vx;
  sb[2] = det_pow(t_coeff, 1.5) * P->h_source * P->vel_source * vy;
  sb[3] = P->T_source;
  for (int i = 0; i < ns; i++)
    sb[4 + i] = P->alpha_flag ? P->alphas_source[i]
                              : t_coeff * P->h_source * P->alphas_source[i];
  sb[o->idx_u - 1] = det_pow(t_coeff, 0.5) * P->vel_source * vx;
  sb[o->idx_v - 1] = det_pow(t_coeff, 0.5) * P->vel_source * vy;
}
void orc_source_bdry(const orc_t *o, double time, double vx, double vy,
                     double *out) {
  source_bdry(o, time, vx, vy, out);
}

/* ------------------------------------------------------------------------ */
/* context creation: allocate_solver_variables + derived constants          */
/* ------------------------------------------------------------------------ */
orc_t *orc_create(const imexsw_params *p) {
  orc_t *o = (orc_t *)xcalloc(1, sizeof(orc_t));
  o->p = *p;
  const imexsw_params *P = &o->p;
  o->nx = P->nx;
  o->ny = P->ny;
  o->nix = o->nx + 1;
  o->niy = o->ny + 1;
  o->ns = P->n_solid;
  o->nv = 4 + o->ns; /* src/inpout_2d.f90:1429, 1852-1853 */
  o->ne = o->nv;
  o->idx_u = o->nv + 1;
  o->idx_v = o->nv + 2;
  o->nrk = P->n_rk;
  /* init_grid (src/geometry_2d.f90:251-300) */
  o->dx = (o->nx > 1) ? P->cell_size : 1.0;
  o->dy = (o->ny > 1) ? P->cell_size : 1.0;
  o->dx2 = o->dx / 2.0;
  o->dy2 = o->dy / 2.0;
  o->one_by_dx = 1.0 / o->dx;
  o->one_by_dy = 1.0 / o->dy;
  o->eps_sing = fmin(fmin(pow(o->dx, 4.0), pow(o->dy, 4.0)), 1.0e-10);
  {
    double e2 = o->eps_sing * o->eps_sing;
    o->eps_sing4 = e2 * e2;
  }
  o->pi_g = 4.0 * atan(1.0);
  o->x_comp = (double *)xcalloc(o->nx + 1, sizeof(double));
  o->x_stag = (double *)xcalloc(o->nix + 1, sizeof(double));
  o->y_comp = (double *)xcalloc(o->ny + 1, sizeof(double));
  o->y_stag = (double *)xcalloc(o->niy + 1, sizeof(double));
  for (int j = 1; j <= o->nix; j++)
    o->x_stag[j] = P->x0 + (j - 1) * o->dx;
  for (int k = 1; k <= o->niy; k++)
    o->y_stag[k] = P->y0 + (k - 1) * o->dy;
  for (int j = 1; j <= o->nx; j++)
    o->x_comp[j] = 0.5 * (o->x_stag[j] + o->x_stag[j + 1]);
  for (int k = 1; k <= o->ny; k++)
    o->y_comp[k] = 0.5 * (o->y_stag[k] + o->y_stag[k + 1]);
  /* read_param derived constants */
  o->inv_pres = 1.0 / P->pres;
  o->rho_a_amb = P->pres / (P->sp_gas_const_a * P->T_ambient);
  for (int i = 0; i < o->ns; i++)
    o->inv_rho_s[i] = 1.0 / P->rho_s[i];
  o->T_erodible = P->T_erodible;
  o->dt_old = -1.0;
  o->dt_old_old = -1.0;
  o->rho_c_sub = P->pres / (P->sp_gas_const_a * P->T_erodible);
  o->coeff_porosity = (P->erosion_coeff > 0.0)
                          ? P->erodible_porosity / (1.0 - P->erodible_porosity)
                          : 0.0;
  /* vertical profiles: quadrature and H_crit_rel (src/inpout_2d.f90:
   * 3761-3791); the shear velocity reads SQRT(friction_factor) */
  o->vp = P->vertical_profiles_flag;
  o->sqrt_ff = sqrt(P->friction_factor);
  o->nq = 0;
  o->h_crit_rel = 0.0;
  if (o->vp) {
    o->nq = P->n_quad;
    orc_gauss_legendre(o->nq, o->zq, o->wq);
    double a = P->vonK / o->sqrt_ff + 1.0;
    double arg = -a * det_exp(-a);
    double l0 = orc_lambertw0(arg), l1 = orc_lambertwm1(arg);
    o->h_crit_rel = 1.0 / 30.0 * (-(a / fmax(l0, l1)) - 1.0);
  }
  /* complex step (src/solver_2d.f90:273-274) */
  o->h_cs = o->nv * EPS_WP;
  o->one_by_h = 1.0 / o->h_cs;
  /* init_problem_param (src/constitutive_2d.f90:469-518) */
  for (int i = 1; i <= o->ne; i++)
    o->implicit_flag[i] = 0;
  o->implicit_flag[2] = 1;
  o->implicit_flag[3] = 1;
  o->n_nh = 2;
  /* IMEX tableaux (src/solver_2d.f90:398-486) */
  memset(o->a_tilde_ij, 0, sizeof(o->a_tilde_ij));
  memset(o->a_dirk_ij, 0, sizeof(o->a_dirk_ij));
  memset(o->omega_tilde, 0, sizeof(o->omega_tilde));
  memset(o->omega, 0, sizeof(o->omega));
  if (o->nrk == 1) {
    o->a_tilde_ij[0][0] = 1.0;
    o->omega_tilde[0] = 1.0;
  } else if (o->nrk == 2) {
    o->a_tilde_ij[1][0] = 1.0;
    o->omega_tilde[0] = 1.0;
    o->a_dirk_ij[1][1] = 1.0;
    o->omega[1] = 1.0;
  } else if (o->nrk == 3) {
    o->a_tilde_ij[1][0] = 0.5;
    o->a_tilde_ij[2][0] = 0.5;
    o->a_tilde_ij[2][1] = 0.5;
    o->omega_tilde[0] = o->omega_tilde[1] = o->omega_tilde[2] = 1.0 / 3.0;
    o->a_dirk_ij[0][0] = 0.25;
    o->a_dirk_ij[1][1] = 0.25;
    o->a_dirk_ij[2][0] = 1.0 / 3.0;
    o->a_dirk_ij[2][1] = 1.0 / 3.0;
    o->a_dirk_ij[2][2] = 1.0 / 3.0;
    o->omega[0] = o->omega[1] = o->omega[2] = 1.0 / 3.0;
  } else if (o->nrk == 4) {
    o->a_tilde_ij[1][0] = 0.5;
    o->a_tilde_ij[2][0] = 1.0 / 3.0;
    o->a_tilde_ij[3][1] = 1.0;
    o->omega_tilde[1] = 1.0;
    o->a_dirk_ij[1][1] = 0.5;
    o->a_dirk_ij[2][2] = 1.0 / 3.0;
    o->a_dirk_ij[3][2] = 0.75;
    o->a_dirk_ij[3][3] = 0.25;
    o->omega[2] = 0.75;
    o->omega[3] = 0.25;
  }
  size_t N = (size_t)o->nx * o->ny, nv = o->nv, nvp = nv + 2, R = o->nrk;
  size_t NFX = (size_t)o->nix * o->ny, NFY = (size_t)o->nx * o->niy;
  o->q = xcalloc(nv * N, 8);
  o->q0 = xcalloc(nv * N, 8);
  o->q_fv = xcalloc(nv * N, 8);
  o->qp = xcalloc(nvp * N, 8);
  for (size_t c = 0; c < N; c++)
    o->qp[3 + nvp * c] = P->T_ambient;
  o->q_rk = xcalloc(nv * N * R, 8);
  o->qp_rk = xcalloc(nvp * N * R, 8);
  o->divFlux = xcalloc(nv * N * R, 8);
  o->NH = xcalloc(nv * N * R, 8);
  o->SI_NH = xcalloc(nv * N * R, 8);
  o->expl_terms = xcalloc(nv * N * R, 8);
  o->residual_term = xcalloc(nv * N, 8);
  o->qL = xcalloc(nv * NFX, 8);
  o->qR = xcalloc(nv * NFX, 8);
  o->qpL = xcalloc(nvp * NFX, 8);
  o->qpR = xcalloc(nvp * NFX, 8);
  o->qB = xcalloc(nv * NFY, 8);
  o->qT = xcalloc(nv * NFY, 8);
  o->qpB = xcalloc(nvp * NFY, 8);
  o->qpT = xcalloc(nvp * NFY, 8);
  o->axNeg = xcalloc(NFX, 8);
  o->axPos = xcalloc(NFX, 8);
  o->axmax = xcalloc(NFX, 8);
  o->Hx = xcalloc(nv * NFX, 8);
  o->ayNeg = xcalloc(NFY, 8);
  o->ayPos = xcalloc(NFY, 8);
  o->aymax = xcalloc(NFY, 8);
  o->Hy = xcalloc(nv * NFY, 8);
  o->solve_mask = xcalloc(N, 1);
  o->solve_mask_temp = xcalloc(N, 1);
  o->smx = xcalloc(NFX, 1);
  o->smy = xcalloc(NFY, 1);
  o->j_cent = xcalloc(N, sizeof(int));
  o->k_cent = xcalloc(N, sizeof(int));
  o->j_stag_x = xcalloc(NFX, sizeof(int));
  o->k_stag_x = xcalloc(NFX, sizeof(int));
  o->j_stag_y = xcalloc(NFY, sizeof(int));
  o->k_stag_y = xcalloc(NFY, sizeof(int));
  o->hmax = xcalloc(N, 8);
  o->pdynmax = xcalloc(N, 8);
  o->mod_vel_max = xcalloc(N, 8);
  int nl = P->n_thickness_levels * P->n_dyn_pres_levels;
  o->vuln_table = xcalloc((size_t)(nl > 0 ? nl : 1) * N, 1);
  o->thck_table = xcalloc(N, 1);
  o->pdyn_table = xcalloc(N, 1);
  o->B_cent = xcalloc(N, 8);
  o->Bx = xcalloc(N, 8);
  o->By = xcalloc(N, 8);
  o->Bxx = xcalloc(N, 8);
  o->Bxy = xcalloc(N, 8);
  o->Byy = xcalloc(N, 8);
  o->grav_coeff = xcalloc(N, 8);
  o->dgdx = xcalloc(N, 8);
  o->gsx = xcalloc(NFX, 8);
  o->gsy = xcalloc(NFY, 8);
  for (size_t c = 0; c < N; c++)
    o->grav_coeff[c] = 1.0;
  for (size_t c = 0; c < NFX; c++)
    o->gsx[c] = 1.0;
  for (size_t c = 0; c < NFY; c++)
    o->gsy[c] = 1.0;
  o->B_nodata = xcalloc(N, 1);
  o->source_cell = xcalloc(N, sizeof(int));
  o->sE = xcalloc(N, 1);
  o->sW = xcalloc(N, 1);
  o->sS = xcalloc(N, 1);
  o->sN = xcalloc(N, 1);
  o->sEx = xcalloc(N, 8);
  o->sEy = xcalloc(N, 8);
  o->sWx = xcalloc(N, 8);
  o->sWy = xcalloc(N, 8);
  o->sSx = xcalloc(N, 8);
  o->sSy = xcalloc(N, 8);
  o->sNx = xcalloc(N, 8);
  o->sNy = xcalloc(N, 8);
  o->deposit = xcalloc(N * (o->ns ? o->ns : 1), 8);
  o->erosion = xcalloc(N * (o->ns ? o->ns : 1), 8);
  o->erodible = xcalloc(N * (o->ns ? o->ns : 1), 8);
  /* boundary cells are always in the mask (src/solver_2d.f90:346-349) */
  for (int k = 1; k <= o->ny; k++)
    for (int j = 1; j <= o->nx; j++)
      if (j == 1 || j == o->nx || k == 1 || k == o->ny)
        o->solve_mask[I2(j, k)] = 1;
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o)
    return;
  void *ptrs[] = {o->x_comp, o->x_stag, o->y_comp, o->y_stag, o->q, o->q0,
                  o->q_fv, o->qp, o->q_rk, o->qp_rk, o->divFlux, o->NH,
                  o->SI_NH, o->expl_terms, o->residual_term, o->qL, o->qR,
                  o->qpL, o->qpR, o->qB, o->qT, o->qpB, o->qpT, o->axNeg,
                  o->axPos, o->axmax, o->Hx, o->ayNeg, o->ayPos, o->aymax,
                  o->Hy, o->solve_mask, o->solve_mask_temp, o->smx, o->smy,
                  o->j_cent, o->k_cent, o->j_stag_x, o->k_stag_x, o->j_stag_y,
                  o->k_stag_y, o->hmax, o->pdynmax, o->mod_vel_max,
                  o->vuln_table, o->thck_table, o->pdyn_table, o->B_cent,
                  o->Bx, o->By, o->Bxx, o->Bxy, o->Byy, o->grav_coeff,
                  o->dgdx, o->gsx, o->gsy, o->B_nodata, o->source_cell, o->sE,
                  o->sW, o->sS, o->sN, o->sEx, o->sEy, o->sWx, o->sWy, o->sSx,
                  o->sSy, o->sNx, o->sNy, o->deposit, o->erosion, o->erodible};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++)
    free(ptrs[i]);
  free(o);
}

void orc_set_fresh(orc_t *o, int fresh) { o->fresh = fresh; }
int orc_n_vars(const orc_t *o) { return o->nv; }

/* ------------------------------------------------------------------------ */
/* geometry_2d: topography_reconstruction (src/geometry_2d.f90:535-860),    */
/* restricted to the fields the hot path reads (*_geom, grav_coeff family). */
/* For nx < 5 or ny < 5 the reference reads and writes outside the arrays   */
/* (:699-795); the oracle defines those fields as zero there.               */
/* ------------------------------------------------------------------------ */
void orc_topography_reconstruction(orc_t *o) {
  const int nx = o->nx, ny = o->ny;
  const double c1[5] = {-2.0, -1.0, 0.0, 1.0, 2.0};
  const double c2[5] = {2.0, -1.0, -2.0, -1.0, 2.0};
  double norm1_x = 10.0 * o->dx, norm1_y = 10.0 * o->dy;
  double norm2_x = 7.0 * (o->dx * o->dx), norm2_y = 7.0 * (o->dy * o->dy);
  double norm_xy = (10.0 * o->dx) * (10.0 * o->dy);
  const double *B = o->B_cent;
  size_t N = (size_t)nx * ny;
  if (nx < 5 || ny < 5) {
    memset(o->Bx, 0, N * 8);
    memset(o->By, 0, N * 8);
    memset(o->Bxx, 0, N * 8);
    memset(o->Bxy, 0, N * 8);
    memset(o->Byy, 0, N * 8);
  } else {
#pragma omp parallel for schedule(static)
    for (int k = 3; k <= ny - 2; k++) {
      for (int j = 3; j <= nx - 2; j++) {
        double px = 0.0, py = 0.0, sxx = 0.0, syy = 0.0;
        for (int kj = 1; kj <= 5; kj++) {
          px = px + c1[kj - 1] * B[I2(j + kj - 3, k)];
          sxx = sxx + c2[kj - 1] * B[I2(j + kj - 3, k)];
        }
        for (int kk = 1; kk <= 5; kk++) {
          py = py + c1[kk - 1] * B[I2(j, k + kk - 3)];
          syy = syy + c2[kk - 1] * B[I2(j, k + kk - 3)];
        }
        o->Bx[I2(j, k)] = px / norm1_x;
        o->By[I2(j, k)] = py / norm1_y;
        o->Bxx[I2(j, k)] = sxx / norm2_x;
        o->Byy[I2(j, k)] = syy / norm2_y;
        double ws = 0.0;
        for (int kk = 1; kk <= 5; kk++)
          for (int kj = 1; kj <= 5; kj++)
            ws = ws + c1[kj - 1] * c1[kk - 1] * B[I2(j + kj - 3, k + kk - 3)];
        o->Bxy[I2(j, k)] = ws / norm_xy;
      }
    }
    double *F[5] = {o->Bx, o->By, o->Bxx, o->Byy, o->Bxy};
    for (int k = 1; k <= ny; k++)
      for (int f = 0; f < 5; f++) {
        F[f][I2(1, k)] = F[f][I2(3, k)];
        F[f][I2(2, k)] = F[f][I2(3, k)];
        F[f][I2(nx - 1, k)] = F[f][I2(nx - 2, k)];
        F[f][I2(nx, k)] = F[f][I2(nx - 2, k)];
      }
    for (int j = 1; j <= nx; j++)
      for (int f = 0; f < 5; f++) {
        F[f][I2(j, 1)] = F[f][I2(j, 3)];
        F[f][I2(j, 2)] = F[f][I2(j, 3)];
        F[f][I2(j, ny - 1)] = F[f][I2(j, ny - 2)];
        F[f][I2(j, ny)] = F[f][I2(j, ny - 2)];
      }
  }
  if (o->p.slope_correction_flag) {
    for (size_t c = 0; c < N; c++) {
      double g = 1.0 / (1.0 + o->Bx[c] * o->Bx[c] + o->By[c] * o->By[c]);
      o->grav_coeff[c] = g;
      o->dgdx[c] = -(2.0 * (g * g) *
                     (o->Bx[c] * o->Bxx[c] + o->By[c] * o->Bxy[c]));
    }
    for (int k = 1; k <= ny; k++) {
      o->gsx[IFX(1, k)] = o->grav_coeff[I2(1, k)];
      for (int j = 2; j <= o->nix - 1; j++)
        o->gsx[IFX(j, k)] =
            0.5 * (o->grav_coeff[I2(j - 1, k)] + o->grav_coeff[I2(j, k)]);
      o->gsx[IFX(o->nix, k)] = o->grav_coeff[I2(nx, k)];
    }
    for (int j = 1; j <= nx; j++) {
      o->gsy[IFY(j, 1)] = o->grav_coeff[I2(j, 1)];
      for (int k = 2; k <= o->niy - 1; k++)
        o->gsy[IFY(j, k)] =
            0.5 * (o->grav_coeff[I2(j, k - 1)] + o->grav_coeff[I2(j, k)]);
      o->gsy[IFY(j, o->niy)] = o->grav_coeff[I2(j, ny)];
    }
  } else {
    for (size_t c = 0; c < N; c++) {
      o->grav_coeff[c] = 1.0;
      o->dgdx[c] = 0.0;
    }
    for (size_t c = 0; c < (size_t)o->nix * ny; c++)
      o->gsx[c] = 1.0;
    for (size_t c = 0; c < (size_t)nx * o->niy; c++)
      o->gsy[c] = 1.0;
  }
}

void orc_set_topography(orc_t *o, const double *B_cent,
                        const uint8_t *nodata) {
  size_t N = (size_t)o->nx * o->ny;
  memcpy(o->B_cent, B_cent, N * 8);
  if (nodata)
    memcpy(o->B_nodata, nodata, N);
  orc_topography_reconstruction(o);
}

void orc_get_geometry(const orc_t *o, double *B, double *Bx, double *By,
                      double *Bxx, double *Bxy, double *Byy) {
  size_t N = (size_t)o->nx * o->ny * 8;
  if (B) memcpy(B, o->B_cent, N);
  if (Bx) memcpy(Bx, o->Bx, N);
  if (By) memcpy(By, o->By, N);
  if (Bxx) memcpy(Bxx, o->Bxx, N);
  if (Bxy) memcpy(Bxy, o->Bxy, N);
  if (Byy) memcpy(Byy, o->Byy, N);
}

/* init_source, radial branch (src/geometry_2d.f90:1005-1100) */
void orc_init_source(orc_t *o) {
  const imexsw_params *P = &o->p;
  size_t N = (size_t)o->nx * o->ny;
  memset(o->source_cell, 0, N * sizeof(int));
  memset(o->sE, 0, N);
  memset(o->sW, 0, N);
  memset(o->sS, 0, N);
  memset(o->sN, 0, N);
  if (!P->radial_source_flag)
    return;
  double xs = P->x_source, ys = P->y_source, rs = P->r_source;
  double r2 = rs * rs;
#define SQ(a) ((a) * (a))
  for (int k = 2; k <= o->ny - 1; k++) {
    for (int j = 2; j <= o->nx - 1; j++) {
      if (SQ(o->x_comp[j] - xs) + SQ(o->y_comp[k] - ys) <= r2) {
        o->source_cell[I2(j, k)] = 1;
        if (SQ(o->x_comp[j - 1] - xs) + SQ(o->y_comp[k] - ys) >= r2) {
          o->source_cell[I2(j - 1, k)] = 2;
          o->sE[I2(j - 1, k)] = 1;
          double d = sqrt(SQ(o->x_stag[j] - xs) + SQ(o->y_comp[k] - ys));
          o->sEx[I2(j - 1, k)] = (o->x_stag[j] - xs) * rs / SQ(d);
          o->sEy[I2(j - 1, k)] = (o->y_comp[k] - ys) * rs / SQ(d);
        } else if (SQ(o->x_comp[j + 1] - xs) + SQ(o->y_comp[k] - ys) >= r2) {
          o->source_cell[I2(j + 1, k)] = 2;
          o->sW[I2(j + 1, k)] = 1;
          double d = sqrt(SQ(o->x_stag[j + 1] - xs) + SQ(o->y_comp[k] - ys));
          o->sWx[I2(j + 1, k)] = (o->x_stag[j + 1] - xs) * rs / SQ(d);
          o->sWy[I2(j + 1, k)] = (o->y_comp[k] - ys) * rs / SQ(d);
        }
        if (SQ(o->x_comp[j] - xs) + SQ(o->y_comp[k - 1] - ys) >= r2) {
          o->source_cell[I2(j, k - 1)] = 2;
          o->sN[I2(j, k - 1)] = 1;
          double d = sqrt(SQ(o->x_comp[j] - xs) + SQ(o->y_stag[k] - ys));
          o->sNx[I2(j, k - 1)] = (o->x_comp[j] - xs) * rs / SQ(d);
          o->sNy[I2(j, k - 1)] = (o->y_stag[k] - ys) * rs / SQ(d);
        } else if (SQ(o->x_comp[j] - xs) + SQ(o->y_comp[k + 1] - ys) >= r2) {
          o->source_cell[I2(j, k + 1)] = 2;
          o->sS[I2(j, k + 1)] = 1;
          double d = sqrt(SQ(o->x_comp[j] - xs) + SQ(o->y_stag[k + 1] - ys));
          o->sSx[I2(j, k + 1)] = (o->x_comp[j] - xs) * rs / SQ(d);
          o->sSy[I2(j, k + 1)] = (o->y_stag[k + 1] - ys) * rs / SQ(d);
        }
      }
    }
  }
#undef SQ
}

void orc_get_source(const orc_t *o, int32_t *sc, uint8_t *bits, double *v) {
  size_t N = (size_t)o->nx * o->ny;
  for (size_t c = 0; c < N; c++) {
    if (sc) sc[c] = o->source_cell[c];
    if (bits)
      bits[c] = (o->sE[c] ? IMEXSW_SRC_E : 0) | (o->sW[c] ? IMEXSW_SRC_W : 0) |
                (o->sS[c] ? IMEXSW_SRC_S : 0) | (o->sN[c] ? IMEXSW_SRC_N : 0);
    if (v) {
      v[8 * c + 0] = o->sEx[c];
      v[8 * c + 1] = o->sEy[c];
      v[8 * c + 2] = o->sWx[c];
      v[8 * c + 3] = o->sWy[c];
      v[8 * c + 4] = o->sSx[c];
      v[8 * c + 5] = o->sSy[c];
      v[8 * c + 6] = o->sNx[c];
      v[8 * c + 7] = o->sNy[c];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* state I/O                                                                */
/* ------------------------------------------------------------------------ */
/* a new state starts a new run: the next orc_run takes its step history
   from dt0 (src/IMEX_SfloW2D.f90:248-249) */
void orc_set_q(orc_t *o, const double *q) {
  memcpy(o->q, q, (size_t)o->nv * o->nx * o->ny * 8);
  o->dt_old = -1.0;
}
void orc_get_q(const orc_t *o, double *q) {
  memcpy(q, o->q, (size_t)o->nv * o->nx * o->ny * 8);
}
void orc_get_qp(const orc_t *o, double *qp) {
  memcpy(qp, o->qp, (size_t)(o->nv + 2) * o->nx * o->ny * 8);
}
void orc_set_erodible(orc_t *o, const double *e) {
  memcpy(o->erodible, e, (size_t)o->ns * o->nx * o->ny * 8);
}
void orc_get_erosion(const orc_t *o, double *erodible, double *deposit,
                     double *erosion, double *B) {
  size_t N = (size_t)o->nx * o->ny;
  if (erodible) memcpy(erodible, o->erodible, N * o->ns * 8);
  if (deposit) memcpy(deposit, o->deposit, N * o->ns * 8);
  if (erosion) memcpy(erosion, o->erosion, N * o->ns * 8);
  if (B) memcpy(B, o->B_cent, N * 8);
}
void orc_get_post(const orc_t *o, double *hmax, double *pdynmax, double *mv,
                  uint8_t *vuln) {
  size_t N = (size_t)o->nx * o->ny;
  if (hmax) memcpy(hmax, o->hmax, N * 8);
  if (pdynmax) memcpy(pdynmax, o->pdynmax, N * 8);
  if (mv) memcpy(mv, o->mod_vel_max, N * 8);
  int nl = o->p.n_thickness_levels * o->p.n_dyn_pres_levels;
  if (vuln && nl > 0) memcpy(vuln, o->vuln_table, N * nl);
}
void orc_get_mask(const orc_t *o, uint8_t *m) {
  memcpy(m, o->solve_mask, (size_t)o->nx * o->ny);
}
void orc_counters(const orc_t *o, int64_t *c10) {
  memcpy(c10, o->cnt, sizeof(o->cnt));
}

/* ------------------------------------------------------------------------ */
/* check_solve (src/solver_2d.f90:666-857)                                  */
/* ------------------------------------------------------------------------ */
int orc_check_solve(orc_t *o, int solve_all) {
  const int nx = o->nx, ny = o->ny;
  uint8_t *m = o->solve_mask, *mt = o->solve_mask_temp;
  for (int k = 2; k <= ny - 1; k++)
    for (int j = 2; j <= nx - 1; j++)
      m[I2(j, k)] = solve_all ? 1 : 0;
  for (int k = 2; k <= ny - 1; k++)
    for (int j = 2; j <= nx - 1; j++)
      if (o->q[IQ(1, j, k)] > 0.0) /* solve_mask_time(j,k) = 0 <= t */
        m[I2(j, k)] = 1;
  if (o->p.radial_source_flag)
    for (int k = 2; k <= ny - 1; k++)
      for (int j = 2; j <= nx - 1; j++)
        if (o->source_cell[I2(j, k)] == 2)
          m[I2(j, k)] = 1;
  for (int it = 1; it <= o->nrk; it++) {
    memcpy(mt, m, (size_t)nx * ny);
    for (int k = 2; k <= ny - 1; k++)
      for (int j = 2; j <= nx - 1; j++)
        m[I2(j, k)] = m[I2(j, k)] || mt[I2(j - 1, k)];
    for (int k = 2; k <= ny - 1; k++)
      for (int j = 2; j <= nx - 1; j++)
        m[I2(j, k)] = m[I2(j, k)] || mt[I2(j + 1, k)];
    for (int k = 2; k <= ny - 1; k++)
      for (int j = 2; j <= nx - 1; j++)
        m[I2(j, k)] = m[I2(j, k)] || mt[I2(j, k - 1)];
    for (int k = 2; k <= ny - 1; k++)
      for (int j = 2; j <= nx - 1; j++)
        m[I2(j, k)] = m[I2(j, k)] || mt[I2(j, k + 1)];
  }
  if (o->p.radial_source_flag)
    for (size_t c = 0; c < (size_t)nx * ny; c++)
      if (o->source_cell[c] == 1)
        m[c] = 0;
  memset(o->smx, 0, (size_t)o->nix * ny);
  memset(o->smy, 0, (size_t)nx * o->niy);
  int i = 0;
  for (int k = 1; k <= ny; k++)
    for (int j = 1; j <= nx; j++)
      if (m[I2(j, k)]) {
        o->j_cent[i] = j;
        o->k_cent[i] = k;
        i++;
        o->smx[IFX(j, k)] = 1;
        o->smx[IFX(j + 1, k)] = 1;
        o->smy[IFY(j, k)] = 1;
        o->smy[IFY(j, k + 1)] = 1;
      }
  o->solve_cells = i;
  i = 0;
  for (int k = 1; k <= ny; k++)
    for (int j = 1; j <= o->nix; j++)
      if (o->smx[IFX(j, k)]) {
        o->j_stag_x[i] = j;
        o->k_stag_x[i] = k;
        i++;
      }
  o->sif_x = i;
  i = 0;
  for (int k = 1; k <= o->niy; k++)
    for (int j = 1; j <= nx; j++)
      if (o->smy[IFY(j, k)]) {
        o->j_stag_y[i] = j;
        o->k_stag_y[i] = k;
        i++;
      }
  o->sif_y = i;
  return o->solve_cells;
}

/* ------------------------------------------------------------------------ */
/* reconstruction (src/solver_2d.f90:3093-3918)                             */
/* ------------------------------------------------------------------------ */
static void reconstruct_cell(orc_t *o, int j, int k, const double *q_expl,
                             const double *qp_expl) {
  const imexsw_params *P = &o->p;
  const int nv = o->nv, nvp = nv + 2, nx = o->nx, ny = o->ny;
  const int src2 = (o->source_cell[I2(j, k)] == 2);
  const int radial = P->radial_source_flag;
  double qrecW[MAXV + 2], qrecE[MAXV + 2], qrecS[MAXV + 2], qrecN[MAXV + 2];
  double qpx[MAXV + 2], qpy[MAXV + 2], sb[MAXV + 2];
  double xs[3], ys[3], qs[3], p2W[3], p2E[3], p2S[3], p2N[3];
#define QPE(i, jj, kk) qp_expl[IQP(i, jj, kk)]
  for (int i = 1; i <= nvp; i++) {
    qrecW[i - 1] = qrecE[i - 1] = qrecS[i - 1] = qrecN[i - 1] =
        QPE(i, j, k);
    sb[i - 1] = 0.0;
  }
  xs[1] = o->x_comp[j];
  ys[1] = o->y_comp[k];
  xs[0] = xs[2] = ys[0] = ys[2] = 0.0;
  if (radial && src2) {
    size_t c = I2(j, k);
    if (o->sE[c])
      source_bdry(o, o->t, o->sEx[c], o->sEy[c], sb);
    else if (o->sW[c])
      source_bdry(o, o->t, o->sWx[c], o->sWy[c], sb);
    else if (o->sS[c])
      source_bdry(o, o->t, o->sSx[c], o->sSy[c], sb);
    else if (o->sN[c])
      source_bdry(o, o->t, o->sNx[c], o->sNy[c], sb);
  }
  const int sE = o->sE[I2(j, k)], sW = o->sW[I2(j, k)];
  const int sS = o->sS[I2(j, k)], sN = o->sN[I2(j, k)];
  for (int i = 1; i <= nv; i++) {
    qs[1] = QPE(i, j, k);
    if (nx > 1) {
      if (j == 1) {
        xs[0] = o->x_stag[1];
        xs[2] = o->x_comp[j + 1];
        if (src2) {
          qs[0] = sb[i - 1];
          qs[2] = QPE(i, j + 1, k);
          qpx[i - 1] = limit(o, qs, xs, P->limiter[i - 1]);
        } else if (P->bc_flag[IMEXSW_W][i - 1] == 0) {
          qs[0] = P->bc_value[IMEXSW_W][i - 1];
          qs[2] = QPE(i, j + 1, k);
          qpx[i - 1] = limit(o, qs, xs, P->limiter[i - 1]);
        } else if (P->bc_flag[IMEXSW_W][i - 1] == 1) {
          qpx[i - 1] = P->bc_value[IMEXSW_W][i - 1];
        } else if (P->bc_flag[IMEXSW_W][i - 1] == 2) {
          qpx[i - 1] = (QPE(i, 2, k) - QPE(i, 1, k)) * o->one_by_dx;
        }
      } else if (j == nx) {
        xs[2] = o->x_stag[o->nix];
        xs[0] = o->x_comp[j - 1];
        if (src2) {
          qs[2] = sb[i - 1];
          qs[0] = QPE(i, j - 1, k);
          qpx[i - 1] = limit(o, qs, xs, P->limiter[i - 1]);
        } else if (P->bc_flag[IMEXSW_E][i - 1] == 0) {
          qs[2] = P->bc_value[IMEXSW_E][i - 1];
          qs[0] = QPE(i, j - 1, k);
          qpx[i - 1] = limit(o, qs, xs, P->limiter[i - 1]);
        } else if (P->bc_flag[IMEXSW_E][i - 1] == 1) {
          qpx[i - 1] = P->bc_value[IMEXSW_E][i - 1];
        } else if (P->bc_flag[IMEXSW_E][i - 1] == 2) {
          qpx[i - 1] = (QPE(i, nx, k) - QPE(i, nx - 1, k)) * o->one_by_dx;
        }
      } else {
        xs[0] = o->x_comp[j - 1];
        xs[2] = o->x_comp[j + 1];
        qs[0] = QPE(i, j - 1, k);
        qs[2] = QPE(i, j + 1, k);
        if (radial && src2) {
          if (sE) {
            xs[2] = o->x_stag[j + 1];
            qs[2] = sb[i - 1];
          } else if (sW) {
            xs[0] = o->x_stag[j];
            qs[0] = sb[i - 1];
          }
        }
        qpx[i - 1] = limit(o, qs, xs, P->limiter[i - 1]);
      }
      double dq = P->reconstr_coeff * o->dx2 * qpx[i - 1];
      qrecW[i - 1] = qs[1] - dq;
      qrecE[i - 1] = qs[1] + dq;
      if (j == 1) {
        if (src2)
          qrecW[i - 1] = sb[i - 1];
        else if (P->bc_flag[IMEXSW_W][i - 1] == 0)
          qrecW[i - 1] = P->bc_value[IMEXSW_W][i - 1];
        else if (i == 2)
          qrecW[i - 1] = fmin(qrecW[i - 1], 0.0);
      }
      if (j == nx) {
        if (src2)
          qrecE[i - 1] = sb[i - 1];
        else if (P->bc_flag[IMEXSW_E][i - 1] == 0)
          qrecE[i - 1] = P->bc_value[IMEXSW_E][i - 1];
        else if (i == 2)
          qrecE[i - 1] = fmax(qrecE[i - 1], 0.0);
      }
    }
    if (ny > 1) {
      if (k == 1) {
        ys[0] = o->y_stag[1];
        ys[2] = o->y_comp[k + 1];
        if (P->bc_flag[IMEXSW_S][i - 1] == 0) {
          qs[0] = P->bc_value[IMEXSW_S][i - 1];
          qs[2] = QPE(i, j, k + 1);
          qpy[i - 1] = limit(o, qs, ys, P->limiter[i - 1]);
        } else if (P->bc_flag[IMEXSW_S][i - 1] == 1) {
          qpy[i - 1] = P->bc_value[IMEXSW_S][i - 1];
        } else if (P->bc_flag[IMEXSW_S][i - 1] == 2) {
          qpy[i - 1] = (QPE(i, j, 2) - QPE(i, j, 1)) * o->one_by_dy;
        }
      } else if (k == ny) {
        ys[0] = o->y_comp[k - 1];
        ys[2] = o->y_stag[o->niy];
        if (P->bc_flag[IMEXSW_N][i - 1] == 0) {
          qs[0] = QPE(i, j, k - 1);
          qs[2] = P->bc_value[IMEXSW_N][i - 1];
          qpy[i - 1] = limit(o, qs, ys, P->limiter[i - 1]);
        } else if (P->bc_flag[IMEXSW_N][i - 1] == 1) {
          qpy[i - 1] = P->bc_value[IMEXSW_N][i - 1];
        } else if (P->bc_flag[IMEXSW_N][i - 1] == 2) {
          qpy[i - 1] = (QPE(i, j, ny) - QPE(i, j, ny - 1)) * o->one_by_dy;
        }
      } else {
        ys[0] = o->y_comp[k - 1];
        ys[2] = o->y_comp[k + 1];
        qs[0] = QPE(i, j, k - 1);
        qs[2] = QPE(i, j, k + 1);
        if (radial && src2) {
          if (sS) {
            ys[0] = o->y_stag[k];
            qs[0] = sb[i - 1];
          } else if (sN) {
            ys[2] = o->y_stag[k + 1];
            qs[2] = sb[i - 1];
          }
        }
        qpy[i - 1] = limit(o, qs, ys, P->limiter[i - 1]);
      }
      double dq = P->reconstr_coeff * o->dy2 * qpy[i - 1];
      qrecS[i - 1] = qs[1] - dq;
      qrecN[i - 1] = qs[1] + dq;
      if (k == 1) {
        if (P->bc_flag[IMEXSW_S][i - 1] == 0)
          qrecS[i - 1] = P->bc_value[IMEXSW_S][i - 1];
        else if (i == 3)
          qrecS[i - 1] = fmin(qrecS[i - 1], 0.0);
      }
      if (k == ny) {
        if (P->bc_flag[IMEXSW_N][i - 1] == 0)
          qrecN[i - 1] = P->bc_value[IMEXSW_N][i - 1];
        else if (i == 3)
          qrecN[i - 1] = fmax(qrecN[i - 1], 0.0);
      }
    }
  }
  /* u, v with the limiters of hu, hv (:3496-3661) */
  for (int i = nv + 1; i <= nv + 2; i++) {
    if (nx > 1) {
      qs[1] = QPE(i, j, k);
      if (j == 1) {
        qp_to_qp2(qrecW, o->B_cent[I2(j, k)], p2W);
        qs[0] = p2W[i - nv];
        qs[2] = QPE(i, j + 1, k);
      } else if (j == nx) {
        qp_to_qp2(qrecE, o->B_cent[I2(j, k)], p2E);
        qs[0] = QPE(i, j - 1, k);
        qs[2] = p2E[i - nv];
      } else {
        qs[0] = QPE(i, j - 1, k);
        qs[2] = QPE(i, j + 1, k);
        if (radial && src2) {
          if (sE) {
            xs[2] = o->x_stag[j + 1];
            qs[2] = sb[i - 1];
          } else if (sW) {
            xs[0] = o->x_stag[j];
            qs[0] = sb[i - 1];
          }
        }
      }
      qpx[i - 1] = limit(o, qs, xs, P->limiter[i - 1]);
      double dq = P->reconstr_coeff * o->dx2 * qpx[i - 1];
      qrecW[i - 1] = qs[1] - dq;
      qrecE[i - 1] = qs[1] + dq;
      if (j == 1) {
        qp_to_qp2(qrecW, o->B_cent[I2(j, k)], p2W);
        qrecW[i - 1] = p2W[i - nv];
      } else if (j == nx) {
        qp_to_qp2(qrecE, o->B_cent[I2(j, k)], p2E);
        qrecE[i - 1] = p2E[i - nv];
      } else if (radial && src2) {
        if (sE) {
          for (int m = 0; m < nvp; m++)
            qrecE[m] = sb[m];
        } else if (sW) {
          for (int m = 0; m < nvp; m++)
            qrecW[m] = sb[m];
        }
      }
    }
    if (ny > 1) {
      qs[1] = QPE(i, j, k);
      if (k == 1) {
        qp_to_qp2(qrecS, o->B_cent[I2(j, k)], p2S);
        qs[0] = p2S[i - nv];
        qs[2] = QPE(i, j, k + 1);
      } else if (k == ny) {
        qp_to_qp2(qrecN, o->B_cent[I2(j, k)], p2N);
        qs[0] = QPE(i, j, k - 1);
        qs[2] = p2N[i - nv];
      } else {
        qs[0] = QPE(i, j, k - 1);
        qs[2] = QPE(i, j, k + 1);
        if (radial && src2) {
          if (sS) {
            ys[0] = o->y_stag[k];
            qs[0] = sb[i - 1];
          } else if (sN) {
            /* the reference writes x_stencil here (src/solver_2d.f90:3616) */
            xs[2] = o->y_stag[k + 1];
            qs[2] = sb[i - 1];
          }
        }
      }
      qpy[i - 1] = limit(o, qs, ys, P->limiter[i - 1]);
      double dq = P->reconstr_coeff * o->dy2 * qpy[i - 1];
      qrecS[i - 1] = qs[1] - dq;
      qrecN[i - 1] = qs[1] + dq;
      if (k == 1) {
        qp_to_qp2(qrecS, o->B_cent[I2(j, k)], p2S);
        qrecS[i - 1] = p2S[i - nv];
      } else if (k == ny) {
        qp_to_qp2(qrecN, o->B_cent[I2(j, k)], p2N);
        qrecN[i - 1] = p2N[i - nv];
      } else if (radial && src2) {
        if (sS) {
          for (int m = 0; m < nvp; m++)
            qrecS[m] = sb[m];
        } else if (sN) {
          for (int m = 0; m < nvp; m++)
            qrecN[m] = sb[m];
        }
      }
    }
  }
  /* interface values (:3664-3912); the dry-cell shortcut of :3677-3701 and
   * :3790-3817 is always overwritten by the conversion that follows it. */
  if (nx > 1) {
    qp_to_qc(o, qrecW, &o->qR[(size_t)nv * IFX(j, k)]);
    qp_to_qc(o, qrecE, &o->qL[(size_t)nv * IFX(j + 1, k)]);
    memcpy(&o->qpR[(size_t)nvp * IFX(j, k)], qrecW, nvp * 8);
    memcpy(&o->qpL[(size_t)nvp * IFX(j + 1, k)], qrecE, nvp * 8);
    if (j == 1) {
      memcpy(&o->qL[(size_t)nv * IFX(j, k)], &o->qR[(size_t)nv * IFX(j, k)],
             nv * 8);
      memcpy(&o->qpL[(size_t)nvp * IFX(j, k)],
             &o->qpR[(size_t)nvp * IFX(j, k)], nvp * 8);
    } else if (j == nx) {
      memcpy(&o->qR[(size_t)nv * IFX(j + 1, k)],
             &o->qL[(size_t)nv * IFX(j + 1, k)], nv * 8);
      memcpy(&o->qpR[(size_t)nvp * IFX(j + 1, k)],
             &o->qpL[(size_t)nvp * IFX(j + 1, k)], nvp * 8);
    } else if (radial && src2) {
      if (sE) {
        memcpy(&o->qR[(size_t)nv * IFX(j + 1, k)],
               &o->qL[(size_t)nv * IFX(j + 1, k)], nv * 8);
        memcpy(&o->qpR[(size_t)nvp * IFX(j + 1, k)],
               &o->qpL[(size_t)nvp * IFX(j + 1, k)], nvp * 8);
      } else if (sW) {
        memcpy(&o->qL[(size_t)nv * IFX(j, k)], &o->qR[(size_t)nv * IFX(j, k)],
               nv * 8);
        memcpy(&o->qpL[(size_t)nvp * IFX(j, k)],
               &o->qpR[(size_t)nvp * IFX(j, k)], nvp * 8);
      }
    }
  } else {
    memcpy(&o->qR[(size_t)nv * IFX(j, k)], &q_expl[IQ(1, j, k)], nv * 8);
    memcpy(&o->qL[(size_t)nv * IFX(j + 1, k)], &q_expl[IQ(1, j, k)], nv * 8);
    memcpy(&o->qpR[(size_t)nvp * IFX(j, k)], &qp_expl[IQP(1, j, k)], nvp * 8);
    memcpy(&o->qpL[(size_t)nvp * IFX(j + 1, k)], &qp_expl[IQP(1, j, k)],
           nvp * 8);
  }
  if (ny > 1) {
    qp_to_qc(o, qrecS, &o->qT[(size_t)nv * IFY(j, k)]);
    qp_to_qc(o, qrecN, &o->qB[(size_t)nv * IFY(j, k + 1)]);
    memcpy(&o->qpT[(size_t)nvp * IFY(j, k)], qrecS, nvp * 8);
    memcpy(&o->qpB[(size_t)nvp * IFY(j, k + 1)], qrecN, nvp * 8);
    if (k == 1) {
      memcpy(&o->qB[(size_t)nv * IFY(j, k)], &o->qT[(size_t)nv * IFY(j, k)],
             nv * 8);
      memcpy(&o->qpB[(size_t)nvp * IFY(j, k)],
             &o->qpT[(size_t)nvp * IFY(j, k)], nvp * 8);
    } else if (k == ny) {
      memcpy(&o->qT[(size_t)nv * IFY(j, k + 1)],
             &o->qB[(size_t)nv * IFY(j, k + 1)], nv * 8);
      memcpy(&o->qpT[(size_t)nvp * IFY(j, k + 1)],
             &o->qpB[(size_t)nvp * IFY(j, k + 1)], nvp * 8);
    } else if (radial && src2) {
      if (sS) {
        memcpy(&o->qB[(size_t)nv * IFY(j, k)], &o->qT[(size_t)nv * IFY(j, k)],
               nv * 8);
        memcpy(&o->qpB[(size_t)nvp * IFY(j, k)],
               &o->qpT[(size_t)nvp * IFY(j, k)], nvp * 8);
      } else if (sN) {
        memcpy(&o->qT[(size_t)nv * IFY(j, k + 1)],
               &o->qB[(size_t)nv * IFY(j, k + 1)], nv * 8);
        memcpy(&o->qpT[(size_t)nvp * IFY(j, k + 1)],
               &o->qpB[(size_t)nvp * IFY(j, k + 1)], nvp * 8);
      }
    }
  } else {
    const double *qe = &q_expl[IQ(1, j, k)], *pe = &qp_expl[IQP(1, j, k)];
    memcpy(&o->qB[(size_t)nv * IFY(j, k)], qe, nv * 8);
    memcpy(&o->qT[(size_t)nv * IFY(j, k)], qe, nv * 8);
    memcpy(&o->qB[(size_t)nv * IFY(j, k + 1)], qe, nv * 8);
    memcpy(&o->qT[(size_t)nv * IFY(j, k + 1)], qe, nv * 8);
    memcpy(&o->qpB[(size_t)nvp * IFY(j, k)], pe, nvp * 8);
    memcpy(&o->qpT[(size_t)nvp * IFY(j, k)], pe, nvp * 8);
    memcpy(&o->qpB[(size_t)nvp * IFY(j, k + 1)], pe, nvp * 8);
    memcpy(&o->qpT[(size_t)nvp * IFY(j, k + 1)], pe, nvp * 8);
  }
#undef QPE
}

static void reconstruction(orc_t *o, const double *q_expl,
                           const double *qp_expl) {
  if (o->fresh) {
    /* every cell outside the source interior, from the current state */
    const int N = o->nx * o->ny;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < N; c++) {
      if (o->source_cell[c] == 1)
        continue;
      reconstruct_cell(o, c % o->nx + 1, c / o->nx + 1, q_expl, qp_expl);
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int l = 0; l < o->solve_cells; l++)
      reconstruct_cell(o, o->j_cent[l], o->k_cent[l], q_expl, qp_expl);
  }
}

/* eval_speeds (src/solver_2d.f90:3931-4008); all n_eqns entries of
 * a_interface_* are equal, so one value per face is kept */
static void eval_speeds(orc_t *o) {
  const int nvp = o->nv + 2;
  if (o->nx > 1) {
#pragma omp parallel for schedule(static)
    for (int l = 0; l < o->sif_x; l++) {
      int j = o->j_stag_x[l], k = o->k_stag_x[l];
      double lmin, lmax, rmin, rmax;
      local_speeds(o, &o->qpL[(size_t)nvp * IFX(j, k)], o->gsx[IFX(j, k)], 1,
                   &lmin, &lmax);
      local_speeds(o, &o->qpR[(size_t)nvp * IFX(j, k)], o->gsx[IFX(j, k)], 1,
                   &rmin, &rmax);
      o->axNeg[IFX(j, k)] = fmin(fmin(lmin, rmin), 0.0);
      o->axPos[IFX(j, k)] = fmax(fmax(lmax, rmax), 0.0);
    }
  }
  if (o->ny > 1) {
#pragma omp parallel for schedule(static)
    for (int l = 0; l < o->sif_y; l++) {
      int j = o->j_stag_y[l], k = o->k_stag_y[l];
      double bmin, bmax, tmin, tmax;
      local_speeds(o, &o->qpB[(size_t)nvp * IFY(j, k)], o->gsy[IFY(j, k)], 2,
                   &bmin, &bmax);
      local_speeds(o, &o->qpT[(size_t)nvp * IFY(j, k)], o->gsy[IFY(j, k)], 2,
                   &tmin, &tmax);
      o->ayNeg[IFY(j, k)] = fmin(fmin(bmin, tmin), 0.0);
      o->ayPos[IFY(j, k)] = fmax(fmax(bmax, tmax), 0.0);
    }
  }
}

/* count active-face sides whose owner cell is inactive and whose stored
 * value is not the dry state a fresh reconstruction would give */
static void count_stale(orc_t *o) {
  const int nv = o->nv, nvp = nv + 2;
  int64_t n = 0;
  for (int l = 0; l < o->sif_x; l++) {
    int j = o->j_stag_x[l], k = o->k_stag_x[l];
    for (int side = 0; side < 2; side++) {
      int jc = side == 0 ? j - 1 : j;
      if (jc < 1 || jc > o->nx || o->solve_mask[I2(jc, k)] ||
          o->source_cell[I2(jc, k)] == 1)
        continue;
      const double *qp = side == 0 ? &o->qpL[(size_t)nvp * IFX(j, k)]
                                   : &o->qpR[(size_t)nvp * IFX(j, k)];
      if (qp[0] != 0.0 || qp[1] != 0.0 || qp[o->idx_u - 1] != 0.0)
        n++;
    }
  }
  for (int l = 0; l < o->sif_y; l++) {
    int j = o->j_stag_y[l], k = o->k_stag_y[l];
    for (int side = 0; side < 2; side++) {
      int kc = side == 0 ? k - 1 : k;
      if (kc < 1 || kc > o->ny || o->solve_mask[I2(j, kc)] ||
          o->source_cell[I2(j, kc)] == 1)
        continue;
      const double *qp = side == 0 ? &o->qpB[(size_t)nvp * IFY(j, k)]
                                   : &o->qpT[(size_t)nvp * IFY(j, k)];
      if (qp[0] != 0.0 || qp[2] != 0.0 || qp[o->idx_v - 1] != 0.0)
        n++;
    }
  }
  o->cnt[3] += n;
}

/* eval_flux_KT + average_KT (src/solver_2d.f90:2832-3047) */
static void eval_flux_KT(orc_t *o) {
  const int nv = o->nv, nvp = nv + 2, ns = o->ns;
  if (o->nx > 1) {
#pragma omp parallel for schedule(static)
    for (int l = 0; l < o->sif_x; l++) {
      int j = o->j_stag_x[l], k = o->k_stag_x[l];
      size_t f = IFX(j, k);
      double fL[MAXV], fR[MAXV], avg[MAXV];
      double *H = &o->Hx[(size_t)nv * f];
      const double *qL = &o->qL[(size_t)nv * f], *qR = &o->qR[(size_t)nv * f];
      size_t cl = I2(j > 1 ? j - 1 : 1, k < o->ny ? k : o->ny);
      size_t cr = I2(j < o->nx ? j : o->nx, k < o->ny ? k : o->ny);
      eval_fluxes(o, qL, &o->qpL[(size_t)nvp * f], o->Bx[cl], o->By[cl],
                  o->gsx[f], 1, fL);
      eval_fluxes(o, qR, &o->qpR[(size_t)nvp * f], o->Bx[cr], o->By[cr],
                  o->gsx[f], 1, fR);
      double aN = o->axNeg[f], aP = o->axPos[f];
      for (int i = 0; i < nv; i++)
        avg[i] = (aN == aP) ? 0.0 : (aP * fL[i] - aN * fR[i]) / (aP - aN);
      for (int i = 0; i < nv; i++) {
        if (aN == aP)
          H[i] = 0.0;
        else
          H[i] = avg[i] + (aP * aN) / (aP - aN) * (qR[i] - qL[i]);
      }
      double s = 0.0;
      for (int i = 0; i < ns; i++)
        s += H[4 + i];
      if (s / H[0] >= 1.0) {
        double sr = 0.0;
        for (int i = 0; i < ns; i++)
          sr += H[4 + i] / H[0];
        for (int i = 0; i < ns; i++)
          H[4 + i] = H[4 + i] / sr;
      }
      if ((o->qpL[(size_t)nvp * f + 1] == 0.0) &&
          (o->qpR[(size_t)nvp * f + 1] == 0.0)) {
        H[0] = 0.0;
        for (int i = 3; i < nv; i++)
          H[i] = 0.0;
      }
    }
  }
  if (o->ny > 1) {
#pragma omp parallel for schedule(static)
    for (int l = 0; l < o->sif_y; l++) {
      int j = o->j_stag_y[l], k = o->k_stag_y[l];
      size_t f = IFY(j, k);
      double fB[MAXV], fT[MAXV], avg[MAXV];
      double *H = &o->Hy[(size_t)nv * f];
      const double *qB = &o->qB[(size_t)nv * f], *qT = &o->qT[(size_t)nv * f];
      size_t cb = I2(j < o->nx ? j : o->nx, k > 1 ? k - 1 : 1);
      size_t ct = I2(j < o->nx ? j : o->nx, k < o->ny ? k : o->ny);
      eval_fluxes(o, qB, &o->qpB[(size_t)nvp * f], o->Bx[cb], o->By[cb],
                  o->gsy[f], 2, fB);
      eval_fluxes(o, qT, &o->qpT[(size_t)nvp * f], o->Bx[ct], o->By[ct],
                  o->gsy[f], 2, fT);
      double aN = o->ayNeg[f], aP = o->ayPos[f];
      for (int i = 0; i < nv; i++)
        avg[i] = (aN == aP) ? 0.0 : (aP * fB[i] - aN * fT[i]) / (aP - aN);
      for (int i = 0; i < nv; i++) {
        if (aN == aP)
          H[i] = 0.0;
        else
          H[i] = avg[i] + (aP * aN) / (aP - aN) * (qT[i] - qB[i]);
      }
      double s = 0.0;
      for (int i = 0; i < ns; i++)
        s += H[4 + i];
      if (s / H[0] > 1.0) {
        for (int i = 0; i < ns; i++)
          H[4 + i] = H[4 + i] / (s / H[0]);
      }
      if ((qB[2] == 0.0) && (qT[2] == 0.0)) {
        H[0] = 0.0;
        for (int i = 3; i < nv; i++)
          H[i] = 0.0;
      }
    }
  }
}

/* eval_hyperbolic_terms (src/solver_2d.f90:2614-2691) */
static void eval_hyperbolic_terms(orc_t *o, const double *q_expl,
                                  const double *qp_expl, double *div) {
  const int nv = o->nv;
  reconstruction(o, q_expl, qp_expl);
  eval_speeds(o);
  if (!o->fresh)
    count_stale(o);
  eval_flux_KT(o);
#pragma omp parallel for schedule(static)
  for (int l = 0; l < o->solve_cells; l++) {
    int j = o->j_cent[l], k = o->k_cent[l];
    for (int i = 1; i <= nv; i++) {
      double d = 0.0;
      if (o->nx > 1)
        d = d + (o->Hx[(size_t)nv * IFX(j + 1, k) + i - 1] -
                 o->Hx[(size_t)nv * IFX(j, k) + i - 1]) *
                    o->one_by_dx;
      if (o->ny > 1)
        d = d + (o->Hy[(size_t)nv * IFY(j, k + 1) + i - 1] -
                 o->Hy[(size_t)nv * IFY(j, k) + i - 1]) *
                    o->one_by_dy;
      div[IQ(i, j, k)] = d;
    }
  }
}

/* phys of the current state for cells outside the active set (fresh mode) */
static void fill_inactive(orc_t *o, double *qdst, double *qpdst) {
  const int nv = o->nv, nvp = nv + 2, N = o->nx * o->ny;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < N; c++) {
    if (o->solve_mask[c])
      continue;
    const double *qc = &o->q[(size_t)nv * c];
    if (qdst)
      memcpy(&qdst[(size_t)nv * c], qc, nv * 8);
    double *qp = &qpdst[(size_t)nvp * c], pd;
    if (qc[0] > 0.0) {
      qc_to_qp(o, qc, qp, &pd);
    } else {
      for (int i = 0; i < nvp; i++)
        qp[i] = 0.0;
      qp[3] = o->p.T_ambient;
    }
  }
}

/* timestep (src/solver_2d.f90:872-1005) */
double orc_timestep(orc_t *o) {
  const imexsw_params *P = &o->p;
  const int nv = o->nv, nvp = nv + 2;
  double dt = P->max_dt;
  if (P->cfl != -1.0) {
#pragma omp parallel for schedule(static)
    for (int l = 0; l < o->solve_cells; l++) {
      int j = o->j_cent[l], k = o->k_cent[l];
      double pd;
      if (o->q[IQ(1, j, k)] > 0.0) {
        qc_to_qp(o, &o->q[IQ(1, j, k)], &o->qp[IQP(1, j, k)], &pd);
      } else {
        for (int i = 1; i <= nvp; i++)
          o->qp[IQP(i, j, k)] = 0.0;
        o->qp[IQP(4, j, k)] = P->T_ambient;
      }
    }
    if (o->fresh)
      fill_inactive(o, NULL, o->qp);
    reconstruction(o, o->q, o->qp);
    eval_speeds(o);
    for (int l = 0; l < o->sif_x; l++) {
      size_t f = IFX(o->j_stag_x[l], o->k_stag_x[l]);
      o->axmax[f] = fmax(o->axPos[f], -o->axNeg[f]);
    }
    for (int l = 0; l < o->sif_y; l++) {
      size_t f = IFY(o->j_stag_y[l], o->k_stag_y[l]);
      o->aymax[f] = fmax(o->ayPos[f], -o->ayNeg[f]);
    }
    double dtmin = dt;
#pragma omp parallel for schedule(static) reduction(min : dtmin)
    for (int l = 0; l < o->solve_cells; l++) {
      int j = o->j_cent[l], k = o->k_cent[l];
      double max_a = fmax(o->axmax[IFX(j, k)], o->axmax[IFX(j + 1, k)]);
      double dtx = (max_a > 0.0) ? P->cfl * o->dx / max_a : dt;
      max_a = (o->ny > 1) ? fmax(o->aymax[IFY(j, k)], o->aymax[IFY(j, k + 1)])
                          : 0.0;
      double dty = (max_a > 0.0) ? P->cfl * o->dy / max_a : dt;
      double dtc = fmin(dtx, dty);
      dtmin = fmin(dtmin, dtc);
    }
    dt = dtmin;
  }
  return dt;
}

/* ------------------------------------------------------------------------ */
/* the implicit solve: eval_f, eval_jacobian, lnsrch, solve_rk_step         */
/* (src/solver_2d.f90:1660-2365)                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  double qj_old[MAXV], Rj[MAXV];
  double a_diag, dt, Bx, By;
} nl_ctx;

static void eval_f(const orc_t *o, const nl_ctx *c, const double *qj,
                   double *f, double *scal_f) {
  double nh[MAXV];
  implicit_terms_r(o, qj, c->Bx, c->By, nh);
  double s = 0.0;
  for (int i = 0; i < o->ne; i++) {
    double Rj = c->Rj[i] - c->a_diag * nh[i];
    f[i] = qj[i] - c->qj_old[i] + c->dt * Rj;
    f[i] = 1.0 * f[i];
  }
  for (int i = 0; i < o->ne; i++)
    s += f[i] * f[i];
  *scal_f = 0.5 * s;
}

static void eval_jacobian(const orc_t *o, const nl_ctx *c,
                          const double *qj_rel, const double *qj_org,
                          double L[MAXV][MAXV] /* L[row][col] */) {
  const int nv = o->nv;
  for (int r = 0; r < nv; r++)
    for (int m = 0; m < nv; m++)
      L[r][m] = 0.0;
  for (int i = 1; i <= nv; i++) {
    L[i - 1][i - 1] = 1.0 * qj_org[i - 1];
    if (o->implicit_flag[i]) {
      cx cq[MAXV], nh[MAXV];
      for (int m = 0; m < nv; m++)
        cq[m] = C(qj_rel[m], 0.0);
      cq[i - 1] = C(qj_rel[i - 1], o->h_cs);
      for (int m = 0; m < nv; m++)
        cq[m] = cmulr(cq[m], qj_org[m]);
      implicit_terms_cs(o, cq, c->Bx, c->By, nh);
      for (int r = 0; r < o->ne; r++) {
        double J = 1.0 * nh[r].im * o->one_by_h;
        L[r][i - 1] = L[r][i - 1] - c->dt * c->a_diag * J;
      }
    }
  }
}

/* gfortran NORM2: scaled sum of squares */
static double norm2_f(const double *x, int n) {
  double scale = 0.0, res = 1.0;
  for (int i = 0; i < n; i++) {
    if (x[i] != 0.0) {
      double ax = fabs(x[i]);
      if (scale < ax) {
        double r = scale / ax;
        res = 1.0 + res * (r * r);
        scale = ax;
      } else {
        double r = ax / scale;
        res = res + r * r;
      }
    }
  }
  return scale * sqrt(res);
}

/* lnsrch (src/solver_2d.f90:2019-2224).  Returns the number of function
 * evaluations.  MAX(tmplam, 0.5*alam) is taken with fmax (a NaN tmplam
 * falls back to halving). */
static int lnsrch(const orc_t *o, const nl_ctx *c, const double *old,
                  const double *org, double scal_f_old, const double *grad_f,
                  double *desc, double *qj_rel, double *scal_f,
                  double *right_term, double stpmax) {
  const int n = o->nv;
  const double TOLX = EPS_WP;
  int evals = 0;
  double dabs = norm2_f(desc, n);
  if (dabs > stpmax)
    for (int i = 0; i < n; i++)
      desc[i] = desc[i] * stpmax / dabs;
  double slope = 0.0;
  for (int i = 0; i < n; i++)
    slope += grad_f[i] * desc[i];
  double mx = -INFINITY;
  for (int i = 0; i < n; i++) {
    double v = fabs(desc[i]) / fmax(fabs(old[i]), 1.0);
    if (v > mx)
      mx = v;
  }
  double alamin = TOLX / mx;
  if (alamin == 0.0) {
    for (int i = 0; i < n; i++)
      qj_rel[i] = old[i];
    return evals;
  }
  double alam = 1.0, alam2 = 0.0, scal_f2 = 0.0, tmplam = 0.0;
  double scal_f_min = scal_f_old, alam_min = 0.0;
  (void)alam_min;
  for (int it = 0; it < 10000; it++) {
    double qj[MAXV];
    for (int i = 0; i < n; i++)
      qj_rel[i] = old[i] + alam * desc[i];
    for (int i = 0; i < n; i++)
      qj[i] = qj_rel[i] * org[i];
    eval_f(o, c, qj, right_term, scal_f);
    evals++;
    if (*scal_f < scal_f_min) {
      scal_f_min = *scal_f;
      alam_min = alam;
    }
    if (*scal_f <= 0.9 * scal_f_old) {
      break;
    } else if (alam < alamin) {
      for (int i = 0; i < n; i++)
        qj_rel[i] = old[i];
      *scal_f = scal_f_old;
      break;
    } else {
      if (alam == 1.0) {
        tmplam = -(slope / (2.0 * (*scal_f - scal_f_old - slope)));
      } else {
        double rhs1 = *scal_f - scal_f_old - alam * slope;
        double rhs2 = scal_f2 - scal_f_old - alam2 * slope;
        double a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) /
                   (alam - alam2);
        double b = (-(alam2 * rhs1 / (alam * alam)) +
                    alam * rhs2 / (alam2 * alam2)) /
                   (alam - alam2);
        if (a == 0.0) {
          tmplam = -(slope / (2.0 * b));
        } else {
          double disc = b * b - 3.0 * a * slope;
          if (disc < 0.0)
            tmplam = 0.5 * alam;
          else if (b <= 0.0)
            tmplam = (-b + sqrt(disc)) / (3.0 * a);
          else
            tmplam = -(slope / (b + sqrt(disc)));
        }
        if (tmplam > 0.5 * alam)
          tmplam = 0.5 * alam;
      }
    }
    alam2 = alam;
    scal_f2 = *scal_f;
    alam = fmax(tmplam, 0.5 * alam);
  }
  return evals;
}

/* solve_rk_step (src/solver_2d.f90:1660-1994); returns iterations, sets
 * *conv = 0 when max_nl_iter was exhausted */
static int solve_rk_step(const orc_t *o, const nl_ctx *c, double *qj,
                         int *conv) {
  const int n = o->nv;
  const double TOLF = 1.0e-10;
  double qj_org[MAXV], qj_rel[MAXV], right_term[MAXV], old[MAXV];
  double desc[MAXV], grad_f[MAXV], L[MAXV][MAXV], scal_f, scal_f_old;
  int nimpl[MAXV], impl[MAXV], n1 = 0, n2 = 0;
  for (int i = 1; i <= n; i++) {
    if (o->implicit_flag[i])
      impl[n2++] = i - 1;
    else
      nimpl[n1++] = i - 1;
  }
  for (int i = 0; i < n; i++)
    grad_f[i] = 0.0;
  for (int i = 0; i < n; i++)
    qj_org[i] = fmax(fabs(qj[i]), 1.0e-3);
  for (int i = 0; i < n; i++)
    qj_rel[i] = qj[i] / qj_org[i];
  *conv = 0;
  int nl_iter;
  for (nl_iter = 1; nl_iter <= 100; nl_iter++) {
    const double TOLX = EPS_WP;
    eval_f(o, c, qj, right_term, &scal_f);
    double mf = 0.0;
    for (int i = 0; i < n; i++)
      mf = fmax(mf, fabs(right_term[i]));
    if (mf < TOLF) {
      *conv = 1;
      break;
    }
    eval_jacobian(o, c, qj_rel, qj_org, L);
    /* partitioned solve (:1846-1925) */
    double d1[MAXV], d2[MAXV], s21[MAXV * MAXV];
    for (int r = 0; r < n1; r++)
      d1[r] = right_term[nimpl[r]];
    for (int r = 0; r < n2; r++)
      d2[r] = right_term[impl[r]];
    for (int r = 0; r < n1; r++)
      d1[r] = d1[r] / L[nimpl[r]][nimpl[r]];
    /* small21 = reshape(pack(L, mask21), [n1, n2]) */
    {
      int m = 0;
      for (int ci = 0; ci < n1; ci++)
        for (int rj = 0; rj < n2; rj++)
          s21[m++] = L[impl[rj]][nimpl[ci]];
    }
    for (int cc = 0; cc < n2; cc++) {
      double s = 0.0;
      for (int r = 0; r < n1; r++)
        s += d1[r] * s21[r + n1 * cc];
      d2[cc] = d2[cc] - s;
    }
    {
      double s11 = L[impl[0]][impl[0]], s21v = L[impl[1]][impl[0]];
      double s12 = L[impl[0]][impl[1]], s22 = L[impl[1]][impl[1]];
      double inv_det = 1.0 / (s11 * s22 - s21v * s12);
      double x1 = (d2[0] * s22 - d2[1] * s12) * inv_det;
      double x2 = (s11 * d2[1] - s21v * d2[0]) * inv_det;
      d2[0] = x1;
      d2[1] = x2;
    }
    {
      double u2[MAXV], u1[MAXV];
      for (int i = 0; i < n; i++)
        u2[i] = u1[i] = 0.0;
      for (int r = 0; r < n2; r++)
        u2[impl[r]] = -d2[r];
      for (int r = 0; r < n1; r++)
        u1[nimpl[r]] = -d1[r];
      for (int i = 0; i < n; i++)
        desc[i] = u2[i] + u1[i];
    }
    for (int i = 0; i < n; i++)
      old[i] = qj_rel[i];
    scal_f_old = scal_f;
    if (nl_iter > 1) {
      double dd = 0.0;
      for (int i = 0; i < n; i++)
        dd += qj_rel[i] * qj_rel[i];
      double stpmax = 100.0 * fmax(sqrt(dd), (double)n);
      for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int r = 0; r < n; r++)
          s += right_term[r] * L[r][i];
        grad_f[i] = s;
      }
      lnsrch(o, c, old, qj_org, scal_f_old, grad_f, desc, qj_rel, &scal_f,
             right_term, stpmax);
    } else {
      for (int i = 0; i < n; i++)
        qj_rel[i] = old[i] + desc[i];
      for (int i = 0; i < n; i++)
        qj[i] = qj_rel[i] * qj_org[i];
      eval_f(o, c, qj, right_term, &scal_f);
    }
    for (int i = 0; i < n; i++)
      qj[i] = qj_rel[i] * qj_org[i];
    mf = 0.0;
    for (int i = 0; i < n; i++)
      mf = fmax(mf, fabs(right_term[i]));
    if (mf < TOLF) {
      *conv = 1;
      break;
    }
    double mx = 0.0;
    for (int i = 0; i < n; i++)
      mx = fmax(mx, fabs(qj_rel[i] - old[i]) / fmax(fabs(qj_rel[i]), 1.0));
    if (mx < TOLX) {
      *conv = 1;
      break;
    }
  }
  return nl_iter > 100 ? 100 : nl_iter;
}

/* ------------------------------------------------------------------------ */
/* imex_RK_solver (src/solver_2d.f90:1021-1632)                             */
/* ------------------------------------------------------------------------ */
static inline double *rk(orc_t *o, double *a, int i, int j, int k, int r) {
  return &a[IRK(i, j, k, r, o->nv)];
}

void orc_imex_rk(orc_t *o, double t, double dt) {
  const imexsw_params *P = &o->p;
  const int nv = o->nv, nvp = nv + 2, ne = o->ne, R = o->nrk;
  o->t = t;
#pragma omp parallel for schedule(static)
  for (int l = 0; l < o->solve_cells; l++) {
    int j = o->j_cent[l], k = o->k_cent[l];
    for (int i = 1; i <= nv; i++)
      o->q0[IQ(i, j, k)] = o->q[IQ(i, j, k)];
    for (int r = 1; r <= R; r++) {
      for (int i = 1; i <= nv; i++) {
        *rk(o, o->q_rk, i, j, k, r) = 0.0;
        *rk(o, o->divFlux, i, j, k, r) = 0.0;
        *rk(o, o->NH, i, j, k, r) = 0.0;
        *rk(o, o->SI_NH, i, j, k, r) = 0.0;
        *rk(o, o->expl_terms, i, j, k, r) = 0.0;
      }
      for (int i = 1; i <= nvp; i++)
        o->qp_rk[IRK(i, j, k, r, nvp)] = 0.0;
      o->qp_rk[IRK(4, j, k, r, nvp)] = P->T_ambient;
    }
  }
  int64_t n_cells = 0, n_iters = 0, n_nc = 0;
  for (int i_RK = 1; i_RK <= R; i_RK++) {
    double a_tilde[4] = {0, 0, 0, 0}, a_dirk[4] = {0, 0, 0, 0};
    for (int m = 0; m < i_RK - 1; m++) {
      a_tilde[m] = o->a_tilde_ij[i_RK - 1][m];
      a_dirk[m] = o->a_dirk_ij[i_RK - 1][m];
    }
    double a_diag = o->a_dirk_ij[i_RK - 1][i_RK - 1];
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : n_cells, n_iters, n_nc)
    for (int l = 0; l < o->solve_cells; l++) {
      int j = o->j_cent[l], k = o->k_cent[l];
      double q_guess[MAXV], q_si[MAXV], Rj[MAXV], pd;
      if (i_RK == 1)
        for (int i = 1; i <= nv; i++)
          q_guess[i - 1] = o->q0[IQ(i, j, k)];
      for (int i = 1; i <= nv; i++) {
        double s1 = 0.0, s2 = 0.0;
        for (int r = 1; r <= i_RK; r++)
          s1 = s1 + (*rk(o, o->divFlux, i, j, k, r) -
                     *rk(o, o->expl_terms, i, j, k, r)) *
                        a_tilde[r - 1];
        for (int r = 1; r <= i_RK; r++)
          s2 = s2 + (*rk(o, o->NH, i, j, k, r) + *rk(o, o->SI_NH, i, j, k, r)) *
                        a_dirk[r - 1];
        o->q_fv[IQ(i, j, k)] = o->q0[IQ(i, j, k)] - dt * (s1 - s2);
      }
      qc_to_qp(o, &o->q_fv[IQ(1, j, k)], &o->qp[IQP(1, j, k)], &pd);
      if (a_diag != 0.0) {
        if (o->q_fv[IQ(1, j, k)] > 0.0) {
          size_t c = I2(j, k);
          double si[MAXV];
          semi_impl_terms(o, o->Bx[c], o->By[c], o->Bxx[c], o->Bxy[c],
                          o->Byy[c], o->grav_coeff[c], &o->q_fv[IQ(1, j, k)],
                          &o->qp[IQP(1, j, k)], si);
          for (int i = 1; i <= ne; i++)
            *rk(o, o->SI_NH, i, j, k, i_RK) = si[i - 1];
          for (int i = 1; i <= nv; i++)
            q_si[i - 1] = o->q_fv[IQ(i, j, k)] + dt * a_diag * si[i - 1];
          double f2 = o->q_fv[IQ(2, j, k)], f3 = o->q_fv[IQ(3, j, k)];
          if ((f2 * f2 + f3 * f3) == 0.0) {
            q_si[1] = 0.0;
            q_si[2] = 0.0;
          } else if ((q_si[1] * f2 < 0.0) || (q_si[2] * f3 < 0.0)) {
            q_si[1] = 0.0;
            q_si[2] = 0.0;
          } else {
            double m = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
            double den = sqrt(f2 * f2 + f3 * f3);
            q_si[1] = m * f2 / den;
            q_si[2] = m * f3 / den;
          }
          for (int i = 1; i <= ne; i++)
            *rk(o, o->SI_NH, i, j, k, i_RK) =
                (q_si[i - 1] - o->q_fv[IQ(i, j, k)]) / (dt * a_diag);
          for (int i = 1; i <= nv; i++)
            q_guess[i - 1] = q_si[i - 1];
          for (int i = 1; i <= ne; i++) {
            double s1 = 0.0, s2 = 0.0;
            for (int r = 1; r <= i_RK - 1; r++)
              s1 = s1 + (*rk(o, o->divFlux, i, j, k, r) -
                         *rk(o, o->expl_terms, i, j, k, r)) *
                            a_tilde[r - 1];
            for (int r = 1; r <= i_RK - 1; r++)
              s2 = s2 +
                   (*rk(o, o->NH, i, j, k, r) + *rk(o, o->SI_NH, i, j, k, r)) *
                       a_dirk[r - 1];
            Rj[i - 1] = (s1 - s2) - a_diag * *rk(o, o->SI_NH, i, j, k, i_RK);
          }
          nl_ctx nc;
          for (int i = 0; i < nv; i++) {
            nc.qj_old[i] = o->q0[IQ(i + 1, j, k)];
            nc.Rj[i] = Rj[i];
          }
          nc.a_diag = a_diag;
          nc.dt = dt;
          nc.Bx = o->Bx[c];
          nc.By = o->By[c];
          const int fric8 = P->rheology_flag && P->rheology_model == 8;
          if (fric8) {
            /* solve_rk_step returns after integrate_friction_term
             * (src/solver_2d.f90:1734-1739) */
            integrate_friction(o, q_guess, dt);
          } else {
            int conv;
            n_iters += solve_rk_step(o, &nc, q_guess, &conv);
            n_cells++;
            if (!conv)
              n_nc++;
          }
          if (o->ny == 1)
            q_guess[2] = 0.0;
          if (o->nx == 1)
            q_guess[1] = 0.0;
          /* model 8: NH from the update alone, no realignment (:1226-1230) */
          if (!fric8) {
            double nh[MAXV];
            implicit_terms_r(o, q_guess, o->Bx[c], o->By[c], nh);
            for (int i = 1; i <= ne; i++)
              *rk(o, o->NH, i, j, k, i_RK) = nh[i - 1];
            if (q_si[1] * q_si[1] + q_si[2] * q_si[2] == 0.0) {
              q_guess[1] = 0.0;
              q_guess[2] = 0.0;
            } else if ((q_guess[1] * q_si[1] <= 0.0) &&
                       (q_guess[2] * q_si[2] <= 0.0)) {
              q_guess[1] = 0.0;
              q_guess[2] = 0.0;
            } else {
              double m =
                  sqrt(q_guess[1] * q_guess[1] + q_guess[2] * q_guess[2]);
              double den = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
              q_guess[1] = m * q_si[1] / den;
              q_guess[2] = m * q_si[2] / den;
            }
          }
        } else {
          for (int i = 1; i <= nv; i++) {
            q_guess[i - 1] = o->q_fv[IQ(i, j, k)];
            q_si[i - 1] = o->q_fv[IQ(i, j, k)];
            *rk(o, o->SI_NH, i, j, k, i_RK) = 0.0;
            *rk(o, o->NH, i, j, k, i_RK) = 0.0;
          }
        }
      }
      if (a_diag != 0.0)
        for (int i = 1; i <= nv; i++)
          *rk(o, o->NH, i, j, k, i_RK) =
              (q_guess[i - 1] - q_si[i - 1]) / (dt * a_diag);
      for (int i = 1; i <= nv; i++)
        *rk(o, o->q_rk, i, j, k, i_RK) = q_guess[i - 1];
      if (o->omega_tilde[i_RK - 1] > 0.0) {
        double *qprk = &o->qp_rk[IRK(1, j, k, i_RK, nvp)];
        if (*rk(o, o->q_rk, 1, j, k, i_RK) > 0.0) {
          qc_to_qp(o, rk(o, o->q_rk, 1, j, k, i_RK), qprk, &pd);
        } else {
          for (int i = 0; i < nvp; i++)
            qprk[i] = 0.0;
          qprk[3] = P->T_ambient;
        }
        size_t c = I2(j, k);
        double ex[MAXV];
        eval_expl_terms(o, o->Bx[c], o->By[c], o->Bxx[c], o->Bxy[c],
                        o->Byy[c], o->grav_coeff[c], o->dgdx[c], o->dgdx[c],
                        qprk, ex);
        for (int i = 1; i <= ne; i++)
          *rk(o, o->expl_terms, i, j, k, i_RK) = ex[i - 1];
      }
    }
    if (o->omega_tilde[i_RK - 1] > 0.0) {
      double *qrk = &o->q_rk[IRK(1, 1, 1, i_RK, nv)];
      double *qprk = &o->qp_rk[IRK(1, 1, 1, i_RK, nvp)];
      if (o->fresh)
        fill_inactive(o, qrk, qprk);
      eval_hyperbolic_terms(o, qrk, qprk, &o->divFlux[IRK(1, 1, 1, i_RK, nv)]);
    }
  }
  o->cnt[0] += n_cells;
  o->cnt[1] += n_iters;
  o->cnt[2] += n_nc;
  /* assemble_sol (:1341-1626) */
  int stiff = 1;
  {
    double s1 = 0.0, s2 = 0.0;
    for (int r = 0; r < R; r++) {
      s1 += fabs(o->omega_tilde[r] - o->a_tilde_ij[R - 1][r]);
      s2 += fabs(o->omega[r] - o->a_dirk_ij[R - 1][r]);
    }
    stiff = (s1 == 0.0) && (s2 == 0.0);
  }
  int64_t e_nan = 0, e_negh = 0, e_negs = 0, e_lowT = 0, e_sum = 0;
#pragma omp parallel for schedule(static) reduction(+ : e_nan, e_negh, e_negs, e_lowT, e_sum)
  for (int l = 0; l < o->solve_cells; l++) {
    int j = o->j_cent[l], k = o->k_cent[l];
    double *q = &o->q[IQ(1, j, k)], pd;
    for (int i = 1; i <= nv; i++) {
      double s1 = 0.0, s2 = 0.0;
      for (int r = 1; r <= R; r++)
        s1 = s1 + (*rk(o, o->divFlux, i, j, k, r) -
                   *rk(o, o->expl_terms, i, j, k, r)) *
                      o->omega_tilde[r - 1];
      for (int r = 1; r <= R; r++)
        s2 = s2 + (*rk(o, o->NH, i, j, k, r) + *rk(o, o->SI_NH, i, j, k, r)) *
                      o->omega[r - 1];
      o->residual_term[IQ(i, j, k)] = s1 - s2;
    }
    if (stiff) {
      for (int i = 1; i <= nv; i++)
        q[i - 1] = *rk(o, o->q_rk, i, j, k, R);
    } else {
      for (int i = 1; i <= nv; i++)
        q[i - 1] = o->q0[IQ(i, j, k)] - dt * o->residual_term[IQ(i, j, k)];
    }
    for (int i = 0; i < nv; i++)
      if (isnan(q[i])) {
        e_nan++;
        break;
      }
    if (q[0] < EPS_WP) {
      if (q[0] > -1.0e-7) {
        for (int i = 0; i < nv; i++)
          q[i] = 0.0;
      } else {
        e_negh++;
      }
    }
    {
      int anyneg = 0, anybig = 0;
      for (int i = 0; i < o->ns; i++) {
        if (q[4 + i] < 0.0)
          anyneg = 1;
        if (q[4 + i] <= -1.0e-7)
          anybig = 1;
      }
      if (anyneg) {
        if (anybig)
          e_negs++;
        else
          for (int i = 0; i < o->ns; i++)
            if (q[4 + i] > -1.0e-7)
              q[4 + i] = 0.0;
      }
    }
    qc_to_qp(o, q, &o->qp[IQP(1, j, k)], &pd);
    if (o->qp[IQP(4, j, k)] < 273.0)
      e_lowT++;
    double ssol = 0.0;
    for (int i = 0; i < o->ns; i++)
      ssol += q[4 + i];
    if (ssol > q[0]) {
      if (((ssol - q[0]) / q[0] < 1.0e-10) || (q[0] < EPS_WP)) {
        qc_to_qp(o, &o->q0[IQ(1, j, k)], &o->qp[IQP(1, j, k)], &pd);
        for (int i = 0; i < o->ns; i++)
          q[4 + i] = q[4 + i] / ssol * q[0];
      } else {
        e_sum++;
      }
    }
    if (o->B_nodata[I2(j, k)])
      for (int i = 0; i < nv; i++)
        q[i] = 0.0;
  }
  o->cnt[4] += e_nan;
  o->cnt[5] += e_negh;
  o->cnt[6] += e_negs;
  o->cnt[7] += e_sum;
  o->cnt[8] += e_lowT;
}

/* one stage-2 update of an isolated cell (unit-test entry point): the body
 * of the i_RK = 2 iteration of imex_RK_solver for N_RK = 2 on a flat bed */
int orc_local_update(const orc_t *o, const double *q0, const double *D,
                     double dt, double *q_out) {
  const int nv = o->nv;
  double q_fv[MAXV], qp[MAXV + 2], q_si[MAXV], q_guess[MAXV], si[MAXV], pd;
  for (int i = 0; i < nv; i++)
    q_fv[i] = q0[i] - dt * (((0.0 + D[i] * 1.0) + 0.0 * 0.0) - 0.0);
  qc_to_qp(o, q_fv, qp, &pd);
  if (!(q_fv[0] > 0.0)) {
    memcpy(q_out, q_fv, nv * 8);
    return 0;
  }
  semi_impl_terms(o, 0, 0, 0, 0, 0, 1.0, q_fv, qp, si);
  for (int i = 0; i < nv; i++)
    q_si[i] = q_fv[i] + dt * 1.0 * si[i];
  double f2 = q_fv[1], f3 = q_fv[2];
  if ((f2 * f2 + f3 * f3) == 0.0) {
    q_si[1] = q_si[2] = 0.0;
  } else if ((q_si[1] * f2 < 0.0) || (q_si[2] * f3 < 0.0)) {
    q_si[1] = q_si[2] = 0.0;
  } else {
    double m = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
    double den = sqrt(f2 * f2 + f3 * f3);
    q_si[1] = m * f2 / den;
    q_si[2] = m * f3 / den;
  }
  nl_ctx nc;
  for (int i = 0; i < nv; i++) {
    si[i] = (q_si[i] - q_fv[i]) / (dt * 1.0);
    q_guess[i] = q_si[i];
    nc.qj_old[i] = q0[i];
    nc.Rj[i] = (((0.0 + D[i] * 1.0)) - (0.0 + 0.0 * 0.0)) - 1.0 * si[i];
  }
  nc.a_diag = 1.0;
  nc.dt = dt;
  nc.Bx = nc.By = 0.0;
  int conv;
  int it = solve_rk_step(o, &nc, q_guess, &conv);
  if (o->ny == 1)
    q_guess[2] = 0.0;
  if (o->nx == 1)
    q_guess[1] = 0.0;
  if (q_si[1] * q_si[1] + q_si[2] * q_si[2] == 0.0) {
    q_guess[1] = q_guess[2] = 0.0;
  } else if ((q_guess[1] * q_si[1] <= 0.0) && (q_guess[2] * q_si[2] <= 0.0)) {
    q_guess[1] = q_guess[2] = 0.0;
  } else {
    double m = sqrt(q_guess[1] * q_guess[1] + q_guess[2] * q_guess[2]);
    double den = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
    q_guess[1] = m * q_si[1] / den;
    q_guess[2] = m * q_si[2] / den;
  }
  memcpy(q_out, q_guess, nv * 8);
  return it;
}

/* ------------------------------------------------------------------------ */
/* update_erosion_deposition_cell (src/solver_2d.f90:2380-2595)             */
/* ------------------------------------------------------------------------ */
void orc_erosion_deposition(orc_t *o, double dt) {
  const imexsw_params *P = &o->p;
  const int nv = o->nv, nvp = nv + 2, ns = o->ns;
  if ((P->erosion_coeff == 0.0) && !P->settling_flag && !P->entrainment_flag)
    return;
  const size_t N = (size_t)o->nx * o->ny;
#pragma omp parallel for schedule(static)
  for (int l = 0; l < o->solve_cells; l++) {
    int j = o->j_cent[l], k = o->k_cent[l];
    size_t c = I2(j, k);
    double *q = &o->q[IQ(1, j, k)], *qp = &o->qp[IQP(1, j, k)], pd;
    if (q[0] > 0.0) {
      qc_to_qp(o, q, qp, &pd);
    } else {
      for (int i = 0; i < nvp; i++)
        qp[i] = 0.0;
      qp[3] = P->T_ambient;
    }
    double er[MAXS], dep[MAXS], eq[MAXV], topo;
    mass_exchange_terms(o, qp, o->Bx[c], o->By[c], &o->erodible[ns * c], dt,
                        er, dep, eq, &topo);
    for (int i = 0; i < o->ne; i++)
      q[i] = q[i] + dt * eq[i];
    for (int i = 0; i < ns; i++)
      q[4 + i] = fmax(0.0, q[4 + i]);
    for (int i = 0; i < ns; i++) {
      o->deposit[c + N * i] = o->deposit[c + N * i] + dt * dep[i];
      o->erosion[c + N * i] = o->erosion[c + N * i] + dt * er[i];
      o->erodible[ns * c + i] = o->erodible[ns * c + i] - dt * er[i];
    }
    if (P->erodible_deposit_flag)
      for (int i = 0; i < ns; i++)
        o->erodible[ns * c + i] = o->erodible[ns * c + i] + dt * dep[i];
    if (P->topo_change_flag)
      o->B_cent[c] = o->B_cent[c] + dt * topo;
    if (q[0] <= 0.0) {
      if (q[0] > -1.0e-10)
        for (int i = 0; i < nv; i++)
          q[i] = 0.0;
    }
    double ssol = 0.0;
    for (int i = 0; i < ns; i++)
      ssol += q[4 + i];
    if (ssol > q[0]) {
      if (q[0] < 1.0e-10)
        for (int i = 0; i < ns; i++)
          q[4 + i] = q[4 + i] / ssol * q[0];
    }
    double rg;
    if (q[0] > 0.0) {
      double ri, rm, rc, c1, c2;
      qc_to_qp(o, q, qp, &pd);
      mixt_var(o, qp, &ri, &rm, &rc, &rg, 0, &c1, &c2);
    } else {
      for (int i = 0; i < nvp; i++)
        qp[i] = 0.0;
      qp[3] = P->T_ambient;
      rg = 0.0;
    }
    if (rg <= 0.0)
      for (int i = 0; i < nv; i++)
        q[i] = 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* host loop pieces (src/IMEX_SfloW2D.f90)                                  */
/* ------------------------------------------------------------------------ */
static void post_cell(orc_t *o, int j, int k, int init) {
  const imexsw_params *P = &o->p;
  const int nv = o->nv, nvp = nv + 2;
  size_t c = I2(j, k);
  double *q = &o->q[IQ(1, j, k)], *qp = &o->qp[IQP(1, j, k)], pd;
  int nl = P->n_thickness_levels * P->n_dyn_pres_levels;
  if (q[0] > 0.0) {
    qc_to_qp(o, q, qp, &pd);
    o->hmax[c] = init ? qp[0] : fmax(o->hmax[c], qp[0]);
    double r_u = qp[nv], r_v = qp[nv + 1];
    double mv = sqrt(r_u * r_u + r_v * r_v);
    if (qp[0] > 0.001) {
      o->pdynmax[c] = init ? pd : fmax(o->pdynmax[c], pd);
      o->mod_vel_max[c] = init ? mv : fmax(o->mod_vel_max[c], mv);
    }
    int it = 0;
    for (int a = 0; a < P->n_thickness_levels; a++) {
      o->thck_table[c] = qp[0] >= P->thickness_levels[a];
      for (int b = 0; b < P->n_dyn_pres_levels; b++) {
        o->pdyn_table[c] = pd >= P->dyn_pres_levels[b];
        uint8_t v = o->thck_table[c] && o->pdyn_table[c];
        size_t ix = (size_t)it + (size_t)nl * c;
        o->vuln_table[ix] = init ? v : (o->vuln_table[ix] || v);
        it++;
      }
    }
  } else {
    if (init) {
      for (int i = 0; i < nv; i++)
        qp[i] = 0.0;
      qp[3] = P->T_ambient;
      o->hmax[c] = 0.0;
      o->pdynmax[c] = 0.0;
    } else {
      for (int i = 0; i < nvp; i++)
        qp[i] = 0.0;
      qp[3] = P->T_ambient;
    }
  }
}

void orc_init_post(orc_t *o) {
#pragma omp parallel for schedule(static)
  for (int l = 0; l < o->solve_cells; l++)
    post_cell(o, o->j_cent[l], o->k_cent[l], 1);
}

void orc_post_step(orc_t *o, double sums[4]) {
#pragma omp parallel for schedule(static)
  for (int l = 0; l < o->solve_cells; l++)
    post_cell(o, o->j_cent[l], o->k_cent[l], 0);
  if (sums) {
    const int nv = o->nv, nvp = nv + 2;
    size_t N = (size_t)o->nx * o->ny;
    double m = 0.0, v = 0.0, a = 0.0, s = 0.0;
    for (size_t c = 0; c < N; c++) {
      m += o->q[(size_t)nv * c];
      v += o->qp[(size_t)nvp * c];
      if (o->q[(size_t)nv * c] > 1.0e-7)
        a += 1.0;
      for (int i = 0; i < o->ns; i++)
        s += o->q[(size_t)nv * c + 4 + i];
    }
    sums[0] = m;
    sums[1] = v;
    sums[2] = a;
    sums[3] = s;
  }
}

int orc_run(orc_t *o, int nsteps, double dt0, double *t_io, double t_end,
            double t_output, double *dt_last) {
  if (g_threads > 0) {
#ifdef _OPENMP
    omp_set_num_threads(g_threads);
#endif
  }
  double t = *t_io;
  if (o->dt_old < 0.0) {
    o->dt_old = dt0;
    o->dt_old_old = o->dt_old;
  }
  int s = 0;
  if (t_end - t_output < 1.0e-7)
    t_output = t_end;
  /* the loop pauses where the reference writes output (t >= t_output,
     src/IMEX_SfloW2D.f90:510-527); the caller raises t_output and calls
     again, as output_solution does (src/inpout_2d.f90:6166) */
  while (s < nsteps && t < t_end && t < t_output) {
    orc_check_solve(o, 0);
    o->t = t;
    double dt = orc_timestep(o);
    if (t + dt > t_end)
      dt = t_end - t;
    if (t + dt > t_output)
      dt = t_output - t;
    dt = fmin(dt, 1.1 * 0.5 * (o->dt_old + o->dt_old_old));
    o->dt_old_old = o->dt_old;
    o->dt_old = dt;
    orc_imex_rk(o, t, dt);
    orc_erosion_deposition(o, dt);
    if (o->p.topo_change_flag)
      orc_topography_reconstruction(o);
    t = t + dt;
    orc_post_step(o, NULL);
    o->cnt[9] += o->solve_cells;
    s++;
    if (dt_last)
      *dt_last = dt;
  }
  *t_io = t;
  return s;
}
