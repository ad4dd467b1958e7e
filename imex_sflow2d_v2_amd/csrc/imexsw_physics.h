// imexsw_physics.h -- per-cell / per-face physics of the IMEX_SfloW2D hot
// path as gfx950 device functions.  One cell lives in registers: every
// routine is templated on the number of solid classes NS so that the
// n_vars = 4 + NS state vectors are fully unrolled register arrays.
//
// Arithmetic follows the reference expression by expression (operands in
// the Fortran order, reductions sequential from zero, IEEE division and
// square root, no FMA contraction: the library is built with
// -ffp-contract=off), so a device result equals the reference's to the last
// bit wherever the math library is not involved (pow, cos).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/imexsw.h"
#include "imexsw_math.h"

namespace imexsw {

constexpr double kEps = 2.220446049250313e-16;  // EPSILON(1.0_wp)

// Constants of one grid: the module variables the hot path reads, derived
// the way read_param / init_grid / allocate_solver_variables derive them.
struct DevParams {
  int nx, ny, pitch, ns, nv, nrk;
  int energy, alpha, slope, curv, topo, radial, rheo, rheo_model, settling,
      entrain, erod_dep;
  int n_thk, n_pdyn;
  double x0, y0, dx, dy, dx2, dy2, one_by_dx, one_by_dy;
  double eps_sing, eps_sing4;
  double cfl, max_dt, reconstr_coeff, theta;
  int limiter[IMEXSW_MAX_VARS + 2];
  int lim_c;  // some limiter (3, 4, 5) uses the centred difference
  int bc_flag[4][IMEXSW_MAX_VARS];
  double bc_value[4][IMEXSW_MAX_VARS];
  // the side's condition on h keeps a dry neighbourhood dry: zero-gradient
  // extrapolation, or a prescribed value / slope of exactly zero
  int bc_quiet[4];
  double grav, pres, inv_pres, T_ambient, sp_heat_a, sp_gas_const_a,
      rho_a_amb, kin_visc_c;
  double rho_s[IMEXSW_MAX_SOLID], inv_rho_s[IMEXSW_MAX_SOLID],
      diam_s[IMEXSW_MAX_SOLID], sp_heat_s[IMEXSW_MAX_SOLID];
  double mu, xi;
  double erosion_coeff, erodible_porosity, erodible_fract[IMEXSW_MAX_SOLID],
      T_erodible, rho_c_sub, coeff_porosity, alphastot_min, max_packing;
  double x_source, y_source, r_source, h_source, vel_source, T_source,
      alphas_source[IMEXSW_MAX_SOLID], time_param[4], pi_g;
  double thickness_levels[IMEXSW_MAX_LEVELS], dyn_pres_levels[IMEXSW_MAX_LEVELS];
  double h_cs, one_by_h;  // complex step n_vars*EPSILON and its reciprocal
  // vertical profiles: Gauss-Legendre rule on [-1, 1], the closure constants,
  // H_crit_rel (src/inpout_2d.f90:3761-3791); SQRT(friction_factor)
  int vp, bottom_conc, nq;
  double vonK, k_s, Sc, sqrt_ff, ff, h_crit_rel;
  double zq[IMEXSW_MAX_QUAD], wq[IMEXSW_MAX_QUAD];
};

// ---------------------------------------------------------------------------
// complex arithmetic with gfortran's lowering rules
// ---------------------------------------------------------------------------
struct cpx {
  double re, im;
};
__device__ __forceinline__ cpx mk(double r, double i) { return cpx{r, i}; }
__device__ __forceinline__ cpx cadd(cpx a, cpx b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cpx csub(cpx a, cpx b) { return mk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cpx cmul(cpx a, cpx b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cpx cmulr(cpx a, double b) { return mk(a.re * b, a.im * b); }
__device__ __forceinline__ cpx cdivr(cpx a, double b) { return mk(a.re / b, a.im / b); }
__device__ __forceinline__ cpx cconj(cpx a) { return mk(a.re, -a.im); }
// Smith's range-reduced quotient
__device__ __forceinline__ cpx cdiv(cpx a, cpx b) {
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
  return mk(tr / div, ti / div);
}
// principal square root (glibc csqrt formulas for finite arguments)
__device__ __forceinline__ cpx csqrt_(cpx x) {
  if (x.im == 0.0) {
    if (x.re < 0.0) return mk(0.0, copysign(sqrt(-x.re), x.im));
    return mk(fabs(sqrt(x.re)), copysign(0.0, x.im));
  }
  if (x.re == 0.0) {
    double r = (fabs(x.im) >= 2.0 * 2.2250738585072014e-308)
                   ? sqrt(0.5 * fabs(x.im))
                   : 0.5 * sqrt(2.0 * fabs(x.im));
    return mk(r, copysign(r, x.im));
  }
  double d = det_hypot(x.re, x.im), r, s;
  if (x.re > 0.0) {
    r = sqrt(0.5 * (d + x.re));
    s = 0.5 * (x.im / r);
  } else {
    s = sqrt(0.5 * (d - x.re));
    r = fabs(0.5 * (x.im / s));
  }
  return mk(r, copysign(s, x.im));
}

// ---------------------------------------------------------------------------
// limiters (src/geometry_2d.f90:1555-1676)
// ---------------------------------------------------------------------------
// minmod(a,b) = 0.5*(sign a + sign b)*min(|a|,|b|); identical for all
// finite arguments to the reference's division-based form.
__device__ __forceinline__ double minmod_(double a, double b) {
  double m = fmin(fabs(a), fabs(b));
  if (!(m > 0.0)) return 0.0;
  if (a > 0.0 && b > 0.0) return m;
  if (a < 0.0 && b < 0.0) return -m;
  return 0.0;
}
__device__ __forceinline__ double maxmod_(double a, double b) {
  if (fabs(a * b) <= 0.0) return 0.0;
  double m = fmax(fabs(a), fabs(b));
  if (a > 0.0 && b > 0.0) return m;
  if (a < 0.0 && b < 0.0) return -m;
  return 0.0;
}
__device__ __forceinline__ double limit_(const DevParams& P, double v0, double v1,
                                         double v2, double z0, double z1,
                                         double z2, int lim) {
  if (lim == 0) return 0.0;
  double a = (v2 - v1) / (z2 - z1);
  double b = (v1 - v0) / (z1 - z0);
  switch (lim) {
    case 1: return minmod_(a, b);
    case 2: return maxmod_(minmod_(a, 2.0 * b), minmod_(2.0 * a, b));
    case 6: return a;
    case 7: return b;
    default: break;
  }
  double c = (v2 - v0) / (z2 - z0);
  if (lim == 3) return minmod_(c, P.theta * minmod_(a, b));
  if (lim == 4) return minmod_(c, 2.0 * minmod_(a, b));
  return c;  // 5: centred
}
// limit_ from its differences a, b and (when P.lim_c) c, already divided:
// the same value, with the divisions of all variables of a reconstruction
// issued together rather than one variable's after another's branches.
__device__ __forceinline__ double limit_pick(const DevParams& P, double a, double b, double c,
                                             int lim) {
  if (lim == 0) return 0.0;
  switch (lim) {
    case 1: return minmod_(a, b);
    case 2: return maxmod_(minmod_(a, 2.0 * b), minmod_(2.0 * a, b));
    case 6: return a;
    case 7: return b;
    default: break;
  }
  if (lim == 3) return minmod_(c, P.theta * minmod_(a, b));
  if (lim == 4) return minmod_(c, 2.0 * minmod_(a, b));
  return c;
}

// settling_velocity (src/constitutive_2d.f90:5148-5214)
__device__ __forceinline__ double settling_velocity(const DevParams& P, double diam,
                                                    double rhos, double rhoc,
                                                    double inv_kin_visc) {
  if (rhos <= rhoc) return 0.0;
  double cp = sqrt(0.75 * (rhos / rhoc - 1.0) * diam * P.grav);
  double sv = cp * 1.0;
  double Rey = diam * sv * inv_kin_visc;
  if (Rey <= 1000.0) {
    for (int i = 1; i <= 20; i++) {
      double old = sv;
      double isc = sqrt(Rey / (24.0 * (1.0 + 0.15 * det_pow(Rey, 0.687))));
      sv = cp * isc;
      if (fabs(old - sv) / old < 1.0e-6) return sv;
      Rey = diam * sv * inv_kin_visc;
    }
  }
  return sv;
}

}  // namespace imexsw

#include "imexsw_profiles.h"

namespace imexsw {

// ---------------------------------------------------------------------------
// variable transforms (src/constitutive_2d.f90:541-967, 1333-2179)
// ---------------------------------------------------------------------------
// qc_to_qp (r_phys_var, gas carrier; VP: the profile branch's mean speed).
// qp = h, hu, hv, T, (h*)alphas, u, v; returns p_dyn.
template <int NS, bool VP = false>
__device__ __forceinline__ double phys(const DevParams& P, const double* qj,
                                       double* qp) {
  constexpr int NV = 4 + NS;
  if (!(qj[0] > kEps)) {
#pragma unroll
    for (int i = 0; i < NV + 2; i++) qp[i] = 0.0;
    qp[3] = P.T_ambient;
    return 0.0;
  }
  double inv_qj1 = 1.0 / qj[0];
  double xs[NS > 0 ? NS : 1];
  double sq = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) xs[i] = qj[4 + i] * inv_qj1;
#pragma unroll
  for (int i = 0; i < NS; i++) sq += qj[4 + i];
  if (sq == qj[0]) {
    double sx = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) sx += xs[i];
#pragma unroll
    for (int i = 0; i < NS; i++) xs[i] = xs[i] / sx;
  }
  double xs_tot = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) xs_tot += xs[i];
  double xc = 1.0 - xs_tot;
  double shc_by_xc = (xc - 0.0) * P.sp_heat_a + 0.0;
  double dsh = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) dsh += xs[i] * P.sp_heat_s[i];
  double shm = dsh + shc_by_xc;
  double T;
  if (qj[0] > P.eps_sing) {
    if (P.energy)
      T = (qj[3] - 0.5 * (qj[1] * qj[1] + qj[2] * qj[2]) * inv_qj1) / (qj[0] * shm);
    else
      T = qj[3] / (qj[0] * shm);
    if (T <= 0.0) T = P.T_ambient;
  } else {
    T = P.T_ambient;
  }
  double rgc = (xc > kEps) ? ((xc - 0.0) * P.sp_gas_const_a + 0.0) / xc
                           : P.sp_gas_const_a;
  double inv_rho_c = rgc * T * P.inv_pres;
  double dinv = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) dinv += xs[i] * P.inv_rho_s[i];
  double inv_rhom = dinv + xc * inv_rho_c;
  double rho_m = 1.0 / inv_rhom;
  double h = qj[0] * inv_rhom;
  double u, v;
  if (VP) {
    double al[NS > 0 ? NS : 1];
#pragma unroll
    for (int i = 0; i < NS; i++) al[i] = xs[i] * rho_m * P.inv_rho_s[i];
    const double rho_c = P.pres / (rgc * T);
    const double ua = vp_velocity<NS>(P, h, qj[1], qj[2], rho_m, rho_c, al);
    u = ua * qj[1] / (sqrt(qj[1] * qj[1] + qj[2] * qj[2]));
    v = ua * qj[2] / (sqrt(qj[1] * qj[1] + qj[2] * qj[2]));
  } else if (qj[0] > P.eps_sing) {
    u = qj[1] * inv_qj1;
    v = qj[2] * inv_qj1;
  } else {
    double q1sq = qj[0] * qj[0];
    u = sqrt(2.0) * qj[0] * qj[1] / sqrt(q1sq * q1sq + P.eps_sing4);
    v = sqrt(2.0) * qj[0] * qj[2] / sqrt(q1sq * q1sq + P.eps_sing4);
  }
  qp[0] = h;
  qp[1] = h * u;
  qp[2] = h * v;
  qp[3] = T;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    double al = xs[i] * rho_m * P.inv_rho_s[i];
    qp[4 + i] = P.alpha ? al : al * h;
  }
  qp[NV] = u;
  qp[NV + 1] = v;
  return 0.5 * rho_m * (u * u + v * v);
}

// "phys of q if q1 > 0 else the dry state" as the solver loops write it
template <int NS, bool VP = false>
__device__ __forceinline__ double phys_or_dry(const DevParams& P, const double* q,
                                              double* qp) {
  constexpr int NV = 4 + NS;
  if (q[0] > 0.0) return phys<NS, VP>(P, q, qp);
#pragma unroll
  for (int i = 0; i < NV + 2; i++) qp[i] = 0.0;
  qp[3] = P.T_ambient;
  return 0.0;
}

struct Mixt {
  double ri, rho_m, rho_c, rg, shc, shm;
};

// mixt_var (src/constitutive_2d.f90:1333-1483) incl. eval_sp_heat
template <int NS, bool HEAT>
__device__ __forceinline__ Mixt mixt(const DevParams& P, const double* qp) {
  constexpr int NV = 4 + NS;
  Mixt m;
  if (qp[0] <= kEps) {
    m.rg = 0.0;
    m.rho_m = P.rho_a_amb;
    m.rho_c = P.rho_a_amb;
    m.shc = P.sp_heat_a;
    m.shm = P.sp_heat_a;
    m.ri = 0.0;
    return m;
  }
  double r_h = qp[0], r_u = qp[NV], r_v = qp[NV + 1], r_T = qp[3];
  double al[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) al[i] = P.alpha ? qp[4 + i] : qp[4 + i] / qp[0];
  double at = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) at += al[i];
  double rho_a = P.pres / (P.sp_gas_const_a * r_T);
  double rho_c = ((1.0 - 0.0 - at - 0.0) * rho_a + 0.0) / (1.0 - 0.0 - at);
  double d = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) d += al[i] * P.rho_s[i];
  double rho_m = (1.0 - at) * rho_c + d;
  double inv_rhom = 1.0 / rho_m;
  double rg = (rho_m - P.rho_a_amb) / rho_m * P.grav;
  double u2 = r_u * r_u + r_v * r_v;
  m.ri = (u2 > 0.0) ? fmin(1.0e15, rg * r_h / u2) : 0.0;
  m.rho_m = rho_m;
  m.rho_c = rho_c;
  m.rg = rg;
  m.shc = 0.0;
  m.shm = 0.0;
  if (HEAT) {
    double xs[NS > 0 ? NS : 1], sxs = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) xs[i] = al[i] * P.rho_s[i] * inv_rhom;
#pragma unroll
    for (int i = 0; i < NS; i++) sxs += xs[i];
    double xc = 1.0 - sxs;
    double c = (xc > kEps) ? ((xc - 0.0) * P.sp_heat_a + 0.0) / xc : P.sp_heat_a;
    double dd = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) dd += xs[i] * P.sp_heat_s[i];
    m.shc = c;
    m.shm = dd + xc * c;
  }
  return m;
}

// qp_to_qc (src/constitutive_2d.f90:1709-2140, no liquid; VP: the profile
// momentum, with the state's h0 and settling speeds from pre when given)
template <int NS, bool VP = false>
__device__ __forceinline__ void qp2qc(const DevParams& P, const double* qp,
                                      double* qc, const VpPre<NS>* pre = nullptr) {
  constexpr int NV = 4 + NS;
  double r_h = qp[0];
  if (!(r_h > kEps)) {
#pragma unroll
    for (int i = 0; i < NV; i++) qc[i] = 0.0;
    return;
  }
  double r_hu = qp[1], r_hv = qp[2], r_u = qp[NV], r_v = qp[NV + 1], r_T = qp[3];
  double ars[NS > 0 ? NS : 1], al[NS > 0 ? NS : 1], at = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    al[i] = P.alpha ? qp[4 + i] : qp[4 + i] / qp[0];
    at += al[i];
    ars[i] = al[i] * P.rho_s[i];
  }
  double rho_a = P.pres / (P.sp_gas_const_a * r_T);
  double rho_c = ((1.0 - 0.0 - at - 0.0) * rho_a + 0.0) / (1.0 - 0.0 - at);
  double alphac = 1.0 - at;
  double sar = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) sar += ars[i];
  double rho_m = alphac * rho_c + sar;
  double inv_rhom = 1.0 / rho_m;
  double xs[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) xs[i] = ars[i] * inv_rhom;
  double xc = alphac * rho_c * inv_rhom;
  double shc = (xc > kEps) ? ((xc - 0.0) * P.sp_heat_a + 0.0) / xc : P.sp_heat_a;
  double dsh = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) dsh += xs[i] * P.sp_heat_s[i];
  double shm = dsh + xc * shc;
  qc[0] = rho_m * r_h;
  if (VP) {
    // the alphas corrections of :1911-1919 (only the profiles read them)
    double ra[NS > 0 ? NS : 1];
#pragma unroll
    for (int i = 0; i < NS; i++) ra[i] = (at > 1.0) ? al[i] / at : (at < 0.0) ? 0.0 : al[i];
    const VpPre<NS> pr = pre ? *pre : vp_pre<NS>(P, r_h, rho_c);
    vp_momentum<NS>(P, r_h, r_u, r_v, pr, rho_c, ra, qc[1], qc[2]);
  } else {
    qc[1] = rho_m * r_hu;
    qc[2] = rho_m * r_hv;
  }
  if (P.energy)
    qc[3] = (r_h > 0.0) ? r_h * rho_m * (shm * r_T + 0.5 * (r_u * r_u + r_v * r_v))
                        : 0.0;
  else
    qc[3] = r_h * rho_m * shm * r_T;
#pragma unroll
  for (int i = 0; i < NS; i++) qc[4 + i] = xs[i] * qc[0];
}

// ---------------------------------------------------------------------------
// wave speeds and physical fluxes (src/constitutive_2d.f90:2194-2509)
// ---------------------------------------------------------------------------
template <int NS, int DIR>
__device__ __forceinline__ void speeds(const DevParams& P, const double* qp,
                                       double gc, double& vmin, double& vmax) {
  constexpr int NV = 4 + NS;
  Mixt m = mixt<NS, false>(P, qp);
  double r_h = qp[0];
  double vel = (DIR == 1) ? qp[NV] : qp[NV + 1];
  if (m.rg * r_h < 0.0) {
    vmin = vel;
    vmax = vel;
  } else {
    double cel = (DIR == 1) ? sqrt(m.rg * r_h * gc) : sqrt(gc * m.rg * r_h);
    vmin = vel - cel;
    vmax = vel + cel;
  }
}

// eval_fluxes; VP: with the profile shape coefficients, which read the slope
// (Bx, By) of the cell the face state belongs to
template <int NS, int DIR, bool VP = false>
__device__ __forceinline__ void fluxes(const DevParams& P, const double* qc,
                                       const double* qp, double gc, double* f,
                                       double Bx = 0.0, double By = 0.0,
                                       const VpPre<NS>* pre = nullptr) {
  constexpr int NV = 4 + NS;
  if (!(qp[0] > kEps)) {
#pragma unroll
    for (int i = 0; i < NV; i++) f[i] = 0.0;
    return;
  }
  double r_h = qp[0], r_u = qp[NV], r_v = qp[NV + 1];
  Mixt m = mixt<NS, false>(P, qp);
  double sc[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) sc[i] = 1.0;
  if (VP) {
    const VpPre<NS> pr = pre ? *pre : vp_pre<NS>(P, r_h, m.rho_c);
    vp_flux_coeffs<NS>(P, qp, Bx, By, pr, m.rho_c, m.rho_m, sc);
  }
  double vel = (DIR == 1) ? r_u : r_v;
  double pt = 0.5 * m.rho_m * gc * m.rg * (r_h * r_h);
  f[0] = vel * qc[0] * sc[0];
  if (DIR == 1) {
    f[1] = r_u * qc[1] * sc[1] + pt;
    f[2] = r_u * qc[2] * sc[2];
  } else {
    f[1] = r_v * qc[1] * sc[1];
    f[2] = r_v * qc[2] * sc[2] + pt;
  }
  f[3] = P.energy ? vel * (qc[3] * sc[3] + pt) : vel * qc[3] * sc[3];
#pragma unroll
  for (int i = 0; i < NS; i++) f[4 + i] = vel * qc[4 + i] * sc[4 + i];
  double sf = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) sf += f[4 + i];
  if ((f[0] > 0.0) && (sf / f[0] > 1.0)) {
#pragma unroll
    for (int i = 0; i < NS; i++) f[4 + i] = f[4 + i] / sf * f[0];
  }
}

// Kurganov-Tadmor numerical flux at one face from both sides' fluxes and
// conservative values (src/solver_2d.f90:2862-2922, 3017-3047).  For the x
// direction the zero-velocity test reads the reconstructed hu of both sides,
// for y the conservative rho*h*v (:2914-2920 vs :2982-2988), and the solid
// cap compares with >= in x and > in y with different operand groupings.
template <int NS, int DIR>
__device__ __forceinline__ void kt_flux(const double* fL, const double* fR,
                                        const double* qL, const double* qR,
                                        double aN, double aP, double zvL,
                                        double zvR, double* H) {
  constexpr int NV = 4 + NS;
  const bool same = (aN == aP);
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double avg = same ? 0.0 : (aP * fL[i] - aN * fR[i]) / (aP - aN);
    H[i] = same ? 0.0 : avg + (aP * aN) / (aP - aN) * (qR[i] - qL[i]);
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) s += H[4 + i];
  if (DIR == 1) {
    if (s / H[0] >= 1.0) {
      double sr = 0.0;
#pragma unroll
      for (int i = 0; i < NS; i++) sr += H[4 + i] / H[0];
#pragma unroll
      for (int i = 0; i < NS; i++) H[4 + i] = H[4 + i] / sr;
    }
  } else {
    if (s / H[0] > 1.0) {
      double r = s / H[0];
#pragma unroll
      for (int i = 0; i < NS; i++) H[4 + i] = H[4 + i] / r;
    }
  }
  if (zvL == 0.0 && zvR == 0.0) {
    H[0] = 0.0;
#pragma unroll
    for (int i = 3; i < NV; i++) H[i] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// source terms
// ---------------------------------------------------------------------------
// eval_expl_terms without bottom source (src/constitutive_2d.f90:3015-3245)
template <int NS>
__device__ __forceinline__ void expl_terms(const DevParams& P, double Bx, double By,
                                           double Bxx, double Bxy, double Byy,
                                           double gc, double dgdx,
                                           const double* qp, double* ex) {
  constexpr int NV = 4 + NS;
#pragma unroll
  for (int i = 0; i < NV; i++) ex[i] = 0.0;
  if (qp[0] <= kEps) return;
  double r_h = qp[0], r_u = qp[NV], r_v = qp[NV + 1];
  Mixt m = mixt<NS, true>(P, qp);
  double cft = P.curv ? Bxx * (r_u * r_u) + 2.0 * Bxy * r_u * r_v + Byy * (r_v * r_v)
                      : 0.0;
  double tg = m.rg + cft;
  ex[1] = -(gc * m.rho_m * tg * r_h * Bx) + 0.5 * m.rho_m * m.rg * (r_h * r_h) * dgdx;
  ex[2] = -(gc * m.rho_m * tg * r_h * By) + 0.5 * m.rho_m * m.rg * (r_h * r_h) * dgdx;
  ex[3] = P.energy ? ex[1] * r_u + ex[2] * r_v : 0.0;
}

// Voellmy-Salm Coulomb part, semi-implicit (src/constitutive_2d.f90:3979-4028)
template <int NS>
__device__ __forceinline__ void semi_impl(const DevParams& P, double Bx, double By,
                                          double Bxx, double Bxy, double Byy,
                                          double gc, const double* qp,
                                          double& s1, double& s2) {
  constexpr int NV = 4 + NS;
  s1 = 0.0;
  s2 = 0.0;
  if (!P.rheo) return;
  double r_h = qp[0], r_u = qp[NV], r_v = qp[NV + 1];
  Mixt m = mixt<NS, false>(P, qp);
  double r_w = P.slope ? r_u * Bx + r_v * By : 0.0;
  double mod_vel = sqrt(r_u * r_u + r_v * r_v + r_w * r_w);
  if (P.rheo_model == 1 && mod_vel > 0.0) {
    double cft = P.curv ? Bxx * (r_u * r_u) + 2.0 * Bxy * r_u * r_v + Byy * (r_v * r_v)
                        : 0.0;
    double tt = P.mu * m.rho_m * (m.rg + cft) * r_h * sqrt(gc);
    tt = fmax(0.0, tt);
    s1 = 0.0 - tt * r_u / mod_vel;
    s2 = 0.0 - tt * r_v / mod_vel;
  }
}

// Turbulent drag, implicit, complex arguments (eval_implicit_terms +
// c_phys_var, src/constitutive_2d.f90:3327-3823, 1145-1313): Voellmy-Salm
// (rheology model 1), or with RM = 6 the quadratic drag rho_m f |u|^2 of
// model 6 (:3639-3650).  Only the momentum components are non-zero.
template <int NS, int RM = 1>
__device__ __forceinline__ void implicit_c(const DevParams& P, const cpx* cq,
                                           double Bx, double By, cpx& n1,
                                           cpx& n2) {
  n1 = mk(0.0, 0.0);
  n2 = mk(0.0, 0.0);
  if (!P.rheo) return;
  cpx rho_m, u, v;
  if (cq[0].re > kEps) {
    cpx inv_q1 = cdiv(mk(1.0, 0.0), cq[0]);
    cpx xs[NS > 0 ? NS : 1];
#pragma unroll
    for (int i = 0; i < NS; i++) xs[i] = cmul(cq[4 + i], inv_q1);
    cpx xs_tot = mk(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) xs_tot = cadd(xs_tot, xs[i]);
    cpx xc = csub(mk(1.0, 0.0), xs_tot);
    cpx shc = cdiv(cadd(cmulr(csub(xc, mk(0.0, 0.0)), P.sp_heat_a), mk(0.0, 0.0)), xc);
    cpx dsh = mk(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) dsh = cadd(dsh, cmul(cconj(xs[i]), mk(P.sp_heat_s[i], 0.0)));
    cpx shm = cadd(dsh, cmul(xc, shc));
    cpx T;
    if (cq[0].re > P.eps_sing) {
      if (P.energy) {
        cpx ke = cmul(cmulr(cadd(cmul(cq[1], cq[1]), cmul(cq[2], cq[2])), 0.5), inv_q1);
        T = cdiv(csub(cq[3], ke), cmul(cq[0], shm));
      } else {
        T = cdiv(cq[3], cmul(cq[0], shm));
      }
      if (T.re <= 0.0) T = mk(P.T_ambient, 0.0);
    } else {
      T = mk(P.T_ambient, 0.0);
    }
    cpx rgc = cdiv(cadd(cmulr(csub(xc, mk(0.0, 0.0)), P.sp_gas_const_a), mk(0.0, 0.0)), xc);
    cpx inv_rho_c = cmulr(cmul(rgc, T), P.inv_pres);
    cpx dinv = mk(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NS; i++) dinv = cadd(dinv, cmul(cconj(xs[i]), mk(P.inv_rho_s[i], 0.0)));
    cpx inv_rhom = cadd(dinv, cmul(xc, inv_rho_c));
    rho_m = cdiv(mk(1.0, 0.0), inv_rhom);
    if (cq[0].re > P.eps_sing) {
      u = cmul(cq[1], inv_q1);
      v = cmul(cq[2], inv_q1);
    } else {
      cpx q1sq = cmul(cq[0], cq[0]);
      cpx den = csqrt_(cadd(cmul(q1sq, q1sq), mk(P.eps_sing4, 0.0)));
      u = cdiv(cmul(cmulr(cq[0], sqrt(2.0)), cq[1]), den);
      v = cdiv(cmul(cmulr(cq[0], sqrt(2.0)), cq[2]), den);
    }
  } else {
    u = mk(0.0, 0.0);
    v = mk(0.0, 0.0);
    rho_m = mk(P.rho_a_amb, 0.0);
  }
  cpx rg = cmulr(cdiv(csub(rho_m, mk(P.rho_a_amb, 0.0)), rho_m), P.grav);
  cpx w = P.slope ? cadd(cmulr(u, Bx), cmulr(v, By)) : mk(0.0, 0.0);
  cpx mv2 = cadd(cadd(cmul(u, u), cmul(v, v)), cmul(w, w));
  cpx mv = csqrt_(mv2);
  if (RM == 6) {
    const cpx mvh = csqrt_(cadd(cmul(u, u), cmul(v, v)));
    if (mv.re != 0.0) {
      n1 = csub(n1, cmul(cmulr(cmul(rho_m, cdiv(u, mvh)), P.ff), mv2));
      n2 = csub(n2, cmul(cmulr(cmul(rho_m, cdiv(v, mvh)), P.ff), mv2));
    }
  } else if (P.rheo_model == 1 && mv.re != 0.0) {
    cpx turb = cmul(cdivr(cmul(rho_m, rg), P.xi), mv2);
    n1 = csub(n1, cmul(turb, cdiv(u, mv)));
    n2 = csub(n2, cmul(turb, cdiv(v, mv)));
  }
}

// eval_source_bdry (src/constitutive_2d.f90:5024-5130), split into the
// time factors (the same for every ring cell of a step) and the per-cell
// part that scales them by the cell's direction vector
struct SrcTime {
  bool off;           // time >= time_param(4): the source is switched off
  double tc, p15, p05;  // time coefficient, tc**1.5, tc**0.5
};

__device__ __forceinline__ SrcTime src_time(const DevParams& P, double time) {
  const double* tp = P.time_param;
  SrcTime r{false, 0.0, 0.0, 0.0};
  if (time >= tp[3]) {
    r.off = true;
    return r;
  }
  double t_rem = fmod(time, tp[0]);
  double tc = 0.0;
  if (tp[2] == 0.0) {
    if (t_rem <= tp[1]) tc = 1.0;
  } else {
    if (t_rem < tp[2])
      tc = 0.5 * (1.0 - det_cos(P.pi_g * t_rem / tp[2]));
    else if (t_rem <= (tp[1] - tp[2]))
      tc = 1.0;
    else if (t_rem <= tp[1])
      tc = 0.5 * (1.0 + det_cos(P.pi_g * ((t_rem - tp[1]) / tp[2] + 1.0)));
  }
  r.tc = tc;
  r.p15 = det_pow(tc, 1.5);
  r.p05 = det_pow(tc, 0.5);
  return r;
}

template <int NS>
__device__ __forceinline__ void source_bdry(const DevParams& P, const SrcTime& T, double vx,
                                            double vy, double* sb) {
  constexpr int NV = 4 + NS;
  if (T.off) {
#pragma unroll
    for (int i = 0; i < NV + 2; i++) sb[i] = 0.0;
    sb[3] = P.T_source;
    return;
  }
  const double tc = T.tc;
  sb[0] = tc * P.h_source;
  sb[1] = T.p15 * P.h_source * P.vel_source * vx;
  sb[2] = T.p15 * P.h_source * P.vel_source * vy;
  sb[3] = P.T_source;
#pragma unroll
  for (int i = 0; i < NS; i++)
    sb[4 + i] = P.alpha ? P.alphas_source[i] : tc * P.h_source * P.alphas_source[i];
  sb[NV] = T.p05 * P.vel_source * vx;
  sb[NV + 1] = T.p05 * P.vel_source * vy;
}

}  // namespace imexsw
