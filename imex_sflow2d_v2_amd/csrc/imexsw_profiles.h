// imexsw_profiles.h -- vertical-profile closures (VERTICAL_PROFILES_FLAG) as
// device functions: a log velocity profile u ~ log(30 z / k_s + 1) below a
// boundary-layer top h0 and uniform above it, exponential solid
// concentrations exp(a z) with a = -w_s / epsilon_s (Rouse), their depth
// averages by an n_quad-point Gauss-Legendre rule on [0, h0] plus the
// uniform part on [h0, h].
//   r_phys_var's mean speed (Aitken iteration on avg_profiles_mix),
//     src/constitutive_2d.f90:797-942, 970-1123
//   qp_to_qc's profile momentum, src/constitutive_2d.f90:1976-2107
//   eval_flux_coeffs, src/constitutive_2d.f90:2532-2823
//   eval_dep_coeffs, src/constitutive_2d.f90:2844-2986
//   Lambert W on both real branches, src/geometry_2d.f90:1776-2038
//   integrate_friction_term (rheology model 8), src/constitutive_2d.f90:
//     3252-3306
//
// The quadrature points stay out of arrays: every sum runs over the points
// in order, recomputing z_m, w_m and log(b z_m + 1) per point (one log), and
// advances every solid class's sums together, so a cell holds a few
// accumulators per class instead of n_quad-long arrays.  Each sum is taken
// in the same order as the oracle's, so the results agree bit for bit.
#pragma once

#include "imexsw_math.h"

namespace imexsw {

// ---------------------------------------------------------------------------
// Lambert W for x in [-1/e, 0) (and x >= 0 on the principal branch)
// ---------------------------------------------------------------------------
// W = sum mu_k p^k about the branch point, p = +-sqrt(2 (e x + 1)) (Corless,
// Gonnet, Hare, Jeffrey & Knuth 1996, eqs. 4.22-4.24); used for |p| < 0.1
__device__ __forceinline__ double lw_series(double p) {
  constexpr double mu[13] = {
      -1.0, 1.0, -1.0 / 3.0, 11.0 / 72.0, -43.0 / 540.0, 769.0 / 17280.0,
      -221.0 / 8505.0, 680863.0 / 43545600.0, -1963.0 / 204120.0,
      226287557.0 / 37623398400.0, -5776369.0 / 1515591000.0,
      169709463197.0 / 69528040243200.0, -1118511313.0 / 709296588000.0};
  double s = mu[12];
#pragma unroll
  for (int k = 11; k >= 0; k--) s = s * p + mu[k];
  return s;
}

// Halley's iteration on w e^w = x until the step is below 1e-15 |w|
__device__ __forceinline__ double lw_halley(double x, double w) {
  for (int it = 0; it < 20; it++) {
    const double ew = det_exp(w);
    const double f = w * ew - x;
    const double wp1 = w + 1.0;
    const double dw = f / (ew * wp1 - (w + 2.0) * f / (2.0 * wp1));
    w = w - dw;
    if (!(fabs(dw) > 1.0e-15 * fabs(w))) break;
  }
  return w;
}

constexpr double kE = 2.718281828459045;

// W_0; -1 outside the domain (as lambertw0 returns)
__device__ __forceinline__ double lambertw0(double x) {
  const double p2 = 2.0 * (kE * x + 1.0);
  if (!(p2 > 0.0)) return -1.0;
  if (p2 < 0.01) return lw_series(sqrt(p2));
  double w;
  if (x < -0.25) {
    const double p = sqrt(p2);
    w = -1.0 + p * (1.0 + p * (-1.0 / 3.0 + p * (11.0 / 72.0)));
  } else {
    // Winitzki (2003): W ~ L (1 - log(1 + L) / (2 + L)), L = log(1 + x)
    const double L = det_log(1.0 + x);
    w = (x == 0.0) ? 0.0 : L * (1.0 - det_log(1.0 + L) / (2.0 + L));
  }
  return lw_halley(x, w);
}

// W_{-1}; -1 outside the domain and above -64 e^-64 (as lambertwm1 returns)
__device__ __forceinline__ double lambertwm1(double x) {
  const double p2 = 2.0 * (kE * x + 1.0);
  if (!(p2 > 0.0) || !(x < 0.0) || x > -1.0288423856755186e-26) return -1.0;
  if (p2 < 0.01) return lw_series(-sqrt(p2));
  double w;
  if (x < -0.25) {
    const double p = -sqrt(p2);
    w = -1.0 + p * (1.0 + p * (-1.0 / 3.0 + p * (11.0 / 72.0)));
  } else {
    const double L1 = det_log(-x), L2 = det_log(-L1);
    w = L1 - L2 + L2 / L1;
  }
  return lw_halley(x, w);
}

// ---------------------------------------------------------------------------
// profile geometry
// ---------------------------------------------------------------------------
// boundary-layer top: the smaller root z of b log(c z + 1) - z = a in units
// of k_s when h/k_s > H_crit_rel, else h (src/constitutive_2d.f90:821-854)
__device__ __forceinline__ double vp_h0(const DevParams& P, double r_h) {
  const double h_rel = r_h / P.k_s;
  double h0_rel;
  if (h_rel > P.h_crit_rel) {
    const double a = h_rel * P.vonK / P.sqrt_ff;
    const double b = 1.0 / 30.0 + h_rel, c = 30.0;
    const double d = a / b - 1.0 / (b * c);
    const double arg = -det_exp(d) / (b * c);
    const double h1 = -(b * lambertw0(arg)) - 1.0 / c;
    const double h2 = -(b * lambertwm1(arg)) - 1.0 / c;
    h0_rel = fmin(h1, h2);
  } else {
    h0_rel = h_rel;
  }
  return h0_rel * P.k_s;
}

// point m of the rule on [0, h0]: z, weight and log(b z + 1)
__device__ __forceinline__ void vp_point(const DevParams& P, int m, double h0, double b,
                                         double& z, double& w, double& lz) {
  z = 0.5 * h0 * (P.zq[m] + 1.0);
  w = 0.5 * h0 * P.wq[m];
  lz = det_log(b * z + 1.0);
}

// What the closures of one state share whatever its velocity: the
// boundary-layer top (from h) and the settling speeds (from the carrier
// density), so that a face side's momentum and flux coefficients form
// them once
template <int NS>
struct VpPre {
  double h0;
  double sv[NS > 0 ? NS : 1];
};
template <int NS>
__device__ __forceinline__ VpPre<NS> vp_pre(const DevParams& P, double r_h, double rho_c) {
  VpPre<NS> r;
  r.h0 = vp_h0(P, r_h);
  const double ikv = 1.0 / P.kin_visc_c;
#pragma unroll
  for (int i = 0; i < NS; i++) r.sv[i] = settling_velocity(P, P.diam_s[i], P.rho_s[i], rho_c, ikv);
  return r;
}

// Rouse numbers w_s / (vonK u*) (zero without shear)
template <int NS>
__device__ __forceinline__ void vp_rouse(const DevParams& P, double shear_vel, const double* sv,
                                         double* rouse) {
#pragma unroll
  for (int i = 0; i < NS; i++) rouse[i] = (shear_vel > 0.0) ? sv[i] / (P.vonK * shear_vel) : 0.0;
}

// The sums every closure starts from, for the solid exponents a_i:
//   slog = sum w lz, se_i = sum w e_i, sq_i = sum w (e_i lz),
//   sl_i = sum e_i (w lz^2) and sll = sum w lz^2 (FULL only),
// with e_i = exp(a_i z), over the points in order.
template <int NS, bool FULL>
__device__ __forceinline__ void vp_sums(const DevParams& P, double h0, double b, const double* a,
                                        double& slog, double* se, double* sq, double* sl,
                                        double& sll) {
  slog = 0.0;
  sll = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    se[i] = 0.0;
    sq[i] = 0.0;
    sl[i] = 0.0;
  }
  for (int m = 0; m < P.nq; m++) {
    double z, w, lz;
    vp_point(P, m, h0, b, z, w, lz);
    slog += w * lz;
    const double wll = w * (lz * lz);
    if (FULL) sll += wll;
#pragma unroll
    for (int i = 0; i < NS; i++) {
      const double e = det_exp(a[i] * z);
      se[i] += w * e;
      sq[i] += w * (e * lz);
      if (FULL) sl[i] += e * wll;
    }
  }
}

// avg_profiles_mix (src/constitutive_2d.f90:970-1123): the profile-averaged
// mixture momentum per unit depth at mean speed u_guess (its dynamic
// pressure output is overwritten by the caller and not formed)
template <int NS>
__device__ __forceinline__ double vp_avg_mix(const DevParams& P, double h, const double* sv,
                                             const double* ra, double u_guess, double h0,
                                             double b, double u_rel0, double rho_c) {
  const double shear_vel = u_guess * P.sqrt_ff;
  const double eps_s = P.Sc * shear_vel * P.vonK * ((h0 / 6.0) + (P.k_s / 60.0));
  const double a_coeff = -(P.vonK * shear_vel / eps_s);
  double a[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++)
    a[i] = a_coeff * ((shear_vel > 0.0) ? sv[i] / (P.vonK * shear_vel) : 0.0);
  double slog, sll, se[NS > 0 ? NS : 1], sq[NS > 0 ? NS : 1], sl[NS > 0 ? NS : 1];
  vp_sums<NS, false>(P, h0, b, a, slog, se, sq, sl, sll);
  const double lh0 = det_log(b * h0 + 1.0);
  const double ncu = 1.0 / ((slog + lh0 * (h - h0)) / h);
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double eh0 = det_exp(a[i] * h0);
    const double nca = 1.0 / ((se[i] + eh0 * (h - h0)) / h);
    const double d1 = ncu * nca * sq[i];
    const double d2 = (h - h0) * u_rel0 * (nca * eh0);
    const double rua = ra[i] * u_guess * (d1 + d2) / h;
    s += (P.rho_s[i] - rho_c) / P.rho_s[i] * rua;
  }
  return u_guess * rho_c + s;
}

// r_phys_var's mean speed (src/constitutive_2d.f90:797-942): the fixed point
// of u <- u |q| / (h M(u)) with Aitken's acceleration, at most ten rounds,
// tolerances 1e-8 as the reference's single-precision literals round
template <int NS>
__device__ __forceinline__ double vp_velocity(const DevParams& P, double r_h, double q2,
                                              double q3, double r_rho_m, double r_rho_c,
                                              const double* r_alphas) {
  const double ikv = 1.0 / P.kin_visc_c;
  double sv[NS > 0 ? NS : 1], ra[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    ra[i] = r_alphas[i] * P.rho_s[i];
    sv[i] = settling_velocity(P, P.diam_s[i], P.rho_s[i], r_rho_c, ikv);
  }
  const double h0 = vp_h0(P, r_h);
  const double b = 30.0 / P.k_s;
  double slog = 0.0;
  for (int m = 0; m < P.nq; m++) {
    double z, w, lz;
    vp_point(P, m, h0, b, z, w, lz);
    slog += w * lz;
  }
  const double lh0 = det_log(b * h0 + 1.0);
  const double ncu = 1.0 / ((slog + lh0 * (r_h - h0)) / r_h);
  const double u_rel0 = ncu * lh0;
  const double uRho_avg = sqrt(q2 * q2 + q3 * q3) / r_h;
  double u_guess = uRho_avg / r_rho_m, u_new = u_guess;
  const double rel_tol = (double)1.e-8f, abs_tol = (double)1.e-8f;
  for (int it = 0; it < 10; it++) {
    const double x0 = u_guess;
    double m = vp_avg_mix<NS>(P, r_h, sv, ra, u_guess, h0, b, u_rel0, r_rho_c);
    u_new = u_guess * uRho_avg / m;
    const double x1 = u_new;
    m = vp_avg_mix<NS>(P, r_h, sv, ra, u_new, h0, b, u_rel0, r_rho_c);
    u_new = u_new * uRho_avg / m;
    const double x2 = u_new;
    const double den = (x2 - x1) - (x1 - x0);
    if (fabs(den) < (double)0.1f * abs_tol) break;
    u_new = x2 - ((x2 - x1) * (x2 - x1)) / den;
    if ((fabs(u_guess - u_new) / u_guess < rel_tol) || (fabs(u_guess - u_new) < abs_tol)) break;
    u_guess = u_new;
  }
  return u_new;
}

// qp_to_qc's profile momentum (src/constitutive_2d.f90:1976-2107); pre: the
// state's h0 and settling speeds
template <int NS>
__device__ __forceinline__ void vp_momentum(const DevParams& P, double r_h, double r_u,
                                            double r_v, const VpPre<NS>& pre, double r_rho_c,
                                            const double* r_alphas, double& qc2, double& qc3) {
  const double mod_vel = sqrt(r_u * r_u + r_v * r_v + 0.0 * 0.0);
  const double shear_vel = P.sqrt_ff * mod_vel;
  double rouse[NS > 0 ? NS : 1];
  vp_rouse<NS>(P, shear_vel, pre.sv, rouse);
  const double h0 = pre.h0;
  const double b = 30.0 / P.k_s;
  const double eps_s = P.Sc * shear_vel * P.vonK * ((h0 / 6.0) + (P.k_s / 60.0));
  const double a_coeff = -(P.vonK * shear_vel / eps_s);
  double a[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) a[i] = a_coeff * rouse[i];
  double slog, sll, se[NS > 0 ? NS : 1], sq[NS > 0 ? NS : 1], sl[NS > 0 ? NS : 1];
  vp_sums<NS, false>(P, h0, b, a, slog, se, sq, sl, sll);
  const double lh0 = det_log(b * h0 + 1.0);
  const double ncu = 1.0 / ((slog + lh0 * (r_h - h0)) / r_h);
  const double u_rel0 = ncu * lh0;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double eh0 = det_exp(a[i] * h0);
    const double nca = 1.0 / ((se[i] + eh0 * (r_h - h0)) / r_h);
    const double d1 = (mod_vel * ncu) * (r_alphas[i] * nca) * sq[i];
    const double d2 = (r_h - h0) * (mod_vel * u_rel0) * ((nca * eh0) * r_alphas[i]);
    const double rua = P.rho_s[i] * (d1 + d2) / r_h;
    s += (P.rho_s[i] - r_rho_c) / P.rho_s[i] * rua;
  }
  const double uRho_avg = mod_vel * r_rho_c + s;
  qc2 = r_h * uRho_avg * r_u / mod_vel;
  qc3 = r_h * uRho_avg * r_v / mod_vel;
}

// eval_flux_coeffs (src/constitutive_2d.f90:2532-2823); sc[0..NV-1]; pre:
// the state's h0 and settling speeds
template <int NS>
__device__ __forceinline__ void vp_flux_coeffs(const DevParams& P, const double* qp, double Bx,
                                               double By, const VpPre<NS>& pre, double r_rho_c,
                                               double r_rho_m, double* sc) {
  constexpr int NV = 4 + NS;
  const double r_h = qp[0], r_hu = qp[1], r_hv = qp[2], r_u = qp[NV], r_v = qp[NV + 1];
  double al[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) al[i] = P.alpha ? qp[4 + i] : qp[4 + i] / qp[0];
  const double r_hw = P.slope ? r_hu * Bx + r_hv * By : 0.0;
  const double r_w = P.slope ? r_u * Bx + r_v * By : 0.0;
  const double mod_hvel = sqrt(r_hu * r_hu + r_hv * r_hv + r_hw * r_hw);
  const double mod_vel = sqrt(r_u * r_u + r_v * r_v + r_w * r_w);
  const double shear_vel = P.sqrt_ff * mod_vel;
  double rouse[NS > 0 ? NS : 1];
  vp_rouse<NS>(P, shear_vel, pre.sv, rouse);
  const double h0 = pre.h0;
  const double b = 30.0 / P.k_s;
  const double eps_s = P.Sc * shear_vel * P.vonK * ((h0 / 6.0) + (P.k_s / 60.0));
  const double a_coeff = -(P.vonK * shear_vel / eps_s);
  double a[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) a[i] = a_coeff * rouse[i];
  double slog, sll, se[NS > 0 ? NS : 1], sq[NS > 0 ? NS : 1], sl[NS > 0 ? NS : 1];
  vp_sums<NS, true>(P, h0, b, a, slog, se, sq, sl, sll);
  const double lh0 = det_log(b * h0 + 1.0);
  const double lt_h0 = lh0 * lh0;
  const double ncu = 1.0 / ((slog + lh0 * (r_h - h0)) / r_h);
  const double u_rel0 = ncu * lh0;
  double irhv = 0.0, rua[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double eh0 = det_exp(a[i] * h0);
    const double nca = 1.0 / ((se[i] + eh0 * (r_h - h0)) / r_h);
    const double d1 = (mod_vel * ncu) * (al[i] * nca) * sq[i];
    const double d2 = (r_h - h0) * (mod_vel * u_rel0) * (al[i] * (nca * eh0));
    rua[i] = P.rho_s[i] * (d1 + d2) / r_h;
    const double d3 = (r_h - h0) * eh0 * lt_h0;
    irhv = irhv + (P.rho_s[i] - r_rho_c) * (nca * al[i]) * (mod_hvel * mod_vel * (ncu * ncu)) *
                      (sl[i] + d3);
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) s += (P.rho_s[i] - r_rho_c) / P.rho_s[i] * rua[i];
  const double uRho_avg = mod_vel * r_rho_c + s;
  const double ihv = mod_hvel * mod_vel * (ncu * ncu) * (sll + (r_h - h0) * lt_h0);
  const double rvv = (irhv + r_rho_c * ihv) / (r_h * r_h);
  sc[0] = uRho_avg / (mod_vel * r_rho_m);
  sc[1] = rvv / (uRho_avg * mod_vel);
  sc[2] = sc[1];
  sc[3] = sc[0];
#pragma unroll
  for (int i = 0; i < NS; i++) sc[4 + i] = rua[i] / (P.rho_s[i] * al[i] * mod_vel);
}

// eval_dep_coeffs (src/constitutive_2d.f90:2844-2986): bottom over mean
// concentration of each class
template <int NS>
__device__ __forceinline__ void vp_dep_coeffs(const DevParams& P, double r_h, double mod_vel,
                                              double r_rho_c, double* dep) {
  const double shear_vel = P.sqrt_ff * mod_vel;
  const VpPre<NS> pre = vp_pre<NS>(P, r_h, r_rho_c);
  double rouse[NS > 0 ? NS : 1];
  vp_rouse<NS>(P, shear_vel, pre.sv, rouse);
  const double h0 = pre.h0;
  const double eps_s = P.Sc * shear_vel * P.vonK * ((h0 / 6.0) + (P.k_s / 60.0));
  const double a_coeff = -(P.vonK * shear_vel / eps_s);
  double a[NS > 0 ? NS : 1], se[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    a[i] = a_coeff * rouse[i];
    se[i] = 0.0;
  }
  for (int m = 0; m < P.nq; m++) {
    const double z = 0.5 * h0 * (P.zq[m] + 1.0);
    const double w = 0.5 * h0 * P.wq[m];
#pragma unroll
    for (int i = 0; i < NS; i++) se[i] += w * det_exp(a[i] * z);
  }
#pragma unroll
  for (int i = 0; i < NS; i++) dep[i] = 1.0 / ((se[i] + det_exp(a[i] * h0) * (r_h - h0)) / r_h);
}

}  // namespace imexsw
