// imexsw_newton.h -- the per-cell implicit Runge-Kutta stage (one lane per
// cell, everything in registers): semi-implicit Coulomb predictor, Newton
// with complex-step Jacobian and backtracking line search for the Voellmy
// drag, velocity realignment, and the assembly checks of imex_RK_solver.
#pragma once

#include "imexsw_physics.h"

namespace imexsw {

// eval_f (src/solver_2d.f90:2243-2282) from the implicit terms' real parts
// nh1, nh2 at qj
template <int NS>
__device__ __forceinline__ void finish_f(const double* qj, const double* qold,
                                         const double* Rj, double a_diag, double dt,
                                         double nh1, double nh2, double* f, double& scal) {
  constexpr int NV = 4 + NS;
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double nh = (i == 1) ? nh1 : (i == 2) ? nh2 : 0.0;
    double R = Rj[i] - a_diag * nh;
    f[i] = 1.0 * (qj[i] - qold[i] + dt * R);
  }
#pragma unroll
  for (int i = 0; i < NV; i++) s += f[i] * f[i];
  scal = 0.5 * s;
}
template <int NS, int RM = 1>
__device__ __forceinline__ void eval_f(const DevParams& P, const double* qj,
                                       const double* qold, const double* Rj,
                                       double a_diag, double dt, double Bx,
                                       double By, double* f, double& scal) {
  constexpr int NV = 4 + NS;
  cpx cq[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) cq[i] = mk(qj[i], 0.0);
  cpx n1, n2;
  implicit_c<NS, RM>(P, cq, Bx, By, n1, n2);
  finish_f<NS>(qj, qold, Rj, a_diag, dt, n1.re, n2.re, f, scal);
}

// gfortran NORM2 (scaled sum of squares)
template <int NV>
__device__ __forceinline__ double norm2_(const double* x) {
  double scale = 0.0, res = 1.0;
#pragma unroll
  for (int i = 0; i < NV; i++) {
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

struct NewtonOut {
  int iters;
  int conv;
  int ls_evals;  // residual evaluations inside the line search
};

// solve_rk_step with eval_jacobian and lnsrch (src/solver_2d.f90:1660-2365);
// RM = 6: the implicit terms of rheology model 6.
// The implicit unknowns are the momenta (implicit_flag = F,T,T,F,...,
// src/constitutive_2d.f90:478-480): the Jacobian has the 2x2 momentum block
// plus the diagonal scaling of the other unknowns, which is what the
// reference's packed partitioned solve reduces to.
template <int NS, int RM = 1>
__device__ __forceinline__ NewtonOut solve_rk_step(const DevParams& P, double* qj,
                                                  const double* qold,
                                                  const double* Rj, double a_diag,
                                                  double dt, double Bx, double By,
                                                  const double* rt0 = nullptr,
                                                  double scal0 = 0.0) {
  constexpr int NV = 4 + NS;
  const double TOLF = 1.0e-10;
  const double TOLX = kEps;
  double org[NV], rel[NV], rt[NV], old[NV], desc[NV];
  double scal, scal_old;
#pragma unroll
  for (int i = 0; i < NV; i++) org[i] = fmax(fabs(qj[i]), 1.0e-3);
#pragma unroll
  for (int i = 0; i < NV; i++) rel[i] = qj[i] / org[i];
  NewtonOut out{0, 0, 0};
  int it;
  // rt / scal already hold the residual at qj when the previous iteration
  // ended on an evaluated point (an accepted line-search point or the full
  // first step); the reference evaluates it again at the top of the
  // iteration, with the same operands, so that evaluation is skipped
  bool fresh = false;
  if (rt0) {
    // residual at the starting point already evaluated by the caller
#pragma unroll
    for (int i = 0; i < NV; i++) rt[i] = rt0[i];
    scal = scal0;
    fresh = true;
  }
  for (it = 1; it <= 100; it++) {
    if (!fresh) eval_f<NS, RM>(P, qj, qold, Rj, a_diag, dt, Bx, By, rt, scal);
    bool restored = false;
    double mf = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mf = fmax(mf, fabs(rt[i]));
    if (mf < TOLF) {
      out.conv = 1;
      break;
    }
    // --- Jacobian: complex step on the two momentum columns
    const double dta = dt * a_diag;
    double J1[3], J2[3];  // d nh_{1,2} / d rel_{1}, d nh_{1,2} / d rel_{2}
    {
      cpx cq[NV];
#pragma unroll
      for (int m = 0; m < NV; m++) cq[m] = mk(rel[m], 0.0);
      cq[1] = mk(rel[1], P.h_cs);
#pragma unroll
      for (int m = 0; m < NV; m++) cq[m] = cmulr(cq[m], org[m]);
      cpx n1, n2;
      implicit_c<NS, RM>(P, cq, Bx, By, n1, n2);
      J1[1] = 1.0 * n1.im * P.one_by_h;
      J1[2] = 1.0 * n2.im * P.one_by_h;
    }
    {
      cpx cq[NV];
#pragma unroll
      for (int m = 0; m < NV; m++) cq[m] = mk(rel[m], 0.0);
      cq[2] = mk(rel[2], P.h_cs);
#pragma unroll
      for (int m = 0; m < NV; m++) cq[m] = cmulr(cq[m], org[m]);
      cpx n1, n2;
      implicit_c<NS, RM>(P, cq, Bx, By, n1, n2);
      J2[1] = 1.0 * n1.im * P.one_by_h;
      J2[2] = 1.0 * n2.im * P.one_by_h;
    }
    // left_matrix entries (column 1 and 2 of the momentum block); rows of
    // the other equations of those columns are 0 - dta*(1.0*0*one_by_h) = 0
    const double zJ = 1.0 * 0.0 * P.one_by_h;
    const double L11 = 1.0 * org[1] - dta * J1[1];
    const double L21 = 0.0 - dta * J1[2];
    const double L12 = 0.0 - dta * J2[1];
    const double L22 = 1.0 * org[2] - dta * J2[2];
    const double Lz = 0.0 - dta * zJ;  // off-block rows of the implicit columns
    // partitioned solve (:1846-1925)
    double d1[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) d1[i] = (i == 1 || i == 2) ? 0.0 : rt[i] / (1.0 * org[i]);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++)
      if (i != 1 && i != 2) s += d1[i] * 0.0;
    double b1 = rt[1] - s, b2 = rt[2] - s;
    {
      // the same partial sum for both implicit rows (MATMUL over zeros)
      double inv_det = 1.0 / (L11 * L22 - L21 * L12);
      double x1 = (b1 * L22 - b2 * L12) * inv_det;
      double x2 = (L11 * b2 - L21 * b1) * inv_det;
#pragma unroll
      for (int i = 0; i < NV; i++) {
        if (i == 1) desc[i] = -x1 + 0.0;
        else if (i == 2) desc[i] = -x2 + 0.0;
        else desc[i] = 0.0 + -d1[i];
      }
    }
#pragma unroll
    for (int i = 0; i < NV; i++) old[i] = rel[i];
    scal_old = scal;
    if (it > 1) {
      double dd = 0.0;
#pragma unroll
      for (int i = 0; i < NV; i++) dd += rel[i] * rel[i];
      double stpmax = 100.0 * fmax(sqrt(dd), (double)NV);
      double grad[NV];
#pragma unroll
      for (int i = 0; i < NV; i++) {
        double g = 0.0;
#pragma unroll
        for (int r = 0; r < NV; r++) {
          double Lri;
          if (i == 1) Lri = (r == 1) ? L11 : (r == 2) ? L21 : Lz;
          else if (i == 2) Lri = (r == 1) ? L12 : (r == 2) ? L22 : Lz;
          else Lri = (r == i) ? 1.0 * org[i] : 0.0;
          g += rt[r] * Lri;
        }
        grad[i] = g;
      }
      // ---- lnsrch (:2019-2224)
      double dabs = norm2_<NV>(desc);
      if (dabs > stpmax) {
#pragma unroll
        for (int i = 0; i < NV; i++) desc[i] = desc[i] * stpmax / dabs;
      }
      double slope = 0.0;
#pragma unroll
      for (int i = 0; i < NV; i++) slope += grad[i] * desc[i];
      double mx = -INFINITY;
#pragma unroll
      for (int i = 0; i < NV; i++) {
        double v = fabs(desc[i]) / fmax(fabs(old[i]), 1.0);
        if (v > mx) mx = v;
      }
      double alamin = TOLX / mx;
      if (alamin == 0.0) {
#pragma unroll
        for (int i = 0; i < NV; i++) rel[i] = old[i];
      } else {
        double alam = 1.0, alam2 = 0.0, f2 = 0.0, tmplam = 0.0;
        // alam at least halves every pass after the first, so it falls
        // below alamin (>= ~1e-16 relative) within a few dozen passes; the
        // bound only guards against NaN states
        for (int ls = 0; ls < 256; ls++) {
          double q[NV];
#pragma unroll
          for (int i = 0; i < NV; i++) rel[i] = old[i] + alam * desc[i];
#pragma unroll
          for (int i = 0; i < NV; i++) q[i] = rel[i] * org[i];
          eval_f<NS, RM>(P, q, qold, Rj, a_diag, dt, Bx, By, rt, scal);
          out.ls_evals++;
          if (scal <= 0.9 * scal_old) break;
          if (alam < alamin) {
#pragma unroll
            for (int i = 0; i < NV; i++) rel[i] = old[i];
            scal = scal_old;
            restored = true;  // rt belongs to the rejected trial point
            break;
          }
          if (alam == 1.0) {
            tmplam = -(slope / (2.0 * (scal - scal_old - slope)));
          } else {
            double rhs1 = scal - scal_old - alam * slope;
            double rhs2 = f2 - scal_old - alam2 * slope;
            double a = (rhs1 / (alam * alam) - rhs2 / (alam2 * alam2)) / (alam - alam2);
            double b = (-(alam2 * rhs1 / (alam * alam)) + alam * rhs2 / (alam2 * alam2)) /
                       (alam - alam2);
            if (a == 0.0) {
              tmplam = -(slope / (2.0 * b));
            } else {
              double disc = b * b - 3.0 * a * slope;
              if (disc < 0.0) tmplam = 0.5 * alam;
              else if (b <= 0.0) tmplam = (-b + sqrt(disc)) / (3.0 * a);
              else tmplam = -(slope / (b + sqrt(disc)));
            }
            if (tmplam > 0.5 * alam) tmplam = 0.5 * alam;
          }
          alam2 = alam;
          f2 = scal;
          alam = fmax(tmplam, 0.5 * alam);
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < NV; i++) rel[i] = old[i] + desc[i];
#pragma unroll
      for (int i = 0; i < NV; i++) qj[i] = rel[i] * org[i];
      eval_f<NS, RM>(P, qj, qold, Rj, a_diag, dt, Bx, By, rt, scal);
    }
#pragma unroll
    for (int i = 0; i < NV; i++) qj[i] = rel[i] * org[i];
    fresh = !restored;
    mf = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mf = fmax(mf, fabs(rt[i]));
    if (mf < TOLF) {
      out.conv = 1;
      break;
    }
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mx = fmax(mx, fabs(rel[i] - old[i]) / fmax(fabs(rel[i]), 1.0));
    if (mx < TOLX) {
      out.conv = 1;
      break;
    }
  }
  out.iters = it > 100 ? 100 : it;
  return out;
}

// solve_rk_step on four lanes (4i .. 4i+3, the same cell and operands on
// each).  One implicit-term evaluation per lane and round: lanes 0 and 1 of
// the quad the two complex-step Jacobian columns, lane 2 the residual.  A
// full Newton step (the first iteration) and the line search's first trial
// (alam = 1) evaluate the residual at the trial point and, at the same time,
// the Jacobian there, which the next iteration uses when the trial is
// accepted: the same operands the plain solver would give its Jacobian
// evaluation, so the same bits, one evaluation earlier.  Every value the
// control flow depends on is taken from one lane of the quad, so the four
// lanes stay together.
template <int NS>
__device__ __forceinline__ double quad_from(double v, int role) {
  const int src = (threadIdx.x & 60) | role;
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __shfl(lo, src, 64);
  hi = __shfl(hi, src, 64);
  return __hiloint2double(hi, lo);
}

// One quad round at the point q = rel * org: J1/J2 (columns 1, 2) and the
// residual rt, scal (when `want_f`; else lanes 2, 3 duplicate column 1)
template <int NS>
__device__ __forceinline__ void quad_round(const DevParams& P, const double* q, const double* org,
                                           const double* qold, const double* Rj, double a_diag,
                                           double dt, double Bx, double By, bool want_f,
                                           double* J1, double* J2, double* rt, double& scal) {
  constexpr int NV = 4 + NS;
  const int role = threadIdx.x & 3;
  const int pcol = (role == 1) ? 2 : (role == 0 || !want_f) ? 1 : -1;  // -1: none
  cpx cq[NV];
#pragma unroll
  for (int m = 0; m < NV; m++) cq[m] = mk(q[m], (m == pcol) ? P.h_cs * org[m] : 0.0);
  cpx n1, n2;
  implicit_c<NS>(P, cq, Bx, By, n1, n2);
  const double a = 1.0 * n1.im * P.one_by_h, b = 1.0 * n2.im * P.one_by_h;
  J1[1] = quad_from<NS>(a, 0);
  J1[2] = quad_from<NS>(b, 0);
  J2[1] = quad_from<NS>(a, 1);
  J2[2] = quad_from<NS>(b, 1);
  if (want_f) {
    const double nh1 = quad_from<NS>(n1.re, 2), nh2 = quad_from<NS>(n2.re, 2);
    finish_f<NS>(q, qold, Rj, a_diag, dt, nh1, nh2, rt, scal);
  }
}

template <int NS>
__device__ __forceinline__ NewtonOut solve_rk_step_quad(const DevParams& P, double* qj,
                                                       const double* qold, const double* Rj,
                                                       double a_diag, double dt, double Bx,
                                                       double By, const double* rt0 = nullptr,
                                                       double scal0 = 0.0) {
  constexpr int NV = 4 + NS;
  const double TOLF = 1.0e-10;
  const double TOLX = kEps;
  double org[NV], rel[NV], rt[NV], old[NV], desc[NV];
  double scal, scal_old;
#pragma unroll
  for (int i = 0; i < NV; i++) org[i] = fmax(fabs(qj[i]), 1.0e-3);
#pragma unroll
  for (int i = 0; i < NV; i++) rel[i] = qj[i] / org[i];
  NewtonOut out{0, 0, 0};
  int it;
  bool fresh = false;
  if (rt0) {
    // residual at the starting point already evaluated by the caller
#pragma unroll
    for (int i = 0; i < NV; i++) rt[i] = rt0[i];
    scal = scal0;
    fresh = true;
  }
  bool jnext = false;       // J1n/J2n hold the Jacobian at the current point
  double J1n[3], J2n[3];
  for (it = 1; it <= 100; it++) {
    if (!fresh) eval_f<NS>(P, qj, qold, Rj, a_diag, dt, Bx, By, rt, scal);
    bool restored = false;
    double mf = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mf = fmax(mf, fabs(rt[i]));
    if (mf < TOLF) {
      out.conv = 1;
      break;
    }
    const double dta = dt * a_diag;
    double J1[3], J2[3];
    if (jnext) {
      J1[1] = J1n[1];
      J1[2] = J1n[2];
      J2[1] = J2n[1];
      J2[2] = J2n[2];
    } else {
      double q[NV], rdum[NV], sdum;
#pragma unroll
      for (int m = 0; m < NV; m++) q[m] = rel[m] * org[m];
      quad_round<NS>(P, q, org, qold, Rj, a_diag, dt, Bx, By, false, J1, J2, rdum, sdum);
    }
    jnext = false;
    const double zJ = 1.0 * 0.0 * P.one_by_h;
    const double L11 = 1.0 * org[1] - dta * J1[1];
    const double L21 = 0.0 - dta * J1[2];
    const double L12 = 0.0 - dta * J2[1];
    const double L22 = 1.0 * org[2] - dta * J2[2];
    const double Lz = 0.0 - dta * zJ;
    double d1[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) d1[i] = (i == 1 || i == 2) ? 0.0 : rt[i] / (1.0 * org[i]);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++)
      if (i != 1 && i != 2) s += d1[i] * 0.0;
    double b1 = rt[1] - s, b2 = rt[2] - s;
    {
      double inv_det = 1.0 / (L11 * L22 - L21 * L12);
      double x1 = (b1 * L22 - b2 * L12) * inv_det;
      double x2 = (L11 * b2 - L21 * b1) * inv_det;
#pragma unroll
      for (int i = 0; i < NV; i++) {
        if (i == 1) desc[i] = -x1 + 0.0;
        else if (i == 2) desc[i] = -x2 + 0.0;
        else desc[i] = 0.0 + -d1[i];
      }
    }
#pragma unroll
    for (int i = 0; i < NV; i++) old[i] = rel[i];
    scal_old = scal;
    if (it > 1) {
      double dd = 0.0;
#pragma unroll
      for (int i = 0; i < NV; i++) dd += rel[i] * rel[i];
      double stpmax = 100.0 * fmax(sqrt(dd), (double)NV);
      double grad[NV];
#pragma unroll
      for (int i = 0; i < NV; i++) {
        double g = 0.0;
#pragma unroll
        for (int r = 0; r < NV; r++) {
          double Lri;
          if (i == 1) Lri = (r == 1) ? L11 : (r == 2) ? L21 : Lz;
          else if (i == 2) Lri = (r == 1) ? L12 : (r == 2) ? L22 : Lz;
          else Lri = (r == i) ? 1.0 * org[i] : 0.0;
          g += rt[r] * Lri;
        }
        grad[i] = g;
      }
      double dabs = norm2_<NV>(desc);
      if (dabs > stpmax) {
#pragma unroll
        for (int i = 0; i < NV; i++) desc[i] = desc[i] * stpmax / dabs;
      }
      double slope = 0.0;
#pragma unroll
      for (int i = 0; i < NV; i++) slope += grad[i] * desc[i];
      double mx = -INFINITY;
#pragma unroll
      for (int i = 0; i < NV; i++) {
        double v = fabs(desc[i]) / fmax(fabs(old[i]), 1.0);
        if (v > mx) mx = v;
      }
      double alamin = TOLX / mx;
      if (alamin == 0.0) {
#pragma unroll
        for (int i = 0; i < NV; i++) rel[i] = old[i];
      } else {
        // lnsrch's trials: 1, then a2 = max(quadratic model, 0.5), then
        // halvings (the cubic model is capped at half the last step, so
        // max(tmplam, 0.5 * alam) is always 0.5 * alam).  Lane r of a batch
        // evaluates one trial; the trials are then taken in order with the
        // serial rules, so the accepted point and the count are the serial
        // ones.  First batch: the Jacobian at trial 1 (lanes 0, 1), trial 1
        // (lane 2), and the trial 0.5 that follows whenever a2 = 0.5 (lane 3).
        const int role = threadIdx.x & 3;
        double n1a, n2a;  // this lane's implicit terms (real parts)
        {
          const double al = (role == 3) ? 0.5 : 1.0;
          const int pcol = (role == 0) ? 1 : (role == 1) ? 2 : -1;
          cpx cq[NV];
#pragma unroll
          for (int m = 0; m < NV; m++) {
            const double rm = old[m] + al * desc[m];
            cq[m] = mk(rm * org[m], (m == pcol) ? P.h_cs * org[m] : 0.0);
          }
          cpx n1, n2;
          implicit_c<NS>(P, cq, Bx, By, n1, n2);
          const double a = 1.0 * n1.im * P.one_by_h, b = 1.0 * n2.im * P.one_by_h;
          J1n[1] = quad_from<NS>(a, 0);
          J1n[2] = quad_from<NS>(b, 0);
          J2n[1] = quad_from<NS>(a, 1);
          J2n[2] = quad_from<NS>(b, 1);
          n1a = n1.re;
          n2a = n2.re;
        }
        // trial `alam` from the implicit terms of lane `src`; returns true
        // when the search ends with it (accepted, or below alamin)
        auto take = [&](double alam, int src, bool& accepted) -> bool {
          const double h1 = quad_from<NS>(n1a, src), h2 = quad_from<NS>(n2a, src);
          double q[NV];
#pragma unroll
          for (int i = 0; i < NV; i++) rel[i] = old[i] + alam * desc[i];
#pragma unroll
          for (int i = 0; i < NV; i++) q[i] = rel[i] * org[i];
          finish_f<NS>(q, qold, Rj, a_diag, dt, h1, h2, rt, scal);
          out.ls_evals++;
          accepted = scal <= 0.9 * scal_old;
          if (accepted) return true;
          if (alam < alamin) {
#pragma unroll
            for (int i = 0; i < NV; i++) rel[i] = old[i];
            scal = scal_old;
            restored = true;
            return true;
          }
          return false;
        };
        bool acc = false;
        bool end = take(1.0, 2, acc);
        jnext = acc;
        if (!end) {
          const double tq = -(slope / (2.0 * (scal - scal_old - slope)));
          double alam = fmax(tq, 0.5 * 1.0);
          int trials = 1;
          if (alam == 0.5) {
            end = take(0.5, 3, acc);
            alam = 0.25;
            trials++;
          }
          while (!end && trials < 256) {
            // four halvings at once: lane r evaluates alam * 0.5^r
            double al = alam;
#pragma unroll
            for (int r = 0; r < 3; r++)
              if (role > r) al = al * 0.5;
            {
              cpx cq[NV];
#pragma unroll
              for (int m = 0; m < NV; m++) cq[m] = mk((old[m] + al * desc[m]) * org[m], 0.0);
              cpx n1, n2;
              implicit_c<NS>(P, cq, Bx, By, n1, n2);
              n1a = n1.re;
              n2a = n2.re;
            }
            for (int r = 0; r < 4 && !end && trials < 256; r++) {
              end = take(alam, r, acc);
              alam = alam * 0.5;
              trials++;
            }
          }
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < NV; i++) rel[i] = old[i] + desc[i];
#pragma unroll
      for (int i = 0; i < NV; i++) qj[i] = rel[i] * org[i];
      // the residual and, speculatively, the Jacobian at the full step
      quad_round<NS>(P, qj, org, qold, Rj, a_diag, dt, Bx, By, true, J1n, J2n, rt, scal);
      jnext = true;
    }
#pragma unroll
    for (int i = 0; i < NV; i++) qj[i] = rel[i] * org[i];
    fresh = !restored;
    if (restored) jnext = false;
    mf = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mf = fmax(mf, fabs(rt[i]));
    if (mf < TOLF) {
      out.conv = 1;
      break;
    }
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mx = fmax(mx, fabs(rel[i] - old[i]) / fmax(fabs(rel[i]), 1.0));
    if (mx < TOLX) {
      out.conv = 1;
      break;
    }
  }
  out.iters = it > 100 ? 100 : it;
  return out;
}

// Raw warm-start features of one cell entering Newton: log h, the
// predictor's velocity (u, v), T, three solid volume fractions, log dt and
// the bed slope magnitude.
template <int NS>
__device__ __forceinline__ void mlp_features_raw(const DevParams& P, const double* q_fv,
                                                 const double* qp, const double* q_si,
                                                 double dt, double Bx, double By, float* f) {
  const double h = qp[0];
  f[0] = (float)log(h);
  f[1] = (float)(q_si[1] / q_fv[0]);
  f[2] = (float)(q_si[2] / q_fv[0]);
  f[3] = (float)qp[3];
#pragma unroll
  for (int s = 0; s < 3; s++)
    f[4 + s] = (s < NS) ? (float)(P.alpha ? qp[4 + s] : qp[4 + s] / h) : 0.f;
  f[7] = (float)log(dt);
  f[8] = (float)sqrt(Bx * Bx + By * By);
}

// Mixture density and reduced gravity of a state whose momenta are (mx, my),
// the real part of the chain in implicit_c (c_phys_var and mixt_var,
// src/constitutive_2d.f90:1145-1313, 1333-1483) for q1 > eps_sing.
template <int NS>
__device__ __forceinline__ void rho_rg_of(const DevParams& P, const double* q, double mx,
                                          double my, double& rho_m, double& rg) {
  const double inv_q1 = 1.0 / q[0];
  double xs_tot = 0.0, dsh = 0.0, dinv = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double xs = q[4 + i] * inv_q1;
    xs_tot += xs;
    dsh += xs * P.sp_heat_s[i];
    dinv += xs * P.inv_rho_s[i];
  }
  const double xc = 1.0 - xs_tot;
  const double shm = dsh + xc * P.sp_heat_a;
  double T = P.energy ? (q[3] - 0.5 * (mx * mx + my * my) * inv_q1) / (q[0] * shm)
                      : q[3] / (q[0] * shm);
  if (T <= 0.0) T = P.T_ambient;
  rho_m = 1.0 / (dinv + xc * (P.sp_gas_const_a * T * P.inv_pres));
  rg = (rho_m - P.rho_a_amb) / rho_m * P.grav;
}

// A starting point for the implicit Voellmy solve (no reference
// counterpart: the reference starts Newton at q_si, src/solver_2d.f90:1196).
// The momenta solve m = m_si + dt a NH(m) with NH = -(rho_m g'/xi) |U| u,
// u = m/q1 and |U| = |u| sqrt(1 + (u.grad B)^2/|u|^2) with slope correction.
// m is parallel to m_si, so with the mixture density frozen the speed s =
// |m| solves s + C c s^2 = |m_si| with C = dt a rho_m g'/(xi q1^2), c the
// slope factor: s = 2|m_si| / (1 + sqrt(1 + 4 C c |m_si|)).  The density
// depends on s only through the kinetic energy in the temperature, weakly,
// so three refreshes of rho_m, g' at the current s make the root exact to
// rounding.  Returns false where it does not apply (no Voellmy term, thin
// or lighter-than-air cells, no momentum); Newton then starts at q_si.
template <int NS>
__device__ __forceinline__ bool voellmy_closed_form(const DevParams& P, const double* q_si,
                                                    double a_diag, double dt, double Bx,
                                                    double By, double* m) {
  if (!P.rheo || P.rheo_model != 1 || !(q_si[0] > P.eps_sing)) return false;
  const double mx = q_si[1], my = q_si[2];
  const double ms = sqrt(mx * mx + my * my);
  if (!(ms > 0.0)) return false;
  const double ex = mx / ms, ey = my / ms;
  const double wn = P.slope ? ex * Bx + ey * By : 0.0;
  const double c = sqrt(1.0 + wn * wn);
  const double k0 = dt * a_diag / (P.xi * q_si[0] * q_si[0]) * c;
  double s = ms;
#pragma unroll
  for (int it = 0; it < 4; it++) {
    double rho_m, rg;
    rho_rg_of<NS>(P, q_si, s * ex, s * ey, rho_m, rg);
    const double C = k0 * rho_m * rg;
    if (!(C > 0.0)) return false;
    s = 2.0 * ms / (1.0 + sqrt(1.0 + 4.0 * C * ms));
  }
  m[0] = s * ex;
  m[1] = s * ey;
  return isfinite(m[0]) && isfinite(m[1]);
}

// The stage-2 predictor (src/solver_2d.f90:1132-1196): q_fv = q0 - dt*D, its
// physical variables, and the semi-implicit Coulomb step q_si realigned with
// q_fv.  Returns false when q_fv has no mass (no Newton solve).
template <int NS>
__device__ __forceinline__ bool predictor(const DevParams& P, const double* q0,
                                          const double* D, double dt, double Bx,
                                          double By, double Bxx, double Bxy, double Byy,
                                          double gc, double* q_fv, double* qp,
                                          double* q_si) {
  constexpr int NV = 4 + NS;
#pragma unroll
  for (int i = 0; i < NV; i++) q_fv[i] = q0[i] - dt * (D[i] - 0.0);
  phys<NS>(P, q_fv, qp);
  if (!(q_fv[0] > 0.0)) return false;
  double s1, s2;
  semi_impl<NS>(P, Bx, By, Bxx, Bxy, Byy, gc, qp, s1, s2);
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double si = (i == 1) ? s1 : (i == 2) ? s2 : 0.0;
    q_si[i] = q_fv[i] + dt * 1.0 * si;
  }
  const double f2 = q_fv[1], f3 = q_fv[2];
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
  return true;
}

// Stage 2 of the N_RK = 2 IMEX scheme for one cell (src/solver_2d.f90:
// 1105-1280), from q0 = q^n and D = divFlux - expl_terms of stage 1.  The
// implicit term refresh of :1234-1236 is not evaluated: for the stiffly
// accurate N_RK = 2 tableau it is overwritten at :1271-1277 and never enters
// q.  Returns the Newton iteration count (0: no Newton).
// LANES: lanes per cell, 1 or 4 (the quad solver)
template <int NS, int LANES = 1>
__device__ __forceinline__ int stage2(const DevParams& P, const double* q0,
                                      const double* D, double dt, double Bx,
                                      double By, double Bxx, double Bxy,
                                      double Byy, double gc, double* q,
                                      int& nonconv, int& ls_evals,
                                      const double* warm = nullptr, int* warm_used = nullptr,
                                      float* harvest = nullptr, bool* harvested = nullptr,
                                      bool closed = false) {
  constexpr int NV = 4 + NS;
  double q_fv[NV], qp[NV + 2], q_si[NV];
  nonconv = 0;
  ls_evals = 0;
  if (!predictor<NS>(P, q0, D, dt, Bx, By, Bxx, Bxy, Byy, gc, q_fv, qp, q_si)) {
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = q_fv[i];
    return 0;
  }
  double Rj[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double si = (q_si[i] - q_fv[i]) / (dt * 1.0);
    Rj[i] = (D[i] - 0.0) - 1.0 * si;
    q[i] = q_si[i];
  }
  NewtonOut no;
  double mg[2];
  bool have = false;
  if (closed) {
    have = voellmy_closed_form<NS>(P, q_si, 1.0, dt, Bx, By, mg);
  } else if (warm && warm[0] == warm[0] && warm[1] == warm[1]) {
    mg[0] = warm[0];
    mg[1] = warm[1];
    have = true;
  }
  if (have) {
    // a starting point for the momenta (learned, or the closed form): Newton
    // starts from it when it already meets the tolerance or has the smaller
    // residual of it and the plain predictor (the plain one stays the
    // fallback), and the equation solved is the same
    double qw[NV], rtw[NV], rtp[NV], sw, sp = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) qw[i] = q_si[i];
    qw[1] = mg[0];
    qw[2] = mg[1];
    eval_f<NS>(P, qw, q0, Rj, 1.0, dt, Bx, By, rtw, sw);
    double mfw = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) mfw = fmax(mfw, fabs(rtw[i]));
    bool use = mfw < 1.0e-10;
    if (!use) {
      eval_f<NS>(P, q, q0, Rj, 1.0, dt, Bx, By, rtp, sp);
      use = sw < sp;
    }
    *warm_used = use ? 1 : 0;
    if (use) {
#pragma unroll
      for (int i = 0; i < NV; i++) q[i] = qw[i];
    }
    no = (LANES == 4)
             ? solve_rk_step_quad<NS>(P, q, q0, Rj, 1.0, dt, Bx, By, use ? rtw : rtp,
                                      use ? sw : sp)
             : solve_rk_step<NS>(P, q, q0, Rj, 1.0, dt, Bx, By, use ? rtw : rtp, use ? sw : sp);
  } else {
    if (warm_used) *warm_used = (closed || warm) ? 0 : -1;
    no = (LANES == 4) ? solve_rk_step_quad<NS>(P, q, q0, Rj, 1.0, dt, Bx, By)
                      : solve_rk_step<NS>(P, q, q0, Rj, 1.0, dt, Bx, By);
  }
  nonconv = !no.conv;
  ls_evals = no.ls_evals;
  if (harvest && no.conv) {
    // (state entering Newton, converged momenta) pair for the warm start;
    // a momentum component below 1e-6 of the predictor's magnitude gets a
    // zero target, and pairs with non-finite values are not recorded
    float f[9];
    mlp_features_raw<NS>(P, q_fv, qp, q_si, dt, Bx, By, f);
    const double mag = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
    bool ok = mag > 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
      harvest[i] = f[i];
      ok = ok && isfinite(f[i]);
    }
    harvest[9] = (fabs(q_si[1]) > 1e-6 * mag) ? (float)(q[1] / q_si[1] - 1.0) : 0.f;
    harvest[10] = (fabs(q_si[2]) > 1e-6 * mag) ? (float)(q[2] / q_si[2] - 1.0) : 0.f;
    ok = ok && isfinite(harvest[9]) && isfinite(harvest[10]);
    *harvested = ok;
  }
  if (P.ny == 1) q[2] = 0.0;
  if (P.nx == 1) q[1] = 0.0;
  if (q_si[1] * q_si[1] + q_si[2] * q_si[2] == 0.0) {
    q[1] = 0.0;
    q[2] = 0.0;
  } else if ((q[1] * q_si[1] <= 0.0) && (q[2] * q_si[2] <= 0.0)) {
    q[1] = 0.0;
    q[2] = 0.0;
  } else {
    double m = sqrt(q[1] * q[1] + q[2] * q[2]);
    double den = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
    q[1] = m * q_si[1] / den;
    q[2] = m * q_si[2] / den;
  }
  return no.iters;
}

// eval_mass_exchange_terms, supported branches (src/constitutive_2d.f90:
// 4424-5003).  Returns false when every term is zero.
// LANES > 1: the cell runs on LANES lanes (the same operands on each); the
// solid classes' settling velocities are split between them (class
// LANES*s + r on lane r, the same instructions on every lane) and
// exchanged, so the cell waits for about 1/LANES of them.
template <int NS, int LANES = 1, bool VP = false>
__device__ __forceinline__ void mass_exchange(const DevParams& P, const double* qp,
                                              double Bx, double By,
                                              const double* erodible, double dt,
                                              double* er, double* dep, double* eq,
                                              double& topo) {
  constexpr int NV = 4 + NS;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    er[i] = 0.0;
    dep[i] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < NV; i++) eq[i] = 0.0;
  topo = 0.0;
  if (qp[0] <= kEps) return;
  const double alpha_max = 0.6, hind_exp = 4.65;
  double r_h = qp[0], r_u = qp[NV], r_v = qp[NV + 1];
  double al[NS > 0 ? NS : 1], at = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) al[i] = P.alpha ? qp[4 + i] : qp[4 + i] / qp[0];
#pragma unroll
  for (int i = 0; i < NS; i++) at += al[i];
  double r_w = P.slope ? r_u * Bx + r_v * By : 0.0;
  double mv2 = r_u * r_u + r_v * r_v + r_w * r_w;
  double mv = sqrt(mv2);
  if (P.erosion_coeff > 0.0) {
    double te = P.erosion_coeff * mv * r_h * (1.0 - at);
    double tse = te * (1.0 - P.erodible_porosity);
#pragma unroll
    for (int i = 0; i < NS; i++) er[i] = P.erodible_fract[i] * tse;
  }
#pragma unroll
  for (int i = 0; i < NS; i++) er[i] = fmax(0.0, fmin(er[i], erodible[i] / dt));
  double tse = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) tse += er[i];
  double te = tse / (1.0 - P.erodible_porosity);
  double cont_erosion = te * P.erodible_porosity;
  if (at <= P.alphastot_min) return;
  double r_T = qp[3];
  Mixt m = mixt<NS, true>(P, qp);
  double inv_kin_visc = 1.0 / P.kin_visc_c;
  // bottom over mean concentration with the profiles (:4676-4683)
  double dcf[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; i++) dcf[i] = 1.0;
  if (VP && P.bottom_conc) vp_dep_coeffs<NS>(P, r_h, mv, m.rho_c, dcf);
  if (P.settling) {
    // hindered settling factor, the same for every class
    const double hind = det_pow(1.0 - fmin(1.0, at / alpha_max), hind_exp);
    double svs[NS > 0 ? NS : 1];
    if (LANES > 1) {
      constexpr int NSL = (NS + LANES - 1) / LANES;
      const int p = threadIdx.x & (LANES - 1);
      double mine[NSL > 0 ? NSL : 1];
#pragma unroll
      for (int sl = 0; sl < NSL; sl++) {
        double a = 0.0, diam = 0.0, rhos = 0.0;
        bool has = false;
#pragma unroll
        for (int r = 0; r < LANES; r++) {
          const int c = sl * LANES + r;
          if (c < NS && p == r) {
            a = al[c];
            diam = P.diam_s[c];
            rhos = P.rho_s[c];
            has = true;
          }
        }
        mine[sl] = (has && a > 0.0) ? settling_velocity(P, diam, rhos, m.rho_c, inv_kin_visc)
                                    : 0.0;
      }
#pragma unroll
      for (int i = 0; i < NS; i++) {
        const int src = (threadIdx.x & (64 - LANES)) | (i % LANES);
        int lo = __double2loint(mine[i / LANES]), hi = __double2hiint(mine[i / LANES]);
        lo = __shfl(lo, src, 64);
        hi = __shfl(hi, src, 64);
        svs[i] = __hiloint2double(hi, lo);
      }
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
      if (al[i] > 0.0) {
        const double sv = (LANES > 1) ? svs[i]
                                      : settling_velocity(P, P.diam_s[i], P.rho_s[i], m.rho_c,
                                                          inv_kin_visc);
        double d = al[i] * sv * dcf[i];
        d = d * hind;
        dep[i] = fmin(d, r_h * al[i] / dt);
      }
    }
  }
  double sdep = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) sdep += dep[i];
  double cl = 0.0 + P.coeff_porosity * sdep;
  double sal = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) sal += al[i];
  cl = fmin(cl, r_h * fmax(0.0, P.max_packing - sal) / dt);
  cl = fmin(cl, (r_h * (1.0 - sal) - P.coeff_porosity * (r_h * sal - dt * sdep)) / dt);
  double T_erodible = P.erod_dep ? r_T : P.T_erodible;
  double air = 0.0;
  if (P.entrain && (r_h > 0.0) && (m.ri > 0.0)) {
    double ec = 0.075 / sqrt(1.0 + 718.0 * det_pow(m.ri, 2.4));
    air = ec * sqrt(mv2);
  }
  double dep_tot = 0.0, ers_tot = 0.0, rdep = 0.0, rers = 0.0, sdc = 0.0, sec = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) dep_tot += dep[i];
#pragma unroll
  for (int i = 0; i < NS; i++) ers_tot += er[i];
#pragma unroll
  for (int i = 0; i < NS; i++) rdep += P.rho_s[i] * dep[i];
#pragma unroll
  for (int i = 0; i < NS; i++) rers += P.rho_s[i] * er[i];
  eq[0] = P.rho_a_amb * air + rers - rdep + P.rho_c_sub * cont_erosion - m.rho_c * cl;
  eq[1] = -(r_u * (rdep + m.rho_c * cl));
  eq[2] = -(r_v * (rdep + m.rho_c * cl));
#pragma unroll
  for (int i = 0; i < NS; i++) sdc += P.rho_s[i] * P.sp_heat_s[i] * dep[i];
#pragma unroll
  for (int i = 0; i < NS; i++) sec += P.rho_s[i] * P.sp_heat_s[i] * er[i];
  if (P.energy)
    eq[3] = -(r_T * (sdc + m.rho_c * m.shc * cl)) - 0.5 * mv2 * (rdep + m.rho_c * cl) +
            T_erodible * (sec + P.rho_c_sub * m.shc * cont_erosion) +
            P.T_ambient * P.sp_heat_a * P.rho_a_amb * air;
  else
    eq[3] = -(r_T * (sdc + m.rho_c * m.shc * cl)) +
            T_erodible * (sec + P.rho_c_sub * m.shc * cont_erosion) +
            P.T_ambient * P.sp_heat_a * P.rho_a_amb * air;
#pragma unroll
  for (int i = 0; i < NS; i++) eq[4 + i] = P.rho_s[i] * (er[i] - dep[i]);
  topo = (dep_tot - ers_tot) / (1.0 - P.erodible_porosity);
}

}  // namespace imexsw
