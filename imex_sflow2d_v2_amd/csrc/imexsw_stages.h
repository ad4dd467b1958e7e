// imexsw_stages.h -- the IMEX Runge-Kutta step for the tableaux other than
// N_RK = 2: IMEX-SSP(3,3,2) (N_RK = 3), LRR(3,2,2) (N_RK = 4) and the
// one-stage tableau (N_RK = 1) of src/solver_2d.f90:398-486, driven by the
// tableau itself as imex_RK_solver is (src/solver_2d.f90:1021-1632).
//
// One time step is
//   the mask (seeds, N_RK 4-neighbour dilations, source interior removed),
//   the stage states set to q (inactive cells keep q through the step),
//   the CFL pass over q^n (k_stencil, writing divFlux - expl_terms of q^n),
//   per stage i: kg_stage (predictor, semi-implicit Coulomb step, Newton,
//     realignment; the stage state q_rk(i) and NH + SI_NH of the stage),
//     then, when omega_tilde(i) > 0, k_stencil on q_rk(i) for the stage's
//     divFlux - expl_terms,
//   kg_assemble (assemble_sol and its checks), then the erosion/deposition
//   and diagnostics pass of the N_RK = 2 path.
// Every active cell is computed (no quiet-cell skipping: with several
// stages a dry cell's faces can see a neighbour's stage state).
#pragma once

#include "imexsw_kernels.h"

namespace imexsw {

constexpr int kMaxRK = 4;

// a_tilde_ij, a_dirk_ij, omega_tilde, omega of src/solver_2d.f90:398-486
struct RkTableau {
  int nrk;
  int stiff;  // stiffly accurate: q = q_rk(N_RK) (src/solver_2d.f90:1341-1377)
  double a_tilde[kMaxRK][kMaxRK];
  double a_dirk[kMaxRK][kMaxRK];
  double omega_tilde[kMaxRK], omega[kMaxRK];
};

// stage storage of one context: stage r's planes start at r * NV planes
struct RkPtrs {
  double* qrk;  // q_rk: the stage states
  double* Dx;   // divFlux - expl_terms of each stage (zero when omega_tilde = 0)
  double* Im;   // NH + SI_NH of each stage
  uint8_t* ma;  // two byte planes for the mask dilation
  uint8_t* mb;
};

// ---------------------------------------------------------------------------
// check_solve for N_RK != 2 (src/solver_2d.f90:666-790): seeds, N_RK
// dilation passes over the interior cells, then the final pass that writes
// the mask bytes and the work lists
// ---------------------------------------------------------------------------
// seeds: the border cells (never modified by check_solve), and interior cells
// with solve_all, h > 0 or a ring cell
template <int = 0>
__global__ void kg_mask_seed(DevParams P, DevPtrs D, uint8_t* a, int solve_all) {
  if (D.scal[S_DONE] != 0.0) return;  // a stopped run keeps its last mask
  const size_t n = (size_t)P.pitch * P.ny;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(c % P.pitch), k = (int)(c / P.pitch);
    if (j >= P.nx) {
      a[c] = 0;
      continue;
    }
    const bool bnd = j == 0 || j == P.nx - 1 || k == 0 || k == P.ny - 1;
    const int sc = (P.radial && D.sinfo) ? D.sinfo[c] : 0;
    a[c] = (bnd || solve_all || (D.sflag[c] & SF_POS) || (sc & SI_RING)) ? 1 : 0;
  }
}

// one 4-neighbour dilation of the interior cells (src/solver_2d.f90:732-757)
template <int = 0>
__global__ void kg_mask_dilate(DevParams P, const double* scal, const uint8_t* a, uint8_t* b) {
  if (scal[S_DONE] != 0.0) return;
  const size_t n = (size_t)P.pitch * P.ny;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(c % P.pitch), k = (int)(c / P.pitch);
    uint8_t v = a[c];
    if (j >= 1 && j <= P.nx - 2 && k >= 1 && k <= P.ny - 2)
      v = v | a[c - 1] | a[c + 1] | a[c - P.pitch] | a[c + P.pitch];
    b[c] = v;
  }
}

// The source interior leaves the mask; mask bytes (every active cell is
// computed), active and computed counts, and the tiles and tile rows with
// an active cell appended to the work lists.  One wavefront per work tile,
// kMaskGroup tiles per workgroup, the lane layout of k_mask.
template <int = 0>
__global__ void __launch_bounds__(64 * kMaskGroup) kg_mask_final(DevParams P, DevPtrs D,
                                                                 const uint8_t* a) {
  if (D.scal[S_DONE] != 0.0) return;
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int tiles_y = (P.ny + kTileH - 1) / kTileH;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tyy = blockIdx.y * kMaskGroup + w;
  const int tile = tyy * tiles_x + blockIdx.x;
  const int j = blockIdx.x * kStripW - 2 + lane;
  const bool own = tyy < tiles_y && lane >= 2 && lane < 2 + kStripW && j < P.nx;
  unsigned rows_m = 0;
  int count = 0;
  if (own) {
    for (int rr = 0; rr < kTileH; rr++) {
      const int k = tyy * kTileH + rr;
      if (k >= P.ny) break;
      const size_t c = cidx(P, j, k);
      const bool src = P.radial && D.sinfo && (D.sinfo[c] & SI_SRC);
      const bool act = a[c] && !src;
      D.mask[c] = act ? (uint8_t)(M_ACTIVE | M_COMPUTE) : (uint8_t)0;
      if (act) {
        rows_m |= 1u << rr;
        count++;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    count += __shfl_xor(count, off, 64);
    rows_m |= (unsigned)__shfl_xor((int)rows_m, off, 64);
  }
  if (tyy < tiles_y && lane == 0) {
    const int delta = count - D.tile_nact[tile];
    if (delta) atomicAdd(ctr(D, C_SOLVE_CELLS, tile), (unsigned long long)(long long)delta);
    D.tile_nact[tile] = count;
    D.tile_dirty[tile] = 1;
    if (count) atomicAdd(ctr(D, C_COMPUTE_CELLS, tile), (unsigned long long)count);
    D.tile_flag[tile] = rows_m ? 1 : 0;
  }
  append_group<kMaskGroup>(own_sink(D), tile, tyy < tiles_y ? rows_m : 0u);
}

// q into every stage state (an inactive cell's stage states are its state,
// as its interface values are in the reference's fresh evaluation)
template <int = 0>
__global__ void kg_stage_init(DevParams P, const double* scal, const double* q, double* qrk,
                              int nrk) {
  if (scal[S_DONE] != 0.0) return;
  const size_t per = (size_t)P.nv * P.pitch * P.ny;
  const size_t n = per * nrk;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    qrk[e] = q[e % per];
}

// integrate_friction_term (src/constitutive_2d.f90:3252-3306), rheology
// model 8: the speed after dt of du/dt = -f u^2 / h in closed form, the
// direction kept, through qc_to_qp and back through qp_to_qc
template <int NS, bool VP>
__device__ __forceinline__ void integrate_friction(const DevParams& P, double* qj, double dt) {
  constexpr int NV = 4 + NS;
  if (qj[0] == 0.0) return;
  double qp[NV + 2];
  phys<NS, VP>(P, qj, qp);
  const double r_h = qp[0];
  double r_u = qp[NV], r_v = qp[NV + 1];
  const double mvh = sqrt(r_u * r_u + r_v * r_v);
  const double mv = 1.0 / (1.0 / mvh + P.ff / r_h * dt);
  r_u = r_u * (mv / mvh);
  r_v = r_v * (mv / mvh);
  qp[1] = r_h * r_u;
  qp[2] = r_h * r_v;
  qp[NV] = r_u;
  qp[NV + 1] = r_v;
  qp2qc<NS, VP>(P, qp, qj);
}

// ---------------------------------------------------------------------------
// one stage of imex_RK_solver for one cell (src/solver_2d.f90:1105-1322)
// ---------------------------------------------------------------------------
// q0: q^n; s1p, s2p: (divFlux - expl) . a_tilde and (NH + SI_NH) . a_dirk
// summed over the stages before this one, in stage order from 0.0 (row i_RK
// of the tableau without its diagonal); out: the stage state and NH + SI_NH
// of the stage.  Returns the Newton iteration count (0: no Newton).
template <int NS, bool VP = false>
__device__ __forceinline__ int stage_cell(const DevParams& P, const double* q0,
                                          const double* s1p, const double* s2p,
                                          double a_diag, double dt, double Bx, double By,
                                          double Bxx, double Bxy, double Byy, double gc,
                                          double* q_rk, double* im_out, int& nonconv,
                                          int& ls_evals) {
  constexpr int NV = 4 + NS;
  double q_fv[NV], qp[NV + 2], q_guess[NV], q_si[NV], si_nh[NV], Rj[NV];
  nonconv = 0;
  ls_evals = 0;
  // q_fv = q0 - dt * ((divFlux - expl) . a_tilde - (NH + SI_NH) . a_dirk)
  // over stages 1..i_RK; stage i_RK's own terms are still zero here
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const double s1 = s1p[i] + (0.0 - 0.0) * 0.0;
    const double s2 = s2p[i] + (0.0 + 0.0) * 0.0;
    q_fv[i] = q0[i] - dt * (s1 - s2);
    Rj[i] = s1p[i] - s2p[i];  // the explicit residual before this stage
    q_guess[i] = q0[i];
    si_nh[i] = 0.0;
  }
  // qp of q_fv (:1138) feeds only the semi-implicit Coulomb terms, which
  // rheology model 1 alone has: it is formed only for them
  int iters = 0;
  if (a_diag != 0.0) {
    if (q_fv[0] > 0.0) {
      double s1 = 0.0, s2 = 0.0;
      if (P.rheo && P.rheo_model == 1) {
        phys<NS, VP>(P, q_fv, qp);
        semi_impl<NS>(P, Bx, By, Bxx, Bxy, Byy, gc, qp, s1, s2);
      }
#pragma unroll
      for (int i = 0; i < NV; i++) {
        const double si = (i == 1) ? s1 : (i == 2) ? s2 : 0.0;
        q_si[i] = q_fv[i] + dt * a_diag * si;
      }
      const double f2 = q_fv[1], f3 = q_fv[2];
      if ((f2 * f2 + f3 * f3) == 0.0) {
        q_si[1] = 0.0;
        q_si[2] = 0.0;
      } else if ((q_si[1] * f2 < 0.0) || (q_si[2] * f3 < 0.0)) {
        q_si[1] = 0.0;
        q_si[2] = 0.0;
      } else {
        const double m = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
        const double den = sqrt(f2 * f2 + f3 * f3);
        q_si[1] = m * f2 / den;
        q_si[2] = m * f3 / den;
      }
#pragma unroll
      for (int i = 0; i < NV; i++) {
        si_nh[i] = (q_si[i] - q_fv[i]) / (dt * a_diag);
        q_guess[i] = q_si[i];
        Rj[i] = Rj[i] - a_diag * si_nh[i];
      }
      // rheology model 8: solve_rk_step integrates the friction in closed
      // form and returns (src/solver_2d.f90:1734-1739)
      const bool fric8 = P.rheo && P.rheo_model == 8;
      if (fric8) {
        integrate_friction<NS, VP>(P, q_guess, dt);
      } else {
        const NewtonOut no = (P.rheo && P.rheo_model == 6)
                                 ? solve_rk_step<NS, 6>(P, q_guess, q0, Rj, a_diag, dt, Bx, By)
                                 : solve_rk_step<NS>(P, q_guess, q0, Rj, a_diag, dt, Bx, By);
        iters = no.iters;
        nonconv = !no.conv;
        ls_evals = no.ls_evals;
      }
      if (P.ny == 1) q_guess[2] = 0.0;
      if (P.nx == 1) q_guess[1] = 0.0;
      // (the implicit-term refresh of :1234-1236 is overwritten at
      // :1271-1277 below and is not evaluated; model 8 keeps the update
      // as it is, :1226-1230)
      if (fric8) {
      } else if (q_si[1] * q_si[1] + q_si[2] * q_si[2] == 0.0) {
        q_guess[1] = 0.0;
        q_guess[2] = 0.0;
      } else if ((q_guess[1] * q_si[1] <= 0.0) && (q_guess[2] * q_si[2] <= 0.0)) {
        q_guess[1] = 0.0;
        q_guess[2] = 0.0;
      } else {
        const double m = sqrt(q_guess[1] * q_guess[1] + q_guess[2] * q_guess[2]);
        const double den = sqrt(q_si[1] * q_si[1] + q_si[2] * q_si[2]);
        q_guess[1] = m * q_si[1] / den;
        q_guess[2] = m * q_si[2] / den;
      }
    } else {
#pragma unroll
      for (int i = 0; i < NV; i++) {
        q_guess[i] = q_fv[i];
        q_si[i] = q_fv[i];
      }
    }
#pragma unroll
    for (int i = 0; i < NV; i++) {
      const double nh = (q_guess[i] - q_si[i]) / (dt * a_diag);
      im_out[i] = nh + si_nh[i];
    }
  } else {
    // an explicit stage: the first stage's state is q^n (:1061-1065), its
    // implicit terms stay zero
#pragma unroll
    for (int i = 0; i < NV; i++) im_out[i] = 0.0 + 0.0;
  }
#pragma unroll
  for (int i = 0; i < NV; i++) q_rk[i] = q_guess[i];
  return iters;
}

// Wavefront counters of the stage and assembly kernels, flushed like
// k_local's.
struct StageCounters {
  unsigned long long* e;
  int imax;
};

template <int NS, bool VP = false>
__global__ void __launch_bounds__(256) kg_stage(DevParams P, DevPtrs D, RkPtrs G, RkTableau T,
                                                int st, int policy) {
  constexpr int NV = 4 + NS;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ unsigned long long cnt[4][kLocalEvents];
  __shared__ int itmax[4];
  for (int m = lane; m < kLocalEvents; m += 64) cnt[w][m] = 0;
  __syncthreads();
  bool live;
  double dt;
  if (policy) {
    const StepDt r = step_dt(P, D);
    live = !r.done;
    dt = r.dt;
  } else {
    live = (D.scal[S_DONE] == 0.0);
    dt = D.scal[S_DT];
  }
  const size_t pl = plane(P);
  const size_t sp = (size_t)NV * pl;  // planes of one stage
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int nrows = (int)*ctr(D, C_ROWS, 0);
  const double a_diag = T.a_dirk[st][st];
  const bool flux = T.omega_tilde[st] > 0.0;
  int imax = 0;
  for (int it = blockIdx.x * 4 + w; live && it < nrows; it += gridDim.x * 4) {
    const int row = D.row_list[it];
    const int tile = row / kTileH;
    const int j = (tile % tiles_x) * kStripW - 2 + lane;
    const int k = (tile / tiles_x) * kTileH + row % kTileH;
    if (!(lane >= 2 && lane < 2 + kStripW && j < P.nx && k < P.ny)) continue;
    const size_t c = cidx(P, j, k);
    if (!(D.mask[c] & M_COMPUTE)) continue;
    double q0[NV], s1[NV], s2[NV], qr[NV], im[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
      q0[i] = D.q[i * pl + c];
      s1[i] = 0.0;
      s2[i] = 0.0;
    }
    // the earlier stages' terms with row st of the tableau, in stage order
    for (int r = 0; r < st; r++) {
      const double at = T.a_tilde[st][r], ad = T.a_dirk[st][r];
#pragma unroll
      for (int i = 0; i < NV; i++) {
        s1[i] = s1[i] + G.Dx[r * sp + i * pl + c] * at;
        s2[i] = s2[i] + G.Im[r * sp + i * pl + c] * ad;
      }
    }
    const double bxx = D.Bxx ? D.Bxx[c] : 0.0, bxy = D.Bxy ? D.Bxy[c] : 0.0,
                 byy = D.Byy ? D.Byy[c] : 0.0, gc = D.gcoef ? D.gcoef[c] : 1.0;
    int nonconv, ls;
    const int its = stage_cell<NS, VP>(P, q0, s1, s2, a_diag, dt, D.Bx[c], D.By[c], bxx, bxy,
                                       byy, gc, qr, im, nonconv, ls);
#pragma unroll
    for (int i = 0; i < NV; i++) {
      G.qrk[st * sp + i * pl + c] = qr[i];
      G.Im[st * sp + i * pl + c] = im[i];
      // a stage without a flux evaluation keeps zero terms (:1298-1335)
      if (!flux) G.Dx[st * sp + i * pl + c] = 0.0;
    }
    if (its > 0) {
      atomicAdd(&cnt[w][C_NEWTON_CELLS - C_NEWTON_CELLS], 1ull);
      atomicAdd(&cnt[w][C_NEWTON_ITERS - C_NEWTON_CELLS], (unsigned long long)its);
      if (nonconv) atomicAdd(&cnt[w][C_NONCONV - C_NEWTON_CELLS], 1ull);
      if (ls) atomicAdd(&cnt[w][C_LS_EVALS - C_NEWTON_CELLS], (unsigned long long)ls);
      atomicAdd(&cnt[w][C_HIST0 - C_NEWTON_CELLS + iter_bucket(its)], 1ull);
      imax = max(imax, its);
    }
  }
  for (int off = 32; off > 0; off >>= 1) imax = max(imax, __shfl_xor(imax, off, 64));
  if (lane == 0) itmax[w] = imax;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int slot = blockIdx.x;
    for (int m = 0; m < kLocalEvents; m++) {
      const unsigned long long v = cnt[0][m] + cnt[1][m] + cnt[2][m] + cnt[3][m];
      if (v) atomicAdd(ctr(D, C_NEWTON_CELLS + m, slot), v);
    }
    const int im = max(max(itmax[0], itmax[1]), max(itmax[2], itmax[3]));
    if (im) atomicMax(ctr(D, C_ITMAX, slot), (unsigned long long)im);
  }
}

// assemble_sol (src/solver_2d.f90:1341-1626): q = q_rk(N_RK) for a stiffly
// accurate tableau, else q0 - dt * residual with residual =
// (divFlux - expl) . omega_tilde - (NH + SI_NH) . omega; then the checks
template <int NS, bool VP = false>
__global__ void __launch_bounds__(256) kg_assemble(DevParams P, DevPtrs D, RkPtrs G, RkTableau T,
                                                   int policy) {
  constexpr int NV = 4 + NS;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ unsigned long long cnt[4][kLocalEvents];
  for (int m = lane; m < kLocalEvents; m += 64) cnt[w][m] = 0;
  __syncthreads();
  bool live;
  double dt;
  if (policy) {
    const StepDt r = step_dt(P, D);
    live = !r.done;
    dt = r.dt;
  } else {
    live = (D.scal[S_DONE] == 0.0);
    dt = D.scal[S_DT];
  }
  const size_t pl = plane(P);
  const size_t sp = (size_t)NV * pl;
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int nrows = (int)*ctr(D, C_ROWS, 0);
  for (int it = blockIdx.x * 4 + w; live && it < nrows; it += gridDim.x * 4) {
    const int row = D.row_list[it];
    const int tile = row / kTileH;
    if (lane == 0) mark_dirty(P, D, tile);
    const int j = (tile % tiles_x) * kStripW - 2 + lane;
    const int k = (tile / tiles_x) * kTileH + row % kTileH;
    if (!(lane >= 2 && lane < 2 + kStripW && j < P.nx && k < P.ny)) continue;
    const size_t c = cidx(P, j, k);
    if (!(D.mask[c] & M_COMPUTE)) continue;
    double q[NV], qp[NV + 2];
#pragma unroll
    for (int i = 0; i < NV; i++) {
      double s1 = 0.0, s2 = 0.0;
      for (int r = 0; r < T.nrk; r++)
        s1 = s1 + G.Dx[r * sp + i * pl + c] * T.omega_tilde[r];
      for (int r = 0; r < T.nrk; r++) s2 = s2 + G.Im[r * sp + i * pl + c] * T.omega[r];
      q[i] = T.stiff ? G.qrk[(T.nrk - 1) * sp + i * pl + c] : D.q[i * pl + c] - dt * (s1 - s2);
    }
    bool cur = true;
    assembly_checks<NS, VP>(P, D, c, q, qp, cnt[w], true, &cur);
    bool zero = true;
#pragma unroll
    for (int i = 0; i < NV; i++) {
      D.q[i * pl + c] = q[i];
      zero = zero && (q[i] == 0.0);
    }
    D.sflag[c] = state_flags(q[0], zero);
    if (VP && D.qpp) {
      // the physical variables of the assembled state, for the erosion pass
      if (!cur) phys_or_dry<NS, VP>(P, q, qp);
#pragma unroll
      for (int i = 0; i < NV + 2; i++) D.qpp[i * pl + c] = qp[i];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int slot = blockIdx.x;
    for (int m = 0; m < kLocalEvents; m++) {
      const unsigned long long v = cnt[0][m] + cnt[1][m] + cnt[2][m] + cnt[3][m];
      if (v) atomicAdd(ctr(D, C_NEWTON_CELLS + m, slot), v);
    }
  }
}

}  // namespace imexsw
