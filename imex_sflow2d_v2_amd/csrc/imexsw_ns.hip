// imexsw_ns.hip -- launches of the kernels templated on the number of solid
// classes, compiled once per value with -DIMEXSW_NS=<0..6> so that the
// instantiations build in parallel (the C ABI is in imexsw.hip).
#include "imexsw_ctx.h"

#ifndef IMEXSW_NS
#error "compile with -DIMEXSW_NS=<number of solid classes>"
#endif

namespace imexsw {

template <int NS>
void launch_stencil(imexsw_ctx* c, const DevPtrs& D, int compute_dt, bool rows) {
  // work items: (listed tile, strip) pairs, at most tiles * kTileH / row_unit
  const int items = c->n_tiles * (kTileH / c->row_unit);
  if (c->P.ny > 1 && rows) {
    static int res = 0;
    if (!res) res = resident_grid(k_stencil_rows<NS>, 384, c->n_cu, 1 << 30);
    hipLaunchKernelGGL((k_stencil_rows<NS>), dim3(std::min(c->n_tiles * kTileH, res)),
                       dim3(384), 0, c->stream, c->P, D, compute_dt);
  } else if (c->P.vp) {
    // the profile closures (general path only)
    static int res2 = 0, res1 = 0, res_s = 0, res_v = 0;
    if (c->P.ny > 1 && D.sides && D.qpp) {
      // the state's stored physical variables: its face sides first
      if (!res_s) res_s = resident_grid(k_vp_sides<NS>, 256, c->n_cu, 1 << 30);
      if (!res_v) res_v = resident_grid(k_stencil_vp<NS>, 64, c->n_cu, 1 << 30);
      hipLaunchKernelGGL((k_vp_sides<NS>), dim3(std::min(c->n_tiles, res_s)), dim3(256), 0,
                         c->stream, c->P, D);
      hipLaunchKernelGGL((k_stencil_vp<NS>), dim3(std::min(items, res_v)), dim3(64), 0,
                         c->stream, c->P, D, c->row_unit, c->strip_h, c->target_strips,
                         compute_dt);
    } else if (c->P.ny > 1) {
      if (!res2) res2 = resident_grid(k_stencil<NS, true, true>, 64, c->n_cu, 1 << 30);
      hipLaunchKernelGGL((k_stencil<NS, true, true>), dim3(std::min(items, res2)), dim3(64), 0,
                         c->stream, c->P, D, c->row_unit, c->strip_h, c->target_strips,
                         compute_dt);
    } else {
      if (!res1) res1 = resident_grid(k_stencil<NS, false, true>, 64, c->n_cu, 1 << 30);
      hipLaunchKernelGGL((k_stencil<NS, false, true>), dim3(std::min(items, res1)), dim3(64), 0,
                         c->stream, c->P, D, c->row_unit, c->strip_h, c->target_strips,
                         compute_dt);
    }
  } else if (c->P.ny > 1) {
    static int res = 0;
    if (!res) res = resident_grid(k_stencil<NS, true>, 64, c->n_cu, 1 << 30);
    hipLaunchKernelGGL((k_stencil<NS, true>), dim3(std::min(items, res)), dim3(64), 0,
                       c->stream, c->P, D, c->row_unit, c->strip_h, c->target_strips,
                       compute_dt);
  } else {
    static int res = 0;
    if (!res) res = resident_grid(k_stencil<NS, false>, 64, c->n_cu, 1 << 30);
    hipLaunchKernelGGL((k_stencil<NS, false>), dim3(std::min(items, res)), dim3(64), 0,
                       c->stream, c->P, D, c->row_unit, c->strip_h, c->target_strips,
                       compute_dt);
  }
}
template <int NS>
void launch_local(imexsw_ctx* c, int ops, bool fused) {
  // work items: listed tile rows, four per workgroup
  const int items = (c->n_tiles * kTileH + 3) / 4;
  // a Newton starting point for the stage update: the learned one (its
  // guesses computed first, into the guess planes) or the closed form
  int start = 0;
  if ((ops & OP_RK) && c->warm) {
    static int resg = 0;
    if (!resg) resg = resident_grid(k_mlp_guess<NS>, 64, c->n_cu, 1 << 30);
    hipLaunchKernelGGL((k_mlp_guess<NS>), dim3(std::min(c->n_tiles * kTileH, resg)), dim3(64),
                       0, c->stream, c->P, c->D, c->mlp, (ops & OP_POLICY) ? 1 : 0);
    start = OP_WARM;
  } else if ((ops & OP_RK) && c->closed_start) {
    start = OP_CLOSED;
  }
  if (fused) {
    // listed cells, kCellsPerWave per wavefront, four wavefronts per workgroup
    const long long cells = (long long)c->P.nx * c->P.ny;
    const int groups = (int)std::min<long long>((cells + 4 * kCellsPerWave - 1) /
                                                    (4 * kCellsPerWave), 1 << 30);
    static int res3 = 0, res4 = 0;
    if (!res3) res3 = resident_grid(k_local<NS, 3>, 256, c->n_cu, 1 << 30);
    if (!res4) res4 = resident_grid(k_local<NS, 4>, 256, c->n_cu, 1 << 30);
    if (start)
      hipLaunchKernelGGL((k_local<NS, 4>), dim3(std::min(groups, res4)), dim3(256), 0,
                         c->stream, c->P, c->D, (ops & ~(int)OP_SUMS) | start);
    else
      hipLaunchKernelGGL((k_local<NS, 3>), dim3(std::min(groups, res3)), dim3(256), 0,
                         c->stream, c->P, c->D, ops & ~(int)OP_SUMS);
    return;
  }
  static int res0 = 0, res1 = 0, res2 = 0;
  if (!res0) res0 = resident_grid(k_local<NS, 0>, 256, c->n_cu, 1 << 30);
  if (!res1) res1 = resident_grid(k_local<NS, 1>, 256, c->n_cu, 1 << 30);
  if (!res2) res2 = resident_grid(k_local<NS, 2>, 256, c->n_cu, 1 << 30);
  if (ops & OP_RK) {
    int rk = OP_RK | (ops & OP_POLICY) | start;
    if (c->D.harvest) rk |= OP_HARVEST;
    if (!(rk & (OP_WARM | OP_HARVEST | OP_CLOSED)))
      hipLaunchKernelGGL((k_local<NS, 0>), dim3(std::min(items, res0)), dim3(256), 0, c->stream,
                         c->P, c->D, rk);
    else
      hipLaunchKernelGGL((k_local<NS, 2>), dim3(std::min(items, res2)), dim3(256), 0, c->stream,
                         c->P, c->D, rk);
  }
  if ((ops & (OP_EROSION | OP_POST | OP_SUMS)) && c->P.vp) {
    static int res1v = 0;
    if (!res1v) res1v = resident_grid(k_local<NS, 1, true>, 256, c->n_cu, 1 << 30);
    hipLaunchKernelGGL((k_local<NS, 1, true>), dim3(std::min(items, res1v)), dim3(256), 0,
                       c->stream, c->P, c->D, ops & ~(int)OP_RK);
  } else if (ops & (OP_EROSION | OP_POST | OP_SUMS)) {
    hipLaunchKernelGGL((k_local<NS, 1>), dim3(std::min(items, res1)), dim3(256), 0, c->stream,
                       c->P, c->D, ops & ~(int)OP_RK);
  }
}
template <int NS>
void launch_qp(imexsw_ctx* c, double* dst, int k0, int k1) {
  const dim3 grid(c->grid_tiles.x, (k1 - k0 + kTileY - 1) / kTileY);
  if (c->P.vp)
    hipLaunchKernelGGL((k_qp_out<NS, true>), grid, dim3(256), 0, c->stream, c->P, c->D.q, dst,
                       k0, k1);
  else
    hipLaunchKernelGGL((k_qp_out<NS>), grid, dim3(256), 0, c->stream, c->P, c->D.q, dst, k0, k1);
}
template <int NS>
void launch_qp_planes(imexsw_ctx* c) {
  if (c->P.vp)
    hipLaunchKernelGGL((k_qp_planes<NS, true>), c->grid_tiles, dim3(256), 0, c->stream, c->P,
                       c->D);
  else
    hipLaunchKernelGGL((k_qp_planes<NS, false>), c->grid_tiles, dim3(256), 0, c->stream, c->P,
                       c->D);
}
template <int NS>
void launch_convert(const DevParams& P, hipStream_t s, int n_cu, long long n, const double* in,
                    const double* bx, const double* by, double* out, int mode) {
  const int grid = (int)std::min<long long>((n + 255) / 256, (long long)n_cu * 8);
  if (P.vp)
    hipLaunchKernelGGL((k_convert<NS, true>), dim3(grid), dim3(256), 0, s, P, n, in, bx, by, out,
                       mode);
  else
    hipLaunchKernelGGL((k_convert<NS, false>), dim3(grid), dim3(256), 0, s, P, n, in, bx, by,
                       out, mode);
}

template <int NS, bool VP>
void launch_stage_v(imexsw_ctx* c, int st, int policy) {
  const int items = (c->n_tiles * kTileH + 3) / 4;
  static int res = 0;
  if (!res) res = resident_grid(kg_stage<NS, VP>, 256, c->n_cu, 1 << 30);
  hipLaunchKernelGGL((kg_stage<NS, VP>), dim3(std::min(items, res)), dim3(256), 0, c->stream,
                     c->P, c->D, c->G, c->tab, st, policy);
}
template <int NS>
void launch_stage(imexsw_ctx* c, int st, int policy) {
  if (c->P.vp) launch_stage_v<NS, true>(c, st, policy);
  else launch_stage_v<NS, false>(c, st, policy);
}
template <int NS, bool VP>
void launch_assemble_v(imexsw_ctx* c, int policy) {
  const int items = (c->n_tiles * kTileH + 3) / 4;
  static int res = 0;
  if (!res) res = resident_grid(kg_assemble<NS, VP>, 256, c->n_cu, 1 << 30);
  hipLaunchKernelGGL((kg_assemble<NS, VP>), dim3(std::min(items, res)), dim3(256), 0, c->stream,
                     c->P, c->D, c->G, c->tab, policy);
}
template <int NS>
void launch_assemble(imexsw_ctx* c, int policy) {
  if (c->P.vp) launch_assemble_v<NS, true>(c, policy);
  else launch_assemble_v<NS, false>(c, policy);
}

template <int NS>
void launch_batch_stencil(imexsw_ensemble* e) {
  const BatchLists L = batch_lists(e);
  const int per = kTileH / e->strip_h;
  static int rs = 0;
  if (!rs) rs = resident_grid(kb_stencil<NS>, 64, e->n_cu, 1 << 30);
  hipLaunchKernelGGL((kb_stencil<NS>), dim3(std::min(e->total_tiles * per, rs)), dim3(64), 0,
                     e->stream, e->d_params, e->d_ptrs, L, e->strip_h);
}
template <int NS>
void launch_batch_local(imexsw_ensemble* e) {
  const BatchLists L = batch_lists(e);
  static int rl0 = 0, rl1 = 0;
  if (!rl0) rl0 = resident_grid(kb_local<NS, 0>, 256, e->n_cu, 1 << 30);
  if (!rl1) rl1 = resident_grid(kb_local<NS, 1>, 256, e->n_cu, 1 << 30);
  const int litems = (e->total_tiles * kTileH + 3) / 4;
  hipLaunchKernelGGL((kb_local<NS, 0>), dim3(std::min(litems, rl0)), dim3(256), 0, e->stream,
                     e->d_params, e->d_ptrs, L, (int)OP_RK);
  hipLaunchKernelGGL((kb_local<NS, 1>), dim3(std::min(litems, rl1)), dim3(256), 0, e->stream,
                     e->d_params, e->d_ptrs, L, (int)(OP_EROSION | OP_POST));
}

template void launch_stencil<IMEXSW_NS>(imexsw_ctx*, const DevPtrs&, int, bool);
template void launch_stage<IMEXSW_NS>(imexsw_ctx*, int, int);
template void launch_assemble<IMEXSW_NS>(imexsw_ctx*, int);
template void launch_local<IMEXSW_NS>(imexsw_ctx*, int, bool);
template void launch_qp<IMEXSW_NS>(imexsw_ctx*, double*, int, int);
template void launch_qp_planes<IMEXSW_NS>(imexsw_ctx*);
template void launch_convert<IMEXSW_NS>(const DevParams&, hipStream_t, int, long long,
                                        const double*, const double*, const double*, double*,
                                        int);
#if IMEXSW_NS >= 1 && IMEXSW_NS <= 3
template void launch_batch_stencil<IMEXSW_NS>(imexsw_ensemble*);
template void launch_batch_local<IMEXSW_NS>(imexsw_ensemble*);
#endif

}  // namespace imexsw
