// imexsw_batch.h -- the time step of many independent grids ("members" of an
// ensemble) in one set of launches.
//
// Each member keeps its own device state, constants (DevParams) and
// pointers (DevPtrs), stored in device arrays indexed by member.  k_mask's
// work lists are shared: every entry carries its member in the bits above
// kMemberShift, so the stencil and local kernels walk all members' tiles and
// rows in one grid, whatever the members' sizes.  Per member, the work is
// exactly the single-grid path's (the same device functions), so a member
// advanced in a batch gives the bits it gives alone.
#pragma once

#include "imexsw_kernels.h"

namespace imexsw {

constexpr int kMemberShift = 20;  // tile / row ids below, member above
constexpr int kIdMask = (1 << kMemberShift) - 1;

struct BatchLists {
  int* tiles;                   // tag | tile
  int* rows;                    // tag | (tile * kTileH + row)
  unsigned long long* counts;   // [0] tiles, [1] rows
  const int* tile_off;          // member m's tiles are [tile_off[m], tile_off[m+1])
  int members;
};

template <int = 0>
__global__ void kb_reset(const DevParams* __restrict__ Ps, const DevPtrs* __restrict__ Ds,
                         BatchLists L) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == 0) {
    L.counts[0] = 0;
    L.counts[1] = 0;
  }
  if (m < L.members) reset_member(Ps[m], Ds[m]);
}

// one wavefront per tile of every member (global tile id g), kMaskGroup
// consecutive tiles per workgroup
template <int = 0>
__global__ void __launch_bounds__(64 * kMaskGroup) kb_mask(const DevParams* __restrict__ Ps,
                                                           const DevPtrs* __restrict__ Ds,
                                                           BatchLists L) {
  const int g = blockIdx.x * kMaskGroup + (threadIdx.x >> 6);
  int lo = 0, hi = L.members - 1;  // member with tile_off[m] <= g < tile_off[m+1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.tile_off[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const int m = lo;
  const DevParams& P = Ps[m];
  const DevPtrs& D = Ds[m];
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int t = g - L.tile_off[m];
  unsigned rows_m = 0;
  if (g < L.tile_off[L.members] && D.scal[S_DONE] == 0.0)
    rows_m = mask_tile(P, D, t % tiles_x, t / tiles_x, tiles_x, 0, false);
  const ListSink sink{L.tiles, L.rows, &L.counts[0], &L.counts[1], m << kMemberShift};
  append_group<kMaskGroup>(sink, t, rows_m);
}

// (tile, strip) items of all members, strip height fixed for the batch
template <int NS>
__global__ void __launch_bounds__(64) kb_stencil(const DevParams* __restrict__ Ps,
                                                 const DevPtrs* __restrict__ Ds, BatchLists L,
                                                 int strip_h) {
  const int per = kTileH / strip_h;
  const int n = (int)L.counts[0];
  for (int item = blockIdx.x; item < n * per; item += gridDim.x) {
    const int e = L.tiles[item / per];
    const int m = e >> kMemberShift, tile = e & kIdMask;
    const DevParams& P = Ps[m];
    const DevPtrs& D = Ds[m];
    const int tiles_x = (P.nx + kStripW - 1) / kStripW;
    const int k0 = (tile / tiles_x) * kTileH + (item % per) * strip_h;
    if (k0 >= P.ny) continue;
    stencil_strip<NS, true>(P, D, src_time(P, D.scal[S_T]),
                            (tile % tiles_x) * kStripW - 2 + (int)threadIdx.x, k0,
                            min(k0 + strip_h, P.ny), true, P.cfl != -1.0);
  }
}

template <int = 0>
__global__ void kb_dtpolicy(const DevParams* __restrict__ Ps, const DevPtrs* __restrict__ Ds,
                            int members) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < members) dtpolicy_member(Ps[m], Ds[m]);
}

// rows of all members; each wavefront flushes its event counters into the
// member of the row it just finished
template <int NS, int PH>
__global__ void __launch_bounds__(256) kb_local(const DevParams* __restrict__ Ps,
                                                const DevPtrs* __restrict__ Ds, BatchLists L,
                                                int ops) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ unsigned long long cnt[4][kLocalEvents];
  const int n = (int)L.counts[1];
  for (int it = blockIdx.x * 4 + w; it < n; it += gridDim.x * 4) {
    const int e = L.rows[it];
    const int m = e >> kMemberShift;
    const DevParams& P = Ps[m];
    const DevPtrs& D = Ds[m];
    if (D.scal[S_DONE] != 0.0) continue;
    for (int i = lane; i < kLocalEvents; i += 64) cnt[w][i] = 0;
    __builtin_amdgcn_wave_barrier();
    int imax = 0;
    local_row<NS, PH>(P, D, e & kIdMask, D.scal[S_DT], ops, cnt[w], imax);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (lane < kLocalEvents && cnt[w][lane])
      atomicAdd(ctr(D, C_NEWTON_CELLS + lane, it), cnt[w][lane]);
    for (int off = 32; off > 0; off >>= 1) imax = max(imax, __shfl_xor(imax, off, 64));
    if (lane == 0 && imax) atomicMax(ctr(D, C_ITMAX, it), (unsigned long long)imax);
    __builtin_amdgcn_wave_barrier();
  }
}

template <int = 0>
__global__ void kb_advance(const DevPtrs* __restrict__ Ds, int members) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < members) advance_member(Ds[m]);
}

}  // namespace imexsw
