// imexsw_kernels.h -- gfx950 kernels of the per-timestep path.
//
// Device layout: every per-cell field is a structure-of-arrays plane of
// `pitch * ny` doubles, x fastest (j), so a 64-lane wavefront reading 64
// consecutive cells of a row issues fully coalesced 512-byte loads.  The
// conservative state is NV such planes, the explicit residual D another NV.
//
// Kernels per time step (N_RK = 2):
//   k_step_begin  commit the previous step (host-loop dt policy, t += dt,
//                 counters) and reset the step's counters
//   k_mask        check_solve: active mask (wet cells dilated N_RK times)
//                 and the step's work lists
//   k_stencil, k_stencil_rows
//                 reconstruction, wave speeds, KT fluxes, divergence,
//                 explicit terms -> D = divFlux - expl_terms, and the CFL dt
//                 (an ordered-bits atomicMin); a strip march on dense steps,
//                 six wavefronts per listed row on sparse steps
//   k_local       stage-2 Newton + assembly checks, erosion/deposition,
//                 post-step diagnostics (dense: two kernels; sparse: one
//                 pass, four lanes per cell)
//   k_topo[_list] topography_reconstruction of the slope fields (when the
//                 topography changes)
#pragma once

#include "imexsw_mlp.h"

namespace imexsw {

constexpr int kWave = 64;
constexpr int kStripW = kWave - 4;  // interior columns per stencil wavefront
constexpr int kTileX = 64, kTileY = 4;  // layout-conversion kernels
// Work tiles of the time step: kStripW x kTileH cells.  Lane l of a
// wavefront owns column 60 * tile_x - 2 + l (lanes 2..61 are the tile's
// cells, lanes 0, 1, 62, 63 its two-cell halo), so k_mask, k_stencil,
// k_local and k_topo all see a tile the same way.
constexpr int kTileH = 16;
// lanes per cell in the one-pass local kernel of sparse steps (a divisor of
// kStripW and of the wavefront)
constexpr int kSparseLanes = 4;

// ---------------------------------------------------------------------------
// pointers of one context (device addresses)
// ---------------------------------------------------------------------------
struct DevPtrs {
  double* q;         // NV planes
  double* D;         // NV planes
  uint8_t* mask;     // 1 plane (bytes)
  uint8_t* sflag;    // 1 plane (bytes): StateFlags of the current state
  const uint8_t* sinfo;  // SourceBits plane (source_cell and ring direction) or null
  const double* sring;   // 2 planes: the ring cell's direction vector, or null
  double* B;         // B_cent
  double* Bx;        // B_prime_x_geom
  double* By;
  double* Bxx;       // curvature only (else null)
  double* Bxy;
  double* Byy;
  double* gcoef;     // slope correction only (else null)
  double* dgdx;
  double* gsx;       // (nx+1) x ny, pitch+1
  double* gsy;       // nx x (ny+1)
  const uint8_t* nodata;  // or null
  double* erodible;  // NS planes
  double* deposit;   // NS planes
  double* erosion;   // NS planes
  double* hmax;
  double* pdynmax;
  double* modvel;
  uint32_t* vuln;    // bitmask plane or null
  unsigned long long* dt_bits;  // CFL min (ordered bits of a positive double)
  double* scal;      // device scalars, see ScalarSlot
  unsigned long long* counters;  // see CounterSlot
  double* tile_sums;  // 4 per work tile
  uint8_t* tile_flag; // per work tile: any cell to compute
  uint8_t* tile_dirty;  // per work tile: a state within reach may have changed
  int* tile_nact;       // per work tile: active cells at its last mask scan
  int* tile_list;     // work tiles with a cell to compute (compacted)
  int* row_list;      // tile rows (tile * kTileH + row) with a cell to compute
  int* cell_list;     // cells (plane index) to compute, or null
  double* qpp;        // NV+2 planes: the physical variables of q (with the
                      // vertical profiles, whose conversion iterates), or null
  double* sides;      // 4 x (2 NV + 3) planes: the W, E, S, N face sides of
                      // each cell (vertical profiles, 2-D), or null
  double* guess;      // 2 planes: learned Newton start for the momenta (NaN: none)
  float* harvest;     // harvest_cap pairs of kHarvestWidth floats, or null
  long long harvest_cap;
};

enum ScalarSlot { S_T = 0, S_DT = 1, S_DT_OLD = 2, S_DT_OLD_OLD = 3, S_T_END = 4,
                  S_T_OUT = 5, S_DONE = 6, S_DT_CFL = 7, S_NSLOTS = 8 };
enum CounterSlot { C_SOLVE_CELLS = 0, C_NEWTON_CELLS, C_NEWTON_ITERS, C_NONCONV,
                   C_NAN, C_NEGH, C_NEGS, C_SUMS, C_LOWT, C_LS_EVALS,
                   C_HIST0, C_HIST1, C_HIST2, C_HIST3, C_HIST4, C_HIST5, C_HIST6,
                   C_HIST7, C_WARM_USED, C_WARM_FALLBACK, C_ITMAX, C_STEPS,
                   C_ACTIVE_UPD, C_COMPUTE_CELLS, C_TILES, C_ROWS, C_HARVEST,
                   C_COMPUTE_UPD, C_CELLS, C_NCOUNTERS };
// k_local's per-cell event counters, in CounterSlot order from C_NEWTON_CELLS
constexpr int kLocalEvents = C_ITMAX - C_NEWTON_CELLS;  // summed ones
// Newton iteration histogram buckets: 1, 2, 3, 4, 5-8, 9-16, 17-32, 33-100
__device__ __forceinline__ int iter_bucket(int it) {
  return it <= 4 ? it - 1 : it <= 8 ? 4 : it <= 16 ? 5 : it <= 32 ? 6 : 7;
}
// Every counter has kSlots copies, counters[m * kSlots + slot]; workgroups
// add into slot (workgroup id mod kSlots) so that thousands of workgroups do
// not serialise on one L2 line.  Totals are the sum over the slots.
constexpr int kSlots = 64;
__device__ __forceinline__ unsigned long long* ctr(const DevPtrs& D, int m, int slot) {
  return D.counters + m * kSlots + (slot & (kSlots - 1));
}

// mask byte: bit 0 = check_solve's active flag (solve_mask), bit 1 = the
// cell's update has to be computed.  An active cell whose update is known
// to leave it unchanged is "quiet": it and every cell of its reconstruction
// stencil (two cells along x and along y) hold an all-zero state, no source
// cell is in that stencil and no boundary condition within reach can feed
// mass in.  Its faces then carry dry states (zero
// flux, zero wave speed) and the stage update, erosion/deposition and
// diagnostics return the state and accumulators as they are.
enum MaskBits { M_ACTIVE = 1, M_COMPUTE = 2 };
// source byte of a cell: source_cell == 1 (interior), source_cell == 2
// (ring), the ring's boundary directions and whether one of them is set
// (the direction vector sring is then the one of the first of E, W, S, N)
enum SourceBits { SI_SRC = 1, SI_RING = 2, SI_E = 4, SI_W = 8, SI_S = 16, SI_N = 32,
                  SI_DIR = 64 };
// state flag byte, kept next to the state by every kernel that writes it, so
// that k_mask reads one byte per cell instead of the state planes
enum StateFlags { SF_ZERO = 1,   // every conservative variable is 0.0
                  SF_POS = 2 };  // h > 0
__device__ __forceinline__ uint8_t state_flags(double h, bool all_zero) {
  return (uint8_t)((all_zero ? SF_ZERO : 0) | (h > 0.0 ? SF_POS : 0));
}

__device__ __forceinline__ size_t cidx(const DevParams& P, int j, int k) {
  return (size_t)k * P.pitch + j;
}
__device__ __forceinline__ size_t plane(const DevParams& P) {
  return (size_t)P.pitch * P.ny;
}

// x_stag / x_comp / y_stag / y_comp of init_grid (src/geometry_2d.f90:303-325)
__device__ __forceinline__ double xstag(const DevParams& P, int j) { return P.x0 + (double)j * P.dx; }
__device__ __forceinline__ double ystag(const DevParams& P, int k) { return P.y0 + (double)k * P.dy; }
__device__ __forceinline__ double xcomp(const DevParams& P, int j) {
  return 0.5 * (xstag(P, j) + xstag(P, j + 1));
}
__device__ __forceinline__ double ycomp(const DevParams& P, int k) {
  return 0.5 * (ystag(P, k) + ystag(P, k + 1));
}

// A value that has to be in registers at this point of the program: the
// load that produces it is not moved past it (into a branch further down).
__device__ __forceinline__ void pin(double& x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void pin(int& x) { asm volatile("" : "+v"(x)); }
__device__ __forceinline__ void pin(unsigned long long& x) { asm volatile("" : "+v"(x)); }

// The first BYTES of the kernel's argument block (DevParams and DevPtrs are
// passed by value) touched once per 64-byte line, all loads in flight
// together.  Short of scalar registers, the compiler reloads single
// arguments from that block throughout a large kernel, each reload waited
// for on its own; with the block's lines in the scalar cache from the start
// those reloads no longer go to L2 one at a time.
template <unsigned BYTES>
__device__ __forceinline__ void warm_kernargs() {
  typedef const __attribute__((address_space(4))) uint32_t karg_t;
  karg_t* a = (karg_t*)__builtin_amdgcn_kernarg_segment_ptr();
  constexpr unsigned N = (BYTES + 63) / 64;
  uint32_t x[N];
#pragma unroll
  for (unsigned i = 0; i < N; i++) x[i] = a[i * 16];  // i * 64 < BYTES
#pragma unroll
  for (unsigned i = 0; i < N; i++) asm volatile("" ::"s"(x[i]));
}
constexpr unsigned kArgBytes = sizeof(DevParams) + sizeof(DevPtrs);

// ---------------------------------------------------------------------------
// wave shuffles of doubles (ds_bpermute on 32-bit halves)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double shfl_d(double v, int src) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __shfl(lo, src, kWave);
  hi = __shfl(hi, src, kWave);
  return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------
// k_mask: check_solve (src/solver_2d.f90:666-857), one wavefront per work
// tile and kMaskGroup tiles (stacked along y) per workgroup.  Lane l reads
// the state flags of column 60*tx-2+l over the tile's 16 rows and a two-row
// halo; the N_RK-fold 4-neighbour dilation (an L1 diamond of radius N_RK)
// and the quiet test are bit operations on those 20-bit columns.  Tiles and
// tile rows with a cell to compute are appended to the work lists, one list
// reservation per workgroup.
// ---------------------------------------------------------------------------
constexpr int kMaskGroup = 4;
// Where k_mask appends the tiles and tile rows that have a cell to compute:
// entries are `tag | tile` and `tag | (tile * kTileH + row)`; with `cells`,
// also the cells themselves (plane indices, tile by tile, row-major within a
// tile).
struct ListSink {
  int* tiles;
  int* rows;
  unsigned long long* n_tiles;
  unsigned long long* n_rows;
  int tag;
  int* cells;
  unsigned long long* n_cells;
};

__device__ __forceinline__ ListSink own_sink(const DevPtrs& D) {
  return ListSink{D.tile_list, D.row_list, ctr(D, C_TILES, 0), ctr(D, C_ROWS, 0), 0,
                  D.cell_list, D.cell_list ? ctr(D, C_CELLS, 0) : nullptr};
}

// Bit-parallel check_solve for one tile (tx, ty), one 64-lane wavefront.
// Lane l holds column 60*tx-2+l and keeps its 20 rows (the tile and a
// two-row halo, row r = bit r) as bit words: seeds S (h > 0, boundary,
// source ring), wet W (a non-zero state or any source cell), source
// interior I.  The
// reference's two 4-neighbour dilation passes reach the L1 diamond of
// radius N_RK = 2, which per lane is
//   V2(S) | V1(S of lanes l+-1) | S of lanes l+-2,
// with V1, V2 the vertical dilations by one and two bits; the quiet test
// ORs W over the radius-2 cross.  Everything is 32-bit lane arithmetic plus
// a few cross-lane reads.
__device__ __forceinline__ unsigned vdil(unsigned m, int d) {
  unsigned r = m;
  if (d >= 1) r |= (m << 1) | (m >> 1);
  if (d >= 2) r |= (m << 2) | (m >> 2);
  return r;
}
__device__ __forceinline__ unsigned lane_read(unsigned v, int src) {
  return (unsigned)__shfl((int)v, src & 63, 64);
}

// Writes the tile's mask bytes, counters and tile flag; returns the
// (wavefront-uniform) bit word of the tile rows with a cell to compute.
// `comp_out` (when given) receives this lane's column of cells to compute,
// bit r = tile row r.
// `clean`: a partial scan found the tile not dirty, so it keeps its mask
// bytes and counts (nothing within reach changed state since its last scan,
// so neither can have changed, and it has no cell to compute).  The
// active-cell counter is a running total kept by per-tile differences.
__device__ __forceinline__ unsigned mask_tile(const DevParams& P, const DevPtrs& D, int tx,
                                              int tyy, int tiles_x, int solve_all,
                                              bool clean, unsigned* comp_out = nullptr) {
  constexpr int LH = kTileH + 4;
  const int lane = threadIdx.x & 63;
  const int tile = tyy * tiles_x + tx;
  if (comp_out) *comp_out = 0u;
  if (clean) return 0u;
  // the tile's previous active count, loaded with the flags
  const int nact_old = D.tile_nact[tile];
  const int j = tx * kStripW - 2 + lane;
  const int k0 = tyy * kTileH;
  const bool injx = (j >= 0 && j < P.nx);
  const bool own = injx && lane >= 2 && lane < 2 + kStripW;
  const bool radial = P.radial && D.sinfo;
  unsigned S = 0, W = 0, I = 0, rowin = 0;
  // unconditional loads from clamped (always valid) addresses, selected
  // afterwards: straight-line code lets all loads be in flight at once (a
  // load under a condition is waited for where the branches join, which
  // would serialise the rows); without a source the source bytes are read
  // from the state flags and not used
  const int jc = min(max(j, 0), P.nx - 1);
  const uint8_t* sinfo = radial ? D.sinfo : D.sflag;
  int fv[LH], sv[LH];
#pragma unroll
  for (int r = 0; r < LH; r++) {
    const int kc = min(max(k0 - 2 + r, 0), P.ny - 1);
    fv[r] = D.sflag[cidx(P, jc, kc)];
    sv[r] = sinfo[cidx(P, jc, kc)];
  }
#pragma unroll
  for (int r = 0; r < LH; r++) sv[r] = radial ? sv[r] : 0;
#pragma unroll
  for (int r = 0; r < LH; r++) {
    const int kk = k0 - 2 + r;
    const bool in = injx && kk >= 0 && kk < P.ny;
    const int f = in ? fv[r] : SF_ZERO;
    const int sc = in ? sv[r] : 0;
    const bool bnd = in && (j == 0 || j == P.nx - 1 || kk == 0 || kk == P.ny - 1);
    S |= (in && (bnd || solve_all || (f & SF_POS) || (sc & SI_RING))) ? (1u << r) : 0u;
    // "wet" for the quiet test: any non-zero variable or a source cell
    W |= (in && (!(f & SF_ZERO) || (sc & (SI_SRC | SI_RING)))) ? (1u << r) : 0u;
    I |= (in && (sc & SI_SRC)) ? (1u << r) : 0u;
    rowin |= (kk >= 0 && kk < P.ny) ? (1u << r) : 0u;
  }
  // active: diamond of radius 2 around the seeds, boundary cells, minus the
  // source interior; bits 2..17 are the tile's rows
  const unsigned v1 = vdil(S, 1);
  unsigned act = vdil(S, 2) | lane_read(v1, lane - 1) | lane_read(v1, lane + 1) |
                 lane_read(S, lane - 2) | lane_read(S, lane + 2);
  // lane reads wrap at the wavefront's ends: lanes 0, 1, 62, 63 are halo
  // columns whose results are not used
  unsigned edge_rows = 0;  // rows k == 0 or k == ny-1
  unsigned lo_rows = 0, hi_rows = 0;  // rows k <= 1, k >= ny-2
#pragma unroll
  for (int r = 0; r < LH; r++) {
    const int kk = k0 - 2 + r;
    if (kk == 0 || kk == P.ny - 1) edge_rows |= 1u << r;
    if (kk <= 1) lo_rows |= 1u << r;
    if (kk >= P.ny - 2) hi_rows |= 1u << r;
  }
  if (j == 0 || j == P.nx - 1) act = ~0u;
  act |= edge_rows;
  act &= ~I & rowin;
  if (!own) act = 0;
  // busy: not quiet
  unsigned busy = W | lane_read(W, lane - 1) | lane_read(W, lane + 1) |
                  lane_read(W, lane - 2) | lane_read(W, lane + 2);
  if (P.ny > 1) busy |= vdil(W, 2);
  if (j <= 1 && !P.bc_quiet[IMEXSW_W]) busy = ~0u;
  if (j >= P.nx - 2 && !P.bc_quiet[IMEXSW_E]) busy = ~0u;
  if (P.ny > 1) {
    if (!P.bc_quiet[IMEXSW_S]) busy |= lo_rows;
    if (!P.bc_quiet[IMEXSW_N]) busy |= hi_rows;
  }
  const unsigned rows16 = ((1u << kTileH) - 1u) << 2;
  act &= rows16;
  const unsigned comp = act & busy;
  if (own) {
#pragma unroll
    for (int rr = 0; rr < kTileH; rr++) {
      const int k = k0 + rr;
      if (k < P.ny) {
        const unsigned bit = 1u << (rr + 2);
        D.mask[cidx(P, j, k)] =
            (uint8_t)(((act & bit) ? M_ACTIVE : 0) | ((comp & bit) ? M_COMPUTE : 0));
      }
    }
  }
  if (comp_out) *comp_out = own ? (comp >> 2) : 0u;
  int count = __popc(act), ccount = __popc(comp);
  unsigned rows_m = comp >> 2;
  for (int off = 32; off > 0; off >>= 1) {
    count += __shfl_xor(count, off, 64);
    ccount += __shfl_xor(ccount, off, 64);
    rows_m |= (unsigned)__shfl_xor((int)rows_m, off, 64);
  }
  if (lane == 0) {
    const int delta = count - nact_old;
    if (delta) atomicAdd(ctr(D, C_SOLVE_CELLS, tile), (unsigned long long)(long long)delta);
    D.tile_nact[tile] = count;
    D.tile_dirty[tile] = 0;
    if (ccount) atomicAdd(ctr(D, C_COMPUTE_CELLS, tile), (unsigned long long)ccount);
    D.tile_flag[tile] = rows_m ? 1 : 0;
  }
  return rows_m;
}

// Appends the workgroup's tiles with a row to compute (wavefront w: tile
// `tile`, compute rows `rows_m`, 0 for none or for an idle wavefront) to the
// work lists with one reservation per list for the whole workgroup.  Every
// thread of the workgroup calls it.  With a cell list in the sink, `comp` is
// the lane's column of cells to compute (bit r = tile row r) and `cell0` the
// plane index of its row-0 cell; the cells go to the list tile row by tile
// row, in lane order within a row; entries past `cap` (the list's length)
// are not written.
template <int G>
__device__ __forceinline__ void append_group(const ListSink& L, int tile, unsigned rows_m,
                                             unsigned comp = 0u, int cell0 = 0,
                                             int pitch = 0, long long cap = 0) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ unsigned s_rows[G];
  __shared__ int s_cells[G];
  __shared__ int s_tb, s_rb, s_cb;
  const bool cells = L.cells != nullptr;
  int nc_w = 0;
  if (cells) {
    nc_w = __popc(comp);
    for (int off = 32; off > 0; off >>= 1) nc_w += __shfl_xor(nc_w, off, 64);
  }
  if (lane == 0) {
    s_rows[w] = rows_m;
    s_cells[w] = nc_w;
  }
  __syncthreads();
  if (w == 0) {
    int nt = 0, nr = 0, nc = 0;
    for (int g = 0; g < G; g++) {
      nt += s_rows[g] != 0u;
      nr += __popc(s_rows[g]);
      nc += s_cells[g];
    }
    // the three reservations as one atomic instruction on three lanes
    // (tiles, rows, cells), so that they are in flight together: three
    // single-lane atomics each wait for their result in turn.  nr > 0
    // exactly when nt > 0, and nc > 0 only then.
    // (the three addresses in registers first: a choice between the
    // fields' addresses would move L to memory)
    unsigned long long *c_t = L.n_tiles, *c_r = L.n_rows, *c_c = L.n_cells;
    asm volatile("" : "+v"(c_t), "+v"(c_r), "+v"(c_c));
    unsigned long long* ctr_l = (lane == 0) ? c_t : (lane == 1) ? c_r : c_c;
    const unsigned long long add = (lane == 0) ? nt : (lane == 1) ? nr : nc;
    const bool take = nt && (lane < 2 || (lane == 2 && cells && nc));
    unsigned long long old = 0ull;
    if (take) old = atomicAdd(ctr_l, add);
    if (lane == 0) s_tb = (int)old;
    if (lane == 1) s_rb = (int)old;
    if (lane == 2) s_cb = (int)old;
  }
  __syncthreads();
  if (rows_m) {
    int tb = s_tb, rb = s_rb, cb = s_cb;
    for (int g = 0; g < w; g++) {
      tb += s_rows[g] != 0u;
      rb += __popc(s_rows[g]);
      cb += s_cells[g];
    }
    if (lane == 0) L.tiles[tb] = L.tag | tile;
    if (lane < kTileH && ((rows_m >> lane) & 1u))
      L.rows[rb + __popc(rows_m & ((1u << lane) - 1u))] = L.tag | (tile * kTileH + lane);
    if (cells) {
      const unsigned long long below = (1ull << lane) - 1ull;
      for (int r = 0; r < kTileH; r++) {
        if (!((rows_m >> r) & 1u)) continue;
        const bool mine = (comp >> r) & 1u;
        const unsigned long long b = __ballot(mine);
        const long long at = (long long)cb + __popcll(b & below);
        if (mine && at < cap) L.cells[at] = cell0 + r * pitch;
        cb += __popcll(b);
      }
    }
  }
}

// grid (tiles_x, ceil(tiles_y / kMaskGroup)), kMaskGroup wavefronts
template <int = 0>
__global__ void __launch_bounds__(64 * kMaskGroup) k_mask(DevParams P, DevPtrs D,
                                                          int solve_all, int partial) {
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int tiles_y = (P.ny + kTileH - 1) / kTileH;
  const int w = threadIdx.x >> 6;
  // the group's dirty flags from clamped addresses, in flight together with
  // the done flag
  int fl[kMaskGroup];
#pragma unroll
  for (int g = 0; g < kMaskGroup; g++)
    fl[g] = D.tile_dirty[min((int)blockIdx.y * kMaskGroup + g, tiles_y - 1) * tiles_x +
                         (int)blockIdx.x];
  double done = D.scal[S_DONE];
#pragma unroll
  for (int g = 0; g < kMaskGroup; g++) pin(fl[g]);
  pin(done);
  // a run that has stopped (t_output or t_end reached) keeps the mask of
  // its last step, the solve_mask the reference leaves behind
  if (done != 0.0) return;
  bool clean = false;
  if (partial) {
    // a workgroup whose tiles are all clean has nothing to do (the same
    // test on every wavefront, so none waits at a barrier)
    bool any = false;
#pragma unroll
    for (int g = 0; g < kMaskGroup; g++) {
      const int ty = blockIdx.y * kMaskGroup + g;
      const bool dirty = ty < tiles_y && fl[g];
      any = any || dirty;
      if (g == w) clean = !dirty;
    }
    if (!any) return;
  }
  const int tyy = blockIdx.y * kMaskGroup + w;
  unsigned rows_m = 0, comp = 0;
  if (tyy < tiles_y)
    rows_m = mask_tile(P, D, blockIdx.x, tyy, tiles_x, solve_all, clean, &comp);
  const int j = blockIdx.x * kStripW - 2 + (int)(threadIdx.x & 63);
  append_group<kMaskGroup>(own_sink(D), tyy * tiles_x + blockIdx.x, rows_m, comp,
                           (int)cidx(P, max(j, 0), tyy * kTileH), (int)P.pitch,
                           (long long)P.nx * P.ny);
}

// ---------------------------------------------------------------------------
// reconstruction of one cell in one direction (src/solver_2d.f90:3152-3661)
// ---------------------------------------------------------------------------
struct SrcInfo {
  bool src2;       // source_cell == 2 (ring) and RADIAL_SOURCE_FLAG
  bool is_src;     // source_cell == 1 (source interior)
  bool sE, sW, sS, sN;
  bool dir;        // one of sE, sW, sS, sN
};
__device__ __forceinline__ SrcInfo src_info(int b) {
  SrcInfo S;
  S.is_src = b & SI_SRC;
  S.src2 = b & SI_RING;
  S.sE = b & SI_E;
  S.sW = b & SI_W;
  S.sS = b & SI_S;
  S.sN = b & SI_N;
  S.dir = b & SI_DIR;
  return S;
}

__device__ __forceinline__ void qp2_uv(const double* qp, double& u, double& v) {
  // qp_to_qp2 (src/constitutive_2d.f90:2155-2179), velocity components
  if (qp[0] <= 0.0) {
    u = 0.0;
    v = 0.0;
  } else {
    u = qp[1] / qp[0];
    v = qp[2] / qp[0];
  }
}

// DIR 1: x (lo = west face value qrecW, hi = east qrecE); DIR 2: y (S, N).
// c = own physical state, m / p = neighbours at index-1 / index+1 along DIR
// (not read at a domain boundary), idx / n = position and extent along DIR,
// other = position along the other direction, sb = source boundary state,
// typo = the x_stencil(3) = y_stag(k+1) substitution of src/solver_2d.f90:
// 3616 that reaches the x-slope of v.
template <int NS, int DIR>
__device__ __forceinline__ void recon(const DevParams& P, const double* c,
                                      const double* m, const double* p, int idx,
                                      int n, int other, const SrcInfo S,
                                      const double* sb, bool typo, double* lo,
                                      double* hi) {
  constexpr int NV = 4 + NS;
  const int side_lo = (DIR == 1) ? IMEXSW_W : IMEXSW_S;
  const int side_hi = (DIR == 1) ? IMEXSW_E : IMEXSW_N;
  const double one_by = (DIR == 1) ? P.one_by_dx : P.one_by_dy;
  const double d2 = (DIR == 1) ? P.dx2 : P.dy2;
  const bool first = (idx == 0), last = (idx == n - 1);
  const bool radial2 = P.radial && S.src2;
  // DIR 1 honours the ring at the domain edge (:3185, :3226); DIR 2 does not
  const bool srcb = (DIR == 1) && S.src2;
  const bool ovr_lo = (DIR == 1) ? (radial2 && !S.sE && S.sW) : (radial2 && S.sS);
  const bool ovr_hi = (DIR == 1) ? (radial2 && S.sE) : (radial2 && !S.sS && S.sN);
  auto stag = [&](int i) { return (DIR == 1) ? xstag(P, i) : ystag(P, i); };
  auto comp = [&](int i) { return (DIR == 1) ? xcomp(P, i) : ycomp(P, i); };
  // stencil abscissae left by the loop over the conservative variables
  double z0, z2;
  const double z1 = comp(idx);
  if (first) {
    z0 = stag(0);
    z2 = comp(idx + 1);
  } else if (last) {
    z0 = comp(idx - 1);
    z2 = stag(n);
  } else {
    z0 = ovr_lo ? stag(idx) : comp(idx - 1);
    z2 = ovr_hi ? stag(idx + 1) : comp(idx + 1);
  }
  (void)other;
  // the limiter's operands of every variable, then all their divisions
  // (no branch between them), then the limiter and the boundary cases per
  // variable: the values of limit_ on the same operands
  double v0[NV], v2[NV], A[NV], B[NV], C[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) {
    // the candidates in registers first: a choice between two addresses
    // (of the argument block, or of two of the caller's arrays) would move
    // the arrays to scratch memory
    double blo = P.bc_value[side_lo][i], bhi = P.bc_value[side_hi][i];
    double sbi = sb[i], mi = m[i], pi = p[i];
    pin(blo);
    pin(bhi);
    pin(sbi);
    pin(mi);
    pin(pi);
    if (first) {
      v0[i] = srcb ? sbi : blo;
      v2[i] = pi;
    } else if (last) {
      v0[i] = mi;
      v2[i] = srcb ? sbi : bhi;
    } else {
      v0[i] = ovr_lo ? sbi : mi;
      v2[i] = ovr_hi ? sbi : pi;
    }
  }
#pragma unroll
  for (int i = 0; i < NV; i++) {
    A[i] = (v2[i] - c[i]) / (z2 - z1);
    B[i] = (c[i] - v0[i]) / (z1 - z0);
  }
#pragma unroll
  for (int i = 0; i < NV; i++) C[i] = 0.0;
  if (P.lim_c) {
#pragma unroll
    for (int i = 0; i < NV; i++) C[i] = (v2[i] - v0[i]) / (z2 - z0);
  }
#pragma unroll
  for (int i = 0; i < NV; i++) {
    const double s1 = c[i];
    const double lim_slope = limit_pick(P, A[i], B[i], C[i], P.limiter[i]);
    double slope = lim_slope;
    if (first) {
      const int f = P.bc_flag[side_lo][i];
      const double bv = P.bc_value[side_lo][i];
      if (!srcb && f != 0) slope = (f == 1) ? bv : (f == 2) ? (p[i] - s1) * one_by : 0.0;
    } else if (last) {
      const int f = P.bc_flag[side_hi][i];
      const double bv = P.bc_value[side_hi][i];
      if (!srcb && f != 0) slope = (f == 1) ? bv : (f == 2) ? (s1 - m[i]) * one_by : 0.0;
    }
    const double dq = P.reconstr_coeff * d2 * slope;
    lo[i] = s1 - dq;
    hi[i] = s1 + dq;
    if (first) {
      const int f = P.bc_flag[side_lo][i];
      if (srcb) lo[i] = sb[i];
      else if (f == 0) lo[i] = P.bc_value[side_lo][i];
      else if (i == DIR) lo[i] = fmin(lo[i], 0.0);
    }
    if (last) {
      const int f = P.bc_flag[side_hi][i];
      if (srcb) hi[i] = sb[i];
      else if (f == 0) hi[i] = P.bc_value[side_hi][i];
      else if (i == DIR) hi[i] = fmax(hi[i], 0.0);
    }
  }
  // u and v with the limiters of hu and hv (:3496-3661), from the face
  // values of h, hu, hv just formed when the cell is on the boundary
  double ul = 0.0, wl = 0.0, uh = 0.0, wh = 0.0;
  if (first) qp2_uv(lo, ul, wl);
  if (last) qp2_uv(hi, uh, wh);
  double u0[2], u2v[2], zz2[2], UA[2], UB[2], UC[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int i = NV + r;
    zz2[r] = z2;
    if (DIR == 1 && i == NV + 1 && typo)
      zz2[r] = (!first && !last && radial2 && S.sE) ? stag(idx + 1) : ystag(P, other + 1);
    double sbi = sb[i], mi = m[i], pi = p[i];
    pin(sbi);
    pin(mi);
    pin(pi);
    if (first) {
      u0[r] = (r == 0) ? ul : wl;
      u2v[r] = pi;
    } else if (last) {
      u0[r] = mi;
      u2v[r] = (r == 0) ? uh : wh;
    } else {
      u0[r] = ovr_lo ? sbi : mi;
      u2v[r] = ovr_hi ? sbi : pi;
    }
  }
#pragma unroll
  for (int r = 0; r < 2; r++) {
    UA[r] = (u2v[r] - c[NV + r]) / (zz2[r] - z1);
    UB[r] = (c[NV + r] - u0[r]) / (z1 - z0);
    UC[r] = 0.0;
  }
  if (P.lim_c) {
#pragma unroll
    for (int r = 0; r < 2; r++) UC[r] = (u2v[r] - u0[r]) / (zz2[r] - z0);
  }
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int i = NV + r;
    const double s1 = c[i];
    const double slope = limit_pick(P, UA[r], UB[r], UC[r], P.limiter[i]);
    const double dq = P.reconstr_coeff * d2 * slope;
    lo[i] = s1 - dq;
    hi[i] = s1 + dq;
    if (first) lo[i] = (r == 0) ? ul : wl;
    else if (last) hi[i] = (r == 0) ? uh : wh;
  }
  if (!first && !last) {
    if (ovr_hi) {
#pragma unroll
      for (int i = 0; i < NV + 2; i++) hi[i] = sb[i];
    } else if (ovr_lo) {
#pragma unroll
      for (int i = 0; i < NV + 2; i++) lo[i] = sb[i];
    }
  }
}

// One face side: what the KT flux needs from the state on that side.
template <int NS>
struct Side {
  double f[4 + NS];
  double qc[4 + NS];
  double vmin, vmax, zv;
};

// A face side with the profile closures, kept out of line: the stencil
// march has six call sites, and inlined at all of them the closures' sums
// took the march to 512 registers and 3.5 KB of spills per lane.  The
// side's h0 and settling speeds are shared by its momentum and its flux
// coefficients (the carrier density of both is mixt_var's).
template <int NS, int DIR>
__device__ __forceinline__ void make_side_vp_body(const DevParams& P, const double* qp,
                                                  double gc, double bx, double by,
                                                  Side<NS>* s) {
  if (qp[0] > kEps) {
    const VpPre<NS> pre = vp_pre<NS>(P, qp[0], mixt<NS, false>(P, qp).rho_c);
    qp2qc<NS, true>(P, qp, s->qc, &pre);
    fluxes<NS, DIR, true>(P, s->qc, qp, gc, s->f, bx, by, &pre);
  } else {
    qp2qc<NS, true>(P, qp, s->qc);
    fluxes<NS, DIR, true>(P, s->qc, qp, gc, s->f, bx, by);
  }
  speeds<NS, DIR>(P, qp, gc, s->vmin, s->vmax);
  s->zv = (DIR == 1) ? qp[1] : s->qc[2];
}
template <int NS, int DIR>
__device__ __noinline__ void make_side_vp(const DevParams& P, const double* qp, double gc,
                                          double bx, double by, Side<NS>* s) {
  make_side_vp_body<NS, DIR>(P, qp, gc, bx, by, s);
}

// (VP: bx, by are the slope of the cell the state belongs to, read by the
// profile flux coefficients)
template <int NS, int DIR, bool VP = false>
__device__ __forceinline__ void make_side(const DevParams& P, const double* qp,
                                          double gc, Side<NS>& s, double bx = 0.0,
                                          double by = 0.0) {
  if (VP) {
    make_side_vp<NS, DIR>(P, qp, gc, bx, by, &s);
    return;
  }
  qp2qc<NS>(P, qp, s.qc);
  fluxes<NS, DIR>(P, s.qc, qp, gc, s.f);
  speeds<NS, DIR>(P, qp, gc, s.vmin, s.vmax);
  s.zv = (DIR == 1) ? qp[1] : s.qc[2];
}

template <int NS>
__device__ __forceinline__ void zero_side(Side<NS>& s) {
#pragma unroll
  for (int i = 0; i < 4 + NS; i++) {
    s.f[i] = 0.0;
    s.qc[i] = 0.0;
  }
  s.vmin = s.vmax = s.zv = 0.0;
}

template <int NS>
__device__ __forceinline__ void shfl_side(const Side<NS>& in, Side<NS>& out, int src) {
#pragma unroll
  for (int i = 0; i < 4 + NS; i++) {
    out.f[i] = shfl_d(in.f[i], src);
    out.qc[i] = shfl_d(in.qc[i], src);
  }
  out.vmin = shfl_d(in.vmin, src);
  out.vmax = shfl_d(in.vmax, src);
  out.zv = shfl_d(in.zv, src);
}

template <int NS, int DIR>
__device__ __forceinline__ double face_flux(const Side<NS>& L, const Side<NS>& R,
                                            double* H) {
  const double aN = fmin(fmin(L.vmin, R.vmin), 0.0);
  const double aP = fmax(fmax(L.vmax, R.vmax), 0.0);
  kt_flux<NS, DIR>(L.f, R.f, L.qc, R.qc, aN, aP, L.zv, R.zv, H);
  return fmax(aP, -aN);
}

// Source byte and ring vector of cell (j, k) from a clamped, always valid
// address (zero without a radial source).
__device__ __forceinline__ void load_src(const DevParams& P, const DevPtrs& D, int j, int k,
                                         int& b, double& v0, double& v1) {
  // unconditional loads (from the state planes when there is no source),
  // selected afterwards: a load under a condition is waited for where the
  // branches join, which would serialise it with the row's other loads
  const bool on = P.radial && D.sinfo;
  const uint8_t* bi = on ? D.sinfo : D.sflag;
  const double* vr = on ? D.sring : D.q;
  const size_t c = cidx(P, min(max(j, 0), P.nx - 1), min(max(k, 0), P.ny - 1));
  b = bi[c];
  v0 = vr[c];
  v1 = vr[plane(P) + c];
  b = on ? b : 0;
  v0 = on ? v0 : 0.0;
  v1 = on ? v1 : 0.0;
}

// boundary state of a ring cell (zero elsewhere), src/solver_2d.f90:3879-3889
template <int NS>
__device__ __forceinline__ void ring_state(const DevParams& P, const SrcInfo S,
                                           const SrcTime& T, double v0, double v1,
                                           double* sb) {
  constexpr int NV = 4 + NS;
#pragma unroll
  for (int i = 0; i < NV + 2; i++) sb[i] = 0.0;
  if (!(P.radial && S.src2 && S.dir)) return;
  source_bdry<NS>(P, T, v0, v1, sb);
}

// ---------------------------------------------------------------------------
// k_stencil: the stage-1 stencil pass (timestep + eval_hyperbolic_terms +
// eval_expl_terms at q^n).  One wavefront marches up a strip of rows keeping
// a three-row window of physical states in registers; lane l owns column
// j = 60*tile_x - 2 + l.  Lanes 2..61 are the strip's cells, lanes 1 and
// 62 reconstruct the halo sides of the strip's outer faces, lanes 0 and 63
// only supply states.  x neighbours are exchanged with wave shuffles; rows
// k0-2 .. k1+1 are read once each.
// ---------------------------------------------------------------------------
// Conservative state of cell (j, k) from a clamped, always valid address:
// loads without a branch around them, so that several rows' loads are in
// flight together; finish_row turns them into physical variables (a dry
// state outside the grid).
template <int NS>
__device__ __forceinline__ void load_raw(const DevParams& P, const DevPtrs& D, int j, int k,
                                         double* q) {
  constexpr int NV = 4 + NS;
  const size_t c = cidx(P, min(max(j, 0), P.nx - 1), min(max(k, 0), P.ny - 1)), pl = plane(P);
#pragma unroll
  for (int i = 0; i < NV; i++) q[i] = D.q[i * pl + c];
}
// With the vertical profiles the row's values are the stored physical
// variables of q when there are any (NV+2 planes, kept by every kernel that
// writes q; not for a stage state), so the strip march does not run the
// profile mean-speed iteration again.
template <int NS, bool VP>
__device__ __forceinline__ void load_raw_v(const DevParams& P, const DevPtrs& D, int j, int k,
                                           double* r) {
  constexpr int NV = 4 + NS;
  if (VP && D.qpp) {
    const size_t c = cidx(P, min(max(j, 0), P.nx - 1), min(max(k, 0), P.ny - 1)), pl = plane(P);
#pragma unroll
    for (int i = 0; i < NV + 2; i++) r[i] = D.qpp[i * pl + c];
  } else {
    load_raw<NS>(P, D, j, k, r);
  }
}

// the profile mean-speed iteration of one cell, out of line for the same
// reason as make_side_vp
template <int NS>
__device__ __noinline__ void phys_or_dry_vp(const DevParams& P, const double* q, double* qp) {
  phys_or_dry<NS, true>(P, q, qp);
}

// `stored`: q holds the stored physical variables (load_raw_v)
template <int NS, bool VP = false>
__device__ __forceinline__ void finish_row(const DevParams& P, int j, int k, const double* q,
                                           double* qp, bool stored = false) {
  constexpr int NV = 4 + NS;
  if (j >= 0 && j < P.nx && k >= 0 && k < P.ny) {
    if (VP && stored) {
#pragma unroll
      for (int i = 0; i < NV + 2; i++) qp[i] = q[i];
    } else if (VP) {
      phys_or_dry_vp<NS>(P, q, qp);
    } else {
      phys_or_dry<NS>(P, q, qp);
    }
  } else {
#pragma unroll
    for (int i = 0; i < NV + 2; i++) qp[i] = 0.0;
    qp[3] = P.T_ambient;
  }
}

template <int NS, bool VP = false>
__device__ __forceinline__ void load_row(const DevParams& P, const DevPtrs& D,
                                         int j, int k, double* qp) {
  constexpr int NV = 4 + NS;
  if (j >= 0 && j < P.nx && k >= 0 && k < P.ny) {
    const size_t c = cidx(P, j, k), pl = plane(P);
    if (VP && D.qpp) {
#pragma unroll
      for (int i = 0; i < NV + 2; i++) qp[i] = D.qpp[i * pl + c];
    } else {
      double q[NV];
#pragma unroll
      for (int i = 0; i < NV; i++) q[i] = D.q[i * pl + c];
      if (VP) phys_or_dry_vp<NS>(P, q, qp);
      else phys_or_dry<NS>(P, q, qp);
    }
  } else {
#pragma unroll
    for (int i = 0; i < NV + 2; i++) qp[i] = 0.0;
    qp[3] = P.T_ambient;
  }
}

// the two sides of the y face between rows k-1 (B, owner Sb) and k (T,
// owner St), with the ring overrides of src/solver_2d.f90:3879-3889
// (VP: pb, pt are the slopes (bx, by) of rows k-1 and k; the ring overrides
// keep them by position, which matters only with slope correction, a
// combination the library refuses)
// A face side from one of two value sets (b when use_b), or a zero side
// when `zero`: one make_side on the chosen values, zeroed afterwards.  The
// same side as a make_side call per case; calls per case had left parts of
// the side in scratch memory.
template <int NS, int DIR, bool VP>
__device__ __forceinline__ void side_of(const DevParams& P, const double* a, const double* b,
                                        bool use_b, bool zero, double g, Side<NS>& s,
                                        double bx, double by) {
  constexpr int NV = 4 + NS;
  double q[NV + 2];
#pragma unroll
  for (int i = 0; i < NV + 2; i++) {
    double x = a[i], y = b[i];
    pin(x);
    pin(y);
    q[i] = use_b ? y : x;
  }
  make_side<NS, DIR, VP>(P, q, g, s, bx, by);
#pragma unroll
  for (int i = 0; i < NV; i++) {
    s.f[i] = zero ? 0.0 : s.f[i];
    s.qc[i] = zero ? 0.0 : s.qc[i];
  }
  s.vmin = zero ? 0.0 : s.vmin;
  s.vmax = zero ? 0.0 : s.vmax;
  s.zv = zero ? 0.0 : s.zv;
}

// Bottom side of the y face between rows k-1 (hib, source data Sb) and k
// (lot, St): hib, except next to a source cell (the ring's own value, or
// none).
template <int NS, bool VP = false>
__device__ __forceinline__ void yface_bottom(const DevParams& P, const double* hib,
                                             const double* lot, const SrcInfo Sb,
                                             const SrcInfo St, int k, double g, Side<NS>& Bs,
                                             const double* pb = nullptr) {
  const double bx = VP ? pb[0] : 0.0, by = VP ? pb[1] : 0.0;
  const bool zero = Sb.is_src && !(P.radial && St.src2 && St.sS && k < P.ny - 1);
  side_of<NS, 2, VP>(P, hib, lot, Sb.is_src, zero, g, Bs, bx, by);
}
// Top side of the same face: lot, except next to a source cell.
template <int NS, bool VP = false>
__device__ __forceinline__ void yface_top(const DevParams& P, const double* hib,
                                          const double* lot, const SrcInfo Sb,
                                          const SrcInfo St, int k, double g, Side<NS>& Ts,
                                          const double* pt = nullptr) {
  const double bx = VP ? pt[0] : 0.0, by = VP ? pt[1] : 0.0;
  const bool zero = St.is_src && !(P.radial && Sb.src2 && !Sb.sS && Sb.sN && k - 1 > 0);
  side_of<NS, 2, VP>(P, hib, lot, !St.is_src, zero, g, Ts, bx, by);
}
template <int NS, bool VP = false>
__device__ __forceinline__ void yface_sides(const DevParams& P, const double* hib,
                                            const double* lot, const SrcInfo Sb,
                                            const SrcInfo St, int k, double g,
                                            Side<NS>& Bs, Side<NS>& Ts,
                                            const double* pb = nullptr,
                                            const double* pt = nullptr) {
  yface_bottom<NS, VP>(P, hib, lot, Sb, St, k, g, Bs, pb);
  yface_top<NS, VP>(P, hib, lot, Sb, St, k, g, Ts, pt);
}
// the slope (bx, by) of cell (j, k) from a clamped address
__device__ __forceinline__ void load_bp(const DevParams& P, const DevPtrs& D, int j, int k,
                                        double* b2) {
  const size_t c = cidx(P, min(max(j, 0), P.nx - 1), min(max(k, 0), P.ny - 1));
  b2[0] = D.Bx[c];
  b2[1] = D.By[c];
}

// Per-cell data of row k that the x pass and the explicit terms read, from
// clamped (always valid) addresses; loaded one row ahead of its use.
struct RowData {
  int m;            // mask byte
  double gW, gE;    // slope corrections of the W and E faces
  double bx, by, bxx, bxy, byy, gc, dg;
};
__device__ __forceinline__ void load_rowdata(const DevParams& P, const DevPtrs& D, int j, int k,
                                             RowData& R) {
  const int jc = min(max(j, 0), P.nx - 1), kc = min(max(k, 0), P.ny - 1);
  const size_t c = cidx(P, jc, kc);
  // the optional fields from the state planes when absent (NV planes cover
  // every index here), selected afterwards, so that all loads are in flight
  // together (see load_src)
  const double* gsx = D.gsx ? D.gsx : D.q;
  const double* bxx = D.Bxx ? D.Bxx : D.q;
  const double* bxy = D.Bxy ? D.Bxy : D.q;
  const double* byy = D.Byy ? D.Byy : D.q;
  const double* gco = D.gcoef ? D.gcoef : D.q;
  const double* dgd = D.dgdx ? D.dgdx : D.q;
  R.m = D.mask[c];
  R.gW = gsx[(size_t)kc * (P.pitch + 1) + jc];
  R.gE = gsx[(size_t)kc * (P.pitch + 1) + jc + 1];
  R.bx = D.Bx[c];
  R.by = D.By[c];
  R.bxx = bxx[c];
  R.bxy = bxy[c];
  R.byy = byy[c];
  R.gc = gco[c];
  R.dg = dgd[c];
  R.gW = D.gsx ? R.gW : 1.0;
  R.gE = D.gsx ? R.gE : 1.0;
  R.bxx = D.Bxx ? R.bxx : 0.0;
  R.bxy = D.Bxy ? R.bxy : 0.0;
  R.byy = D.Byy ? R.byy : 0.0;
  R.gc = D.gcoef ? R.gc : 1.0;
  R.dg = D.dgdx ? R.dg : 0.0;
}
// y-face slope correction at the S face of row k (k = ny: the N boundary)
__device__ __forceinline__ double load_gsy(const DevParams& P, const DevPtrs& D, int j, int k) {
  const double* g = D.gsy ? D.gsy : D.q;  // (see load_rowdata)
  const double v = g[cidx(P, min(max(j, 0), P.nx - 1), min(max(k, 0), P.ny))];
  return D.gsy ? v : 1.0;
}

template <int NS>
__device__ __forceinline__ void x_face_fluxes(const DevParams& P, const SrcInfo S, int j,
                                              const Side<NS>& Wself, const Side<NS>& Eself,
                                              const Side<NS>& Eleft, double* HW, double* HE,
                                              double& amaxW, double& amaxE);

// x direction of row k: the W face flux on every lane, the E face by
// shuffle from the right neighbour (a boundary face on the last column).
// S is row k's source data, sb its ring state, qk its physical state.
template <int NS, bool HASY, bool VP = false>
__device__ __forceinline__ void x_fluxes(const DevParams& P, const SrcInfo S, const double* sb,
                                         const double* qk, int j, int k, const RowData& R,
                                         double* HW, double* HE, double& amaxW,
                                         double& amaxE) {
  constexpr int NV = 4 + NS;
  const int lane = threadIdx.x & 63;
  const bool injx = (j >= 0 && j < P.nx);
  double qkm[NV + 2], qkp[NV + 2];
#pragma unroll
  for (int i = 0; i < NV + 2; i++) {
    qkm[i] = shfl_d(qk[i], (lane + 63) & 63);
    qkp[i] = shfl_d(qk[i], (lane + 1) & 63);
  }
  // after a north ring cell the x-slope of v uses y_stag(k+1) as its
  // third abscissa (src/solver_2d.f90:3616)
  const bool typo = HASY && injx && P.radial && S.src2 && k > 0 && k < P.ny - 1 &&
                    !S.sS && S.sN;
  Side<NS> Wself, Eself;
  if (injx) {
    double lo[NV + 2], hi[NV + 2];
    recon<NS, 1>(P, qk, qkm, qkp, j, P.nx, k, S, sb, typo, lo, hi);
    if (S.is_src) {
      zero_side<NS>(Wself);
      zero_side<NS>(Eself);
    } else {
      make_side<NS, 1, VP>(P, lo, R.gW, Wself, R.bx, R.by);
      make_side<NS, 1, VP>(P, hi, R.gE, Eself, R.bx, R.by);
    }
  } else {
    zero_side<NS>(Wself);
    zero_side<NS>(Eself);
  }
  Side<NS> Eleft;
  shfl_side<NS>(Eself, Eleft, (lane + 63) & 63);
  x_face_fluxes<NS>(P, S, j, Wself, Eself, Eleft, HW, HE, amaxW, amaxE);
}

// The x face fluxes of row k from its W and E sides (Eleft: the E side of
// the left neighbour): the W face on every lane, the E face by shuffle from
// the right neighbour (a boundary face on the last column).
template <int NS>
__device__ __forceinline__ void x_face_fluxes(const DevParams& P, const SrcInfo S, int j,
                                              const Side<NS>& Wself, const Side<NS>& Eself,
                                              const Side<NS>& Eleft, double* HW, double* HE,
                                              double& amaxW, double& amaxE) {
  constexpr int NV = 4 + NS;
  const int lane = threadIdx.x & 63;
  const int flags = (S.is_src ? 1 : 0) | (S.src2 ? 2 : 0) | (S.sE ? 4 : 0) | (S.sW ? 8 : 0);
  const int flags_left = __shfl(flags, (lane + 63) & 63, kWave);
  {
    Side<NS> Ls, Rs;
    if (j == 0) {
      Ls = Wself;
      Rs = Wself;
    } else {
      const bool interior_me = (j > 0 && j < P.nx - 1);
      const bool interior_left = (j - 1 > 0 && j - 1 < P.nx - 1);
      if (flags_left & 1) {
        if (P.radial && S.src2 && !S.sE && S.sW && interior_me) Ls = Wself;
        else zero_side<NS>(Ls);
      } else {
        Ls = Eleft;
      }
      if (S.is_src) {
        if (P.radial && (flags_left & 2) && (flags_left & 4) && interior_left) Rs = Eleft;
        else zero_side<NS>(Rs);
      } else {
        Rs = Wself;
      }
    }
    amaxW = face_flux<NS, 1>(Ls, Rs, HW);
  }
  {
    double HEs[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) HEs[i] = shfl_d(HW[i], (lane + 1) & 63);
    double amaxEs = shfl_d(amaxW, (lane + 1) & 63);
    if (j == P.nx - 1) {
      amaxE = face_flux<NS, 1>(Eself, Eself, HE);
    } else {
#pragma unroll
      for (int i = 0; i < NV; i++) HE[i] = HEs[i];
      amaxE = amaxEs;
    }
  }
}

// divergence, explicit terms and the residual D of one computed cell, and
// its CFL time step folded into dtmin
template <int NS, bool HASY>
__device__ __forceinline__ void cell_update(const DevParams& P, const DevPtrs& D, size_t c,
                                            const double* qk, const RowData& R,
                                            const double* HW, const double* HE,
                                            const double* HS, const double* HN, double amaxW,
                                            double amaxE, double amaxS, double amaxN,
                                            int compute_dt, double& dtmin,
                                            const double* ex_in = nullptr) {
  constexpr int NV = 4 + NS;
  const size_t pl = plane(P);
  double ex[NV];
  if (ex_in) {
#pragma unroll
    for (int i = 0; i < NV; i++) ex[i] = ex_in[i];
  } else {
    expl_terms<NS>(P, R.bx, R.by, R.bxx, R.bxy, R.byy, R.gc, R.dg, qk, ex);
  }
#pragma unroll
  for (int i = 0; i < NV; i++) {
    double dv = 0.0;
    dv = dv + (HE[i] - HW[i]) * P.one_by_dx;
    if (HASY) dv = dv + (HN[i] - HS[i]) * P.one_by_dy;
    D.D[i * pl + c] = (0.0 + (dv - ex[i]) * 1.0) + (0.0 - 0.0) * 0.0;
  }
  if (compute_dt) {
    double ma = fmax(amaxW, amaxE);
    const double dtx = (ma > 0.0) ? P.cfl * P.dx / ma : P.max_dt;
    double dty = P.max_dt;
    if (HASY) {
      ma = fmax(amaxS, amaxN);
      dty = (ma > 0.0) ? P.cfl * P.dy / ma : P.max_dt;
    }
    dtmin = fmin(dtmin, fmin(dtx, dty));
  }
}

// wave minimum of the CFL time step, one atomic per wave (dt > 0: IEEE
// order == integer order)
__device__ __forceinline__ void dt_reduce(const DevPtrs& D, const DevParams& P, double dtmin) {
  const int lane = threadIdx.x & 63;
  for (int off = 32; off > 0; off >>= 1) dtmin = fmin(dtmin, shfl_d(dtmin, (lane + off) & 63));
  if (lane == 0 && dtmin < P.max_dt)
    atomicMin(D.dt_bits, (unsigned long long)__double_as_longlong(dtmin));
}

template <int NS, bool HASY, bool VP = false>
__device__ __forceinline__ void stencil_strip(const DevParams& P, const DevPtrs& D,
                                              const SrcTime& T, int j, int k0, int k1,
                                              bool skip_empty, int compute_dt);

// Height of this step's stencil strips: strip_h = unit * 2^m <= max_strip
// (a divisor of the tile height), the tallest that still leaves about
// `target_strips` strips for the step's computed cells (short strips when
// few cells are computed, so the sequential march stays short; tall strips
// on a wet grid, so the two halo rows of each strip are amortised).  w_lane
// is this lane's slot of the computed-cell counter.
__device__ __forceinline__ int strip_height(const DevParams& P, unsigned long long w_lane,
                                            int unit, int max_strip, int target_strips) {
  int strip_h = unit;
  if (P.ny > 1) {
    unsigned long long w = w_lane;
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
    while (2 * strip_h <= max_strip &&
           w / (unsigned long long)(kStripW * 2 * strip_h) >= (unsigned long long)target_strips)
      strip_h *= 2;
  }
  return strip_h;
}

// The stencil pass over the listed work tiles, cut into strips of strip_h
// rows; workgroups take (tile, strip) items in turn until the list is
// exhausted.  With one-row strips every item is a listed row.
template <int NS, bool HASY, bool VP = false>
__global__ void __launch_bounds__(64) k_stencil(DevParams P, DevPtrs D, int unit,
                                                int max_strip, int target_strips,
                                                int compute_dt) {
  const int lane = threadIdx.x;
  // the step's scalars and list sizes in one round of loads
  double done = D.scal[S_DONE], t = D.scal[S_T];
  unsigned long long w = *ctr(D, C_COMPUTE_CELLS, lane);
  int ntiles = (int)*ctr(D, C_TILES, 0), nrows = (int)*ctr(D, C_ROWS, 0);
  pin(done);
  pin(t);
  pin(w);
  pin(ntiles);
  pin(nrows);
  if (done != 0.0) return;
  const int strip_h = strip_height(P, w, unit, max_strip, target_strips);
  // the ring cells' time factors are the same for the whole step
  const SrcTime T = P.radial ? src_time(P, t) : SrcTime{};
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  if (strip_h == 1) {
    for (int item = blockIdx.x; item < nrows; item += gridDim.x) {
      const int row = D.row_list[item];
      const int tile = row / kTileH;
      const int k0 = (tile / tiles_x) * kTileH + row % kTileH;
      stencil_strip<NS, HASY, VP>(P, D, T, (tile % tiles_x) * kStripW - 2 + lane, k0, k0 + 1,
                                  false, compute_dt);
    }
    return;
  }
  const int per = kTileH / strip_h;
  for (int item = blockIdx.x; item < ntiles * per; item += gridDim.x) {
    const int tile = D.tile_list[item / per];
    const int tx = tile % tiles_x, tyy = tile / tiles_x;
    const int k0 = tyy * kTileH + (item % per) * strip_h;
    if (k0 >= P.ny) continue;
    const int k1 = min(k0 + strip_h, P.ny);
    stencil_strip<NS, HASY, VP>(P, D, T, tx * kStripW - 2 + lane, k0, k1, true, compute_dt);
  }
}

// One strip: lane l owns column j = 60 * tile_x - 2 + l, rows k0 .. k1-1.
// Every load of a row is issued one row ahead of its use (the raw states
// three rows ahead), from clamped addresses and without a branch around it,
// so that a strip pays the memory latency about once.
template <int NS, bool HASY, bool VP>
__device__ __forceinline__ void stencil_strip(const DevParams& P, const DevPtrs& D,
                                              const SrcTime& T, int j, int k0, int k1,
                                              bool skip_empty, int compute_dt) {
  constexpr int NV = 4 + NS;
  const int lane = threadIdx.x;
  const bool injx = (j >= 0 && j < P.nx);
  const bool own = (lane >= 2 && lane < 2 + kStripW && injx);

  // skip strips without a cell to compute
  if (skip_empty) {
    int any = 0;
    if (own)
      for (int k = k0; k < k1; k++) any |= D.mask[cidx(P, j, k)] & M_COMPUTE;
    if (!__any(any)) return;
  }
  const size_t pl = plane(P);
  double dtmin = P.max_dt;
  const bool stored = VP && D.qpp;

  double qm[NV + 2], qc[NV + 2], qn[NV + 2];  // rows k-1, k, k+1
  double hiy[NV + 2];                          // N value of row k
  double HS[NV];                               // y flux at the S face of row k
  double amaxS = 0.0;
  double rnext[NV + 2];  // raw values of row k+2, loaded one iteration ahead
  // source data: row k (S, v0, v1) and row k+1 (raw byte bn, vectors), the
  // slope correction of the S face of row k+1, and row k's per-cell data
  int bc = 0, bn = 0;
  double vc0 = 0.0, vc1 = 0.0, vn0 = 0.0, vn1 = 0.0, gn = 1.0;
  RowData R;
  SrcInfo S{};
  if (HASY) {
    double qmm[NV + 2], sb[NV + 2], lo[NV + 2];
    int bm;
    double vm0, vm1, g0 = 1.0;
    {
      double r0[NV + 2], r1[NV + 2], r2[NV + 2], r3[NV + 2];
      load_raw_v<NS, VP>(P, D, j, k0 - 2, r0);
      load_raw_v<NS, VP>(P, D, j, k0 - 1, r1);
      load_raw_v<NS, VP>(P, D, j, k0, r2);
      load_raw_v<NS, VP>(P, D, j, k0 + 1, r3);
      load_raw_v<NS, VP>(P, D, j, k0 + 2, rnext);
      load_src(P, D, j, k0 - 1, bm, vm0, vm1);
      load_src(P, D, j, k0, bc, vc0, vc1);
      load_src(P, D, j, k0 + 1, bn, vn0, vn1);
      g0 = load_gsy(P, D, j, k0);
      gn = load_gsy(P, D, j, k0 + 1);
      load_rowdata(P, D, j, k0, R);
      finish_row<NS, VP>(P, j, k0 - 2, r0, qmm, stored);
      finish_row<NS, VP>(P, j, k0 - 1, r1, qm, stored);
      finish_row<NS, VP>(P, j, k0, r2, qc, stored);
      finish_row<NS, VP>(P, j, k0 + 1, r3, qn, stored);
    }
    if (injx) S = src_info(bc);
    Side<NS> Bs, Ts;
    if (injx) {
      ring_state<NS>(P, S, T, vc0, vc1, sb);
      recon<NS, 2>(P, qc, qm, qn, k0, P.ny, j, S, sb, false, lo, hiy);
      const double g = g0;
      const double pk[2] = {R.bx, R.by};
      if (k0 == 0) {
        // boundary face: both sides carry the S value of row 0
        if (S.is_src) zero_side<NS>(Ts); else make_side<NS, 2, VP>(P, lo, g, Ts, pk[0], pk[1]);
        Bs = Ts;
      } else {
        const SrcInfo Sm = src_info(bm);
        double sbm[NV + 2], lom[NV + 2], him[NV + 2], pm[2] = {0.0, 0.0};
        if (VP) load_bp(P, D, j, k0 - 1, pm);
        ring_state<NS>(P, Sm, T, vm0, vm1, sbm);
        recon<NS, 2>(P, qm, qmm, qc, k0 - 1, P.ny, j, Sm, sbm, false, lom, him);
        yface_sides<NS, VP>(P, him, lo, Sm, S, k0, g, Bs, Ts, pm, pk);
      }
    } else {
      zero_side<NS>(Bs);
      zero_side<NS>(Ts);
#pragma unroll
      for (int i = 0; i < NV + 2; i++) hiy[i] = 0.0;
    }
    amaxS = face_flux<NS, 2>(Bs, Ts, HS);
  } else {
    load_row<NS, VP>(P, D, j, k0, qc);
    load_src(P, D, j, k0, bc, vc0, vc1);
    load_rowdata(P, D, j, k0, R);
    if (injx) S = src_info(bc);
  }

  for (int k = k0; k < k1; k++) {
    const size_t c = injx ? cidx(P, j, k) : 0;
    const bool more = k + 1 < k1;
    // next row's per-cell data
    RowData Rn;
    if (more) load_rowdata(P, D, j, k + 1, Rn);
    double sb[NV + 2];
    if (injx) ring_state<NS>(P, S, T, vc0, vc1, sb);
    else {
#pragma unroll
      for (int i = 0; i < NV + 2; i++) sb[i] = 0.0;
    }
    // ---------------- y: N face of row k ----------------
    double HN[NV];
    double amaxN = 0.0;
    SrcInfo Sn{};
    int bnn = 0;
    double vnn0 = 0.0, vnn1 = 0.0, gnn = 1.0;
    if (HASY) {
      double qnn[NV + 2], hin[NV + 2];
      finish_row<NS, VP>(P, j, k + 2, rnext, qnn, stored);
      if (more) {
        load_raw_v<NS, VP>(P, D, j, k + 3, rnext);  // next iteration's rows
        load_src(P, D, j, k + 2, bnn, vnn0, vnn1);
        gnn = load_gsy(P, D, j, k + 2);
      }
      Side<NS> Bs, Ts;
      const double pk[2] = {R.bx, R.by};
      if (injx) {
        if (k == P.ny - 1) {
          const double g = gn;  // S face of row ny: the N boundary
          if (S.is_src) zero_side<NS>(Bs); else make_side<NS, 2, VP>(P, hiy, g, Bs, pk[0], pk[1]);
          Ts = Bs;
#pragma unroll
          for (int i = 0; i < NV + 2; i++) hin[i] = 0.0;
        } else {
          Sn = src_info(bn);
          double sbn[NV + 2], lon[NV + 2];
          ring_state<NS>(P, Sn, T, vn0, vn1, sbn);
          recon<NS, 2>(P, qn, qc, qnn, k + 1, P.ny, j, Sn, sbn, false, lon, hin);
          const double g = gn;
          double pn[2] = {0.0, 0.0};
          if (VP) load_bp(P, D, j, k + 1, pn);
          yface_sides<NS, VP>(P, hiy, lon, S, Sn, k + 1, g, Bs, Ts, pk, pn);
        }
      } else {
        zero_side<NS>(Bs);
        zero_side<NS>(Ts);
#pragma unroll
        for (int i = 0; i < NV + 2; i++) hin[i] = 0.0;
      }
      amaxN = face_flux<NS, 2>(Bs, Ts, HN);
#pragma unroll
      for (int i = 0; i < NV + 2; i++) {
        qm[i] = qc[i];
        qc[i] = qn[i];
        qn[i] = qnn[i];
        hiy[i] = hin[i];
      }
    }
    // physical state of row k: qm after the shift (2-D) or qc (1-D)
    const double* qk = HASY ? qm : qc;
    double HW[NV], HE[NV], amaxW, amaxE;
    x_fluxes<NS, HASY, VP>(P, S, sb, qk, j, k, R, HW, HE, amaxW, amaxE);
    if (own && (R.m & M_COMPUTE))
      cell_update<NS, HASY>(P, D, c, qk, R, HW, HE, HS, HN, amaxW, amaxE, amaxS, amaxN,
                            compute_dt, dtmin);
    if (more) R = Rn;
    if (HASY) {
#pragma unroll
      for (int i = 0; i < NV; i++) HS[i] = HN[i];
      amaxS = amaxN;
      S = Sn;
      vc0 = vn0;
      vc1 = vn1;
      bn = bnn;
      vn0 = vnn0;
      vn1 = vnn1;
      gn = gnn;
    }
  }
  if (compute_dt) dt_reduce(D, P, dtmin);
}

// ---------------------------------------------------------------------------
// The 2-D stencil pass with the vertical profiles, in two kernels.  Every
// face side the strip march forms is one of the four sides of a cell (its W
// and E values of the x reconstruction, its S and N values of the y
// reconstruction, with the slope correction of that face and the cell's own
// bed slope): k_vp_sides forms them for every cell of the listed tiles and
// a one-cell ring around them, with the profile closures inlined and no
// march around them, and the strip march (stencil_strip_vp) reads them.
// The ring overrides pick the other cell's side of the face, which is the
// side formed from the same state (the slope the closures read differs only
// with slope correction, a combination refused with a radial source).
// ---------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ void side_put(const DevParams& P, const DevPtrs& D, size_t c, int s,
                                         const Side<NS>& v) {
  constexpr int NV = 4 + NS, NSD = 2 * NV + 3;
  const size_t pl = plane(P);
  double* o = D.sides + (size_t)s * NSD * pl + c;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    o[i * pl] = v.f[i];
    o[(NV + i) * pl] = v.qc[i];
  }
  o[(2 * NV) * pl] = v.vmin;
  o[(2 * NV + 1) * pl] = v.vmax;
  o[(2 * NV + 2) * pl] = v.zv;
}
template <int NS>
__device__ __forceinline__ void side_get(const DevParams& P, const DevPtrs& D, int j, int k,
                                         int s, Side<NS>& v) {
  constexpr int NV = 4 + NS, NSD = 2 * NV + 3;
  const size_t pl = plane(P);
  const size_t c = cidx(P, min(max(j, 0), P.nx - 1), min(max(k, 0), P.ny - 1));
  const double* o = D.sides + (size_t)s * NSD * pl + c;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    v.f[i] = o[i * pl];
    v.qc[i] = o[(NV + i) * pl];
  }
  v.vmin = o[(2 * NV) * pl];
  v.vmax = o[(2 * NV + 1) * pl];
  v.zv = o[(2 * NV + 2) * pl];
}
enum SideSlot { SD_W = 0, SD_E = 1, SD_S = 2, SD_N = 3 };

// physical state of cell (j, k) from the stored planes, dry outside the grid
template <int NS>
__device__ __forceinline__ void qpp_get(const DevParams& P, const DevPtrs& D, int j, int k,
                                        double* qp) {
  constexpr int NV = 4 + NS;
  if (j >= 0 && j < P.nx && k >= 0 && k < P.ny) {
    const size_t c = cidx(P, j, k), pl = plane(P);
#pragma unroll
    for (int i = 0; i < NV + 2; i++) qp[i] = D.qpp[i * pl + c];
  } else {
#pragma unroll
    for (int i = 0; i < NV + 2; i++) qp[i] = 0.0;
    qp[3] = P.T_ambient;
  }
}

// the four sides of cell (j, k), with the operands the strip march gives
// make_side (zero sides for a source-interior cell, whose own sides no face
// uses)
template <int NS>
__device__ __forceinline__ void vp_cell_sides(const DevParams& P, const DevPtrs& D,
                                              const SrcTime& T, int j, int k) {
  constexpr int NV = 4 + NS;
  const size_t c = cidx(P, j, k);
  double qk[NV + 2], qkm[NV + 2], qkp[NV + 2], qm[NV + 2], qn[NV + 2];
  qpp_get<NS>(P, D, j, k, qk);
  qpp_get<NS>(P, D, j - 1, k, qkm);
  qpp_get<NS>(P, D, j + 1, k, qkp);
  qpp_get<NS>(P, D, j, k - 1, qm);
  qpp_get<NS>(P, D, j, k + 1, qn);
  int bc;
  double v0, v1;
  load_src(P, D, j, k, bc, v0, v1);
  RowData R;
  load_rowdata(P, D, j, k, R);
  const double gS = load_gsy(P, D, j, k), gN = load_gsy(P, D, j, k + 1);
  const SrcInfo S = src_info(bc);
  Side<NS> a, b;
  if (S.is_src) {
    zero_side<NS>(a);
#pragma unroll
    for (int s = 0; s < 4; s++) side_put<NS>(P, D, c, s, a);
    return;
  }
  // both reconstructions first: the neighbours' states are then dead while
  // the closures run
  double sb[NV + 2], lox[NV + 2], hix[NV + 2], loy[NV + 2], hiy[NV + 2];
  ring_state<NS>(P, S, T, v0, v1, sb);
  const bool typo = P.radial && S.src2 && k > 0 && k < P.ny - 1 && !S.sS && S.sN;
  recon<NS, 1>(P, qk, qkm, qkp, j, P.nx, k, S, sb, typo, lox, hix);
  recon<NS, 2>(P, qk, qm, qn, k, P.ny, j, S, sb, false, loy, hiy);
  make_side_vp_body<NS, 1>(P, lox, R.gW, R.bx, R.by, &a);
  side_put<NS>(P, D, c, SD_W, a);
  make_side_vp_body<NS, 1>(P, hix, R.gE, R.bx, R.by, &b);
  side_put<NS>(P, D, c, SD_E, b);
  make_side_vp_body<NS, 2>(P, loy, gS, R.bx, R.by, &a);
  side_put<NS>(P, D, c, SD_S, a);
  make_side_vp_body<NS, 2>(P, hiy, gN, R.bx, R.by, &b);
  side_put<NS>(P, D, c, SD_N, b);
}

// one workgroup per listed tile in turn: the tile's cells and the ring of
// cells around it
template <int NS>
__global__ void __launch_bounds__(256, 2) k_vp_sides(DevParams P, DevPtrs D) {
  double done = D.scal[S_DONE], t = D.scal[S_T];
  int ntiles = (int)*ctr(D, C_TILES, 0);
  pin(done);
  pin(t);
  pin(ntiles);
  if (done != 0.0) return;
  const SrcTime T = P.radial ? src_time(P, t) : SrcTime{};
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  constexpr int W = kStripW + 2, H = kTileH + 2;
  for (int item = blockIdx.x; item < ntiles; item += gridDim.x) {
    const int tile = D.tile_list[item];
    const int x0 = (tile % tiles_x) * kStripW - 1, y0 = (tile / tiles_x) * kTileH - 1;
    for (int e = threadIdx.x; e < W * H; e += blockDim.x) {
      const int j = x0 + e % W, k = y0 + e / W;
      if (j >= 0 && j < P.nx && k >= 0 && k < P.ny) vp_cell_sides<NS>(P, D, T, j, k);
    }
  }
}

// the sides of the y face between rows k-1 (B) and k (T) from the stored
// sides: Bn, the N side of the cell below, and Ts, the S side of the cell
// above, with the ring overrides of yface_bottom / yface_top
template <int NS>
__device__ __forceinline__ void vp_yface(const DevParams& P, const Side<NS>& Bn,
                                         const Side<NS>& Ts_, const SrcInfo Sb,
                                         const SrcInfo St, int k, Side<NS>& Bs, Side<NS>& Ts) {
  if (Sb.is_src) {
    if (P.radial && St.src2 && St.sS && k < P.ny - 1) Bs = Ts_;
    else zero_side<NS>(Bs);
  } else {
    Bs = Bn;
  }
  if (St.is_src) {
    if (P.radial && Sb.src2 && !Sb.sS && Sb.sN && k - 1 > 0) Ts = Bn;
    else zero_side<NS>(Ts);
  } else {
    Ts = Ts_;
  }
}

// One strip of the profile stencil from the stored sides: lane l owns
// column j = 60 * tile_x - 2 + l, rows k0 .. k1-1.
template <int NS>
__device__ __forceinline__ void stencil_strip_vp(const DevParams& P, const DevPtrs& D, int j,
                                                 int k0, int k1, bool skip_empty,
                                                 int compute_dt) {
  constexpr int NV = 4 + NS;
  const int lane = threadIdx.x;
  const bool injx = (j >= 0 && j < P.nx);
  const bool own = (lane >= 2 && lane < 2 + kStripW && injx);
  if (skip_empty) {
    int any = 0;
    if (own)
      for (int k = k0; k < k1; k++) any |= D.mask[cidx(P, j, k)] & M_COMPUTE;
    if (!__any(any)) return;
  }
  double dtmin = P.max_dt;
  double HS[NV], amaxS = 0.0;
  int bc = 0;
  double v0, v1;
  load_src(P, D, j, k0, bc, v0, v1);
  SrcInfo S = injx ? src_info(bc) : SrcInfo{};
  {
    Side<NS> Bs, Ts;
    if (!injx) {
      zero_side<NS>(Bs);
      zero_side<NS>(Ts);
    } else if (k0 == 0) {
      // boundary face: both sides carry the S side of row 0
      side_get<NS>(P, D, j, 0, SD_S, Ts);
      Bs = Ts;
    } else {
      int bm;
      double w0, w1;
      load_src(P, D, j, k0 - 1, bm, w0, w1);
      Side<NS> Bn, Tn;
      side_get<NS>(P, D, j, k0 - 1, SD_N, Bn);
      side_get<NS>(P, D, j, k0, SD_S, Tn);
      vp_yface<NS>(P, Bn, Tn, src_info(bm), S, k0, Bs, Ts);
    }
    amaxS = face_flux<NS, 2>(Bs, Ts, HS);
  }
  for (int k = k0; k < k1; k++) {
    const size_t c = injx ? cidx(P, j, k) : 0;
    RowData R;
    load_rowdata(P, D, j, k, R);
    double qk[NV + 2];
    qpp_get<NS>(P, D, j, k, qk);
    int bn = 0;
    double n0, n1;
    load_src(P, D, j, k + 1, bn, n0, n1);
    const SrcInfo Sn = (injx && k + 1 < P.ny) ? src_info(bn) : SrcInfo{};
    // N face of row k
    double HN[NV], amaxN;
    {
      Side<NS> Bs, Ts;
      if (!injx) {
        zero_side<NS>(Bs);
        zero_side<NS>(Ts);
      } else if (k == P.ny - 1) {
        side_get<NS>(P, D, j, k, SD_N, Bs);
        Ts = Bs;
      } else {
        Side<NS> Bn, Tn;
        side_get<NS>(P, D, j, k, SD_N, Bn);
        side_get<NS>(P, D, j, k + 1, SD_S, Tn);
        vp_yface<NS>(P, Bn, Tn, S, Sn, k + 1, Bs, Ts);
      }
      amaxN = face_flux<NS, 2>(Bs, Ts, HN);
    }
    // x faces of row k
    Side<NS> Wself, Eself, Eleft;
    if (injx) {
      side_get<NS>(P, D, j, k, SD_W, Wself);
      side_get<NS>(P, D, j, k, SD_E, Eself);
    } else {
      zero_side<NS>(Wself);
      zero_side<NS>(Eself);
    }
    shfl_side<NS>(Eself, Eleft, (lane + 63) & 63);
    double HW[NV], HE[NV], amaxW, amaxE;
    x_face_fluxes<NS>(P, S, j, Wself, Eself, Eleft, HW, HE, amaxW, amaxE);
    if (own && (R.m & M_COMPUTE))
      cell_update<NS, true>(P, D, c, qk, R, HW, HE, HS, HN, amaxW, amaxE, amaxS, amaxN,
                            compute_dt, dtmin);
#pragma unroll
    for (int i = 0; i < NV; i++) HS[i] = HN[i];
    amaxS = amaxN;
    S = Sn;
  }
  if (compute_dt) dt_reduce(D, P, dtmin);
}

// the strip items of k_stencil over the listed tiles, with the stored sides
template <int NS>
__global__ void __launch_bounds__(64) k_stencil_vp(DevParams P, DevPtrs D, int unit,
                                                   int max_strip, int target_strips,
                                                   int compute_dt) {
  const int lane = threadIdx.x;
  double done = D.scal[S_DONE];
  unsigned long long w = *ctr(D, C_COMPUTE_CELLS, lane);
  int ntiles = (int)*ctr(D, C_TILES, 0);
  pin(done);
  pin(w);
  pin(ntiles);
  if (done != 0.0) return;
  const int strip_h = strip_height(P, w, unit, max_strip, target_strips);
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int per = kTileH / strip_h;
  for (int item = blockIdx.x; item < ntiles * per; item += gridDim.x) {
    const int tile = D.tile_list[item / per];
    const int tx = tile % tiles_x, tyy = tile / tiles_x;
    const int k0 = tyy * kTileH + (item % per) * strip_h;
    if (k0 >= P.ny) continue;
    const int k1 = min(k0 + strip_h, P.ny);
    stencil_strip_vp<NS>(P, D, tx * kStripW - 2 + lane, k0, k1, true, compute_dt);
  }
}

// ---------------------------------------------------------------------------
// k_stencil_rows: the stencil pass over the listed rows when few cells are
// computed (the pass is then bound by the latency of one row, not by
// bandwidth).  Six wavefronts share a listed row k and hand their results on
// through LDS in four phases:
//   A  y reconstructions of rows k-1 (w0), k (w1), k+1 (w2), the x
//      reconstruction of row k (w3), row k's explicit terms (w4)
//   B  the sides of the S face (w0 bottom, w1 top), of the N face (w2, w3)
//      and of the x faces (w4 W, w5 E)
//   C  the S (w0), N (w2) and x (w4) face fluxes
//   D  the cell updates and the CFL time step (w4)
// Each face is computed from the same operands, in the same order, as in
// the strip march.
// ---------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ void side_store(double (*buf)[64], const Side<NS>& s, int lane) {
#pragma unroll
  for (int i = 0; i < 4 + NS; i++) {
    buf[i][lane] = s.f[i];
    buf[4 + NS + i][lane] = s.qc[i];
  }
  buf[8 + 2 * NS][lane] = s.vmin;
  buf[9 + 2 * NS][lane] = s.vmax;
  buf[10 + 2 * NS][lane] = s.zv;
}
template <int NS>
__device__ __forceinline__ void side_load(double (*buf)[64], Side<NS>& s, int lane) {
#pragma unroll
  for (int i = 0; i < 4 + NS; i++) {
    s.f[i] = buf[i][lane];
    s.qc[i] = buf[4 + NS + i][lane];
  }
  s.vmin = buf[8 + 2 * NS][lane];
  s.vmax = buf[9 + 2 * NS][lane];
  s.zv = buf[10 + 2 * NS][lane];
}

// One listed row of k_stencil_rows (all six wavefronts of the workgroup).
template <int NS>
__device__ __forceinline__ void stencil_row(const DevParams& P, const DevPtrs& D, int row,
                                            int tiles_x, double t, int compute_dt,
                                            double& dtmin, double (*V)[NS + 6][64],
                                            double (*SD)[2 * NS + 11][64]) {
  constexpr int NV = 4 + NS;
  constexpr int NQ = NV + 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tile = row / kTileH;
  const int k = (tile / tiles_x) * kTileH + row % kTileH;
  const int j = (tile % tiles_x) * kStripW - 2 + lane;
  const bool injx = (j >= 0 && j < P.nx);
  const bool own = (lane >= 2 && lane < 2 + kStripW && injx);
  // row k's source data, the neighbour rows' (k-1 for the S face, k+1
  // for the N face), and row k's per-cell data
  int bc, bo;
  double vc0, vc1, vo0, vo1;
  load_src(P, D, j, k, bc, vc0, vc1);
  load_src(P, D, j, (wv < 2) ? k - 1 : k + 1, bo, vo0, vo1);
  RowData R;
  load_rowdata(P, D, j, k, R);
  const double g = load_gsy(P, D, j, (wv < 2) ? k : k + 1);
  // the raw states this wavefront reads, issued with the loads above
  // (rows kr-1, kr, kr+1; the waves of the x pass use row k alone) and
  // held in registers, so that the row waits for memory once
  const int kr = (wv <= 2) ? k - 1 + wv : k;
  double r0[NV], r1[NV], r2[NV];
  load_raw<NS>(P, D, j, kr - 1, r0);
  load_raw<NS>(P, D, j, kr, r1);
  load_raw<NS>(P, D, j, kr + 1, r2);
  // the source's time factors, evaluated while those loads are in flight
  // (t passes through a register pin, so that the evaluation stays behind
  // the loads)
  double tl = t;
  pin(tl);
  const SrcTime T = P.radial ? src_time(P, tl) : SrcTime{};
#pragma unroll
  for (int i = 0; i < NV; i++) {
    pin(r0[i]);
    pin(r1[i]);
    pin(r2[i]);
  }
  const SrcInfo S = injx ? src_info(bc) : SrcInfo{};
  const SrcInfo So = injx ? src_info(bo) : SrcInfo{};
  double qk[NV + 2], ex[NV];
  // ---- phase A
  if (wv <= 2) {
    // y reconstruction of row kr = k - 1 + wv from rows kr-1, kr, kr+1
    double qa[NV + 2], qb[NV + 2], qc[NV + 2];
    finish_row<NS>(P, j, kr - 1, r0, qa);
    finish_row<NS>(P, j, kr, r1, qb);
    finish_row<NS>(P, j, kr + 1, r2, qc);
    const bool need = injx && (wv == 1 || (wv == 0 && k > 0) || (wv == 2 && k < P.ny - 1));
    double lo[NV + 2], hi[NV + 2];
    if (need) {
      // a copy, not a reference: a reference chosen at run time would put
      // both structures in scratch memory
      const SrcInfo Sr = (wv == 1) ? S : So;
      double sb[NV + 2];
      ring_state<NS>(P, Sr, T, (wv == 1) ? vc0 : vo0, (wv == 1) ? vc1 : vo1, sb);
      recon<NS, 2>(P, qb, qa, qc, kr, P.ny, j, Sr, sb, false, lo, hi);
    } else {
#pragma unroll
      for (int i = 0; i < NV + 2; i++) lo[i] = hi[i] = 0.0;
    }
    if (wv == 0) {
#pragma unroll
      for (int i = 0; i < NQ; i++) V[0][i][lane] = hi[i];  // hib of the S face
    } else if (wv == 1) {
#pragma unroll
      for (int i = 0; i < NQ; i++) {
        V[1][i][lane] = lo[i];  // lot of the S face
        V[2][i][lane] = hi[i];  // hib of the N face
      }
    } else {
#pragma unroll
      for (int i = 0; i < NQ; i++) V[3][i][lane] = lo[i];  // lot of the N face
    }
  } else {
    finish_row<NS>(P, j, k, r1, qk);
    if (wv == 3) {
      double sb[NV + 2], qkm[NV + 2], qkp[NV + 2], lo[NV + 2], hi[NV + 2];
#pragma unroll
      for (int i = 0; i < NV + 2; i++) {
        qkm[i] = shfl_d(qk[i], (lane + 63) & 63);
        qkp[i] = shfl_d(qk[i], (lane + 1) & 63);
      }
      const bool typo = injx && P.radial && S.src2 && k > 0 && k < P.ny - 1 && !S.sS && S.sN;
      if (injx) {
        ring_state<NS>(P, S, T, vc0, vc1, sb);
        recon<NS, 1>(P, qk, qkm, qkp, j, P.nx, k, S, sb, typo, lo, hi);
      } else {
#pragma unroll
        for (int i = 0; i < NV + 2; i++) lo[i] = hi[i] = 0.0;
      }
#pragma unroll
      for (int i = 0; i < NQ; i++) {
        V[4][i][lane] = lo[i];
        V[5][i][lane] = hi[i];
      }
    } else if (wv == 4 && own && (R.m & M_COMPUTE)) {
      expl_terms<NS>(P, R.bx, R.by, R.bxx, R.bxy, R.byy, R.gc, R.dg, qk, ex);
    }
  }
  __syncthreads();
  // ---- phase B
  Side<NS> Wself;
  {
    double va[NQ], vb[NQ];
    const int ia = (wv < 2) ? 0 : (wv < 4) ? 2 : 4 + (wv - 4);
#pragma unroll
    for (int i = 0; i < NQ; i++) va[i] = V[ia][i][lane];
    if (wv < 4) {
#pragma unroll
      for (int i = 0; i < NQ; i++) vb[i] = V[ia + 1][i][lane];
    }
    Side<NS> sd;
    if (wv < 4) {
      // y faces: S face of row k (w0 bottom = row k-1, w1 top = row k) and
      // N face (w2 bottom = row k, w3 top = row k+1).  The cases of
      // yface_bottom / yface_top and the domain boundary only choose the
      // input (the bottom row's value hib = va or the top row's lot = vb)
      // or a zero side, so one make_side on the chosen values, zeroed
      // afterwards where the case says so, gives the same side; distinct
      // calls per case had left parts of the side in scratch memory.
      const bool sface = wv < 2, bottom = (wv == 0 || wv == 2);
      const int kf = sface ? k : k + 1;
      const SrcInfo Sb = sface ? So : S, St = sface ? S : So;
      bool zero, lot;
      if (!injx) {
        zero = true;
        lot = false;
      } else if (sface ? (k == 0) : (k == P.ny - 1)) {
        // boundary face: both sides carry row k's value
        zero = S.is_src;
        lot = sface;
      } else if (bottom) {
        zero = Sb.is_src && !(P.radial && St.src2 && St.sS && kf < P.ny - 1);
        lot = Sb.is_src;
      } else {
        zero = St.is_src && !(P.radial && Sb.src2 && !Sb.sS && Sb.sN && kf - 1 > 0);
        lot = !St.is_src;
      }
      double qs[NQ];
#pragma unroll
      for (int i = 0; i < NQ; i++) {
        double x = va[i], y = vb[i];
        pin(x);
        pin(y);
        qs[i] = lot ? y : x;
      }
      make_side<NS, 2>(P, qs, g, sd);
#pragma unroll
      for (int i = 0; i < NV; i++) {
        sd.f[i] = zero ? 0.0 : sd.f[i];
        sd.qc[i] = zero ? 0.0 : sd.qc[i];
      }
      sd.vmin = zero ? 0.0 : sd.vmin;
      sd.vmax = zero ? 0.0 : sd.vmax;
      sd.zv = zero ? 0.0 : sd.zv;
      side_store<NS>(SD[wv], sd, lane);
    } else {
      // x faces: W side (w4, kept), E side (w5, to LDS)
      if (!injx || S.is_src) zero_side<NS>(sd);
      else make_side<NS, 1>(P, va, (wv == 4) ? R.gW : R.gE, sd);
      if (wv == 4) Wself = sd;
      else side_store<NS>(SD[4], sd, lane);
    }
  }
  __syncthreads();
  // ---- phase C
  double HW[NV], HE[NV], amaxW = 0.0, amaxE = 0.0;
  if (wv == 0 || wv == 2) {
    Side<NS> Bs, Ts;
    side_load<NS>(SD[wv], Bs, lane);
    side_load<NS>(SD[wv + 1], Ts, lane);
    double H[NV];
    const double a = face_flux<NS, 2>(Bs, Ts, H);
    const int o = (wv == 0) ? 0 : 1;
#pragma unroll
    for (int i = 0; i < NV; i++) V[o][i][lane] = H[i];
    V[o][NV][lane] = a;
  } else if (wv == 4) {
    Side<NS> Eself, Eleft;
    side_load<NS>(SD[4], Eself, lane);
    side_load<NS>(SD[4], Eleft, (lane + 63) & 63);
    x_face_fluxes<NS>(P, S, j, Wself, Eself, Eleft, HW, HE, amaxW, amaxE);
  }
  __syncthreads();
  // ---- phase D
  if (wv == 4 && own && (R.m & M_COMPUTE)) {
    double HS[NV], HN[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) {
      HS[i] = V[0][i][lane];
      HN[i] = V[1][i][lane];
    }
    cell_update<NS, true>(P, D, cidx(P, j, k), qk, R, HW, HE, HS, HN, amaxW, amaxE,
                          V[0][NV][lane], V[1][NV][lane], compute_dt, dtmin, ex);
  }
  __syncthreads();  // LDS free for the next row
}

template <int NS>
__global__ void __launch_bounds__(384) k_stencil_rows(DevParams P, DevPtrs D, int compute_dt) {
  warm_kernargs<kArgBytes>();
  constexpr int NV = 4 + NS;
  constexpr int NQ = NV + 2;       // reconstructed values per side
  constexpr int NSD = 2 * NV + 3;  // doubles per face side
  // phase A values (then the y fluxes of phase C), face sides
  __shared__ double V[6][NQ][64];
  __shared__ double SD[5][NSD][64];  // S bottom, S top, N bottom, N top, E of x
  const int wv = threadIdx.x >> 6;
  // the step's scalars, list size and first row in one round of loads (the
  // grid is never larger than the row list, so row_list[blockIdx.x] exists)
  double done = D.scal[S_DONE], t = D.scal[S_T];
  int nrows = (int)*ctr(D, C_ROWS, 0), row = D.row_list[blockIdx.x];
  pin(done);
  pin(t);
  pin(nrows);
  pin(row);
  if (done != 0.0 || (int)blockIdx.x >= nrows) return;
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  double dtmin = P.max_dt;
  // the workgroup's first row outside the loop over further rows: in a loop
  // the compiler hoists the row's uniform operands ahead of it, where, short
  // of scalar registers, it fetches the kernel arguments one at a time
  stencil_row<NS>(P, D, row, tiles_x, t, compute_dt, dtmin, V, SD);
  for (int item = blockIdx.x + gridDim.x; item < nrows; item += gridDim.x)
    stencil_row<NS>(P, D, D.row_list[item], tiles_x, t, compute_dt, dtmin, V, SD);
  if (wv == 4 && compute_dt) dt_reduce(D, P, dtmin);
}

// ---------------------------------------------------------------------------
// the host loop's time-step policy, src/IMEX_SfloW2D.f90:357-381, on the
// device: step_dt evaluates it from the step's scalars and the CFL minimum
// without writing anything (every workgroup of the stage kernel does), and
// dtpolicy_member commits the result
// ---------------------------------------------------------------------------
struct StepDt {
  bool done;       // the run was over, or t >= t_end: no update this step
  double dt, dt_cfl, t_out;
};
__device__ __forceinline__ StepDt step_dt(const DevParams& P, const DevPtrs& D) {
  // every scalar in one round of loads (the tests below would otherwise
  // wait for each load in turn)
  double s[S_NSLOTS];
#pragma unroll
  for (int i = 0; i < S_NSLOTS; i++) s[i] = D.scal[i];
  unsigned long long bits = *D.dt_bits;
#pragma unroll
  for (int i = 0; i < S_NSLOTS; i++) pin(s[i]);
  pin(bits);
  StepDt r{true, 0.0, 0.0, 0.0};
  if (s[S_DONE] != 0.0) return r;
  const double t = s[S_T], t_end = s[S_T_END];
  double t_out = s[S_T_OUT];
  if (!(t < t_end)) return r;
  double dt = (P.cfl != -1.0) ? __longlong_as_double((long long)bits) : P.max_dt;
  r.dt_cfl = dt;
  if (t_end - t_out < 1.0e-7) t_out = t_end;
  // paused at an output time: the host writes its output and raises t_output
  // before the next call (src/IMEX_SfloW2D.f90:510-527, :6166)
  if (!(t < t_out)) return r;
  if (t + dt > t_end) dt = t_end - t;
  if (t + dt > t_out) dt = t_out - t;
  dt = fmin(dt, 1.1 * 0.5 * (s[S_DT_OLD] + s[S_DT_OLD_OLD]));
  r.done = false;
  r.dt = dt;
  r.t_out = t_out;
  return r;
}
__device__ __forceinline__ void dtpolicy_member(const DevParams& P, const DevPtrs& D) {
  double* s = D.scal;
  if (s[S_DONE] != 0.0) return;
  const StepDt r = step_dt(P, D);
  if (r.done) {
    s[S_DONE] = 1.0;
    return;
  }
  s[S_DT_CFL] = r.dt_cfl;
  s[S_DT_OLD_OLD] = s[S_DT_OLD];
  s[S_DT_OLD] = r.dt;
  s[S_DT] = r.dt;
  s[S_T_OUT] = r.t_out;
}

// t = t + dt, step count and active cell-updates
__device__ __forceinline__ void advance_member(const DevPtrs& D) {
  double* s = D.scal;
  if (s[S_DONE] != 0.0) return;
  s[S_T] = s[S_T] + s[S_DT];
  unsigned long long n = 0, nc = 0;
  for (int i = 0; i < kSlots; i++) {
    n += *ctr(D, C_SOLVE_CELLS, i);
    nc += *ctr(D, C_COMPUTE_CELLS, i);
  }
  *ctr(D, C_STEPS, 0) += 1;
  *ctr(D, C_ACTIVE_UPD, 0) += n;
  *ctr(D, C_COMPUTE_UPD, 0) += nc;
  // the run stops once t reaches the output time (t_output, snapped to t_end
  // by the policy), where the reference's loop writes its output
  if (s[S_T] >= s[S_T_OUT]) s[S_DONE] = 1.0;
}

// CFL accumulator and per-step counters / work lists (a stopped run keeps
// the last step's, for the procedures the host calls at the pause)
__device__ __forceinline__ void reset_member(const DevParams& P, const DevPtrs& D) {
  if (D.scal[S_DONE] != 0.0) return;
  *D.dt_bits = (unsigned long long)__double_as_longlong(P.max_dt);
  for (int i = 0; i < kSlots; i++) *ctr(D, C_COMPUTE_CELLS, i) = 0;
  *ctr(D, C_TILES, 0) = 0;
  *ctr(D, C_ROWS, 0) = 0;
  *ctr(D, C_CELLS, 0) = 0;
}

// Between two steps of imexsw_advance: commit the previous step (its time
// step policy, t = t + dt, counters) when `commit`, then reset the per-step
// counters for the next one.  With commit and without reset: the end of the
// last step.
template <int = 0>
__global__ void k_step_begin(DevParams P, DevPtrs D, int commit, int reset) {
  if (commit) {
    dtpolicy_member(P, D);
    advance_member(D);
  }
  if (reset) reset_member(P, D);
}

// ---------------------------------------------------------------------------
// k_local: stage 2 + assembly (imex_RK_solver :1105-1626), erosion /
// deposition (update_erosion_deposition_cell :2380-2595) and the post-step
// loop (src/IMEX_SfloW2D.f90:391-447), fused per cell.
// ---------------------------------------------------------------------------
// OP_POLICY: the step's dt is step_dt (the policy is committed after the
// step), else the scalar S_DT
enum LocalOps { OP_RK = 1, OP_EROSION = 2, OP_POST = 4, OP_SUMS = 8, OP_WARM = 16,
                OP_HARVEST = 32, OP_POLICY = 64, OP_CLOSED = 128 };

// The assembly checks of imex_RK_solver (src/solver_2d.f90:1379-1624) on
// the assembled state q of cell c: NaN trap, small or negative thickness,
// negative solids, low temperature, solids above the total; qp receives the
// physical variables.  Events count into e (C_NEWTON_CELLS-relative) when
// `counts`.
// *qp_cur (when given) tells whether qp is that of the final q (false when
// the solids were rescaled or the cell zeroed after the conversion).
template <int NS, bool VP = false>
__device__ __forceinline__ void assembly_checks(const DevParams& P, const DevPtrs& D, size_t c,
                                                double* q, double* qp, unsigned long long* e,
                                                bool counts, bool* qp_cur = nullptr) {
  constexpr int NV = 4 + NS;
  bool isn = false;
#pragma unroll
  for (int i = 0; i < NV; i++) isn |= isnan(q[i]);
  if (counts && isn) atomicAdd(&e[C_NAN - C_NEWTON_CELLS], 1ull);
  if (q[0] < kEps) {
    if (q[0] > -1.0e-7) {
#pragma unroll
      for (int i = 0; i < NV; i++) q[i] = 0.0;
    } else {
      if (counts) atomicAdd(&e[C_NEGH - C_NEWTON_CELLS], 1ull);
    }
  }
  {
    bool anyneg = false, anybig = false;
#pragma unroll
    for (int i = 0; i < NS; i++) {
      anyneg |= q[4 + i] < 0.0;
      anybig |= q[4 + i] <= -1.0e-7;
    }
    if (anyneg) {
      if (anybig) {
        if (counts) atomicAdd(&e[C_NEGS - C_NEWTON_CELLS], 1ull);
      } else {
#pragma unroll
        for (int i = 0; i < NS; i++)
          if (q[4 + i] > -1.0e-7) q[4 + i] = 0.0;
      }
    }
  }
  phys<NS, VP>(P, q, qp);
  bool cur = true;
  if (counts && qp[3] < 273.0) atomicAdd(&e[C_LOWT - C_NEWTON_CELLS], 1ull);
  double ssol = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) ssol += q[4 + i];
  if (ssol > q[0]) {
    if (((ssol - q[0]) / q[0] < 1.0e-10) || (q[0] < kEps)) {
#pragma unroll
      for (int i = 0; i < NS; i++) q[4 + i] = q[4 + i] / ssol * q[0];
      cur = false;
    } else {
      if (counts) atomicAdd(&e[C_SUMS - C_NEWTON_CELLS], 1ull);
    }
  }
  if (D.nodata && D.nodata[c]) {
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = 0.0;
    cur = false;
  }
  if (qp_cur) *qp_cur = cur;
}

// one cell of k_local; e (the wavefront's counters in LDS) accumulates
// counter C_NEWTON_CELLS + m in e[m]
// PH selects the part compiled in: 0 = stage update and assembly checks
// (OP_RK), 1 = erosion/deposition and diagnostics (OP_EROSION, OP_POST,
// OP_SUMS), 2 = the stage update with a Newton starting point (the learned
// one, OP_WARM, or the closed form, OP_CLOSED) and the pair harvest
// (OP_HARVEST), 3 = parts 0 and 1 in one pass (few computed cells: a row
// goes through both without waiting for every other row's stage update),
// 4 = part 3 with a Newton starting point.  Separate kernels keep the plain
// stage update's register footprint free of the starting-point code, and on
// a wet grid the stage update's free of the erosion part.
// (PH 3 and 4 run each cell on kSparseLanes lanes: every lane computes,
// `counts` is set on one of them)
template <int NS, int PH, bool VP = false>
__device__ __forceinline__ void local_cell(const DevParams& P, const DevPtrs& D, size_t c,
                                           double dt, int ops, unsigned long long* e,
                                           int& imax, double& s_mass, double& s_vol,
                                           double& s_area, double& s_sol, bool counts = true) {
  constexpr int NV = 4 + NS;
  constexpr int NSX = NS > 0 ? NS : 1;
  const size_t pl = plane(P);
  const bool rk = (PH == 0 || PH == 2 || PH >= 3) && (ops & OP_RK);
  const bool ero = (PH == 1 || PH >= 3) && (ops & OP_EROSION) &&
                   !((P.erosion_coeff == 0.0) && !P.settling && !P.entrain);
  const bool post = (PH == 1 || PH >= 3) && (ops & OP_POST);
  // every load of the cell first, the mask byte among them, so that the
  // cell waits for memory once
  double q[NV], qp[NV + 2];
  double Dr[NV], bx = 0.0, by = 0.0, bxx = 0.0, bxy = 0.0, byy = 0.0, gc = 1.0, wg[2];
  double erod[NSX], depo[NSX], eros[NSX], bcell = 0.0, hm = 0.0, pm = 0.0, mvel = 0.0;
  uint32_t vbits = 0;
  const int mk = D.mask[c];
#pragma unroll
  for (int i = 0; i < NV; i++) q[i] = D.q[i * pl + c];
  if (rk) {
#pragma unroll
    for (int i = 0; i < NV; i++) Dr[i] = D.D[i * pl + c];
    bxx = D.Bxx ? D.Bxx[c] : 0.0;
    bxy = D.Bxy ? D.Bxy[c] : 0.0;
    byy = D.Byy ? D.Byy[c] : 0.0;
    gc = D.gcoef ? D.gcoef[c] : 1.0;
    if ((PH == 2 || PH == 4) && (ops & OP_WARM)) {
      wg[0] = D.guess[c];
      wg[1] = D.guess[pl + c];
    }
  }
  if (rk || ero) {
    bx = D.Bx[c];
    by = D.By[c];
  }
  if (ero) {
#pragma unroll
    for (int i = 0; i < NS; i++) {
      erod[i] = D.erodible[i * pl + c];
      depo[i] = D.deposit[i * pl + c];
      eros[i] = D.erosion[i * pl + c];
    }
    if (P.topo) bcell = D.B[c];
  }
  if (post) {
    hm = D.hmax[c];
    pm = D.pdynmax[c];
    mvel = D.modvel[c];
    if (D.vuln) vbits = D.vuln[c];
  }
  // the stored physical variables of q (vertical profiles): the erosion part
  // starts from them instead of converting q again
  const bool qpl = VP && ero && !rk && D.qpp;
  if (qpl) {
#pragma unroll
    for (int i = 0; i < NV + 2; i++) qp[i] = D.qpp[i * pl + c];
  }
  // the loaded values are needed in registers here: no load is moved into
  // the branches below
#pragma unroll
  for (int i = 0; i < NV; i++) pin(q[i]);
  if (rk) {
#pragma unroll
    for (int i = 0; i < NV; i++) pin(Dr[i]);
    pin(bxx);
    pin(bxy);
    pin(byy);
    pin(gc);
  }
  if (rk || ero) {
    pin(bx);
    pin(by);
  }
  if (ero) {
#pragma unroll
    for (int i = 0; i < NS; i++) {
      pin(erod[i]);
      pin(depo[i]);
      pin(eros[i]);
    }
    pin(bcell);
  }
  if (post) {
    pin(hm);
    pin(pm);
    pin(mvel);
  }
  if (qpl) {
#pragma unroll
    for (int i = 0; i < NV + 2; i++) pin(qp[i]);
  }
  if (!(mk & M_COMPUTE)) return;
  // qp holds the physical variables of the current q (and pd their dynamic
  // pressure when set by the last conversion): a later part that needs them
  // does not convert again
  bool qp_cur = qpl, pd_cur = false;
  double pd = 0.0;
  if (rk) {
    double q0[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) q0[i] = q[i];
    int nonconv = 0, ls = 0, wu = -1;
    float hv[kHarvestWidth];
    bool hok = false;
    int it = stage2<NS, (PH >= 3) ? kSparseLanes : 1>(P, q0, Dr, dt, bx, by, bxx, bxy, byy, gc,
                                                       q, nonconv, ls,
                        ((PH == 2 || PH == 4) && (ops & OP_WARM)) ? wg : nullptr, &wu,
                        (PH == 2 && (ops & OP_HARVEST)) ? hv : nullptr, &hok,
                        (PH == 2 || PH == 4) && (ops & OP_CLOSED));
    if (counts && wu == 1) atomicAdd(&e[C_WARM_USED - C_NEWTON_CELLS], 1ull);
    if (counts && wu == 0) atomicAdd(&e[C_WARM_FALLBACK - C_NEWTON_CELLS], 1ull);
    if (PH == 2 && (ops & OP_HARVEST)) {
      // wave-aggregated append of the converged pairs
      const unsigned long long b = __ballot(hok);
      unsigned long long base = 0;
      if ((threadIdx.x & 63) == __ffsll((long long)b) - 1)
        base = atomicAdd(ctr(D, C_HARVEST, 0), (unsigned long long)__popcll(b));
      base = __shfl(base, __ffsll((long long)b) - 1, 64);
      if (hok) {
        const unsigned long long at =
            base + __popcll(b & ((1ull << (threadIdx.x & 63)) - 1ull));
        if ((long long)at < D.harvest_cap)
          for (int i = 0; i < kHarvestWidth; i++) D.harvest[at * kHarvestWidth + i] = hv[i];
      }
    }
    if (it > 0 && counts) {
      atomicAdd(&e[C_NEWTON_CELLS - C_NEWTON_CELLS], 1ull);
      atomicAdd(&e[C_NEWTON_ITERS - C_NEWTON_CELLS], (unsigned long long)it);
      if (nonconv) atomicAdd(&e[C_NONCONV - C_NEWTON_CELLS], 1ull);
      if (ls) atomicAdd(&e[C_LS_EVALS - C_NEWTON_CELLS], (unsigned long long)ls);
      const int b = iter_bucket(it);
      atomicAdd(&e[C_HIST0 - C_NEWTON_CELLS + b], 1ull);
      imax = max(imax, it);
    }
    assembly_checks<NS, VP>(P, D, c, q, qp, e, counts, &qp_cur);
  }
  if (ero) {
    if (!qp_cur) phys_or_dry<NS, VP>(P, q, qp);
    double er[NSX], dep[NSX], eq[NV], topo;
    mass_exchange<NS, (PH >= 3) ? kSparseLanes : 1, VP>(P, qp, bx, by, erod, dt, er, dep, eq,
                                                         topo);
#pragma unroll
    for (int i = 0; i < NV; i++) q[i] = q[i] + dt * eq[i];
#pragma unroll
    for (int i = 0; i < NS; i++) q[4 + i] = fmax(0.0, q[4 + i]);
#pragma unroll
    for (int i = 0; i < NS; i++) {
      D.deposit[i * pl + c] = depo[i] + dt * dep[i];
      D.erosion[i * pl + c] = eros[i] + dt * er[i];
      double ev = erod[i] - dt * er[i];
      if (P.erod_dep) ev = ev + dt * dep[i];
      D.erodible[i * pl + c] = ev;
    }
    if (P.topo) D.B[c] = bcell + dt * topo;
    if (q[0] <= 0.0 && q[0] > -1.0e-10) {
#pragma unroll
      for (int i = 0; i < NV; i++) q[i] = 0.0;
    }
    double ssol = 0.0;
#pragma unroll
    for (int i = 0; i < NS; i++) ssol += q[4 + i];
    if (ssol > q[0] && q[0] < 1.0e-10) {
#pragma unroll
      for (int i = 0; i < NS; i++) q[4 + i] = q[4 + i] / ssol * q[0];
    }
    double rg = 0.0;
    if (q[0] > 0.0) {
      pd = phys<NS, VP>(P, q, qp);
      rg = mixt<NS, false>(P, qp).rg;
    }
    if (rg <= 0.0) {
#pragma unroll
      for (int i = 0; i < NV; i++) q[i] = 0.0;
#pragma unroll
      for (int i = 0; i < NV + 2; i++) qp[i] = 0.0;
      qp[3] = P.T_ambient;
      pd = 0.0;
    }
    qp_cur = true;
    pd_cur = true;
    if (VP && D.qpp) {
#pragma unroll
      for (int i = 0; i < NV + 2; i++) D.qpp[i * pl + c] = qp[i];
    }
  }
  if (ops & (OP_RK | OP_EROSION)) {
    bool zero = true;
#pragma unroll
    for (int i = 0; i < NV; i++) {
      D.q[i * pl + c] = q[i];
      zero = zero && (q[i] == 0.0);
    }
    D.sflag[c] = state_flags(q[0], zero);
  }
  if (post) {
    if (!(qp_cur && pd_cur)) pd = phys_or_dry<NS, VP>(P, q, qp);
    if (q[0] > 0.0) {
      D.hmax[c] = fmax(hm, qp[0]);
      const double mv = sqrt(qp[NV] * qp[NV] + qp[NV + 1] * qp[NV + 1]);
      if (qp[0] > 0.001) {
        D.pdynmax[c] = fmax(pm, pd);
        D.modvel[c] = fmax(mvel, mv);
      }
      if (D.vuln) {
        uint32_t bits = vbits;
        int it = 0;
        for (int a = 0; a < P.n_thk; a++) {
          const bool th = qp[0] >= P.thickness_levels[a];
          for (int b = 0; b < P.n_pdyn; b++) {
            const bool pdb = pd >= P.dyn_pres_levels[b];
            if (th && pdb) bits |= (1u << it);
            it++;
          }
        }
        D.vuln[c] = bits;
      }
    }
    if (ops & OP_SUMS) {
      s_mass += q[0];
      s_vol += qp[0];
      s_area += (q[0] > 1.0e-7) ? 1.0 : 0.0;
#pragma unroll
      for (int i = 0; i < NS; i++) s_sol += q[4 + i];
    }
  }
}

// The stage kernels change states only in the listed rows: their tiles and
// the tiles next to them (the mask reaches two cells) are rescanned by the
// next step's mask.
__device__ __forceinline__ void mark_dirty(const DevParams& P, const DevPtrs& D, int tile) {
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int tiles_y = (P.ny + kTileH - 1) / kTileH;
  const int tx = tile % tiles_x, ty = tile / tiles_x;
  for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
      const int x = tx + dx, y = ty + dy;
      if (x >= 0 && x < tiles_x && y >= 0 && y < tiles_y) D.tile_dirty[y * tiles_x + x] = 1;
    }
}

// One tile row (60 cells, one per lane 2..61) of k_local; per-row partial
// sums go to tile_sums[4 * row ..] when OP_SUMS.
template <int NS, int PH, bool VP = false>
__device__ __forceinline__ void local_row(const DevParams& P, const DevPtrs& D, int row,
                                          double dt, int ops, unsigned long long* e,
                                          int& imax) {
  const int lane = threadIdx.x & 63;
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int tile = row / kTileH;
  const int j = (tile % tiles_x) * kStripW - 2 + lane;
  const int k = (tile / tiles_x) * kTileH + row % kTileH;
  const bool own = lane >= 2 && lane < 2 + kStripW && j < P.nx && k < P.ny;
  double s_mass = 0.0, s_vol = 0.0, s_area = 0.0, s_sol = 0.0;
  if (own)
    local_cell<NS, PH, VP>(P, D, cidx(P, j, k), dt, ops, e, imax, s_mass, s_vol, s_area, s_sol);
  if (PH == 1 && (ops & OP_SUMS)) {
    double v4[4] = {s_mass, s_vol, s_area, s_sol};
#pragma unroll
    for (int m = 0; m < 4; m++) {
      double v = v4[m];
      for (int off = 32; off > 0; off >>= 1) v += shfl_d(v, lane ^ off);
      if (lane == 0) D.tile_sums[4 * (size_t)row + m] = v;
    }
  }
}

// Cells kCellsPerWave * item .. of the step's list of cells to compute, for
// PH 3 and 4: cell i of the item on lanes kSparseLanes*i .. kSparseLanes*i +
// kSparseLanes-1, which also mark its tile and the tiles next to it for the
// next step's mask.  No per-row sums.
constexpr int kCellsPerWave = kWave / kSparseLanes;
template <int NS, int PH>
// (c_pre: the entry of the list for this lane, when already loaded, else -1)
__device__ __forceinline__ void local_cells(const DevParams& P, const DevPtrs& D, int item,
                                            int ncells, double dt, int ops,
                                            unsigned long long* e, int& imax, int c_pre) {
  const int lane = threadIdx.x & 63, i = lane / kSparseLanes;
  const int idx = item * kCellsPerWave + i;
  if (i >= kCellsPerWave || idx >= ncells) return;
  const int c = (c_pre >= 0) ? c_pre : D.cell_list[idx];
  if ((lane % kSparseLanes) == 0) {
    const int tiles_x = (P.nx + kStripW - 1) / kStripW;
    const int k = c / (int)P.pitch, j = c - k * (int)P.pitch;
    mark_dirty(P, D, (k / kTileH) * tiles_x + j / kStripW);
  }
  double s_mass = 0.0, s_vol = 0.0, s_area = 0.0, s_sol = 0.0;
  local_cell<NS, PH>(P, D, (size_t)c, dt, ops & ~(int)OP_SUMS, e, imax, s_mass, s_vol, s_area,
                     s_sol, (lane % kSparseLanes) == 0);
}

// Learned Newton starting points for this step: for every computed cell of
// the listed rows, the stage-2 predictor, its features, and the MLP on the
// matrix cores (the 64 lanes of a wavefront are the 64 samples); the
// starting momenta q_si * (1 + prediction) go to the guess planes (NaN where
// the cell does not enter Newton).
template <int NS>
__global__ void __launch_bounds__(64) k_mlp_guess(DevParams P, DevPtrs D, MlpDev M,
                                                   int policy) {
  constexpr int NV = 4 + NS;
  __shared__ short xs[64][kMlpInPad];
  __shared__ float ys[64][kMlpOut];
  const int lane = threadIdx.x, i = lane & 15, g = lane >> 4;
  double dt;
  if (policy) {
    const StepDt r = step_dt(P, D);
    if (r.done) return;
    dt = r.dt;
  } else {
    if (D.scal[S_DONE] != 0.0) return;
    dt = D.scal[S_DT];
  }
  const int nrows = (int)*ctr(D, C_ROWS, 0);
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const size_t pl = plane(P);
  for (int it = blockIdx.x; it < nrows; it += gridDim.x) {
    const int row = D.row_list[it];
    const int tile = row / kTileH;
    const int j = (tile % tiles_x) * kStripW - 2 + lane;
    const int k = (tile / tiles_x) * kTileH + row % kTileH;
    const bool own = lane >= 2 && lane < 2 + kStripW && j < P.nx && k < P.ny;
    const size_t c = own ? cidx(P, j, k) : 0;
    bool valid = false;
    double s1 = 0.0, s2 = 0.0;
    float f[kMlpIn];
#pragma unroll
    for (int m = 0; m < kMlpIn; m++) f[m] = 0.f;
    if (own && (D.mask[c] & M_COMPUTE)) {
      double q0[NV], Dr[NV], q_fv[NV], qp[NV + 2], q_si[NV];
#pragma unroll
      for (int m = 0; m < NV; m++) {
        q0[m] = D.q[m * pl + c];
        Dr[m] = D.D[m * pl + c];
      }
      const double bx = D.Bx[c], by = D.By[c];
      const double bxx = D.Bxx ? D.Bxx[c] : 0.0, bxy = D.Bxy ? D.Bxy[c] : 0.0,
                   byy = D.Byy ? D.Byy[c] : 0.0;
      const double gc = D.gcoef ? D.gcoef[c] : 1.0;
      if (predictor<NS>(P, q0, Dr, dt, bx, by, bxx, bxy, byy, gc, q_fv, qp, q_si)) {
        valid = true;
        s1 = q_si[1];
        s2 = q_si[2];
        mlp_features_raw<NS>(P, q_fv, qp, q_si, dt, bx, by, f);
      }
    }
#pragma unroll
    for (int m = 0; m < kMlpInPad; m++)
      xs[lane][m] = (valid && m < kMlpIn)
                        ? f2bf((f[m] - M.norm[m]) / M.norm[kMlpIn + m])
                        : (short)0;
    __syncthreads();
    bf16x4 xb[4];
    input_operands(xs, xb);
    f32x4 y[4];
    mlp_forward_wave(M, xb, y, nullptr);
    if (g == 0)
#pragma unroll
      for (int nb = 0; nb < 4; nb++)
#pragma unroll
        for (int o = 0; o < kMlpOut; o++) ys[nb * 16 + i][o] = y[nb][o];
    __syncthreads();
    if (own) {
      double g1 = __builtin_nan(""), g2 = __builtin_nan("");
      if (valid) {
        const float* on = M.norm + 2 * kMlpIn;
        const double a = (double)(ys[lane][0] * on[2] + on[0]);
        const double b = (double)(ys[lane][1] * on[3] + on[1]);
        g1 = s1 * (1.0 + a);
        g2 = s2 * (1.0 + b);
      }
      D.guess[c] = g1;
      D.guess[pl + c] = g2;
    }
    __syncthreads();
  }
}

// One wavefront per listed tile row (per row part for PH 3); the four
// wavefronts of a workgroup take rows in turn until the list is exhausted.
// The plain stage update and the erosion part are held to two workgroups
// per CU (the register budget of two waves per SIMD): on a wet grid the
// second wave hides the first's latency (13% faster stage update measured
// at 2048^2, no change on sparse steps).
template <int NS, int PH, bool VP = false>
__global__ void __launch_bounds__(256, (!VP && (PH == 0 || PH == 1)) ? 2 : 1)
    k_local(DevParams P, DevPtrs D, int ops) {
  warm_kernargs<kArgBytes>();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __shared__ unsigned long long cnt[4][kLocalEvents];
  __shared__ int itmax[4];
  for (int m = lane; m < kLocalEvents; m += 64) cnt[w][m] = 0;
  __syncthreads();
  int imax = 0;
  // PH 3, 4: the list entry of this lane's first item, from a clamped
  // address, in flight with the step's scalars
  int c_first = 0;
  if (PH >= 3) {
    const int idx = (int)min((long long)(blockIdx.x * 4 + w) * kCellsPerWave + lane / kSparseLanes,
                             (long long)P.nx * P.ny - 1);
    c_first = D.cell_list[idx];
  }
  // the list size, issued ahead of the step's scalars (one round of loads)
  unsigned long long nl_raw = (PH >= 3) ? *ctr(D, C_CELLS, 0) : *ctr(D, C_ROWS, 0);
  bool live;
  double dt;
  if (ops & OP_POLICY) {
    const StepDt r = step_dt(P, D);
    live = !r.done;
    dt = r.dt;
  } else {
    live = (D.scal[S_DONE] == 0.0);
    dt = D.scal[S_DT];
  }
  pin(nl_raw);
  // PH 3, 4: items of kCellsPerWave listed cells; else listed tile rows
  const int nlist = (PH >= 3) ? (int)min(nl_raw, (unsigned long long)P.nx * P.ny) : (int)nl_raw;
  const int items = (PH >= 3) ? (nlist + kCellsPerWave - 1) / kCellsPerWave : nlist;
  for (int it = blockIdx.x * 4 + w; live && it < items; it += gridDim.x * 4) {
    if (PH >= 3) {
      local_cells<NS, (PH >= 3 ? PH : 3)>(P, D, it, nlist, dt, ops, cnt[w], imax,
                                           (it == (int)blockIdx.x * 4 + w) ? c_first : -1);
    } else {
      if (lane == 0) mark_dirty(P, D, D.row_list[it] / kTileH);
      local_row<NS, PH, VP>(P, D, D.row_list[it], dt, ops, cnt[w], imax);
    }
  }
  for (int off = 32; off > 0; off >>= 1) imax = max(imax, __shfl_xor(imax, off, 64));
  if (lane == 0) itmax[w] = imax;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int slot = blockIdx.x;
    for (int m = 0; m < kLocalEvents; m++) {
      unsigned long long v = cnt[0][m] + cnt[1][m] + cnt[2][m] + cnt[3][m];
      if (v) atomicAdd(ctr(D, C_NEWTON_CELLS + m, slot), v);
    }
    const int im = max(max(itmax[0], itmax[1]), max(itmax[2], itmax[3]));
    if (im) atomicMax(ctr(D, C_ITMAX, slot), (unsigned long long)im);
  }
}

// ---------------------------------------------------------------------------
// k_topo: topography_reconstruction (src/geometry_2d.f90:688-856) for the
// fields the hot path reads.  Border cells take the value of the nearest
// cell of the 5-point interior (:737-796).  When `sparse`, only tiles within
// four cells of a computed cell are recomputed: B_cent changes only where
// the stage update runs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void topo_cell(const DevParams& P, const DevPtrs& D, int j, int k);

// one workgroup per work tile (lane l: column 60*tx-2+l, four rows per
// wavefront); when `sparse`, only tiles with a computed cell in the tile or
// a neighbour tile (a computed cell within four cells)
__device__ __forceinline__ void topo_tile(const DevParams& P, const DevPtrs& D, int tx,
                                          int ty, int tiles_x, int tiles_y, int sparse) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (sparse) {
    bool any = false;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const int x = tx + dx, y = ty + dy;
        if (x >= 0 && x < tiles_x && y >= 0 && y < tiles_y)
          any = any || D.tile_flag[y * tiles_x + x];
      }
    if (!any) return;
  }
  const int j = tx * kStripW - 2 + lane;
  if (lane < 2 || lane >= 2 + kStripW || j >= P.nx) return;
  for (int rr = w; rr < kTileH; rr += 4) {
    const int k = ty * kTileH + rr;
    if (k < P.ny) topo_cell(P, D, j, k);
  }
}

template <int = 0>
__global__ void __launch_bounds__(256) k_topo(DevParams P, DevPtrs D, int sparse) {
  topo_tile(P, D, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, sparse);
}

__device__ __forceinline__ void topo_cell(const DevParams& P, const DevPtrs& D, int j, int k) {
  const size_t c = cidx(P, j, k);
  double bx = 0.0, by = 0.0, bxx = 0.0, byy = 0.0, bxy = 0.0;
  if (P.nx >= 5 && P.ny >= 5) {
    const int jj = min(max(j, 2), P.nx - 3), kk = min(max(k, 2), P.ny - 3);
    const double c1[5] = {-2.0, -1.0, 0.0, 1.0, 2.0};
    const double c2[5] = {2.0, -1.0, -2.0, -1.0, 2.0};
    double px = 0.0, py = 0.0, sxx = 0.0, syy = 0.0;
#pragma unroll
    for (int t = 0; t < 5; t++) {
      const double b = D.B[cidx(P, jj + t - 2, kk)];
      px = px + c1[t] * b;
      sxx = sxx + c2[t] * b;
    }
#pragma unroll
    for (int t = 0; t < 5; t++) {
      const double b = D.B[cidx(P, jj, kk + t - 2)];
      py = py + c1[t] * b;
      syy = syy + c2[t] * b;
    }
    bx = px / (10.0 * P.dx);
    by = py / (10.0 * P.dy);
    bxx = sxx / (7.0 * (P.dx * P.dx));
    byy = syy / (7.0 * (P.dy * P.dy));
    if (D.Bxy) {
      double ws = 0.0;
      for (int a = 0; a < 5; a++)
        for (int b = 0; b < 5; b++)
          ws = ws + c1[b] * c1[a] * D.B[cidx(P, jj + b - 2, kk + a - 2)];
      bxy = ws / ((10.0 * P.dx) * (10.0 * P.dy));
    }
  }
  D.Bx[c] = bx;
  D.By[c] = by;
  if (D.Bxx) {
    D.Bxx[c] = bxx;
    D.Byy[c] = byy;
    D.Bxy[c] = bxy;
  }
  if (D.gcoef) {
    const double g = 1.0 / (1.0 + bx * bx + by * by);
    D.gcoef[c] = g;
    D.dgdx[c] = -(2.0 * (g * g) * (bx * bxx + by * bxy));
  }
}

// The same recomputation driven by this step's tile list (few computed
// tiles): item (listed tile t, neighbour d of 9) recomputes tile t + d.
// Every tile of the 3x3 neighbourhoods is done once: a listed tile by
// itself, any other by the smallest listed tile next to it.  Persistent
// grid, four wavefronts per workgroup as in topo_tile.
template <int = 0>
__global__ void __launch_bounds__(256) k_topo_list(DevParams P, DevPtrs D) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tiles_x = (P.nx + kStripW - 1) / kStripW;
  const int tiles_y = (P.ny + kTileH - 1) / kTileH;
  const int n = (int)*ctr(D, C_TILES, 0);
  for (int item = blockIdx.x; item < 9 * n; item += gridDim.x) {
    const int t = D.tile_list[item / 9], dd = item % 9;
    const int tx = t % tiles_x + dd % 3 - 1, ty = t / tiles_x + dd / 3 - 1;
    if (tx < 0 || tx >= tiles_x || ty < 0 || ty >= tiles_y) continue;
    const int nt = ty * tiles_x + tx;
    if (nt != t) {
      // the flags of nt and the tiles around it in one round of loads from
      // clamped addresses (conditional loads would be waited for one by one)
      uint8_t f[9];
#pragma unroll
      for (int m = 0; m < 9; m++) {
        const int x = min(max(tx + m % 3 - 1, 0), tiles_x - 1);
        const int y = min(max(ty + m / 3 - 1, 0), tiles_y - 1);
        f[m] = D.tile_flag[y * tiles_x + x];
      }
      if (f[4]) continue;  // listed: done by itself
      int first = t;
#pragma unroll
      for (int m = 0; m < 9; m++) {
        const int x = tx + m % 3 - 1, y = ty + m / 3 - 1;
        if (x >= 0 && x < tiles_x && y >= 0 && y < tiles_y && f[m])
          first = min(first, y * tiles_x + x);
      }
      if (first != t) continue;
    }
    const int j = tx * kStripW - 2 + lane;
    if (lane < 2 || lane >= 2 + kStripW || j >= P.nx) continue;
    for (int rr = w; rr < kTileH; rr += 4) {
      const int k = ty * kTileH + rr;
      if (k < P.ny) topo_cell(P, D, j, k);
    }
  }
}

// staggered grav_coeff (slope correction only), src/geometry_2d.f90:838-848
template <int = 0>
__global__ void k_gstag(DevParams P, DevPtrs D) {
  const int j = blockIdx.x * kTileX + (threadIdx.x & 63);
  const int k = blockIdx.y * kTileY + (threadIdx.x >> 6);
  if (k < P.ny && j <= P.nx) {
    double g;
    if (j == 0) g = D.gcoef[cidx(P, 0, k)];
    else if (j == P.nx) g = D.gcoef[cidx(P, P.nx - 1, k)];
    else g = 0.5 * (D.gcoef[cidx(P, j - 1, k)] + D.gcoef[cidx(P, j, k)]);
    D.gsx[(size_t)k * (P.pitch + 1) + j] = g;
  }
  if (k <= P.ny && j < P.nx) {
    double g;
    if (k == 0) g = D.gcoef[cidx(P, j, 0)];
    else if (k == P.ny) g = D.gcoef[cidx(P, j, P.ny - 1)];
    else g = 0.5 * (D.gcoef[cidx(P, j, k - 1)] + D.gcoef[cidx(P, j, k)]);
    D.gsy[(size_t)k * P.pitch + j] = g;
  }
}

// ---------------------------------------------------------------------------
// layout conversion: host Fortran (var, j, k) <-> device planes
// ---------------------------------------------------------------------------
template <int = 0>
__global__ void k_aos_to_planes(DevParams P, const double* __restrict__ src, int nvar,
                                double* __restrict__ dst) {
  const int j = blockIdx.x * kTileX + (threadIdx.x & 63);
  const int k = blockIdx.y * kTileY + (threadIdx.x >> 6);
  if (j >= P.nx || k >= P.ny) return;
  const size_t pl = plane(P), s = ((size_t)k * P.nx + j) * nvar;
  for (int i = 0; i < nvar; i++) dst[i * pl + cidx(P, j, k)] = src[s + i];
}

// state flags of the current state (after an upload)
template <int = 0>
__global__ void k_sflag(DevParams P, DevPtrs D) {
  const int j = blockIdx.x * kTileX + (threadIdx.x & 63);
  const int k = blockIdx.y * kTileY + (threadIdx.x >> 6);
  if (j >= P.nx || k >= P.ny) return;
  const size_t c = cidx(P, j, k), pl = plane(P);
  bool zero = true;
  for (int i = 0; i < P.nv; i++) zero = zero && (D.q[i * pl + c] == 0.0);
  D.sflag[c] = state_flags(D.q[c], zero);
}

// rows k0 .. k1-1 of nvar planes into the host layout (var, j, k - k0)
template <int = 0>
__global__ void k_planes_to_aos(DevParams P, const double* __restrict__ src, int nvar,
                                double* __restrict__ dst, int k0, int k1) {
  const int j = blockIdx.x * kTileX + (threadIdx.x & 63);
  const int k = k0 + blockIdx.y * kTileY + (threadIdx.x >> 6);
  if (j >= P.nx || k >= k1) return;
  const size_t pl = plane(P), s = ((size_t)(k - k0) * P.nx + j) * nvar;
  for (int i = 0; i < nvar; i++) dst[s + i] = src[i * pl + cidx(P, j, k)];
}

// qp(n_vars+2, j, k) of the current state (the module array qp after the
// post-step loop: src/IMEX_SfloW2D.f90:395-445), rows k0 .. k1-1
template <int NS, bool VP = false>
__global__ void k_qp_out(DevParams P, const double* __restrict__ q, double* __restrict__ dst,
                         int k0, int k1) {
  constexpr int NV = 4 + NS;
  const int j = blockIdx.x * kTileX + (threadIdx.x & 63);
  const int k = k0 + blockIdx.y * kTileY + (threadIdx.x >> 6);
  if (j >= P.nx || k >= k1) return;
  const size_t pl = plane(P), c = cidx(P, j, k), s = ((size_t)(k - k0) * P.nx + j) * (NV + 2);
  double qq[NV], qp[NV + 2];
#pragma unroll
  for (int i = 0; i < NV; i++) qq[i] = q[i * pl + c];
  phys_or_dry<NS, VP>(P, qq, qp);
#pragma unroll
  for (int i = 0; i < NV + 2; i++) dst[s + i] = qp[i];
}

// The stored physical variables of every cell (D.qpp), after the state was
// set from the host; the kernels that change the state keep them after that.
template <int NS, bool VP>
__global__ void k_qp_planes(DevParams P, DevPtrs D) {
  constexpr int NV = 4 + NS;
  const int j = blockIdx.x * kTileX + (threadIdx.x & 63);
  const int k = blockIdx.y * kTileY + (threadIdx.x >> 6);
  if (j >= P.nx || k >= P.ny) return;
  const size_t pl = plane(P), c = cidx(P, j, k);
  double q[NV], qp[NV + 2];
#pragma unroll
  for (int i = 0; i < NV; i++) q[i] = D.q[i * pl + c];
  phys_or_dry<NS, VP>(P, q, qp);
#pragma unroll
  for (int i = 0; i < NV + 2; i++) D.qpp[i * pl + c] = qp[i];
}

// H_crit_rel (src/inpout_2d.f90:3784-3791): with a = vonK / sqrt(f) + 1,
// (-a / max(W_0, W_-1)(-a e^-a) - 1) / 30
template <int = 0>
__global__ void k_vp_hcrit(double vonK, double sqrt_ff, double* out) {
  const double a = vonK / sqrt_ff + 1.0;
  const double arg = -a * det_exp(-a);
  const double l0 = lambertw0(arg), l1 = lambertwm1(arg);
  out[0] = 1.0 / 30.0 * (-(a / fmax(l0, l1)) - 1.0);
}

// Lambert W of n arguments on branch 0 or -1 (the device's)
template <int = 0>
__global__ void k_lambertw(int branch, long long n, const double* x, double* w) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    w[e] = branch == 0 ? lambertw0(x[e]) : lambertwm1(x[e]);
}

// Point conversions of n states (n_vars+2 or n_vars values each, packed):
// mode 0 qp_to_qc, 1 qc_to_qp (src/constitutive_2d.f90:1628-2140), 2
// eval_flux_coeffs of the physical states with the slopes bx, by
// (src/constitutive_2d.f90:2532-2823; 1 without profiles)
template <int NS, bool VP>
__global__ void k_convert(DevParams P, long long n, const double* __restrict__ in,
                          const double* __restrict__ bx, const double* __restrict__ by,
                          double* __restrict__ out, int mode) {
  constexpr int NV = 4 + NS;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    if (mode == 1) {
      double qc[NV], qp[NV + 2];
#pragma unroll
      for (int i = 0; i < NV; i++) qc[i] = in[e * NV + i];
      phys_or_dry<NS, VP>(P, qc, qp);
#pragma unroll
      for (int i = 0; i < NV + 2; i++) out[e * (NV + 2) + i] = qp[i];
      continue;
    }
    double qp[NV + 2], r[NV];
#pragma unroll
    for (int i = 0; i < NV + 2; i++) qp[i] = in[e * (NV + 2) + i];
    if (mode == 0) {
      qp2qc<NS, VP>(P, qp, r);
    } else {
#pragma unroll
      for (int i = 0; i < NV; i++) r[i] = 1.0;
      if (VP && qp[0] > kEps) {
        const Mixt m = mixt<NS, false>(P, qp);
        vp_flux_coeffs<NS>(P, qp, bx[e], by[e], vp_pre<NS>(P, qp[0], m.rho_c), m.rho_c, m.rho_m,
                           r);
      }
    }
#pragma unroll
    for (int i = 0; i < NV; i++) out[e * NV + i] = r[i];
  }
}

}  // namespace imexsw
