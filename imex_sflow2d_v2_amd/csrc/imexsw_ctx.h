// imexsw_ctx.h -- host-side state of a context and of an ensemble, shared by
// the translation units of the library: imexsw.hip (the C ABI) and
// imexsw_ns.hip, which is compiled once per number of solid classes and
// holds the launches of the kernels templated on it.  Splitting the
// instantiations over translation units lets them compile in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "imexsw_batch.h"
#include "imexsw_kernels.h"
#include "imexsw_stages.h"

using imexsw::DevParams;
using imexsw::DevPtrs;
using imexsw::MlpDev;
using imexsw::RkPtrs;
using imexsw::RkTableau;

enum KernelId { K_MASK = 0, K_STENCIL, K_LOCAL, K_TOPO, K_DTPOLICY, K_POST, K_EROSION, K_NK };

struct imexsw_ctx {
  imexsw_params p{};
  DevParams P{};
  DevPtrs D{};
  hipStream_t stream = nullptr;
  std::string err;
  int ns = 0, nv = 0;
  size_t plane = 0;
  double* stage = nullptr;  // layout-conversion staging
  size_t stage_bytes = 0;
  dim3 grid_tiles, grid_mask, grid_stag;
  int n_tiles = 0;  // work tiles (kStripW x kTileH)
  int n_cu = 256;   // compute units of the device
  int device = 0;   // the HIP device current at imexsw_create
  int strip_h = 16;         // tallest stencil strip (rows)
  int row_unit = 1;         // shortest stencil strip (rows)
  int target_strips = 2048; // strips wanted in flight
  // N_RK != 2: the tableau-driven stage pipeline and its stage planes
  bool general = false;
  RkTableau tab{};
  RkPtrs G{};
  bool d_valid = false;   // D holds the residual of the current q
  bool dt_hist = false;   // dt_old / dt_old_old initialised
  bool timing = false;
  double ms[K_NK] = {0};
  int64_t calls[K_NK] = {0};
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<void*> allocs;
  // learned warm start (guess planes in D.guess) and harvest buffer; the
  // closed-form Newton start
  bool warm = false;
  bool closed_start = false;
  MlpDev mlp{};
  // sparse steps (few computed cells): the row stencil and the fused local
  // kernel; chosen on the host from the computed-cell count of an earlier
  // step (both ways give the same bits)
  bool sparse = false;
  int64_t sparse_steps = 0;
  unsigned long long* h_cnt = nullptr;  // pinned: 2 x kSlots computed-cell counters
  double* h_done = nullptr;             // pinned: 2 copies of the run's done flag
  hipEvent_t cnt_ev[2] = {nullptr, nullptr};
  // captured kGraphSteps-step graphs of imexsw_advance (dense, sparse) and
  // what they captured
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  DevParams graph_P{};
  DevPtrs graph_D{};
  long graph_shape = 0;
  MlpDev graph_mlp{};
  bool graph_warm = false;
  bool graph_closed = false;
};

// time steps per captured graph of imexsw_advance
constexpr int kGraphSteps = 20;

struct imexsw_ensemble {
  std::vector<imexsw_ctx*> members;
  int ns = 0;
  hipStream_t stream = nullptr;
  std::string err;
  DevParams* d_params = nullptr;
  DevPtrs* d_ptrs = nullptr;
  int* d_tile_off = nullptr;
  int* d_tiles = nullptr;
  int* d_rows = nullptr;
  unsigned long long* d_counts = nullptr;
  int total_tiles = 0;
  int n_cu = 256;
  int strip_h = 8;
  int graph_steps = 25;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_strip_h = 0;
  double last_ms = 0.0;
};

namespace imexsw {

// The stencil and local kernels walk device-side work lists (listed tiles
// and tile rows) with a grid-stride loop, so the grid is sized for the
// device, not for the work: as many workgroups as fit on the device at once
// (one round, no queue of workgroups that find nothing to do), never more
// than the largest possible list needs.
template <typename K>
int resident_grid(K kernel, int threads, int n_cu, int items) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return std::max(1, std::min(items, per_cu * n_cu));
}

inline BatchLists batch_lists(imexsw_ensemble* e) {
  return BatchLists{e->d_tiles, e->d_rows, e->d_counts, e->d_tile_off, (int)e->members.size()};
}

// Launches of the kernels templated on the number of solid classes NS,
// defined in imexsw_ns.hip (one object per NS).
// rows: the six-wavefront row stencil (sparse steps); else strips.  D: the
// context's pointers, or a copy whose state and residual planes are a
// stage's (N_RK != 2)
template <int NS>
void launch_stencil(imexsw_ctx* c, const DevPtrs& D, int compute_dt, bool rows);
// fused: stage update, erosion and diagnostics of a row in one pass (sparse
// steps without the learned warm start); else up to two kernels
template <int NS>
void launch_local(imexsw_ctx* c, int ops, bool fused);
// qp of rows k0 .. k1-1 into dst in the host layout
template <int NS>
void launch_qp(imexsw_ctx* c, double* dst, int k0, int k1);
template <int NS>
void launch_qp_planes(imexsw_ctx* c);
// point conversions of n packed states (k_convert): mode 0 qp_to_qc,
// 1 qc_to_qp, 2 the profile flux coefficients
template <int NS>
void launch_convert(const DevParams& P, hipStream_t s, int n_cu, long long n, const double* in,
                    const double* bx, const double* by, double* out, int mode);
// the tableau-driven path (N_RK != 2, vertical profiles, rheology model 8):
// stage `st` of the tableau and the assembly (imexsw_stages.h)
template <int NS>
void launch_stage(imexsw_ctx* c, int st, int policy);
template <int NS>
void launch_assemble(imexsw_ctx* c, int policy);
// the ensemble's batched stencil and stage/erosion passes (NS = 1..3)
template <int NS>
void launch_batch_stencil(imexsw_ensemble* e);
template <int NS>
void launch_batch_local(imexsw_ensemble* e);

}  // namespace imexsw
