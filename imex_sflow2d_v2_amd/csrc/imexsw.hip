// imexsw.hip -- the C-ABI of include/imexsw.h on top of the gfx950 kernels.
//
// A context owns one grid's device state (structure-of-arrays planes), one
// HIP stream (default priority) and, when enabled, one pair of HIP events per
// launch for per-kernel timing.  Host arrays cross the ABI in the
// reference's Fortran layout and are converted on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "imexsw_ctx.h"
#include "imexsw_q2d.h"

using namespace imexsw;

namespace {

const char* kKernelNames[K_NK] = {"mask", "stencil", "local", "topo", "dtpolicy", "post", "erosion"};

thread_local std::string g_create_error;

}  // namespace

// Staging memory of imexsw_download_noalloc: one array per GPU, held until
// the process ends.  Host code uses its device address (resolved once per
// device by imexsw_create with hipGetSymbolAddress), never the name.
constexpr size_t kScratchDoubles = (size_t)2 << 20;  // 16 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};

}  // namespace

namespace {

int fail(imexsw_ctx* c, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  else g_create_error = buf;
  return -1;
}

// the device address of `scratch` on device `dev` (the current device),
// resolved once per device; on failure the message goes to c
int resolve_scratch(imexsw_ctx* c, int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(c, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(c, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
  g_scratch_ptr[dev] = static_cast<double*>(ptr);
  return 0;
}

#define HIPCHK(c, call)                                                            \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) return fail((c), "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int dalloc(imexsw_ctx* c, T** ptr, size_t count) {
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return fail(c, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
  e = hipMemsetAsync(p, 0, std::max<size_t>(count, 1) * sizeof(T), c->stream);
  if (e != hipSuccess) return fail(c, "hipMemsetAsync: %s", hipGetErrorString(e));
  c->allocs.push_back(p);
  *ptr = static_cast<T*>(p);
  return 0;
}

// event-bracketed launch for per-kernel timing on the context's stream
struct Timed {
  imexsw_ctx* c;
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(imexsw_ctx* ctx, int kid) : c(ctx), id(kid) {
    if (!c->timing) return;
    if (c->ev_pool.size() < 2) {
      hipEvent_t e1, e2;
      (void)hipEventCreate(&e1);
      (void)hipEventCreate(&e2);
      c->ev_pool.push_back(e1);
      c->ev_pool.push_back(e2);
    }
    a = c->ev_pool.back();
    c->ev_pool.pop_back();
    b = c->ev_pool.back();
    c->ev_pool.pop_back();
    (void)hipEventRecord(a, c->stream);
  }
  ~Timed() {
    if (!c->timing) return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({id, {a, b}});
  }
};

void collect_timing(imexsw_ctx* c) {
  for (auto& pe : c->pending) {
    float ms = 0.f;
    if (hipEventSynchronize(pe.second.second) == hipSuccess &&
        hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
      c->ms[pe.first] += ms;
      c->calls[pe.first] += 1;
    }
    c->ev_pool.push_back(pe.second.first);
    c->ev_pool.push_back(pe.second.second);
  }
  c->pending.clear();
}

int check_launch(imexsw_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, "launch %s: %s", what, hipGetErrorString(e));
  if (c->pending.size() > 4096) collect_timing(c);
  return 0;
}



#define NS_SWITCH(ns, CALL)        \
  switch (ns) {                    \
    case 0: CALL(0); break;        \
    case 1: CALL(1); break;        \
    case 2: CALL(2); break;        \
    case 3: CALL(3); break;        \
    case 4: CALL(4); break;        \
    case 5: CALL(5); break;        \
    default: CALL(6); break;       \
  }

// The stencil pass: CFL minimum (compute_dt) and divFlux - expl_terms of the
// state.  With N_RK != 2 the CFL pass writes its residual into the first
// stage's planes; stage >= 0 runs on that stage's state instead of q.
int run_stencil(imexsw_ctx* c, int compute_dt, bool rows = false, int stage = -1) {
  DevPtrs D = c->D;
  if (c->general) {
    const size_t sp = (size_t)c->nv * c->plane;
    D.D = c->G.Dx + (size_t)std::max(stage, 0) * sp;
    if (stage >= 0) {
      D.q = c->G.qrk + (size_t)stage * sp;
      D.qpp = nullptr;  // the stored physical variables are q's, not the stage's
    }
  }
  {
    Timed tm(c, K_STENCIL);
#define CALL(N) launch_stencil<N>(c, D, compute_dt, rows)
    NS_SWITCH(c->ns, CALL)
#undef CALL
  }
  if (stage < 0) c->d_valid = true;
  return check_launch(c, "k_stencil");
}

int run_local(imexsw_ctx* c, int ops, int kid, bool fused = false) {
  {
    Timed tm(c, kid);
#define CALL(N) launch_local<N>(c, ops, fused)
    NS_SWITCH(c->ns, CALL)
#undef CALL
  }
  return check_launch(c, "k_local");
}

// check_solve for N_RK != 2: seeds, N_RK dilations, the final pass
int run_mask_general(imexsw_ctx* c, int solve_all) {
  {
    Timed tm(c, K_MASK);
    const int grid = (int)std::min<size_t>((c->plane + 255) / 256, (size_t)c->n_cu * 8);
    hipLaunchKernelGGL(kg_mask_seed<>, dim3(grid), dim3(256), 0, c->stream, c->P, c->D,
                       c->G.ma, solve_all);
    uint8_t *a = c->G.ma, *b = c->G.mb;
    for (int i = 0; i < c->P.nrk; i++) {
      hipLaunchKernelGGL(kg_mask_dilate<>, dim3(grid), dim3(256), 0, c->stream, c->P, c->D.scal,
                         a, b);
      std::swap(a, b);
    }
    const dim3 fgrid(c->grid_mask.x, (c->grid_mask.y + kMaskGroup - 1) / kMaskGroup);
    hipLaunchKernelGGL(kg_mask_final<>, fgrid, dim3(64 * kMaskGroup), 0, c->stream, c->P, c->D,
                       a);
  }
  return check_launch(c, "kg_mask");
}

// partial: rescan only the dirty tiles (the step loop)
int run_mask(imexsw_ctx* c, int solve_all, int partial = 0) {
  if (c->general) return run_mask_general(c, solve_all);
  {
    Timed tm(c, K_MASK);
    const dim3 grid(c->grid_mask.x, (c->grid_mask.y + kMaskGroup - 1) / kMaskGroup);
    hipLaunchKernelGGL(k_mask<>, grid, dim3(64 * kMaskGroup), 0, c->stream, c->P, c->D,
                       solve_all, partial);
  }
  return check_launch(c, "k_mask");
}

// sparse: 0 every tile, 1 tiles next to a computed tile (flag test per
// tile), 2 the same tiles found from the tile list (few computed tiles)
int run_topo(imexsw_ctx* c, int sparse) {
  {
    Timed tm(c, K_TOPO);
    if (sparse == 2) {
      static int res = 0;
      if (!res) res = resident_grid(k_topo_list<>, 256, c->n_cu, 1 << 30);
      hipLaunchKernelGGL(k_topo_list<>, dim3(std::min(9 * c->n_tiles, res)), dim3(256), 0,
                         c->stream, c->P, c->D);
    } else {
      hipLaunchKernelGGL(k_topo<>, c->grid_mask, dim3(256), 0, c->stream, c->P, c->D, sparse);
    }
    if (c->D.gcoef)
      hipLaunchKernelGGL(k_gstag<>, c->grid_stag, dim3(256), 0, c->stream, c->P, c->D);
  }
  return check_launch(c, "k_topo");
}

// imex_RK_solver for N_RK != 2 (imexsw_stages.h), after the CFL pass has
// written the residual of q^n into the first stage's planes: the stages,
// each with its flux pass when omega_tilde > 0 (the first stage's is the CFL
// pass's when that stage is explicit, its state then being q^n), and the
// assembly.  policy: dt from the time-step policy, else the scalar dt.
int run_stages(imexsw_ctx* c, int policy) {
  const RkTableau& T = c->tab;
  for (int st = 0; st < T.nrk; st++) {
    {
      Timed tm(c, K_LOCAL);
#define CALL(N) launch_stage<N>(c, st, policy)
      NS_SWITCH(c->ns, CALL)
#undef CALL
    }
    if (check_launch(c, "kg_stage")) return -1;
    const bool same_as_cfl_pass = st == 0 && T.a_dirk[0][0] == 0.0;
    if (T.omega_tilde[st] > 0.0 && !same_as_cfl_pass && run_stencil(c, 0, false, st)) return -1;
  }
  {
    Timed tm(c, K_LOCAL);
#define CALL(N) launch_assemble<N>(c, policy)
    NS_SWITCH(c->ns, CALL)
#undef CALL
  }
  return check_launch(c, "kg_assemble");
}

// every stage state starts as q (what inactive cells keep)
int init_stages(imexsw_ctx* c) {
  const int grid = (int)std::min<size_t>((c->plane * c->nv + 255) / 256, (size_t)c->n_cu * 8);
  hipLaunchKernelGGL(kg_stage_init<>, dim3(grid), dim3(256), 0, c->stream, c->P, c->D.scal,
                     c->D.q, c->G.qrk, c->tab.nrk);
  return check_launch(c, "kg_stage_init");
}

// one time step of the main loop (src/IMEX_SfloW2D.f90:337-531), all on the
// context's stream.  The stage kernel applies the time-step policy itself;
// k_step_begin of the next step (commit_prev) or end_steps commits it and
// advances t.
int enqueue_step(imexsw_ctx* c, bool commit_prev) {
  {
    Timed tm(c, K_DTPOLICY);
    hipLaunchKernelGGL(k_step_begin<>, dim3(1), dim3(1), 0, c->stream, c->P, c->D,
                       commit_prev ? 1 : 0, 1);
  }
  if (c->general) {
    if (run_mask(c, 0) || init_stages(c)) return -1;
    if (run_stencil(c, c->P.cfl != -1.0)) return -1;
    if (run_stages(c, 1)) return -1;
    if (run_local(c, OP_EROSION | OP_POST | OP_POLICY, K_LOCAL)) return -1;
    if (c->P.topo && run_topo(c, 1)) return -1;
    return check_launch(c, "k_step_begin");
  }
  if (run_mask(c, 0, 1)) return -1;
  if (run_stencil(c, c->P.cfl != -1.0, c->sparse)) return -1;
  if (run_local(c, OP_RK | OP_EROSION | OP_POST | OP_POLICY, K_LOCAL, c->sparse)) return -1;
  if (c->P.topo && run_topo(c, c->sparse ? 2 : 1)) return -1;
  return check_launch(c, "k_step_begin");
}
int end_steps(imexsw_ctx* c) {
  hipLaunchKernelGGL(k_step_begin<>, dim3(1), dim3(1), 0, c->stream, c->P, c->D, 1, 0);
  return check_launch(c, "k_step_begin");
}

void drop_ctx_graph(imexsw_ctx* c) {
  for (int m = 0; m < 2; m++) {
    if (c->exec[m]) (void)hipGraphExecDestroy(c->exec[m]);
    if (c->graph[m]) (void)hipGraphDestroy(c->graph[m]);
    c->exec[m] = nullptr;
    c->graph[m] = nullptr;
  }
}

// Sparse steps when the stencil strips would be one row high (the
// strip-height rule of k_stencil applied to w computed cells), with N_RK = 2
// and no harvest (the harvest runs in the dense stage kernel).
bool want_sparse(const imexsw_ctx* c, unsigned long long w) {
  if (c->general || c->D.harvest || c->P.ny < 2) return false;
  return c->strip_h < 2 ||
         w / (unsigned long long)(kStripW * 2) < (unsigned long long)c->target_strips;
}

// Capture kGraphSteps steps of the sparse (m = 1) or dense (m = 0) form
// into a graph and instantiate it.
int capture_steps(imexsw_ctx* c, int m) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const bool sparse = c->sparse;
  c->sparse = (m == 1);
  int rc = 0;
  for (int i = 0; i < kGraphSteps && !rc; i++) rc = enqueue_step(c, i > 0);
  if (!rc) rc = end_steps(c);
  c->sparse = sparse;
  hipGraph_t g = nullptr;
  hipError_t err = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return -1;
  }
  if (err != hipSuccess) return fail(c, "hipStreamEndCapture: %s", hipGetErrorString(err));
  c->graph[m] = g;
  HIPCHK(c, hipGraphInstantiate(&c->exec[m], g, nullptr, nullptr, 0));
  return 0;
}

unsigned long long sum_slots(const unsigned long long* v) {
  unsigned long long s = 0;
  for (int i = 0; i < kSlots; i++) s += v[i];
  return s;
}

// Gauss-Legendre points and weights on [-1, 1] in ascending order (the
// layout gaulegf returns, src/geometry_2d.f90:2407-2449): Newton's method on
// P_n from cos(pi (i - 1/4) / (n + 1/2)), P_n and P_n' by the three-term
// recurrence, weights 2 / ((1 - x^2) P_n'(x)^2).
void gauss_legendre(int n, double* x, double* w) {
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; it++) {
      double pm = 1.0, pc = z;  // P_0, P_1
      for (int m = 2; m <= n; m++) {
        const double pn = ((2.0 * m - 1.0) * z * pc - (m - 1.0) * pm) / m;
        pm = pc;
        pc = pn;
      }
      if (n == 1) {
        pc = z;
        pm = 1.0;
      }
      dp = n * (z * pc - pm) / (z * z - 1.0);
      const double dz = pc / dp;
      z = z - dz;
      if (!(std::fabs(dz) > 2.220446049250313e-15)) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    w[n - 1 - i] = w[i];
  }
}

// H_crit_rel (src/inpout_2d.f90:3784-3791) with the device's Lambert W and
// exponential, so that host and device compare h / k_s against one value
int vp_h_crit(imexsw_ctx* c) {
  double* d = nullptr;
  if (dalloc(c, &d, 1)) return -1;
  hipLaunchKernelGGL(k_vp_hcrit<>, dim3(1), dim3(1), 0, c->stream, c->P.vonK, c->P.sqrt_ff, d);
  if (hipMemcpyAsync(&c->P.h_crit_rel, d, sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(c, "H_crit evaluation failed");
  return 0;
}

// The IMEX Runge-Kutta tableaux of allocate_solver_variables
// (src/solver_2d.f90:398-486): a_tilde_ij and omega_tilde of the explicit
// part, a_dirk_ij and omega of the implicit part.  N_RK = 3 is
// IMEX-SSP(3,3,2) (Pareschi & Russo 2005, Table IV), N_RK = 4 LRR(3,2,2)
// (Pareschi & Russo 2000, Table 3).  The assembly takes the last stage when
// the weights equal the last rows (src/solver_2d.f90:1341-1377).
void fill_tableau(int nrk, RkTableau& T) {
  std::memset(&T, 0, sizeof T);
  T.nrk = nrk;
  if (nrk == 1) {
    T.a_tilde[0][0] = 1.0;
    T.omega_tilde[0] = 1.0;
  } else if (nrk == 2) {
    T.a_tilde[1][0] = 1.0;
    T.omega_tilde[0] = 1.0;
    T.a_dirk[1][1] = 1.0;
    T.omega[1] = 1.0;
  } else if (nrk == 3) {
    T.a_tilde[1][0] = 0.5;
    T.a_tilde[2][0] = 0.5;
    T.a_tilde[2][1] = 0.5;
    for (int r = 0; r < 3; r++) T.omega_tilde[r] = 1.0 / 3.0;
    T.a_dirk[0][0] = 0.25;
    T.a_dirk[1][1] = 0.25;
    T.a_dirk[2][0] = 1.0 / 3.0;
    T.a_dirk[2][1] = 1.0 / 3.0;
    T.a_dirk[2][2] = 1.0 / 3.0;
    for (int r = 0; r < 3; r++) T.omega[r] = 1.0 / 3.0;
  } else if (nrk == 4) {
    T.a_tilde[1][0] = 0.5;
    T.a_tilde[2][0] = 1.0 / 3.0;
    T.a_tilde[3][1] = 1.0;
    T.omega_tilde[1] = 1.0;
    T.a_dirk[1][1] = 0.5;
    T.a_dirk[2][2] = 1.0 / 3.0;
    T.a_dirk[3][2] = 0.75;
    T.a_dirk[3][3] = 0.25;
    T.omega[2] = 0.75;
    T.omega[3] = 0.25;
  }
  double d1 = 0.0, d2 = 0.0;
  for (int r = 0; r < nrk; r++) {
    d1 += std::fabs(T.omega_tilde[r] - T.a_tilde[nrk - 1][r]);
    d2 += std::fabs(T.omega[r] - T.a_dirk[nrk - 1][r]);
  }
  T.stiff = (d1 == 0.0 && d2 == 0.0) ? 1 : 0;
}

int set_scalar(imexsw_ctx* c, int slot, double v) {
  HIPCHK(c, hipMemcpyAsync(c->D.scal + slot, &v, sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int get_scalars(imexsw_ctx* c, double* out) {
  HIPCHK(c, hipMemcpyAsync(out, c->D.scal, S_NSLOTS * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ensure_stage(imexsw_ctx* c, size_t bytes) {
  if (bytes <= c->stage_bytes) return 0;
  if (c->stage) (void)hipFree(c->stage);
  c->stage = nullptr;
  c->stage_bytes = 0;
  HIPCHK(c, hipMalloc(&c->stage, bytes));
  c->stage_bytes = bytes;
  return 0;
}

// host (nvar, nx, ny) Fortran array -> nvar device planes
int upload_aos(imexsw_ctx* c, const double* host, int nvar, double* dst) {
  const size_t n = (size_t)nvar * c->P.nx * c->P.ny;
  if (ensure_stage(c, n * sizeof(double))) return -1;
  HIPCHK(c, hipMemcpyAsync(c->stage, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_aos_to_planes<>, c->grid_tiles, dim3(256), 0, c->stream, c->P, c->stage, nvar, dst);
  if (check_launch(c, "k_aos_to_planes")) return -1;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int download_aos(imexsw_ctx* c, const double* src, int nvar, double* host) {
  const size_t n = (size_t)nvar * c->P.nx * c->P.ny;
  if (ensure_stage(c, n * sizeof(double))) return -1;
  hipLaunchKernelGGL(k_planes_to_aos<>, c->grid_tiles, dim3(256), 0, c->stream, c->P, src, nvar,
                     c->stage, 0, c->P.ny);
  if (check_launch(c, "k_planes_to_aos")) return -1;
  HIPCHK(c, hipMemcpyAsync(host, c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// host (nx, ny) Fortran plane <-> device plane with pitch
template <typename T>
int upload_plane(imexsw_ctx* c, const T* host, T* dst, int w, int h, size_t dpitch_elems) {
  HIPCHK(c, hipMemcpy2DAsync(dst, dpitch_elems * sizeof(T), host, (size_t)w * sizeof(T),
                             (size_t)w * sizeof(T), h, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
template <typename T>
int download_plane(imexsw_ctx* c, const T* src, T* host, int w, int h, size_t spitch_elems) {
  HIPCHK(c, hipMemcpy2DAsync(host, (size_t)w * sizeof(T), src, spitch_elems * sizeof(T),
                             (size_t)w * sizeof(T), h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace

extern "C" {

const char* imexsw_version(void) { return "imexsw 0.1 gfx950"; }

const char* imexsw_last_error(const imexsw_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int imexsw_create(const imexsw_params* params, imexsw_ctx** out) {
  if (!params || !out) return fail(nullptr, "null argument");
  *out = nullptr;
  const imexsw_params& p = *params;
  if (p.nx < 2 || p.ny < 1) return fail(nullptr, "grid must have nx >= 2 and ny >= 1");
  if (p.n_solid < 0 || p.n_solid > IMEXSW_MAX_SOLID) return fail(nullptr, "n_solid out of range");
  if (p.n_rk < 1 || p.n_rk > kMaxRK) return fail(nullptr, "N_RK must be 1, 2, 3 or 4 (got %d)", p.n_rk);
  if (p.rheology_flag && p.rheology_model != 1 && p.rheology_model != 6 && p.rheology_model != 8)
    return fail(nullptr, "rheology model %d is not supported (1, Voellmy-Salm; 6; 8)",
                p.rheology_model);
  // src/inpout_2d.f90:3350-3365
  if (p.rheology_flag && (p.rheology_model == 6 || p.rheology_model == 8) &&
      !(p.friction_factor >= 0.0))
    return fail(nullptr, "rheology model %d needs FRICTION_FACTOR >= 0 (got %g)", p.rheology_model,
                p.friction_factor);
  if (p.vertical_profiles_flag) {
    // src/inpout_2d.f90:3699-3775
    if (p.rheology_flag && p.rheology_model != 8)
      return fail(nullptr, "vertical profiles can be used only with RHEOLOGY_FLAG = F or "
                           "RHEOLOGY_MODEL = 8");
    if (p.energy_flag) return fail(nullptr, "vertical profiles can be used only with ENERGY_FLAG = F");
    if (!(p.vonK > 0.0)) return fail(nullptr, "vertical profiles: vonK must be > 0");
    if (!(p.k_s > 0.0)) return fail(nullptr, "vertical profiles: k_s must be > 0");
    if (!(p.Sc > 0.0)) return fail(nullptr, "vertical profiles: Sc must be > 0");
    if (p.n_quad < 1 || p.n_quad > IMEXSW_MAX_QUAD)
      return fail(nullptr, "vertical profiles: n_quad must be 1..%d (got %d)", IMEXSW_MAX_QUAD,
                  p.n_quad);
    // the shear velocity is SQRT(friction_factor) |u|: without a positive
    // friction factor every profile quantity is NaN
    if (!(p.friction_factor > 0.0))
      return fail(nullptr, "vertical profiles need FRICTION_FACTOR > 0 (got %g)",
                  p.friction_factor);
    // the ring overrides put a neighbour's face state on a face side whose
    // slope the profile coefficients read by position: not reproduced
    if (p.radial_source_flag && p.slope_correction_flag)
      return fail(nullptr, "vertical profiles with a radial source and slope correction are "
                           "not supported");
  }
  if (p.n_thickness_levels < 0 || p.n_dyn_pres_levels < 0 ||
      p.n_thickness_levels * p.n_dyn_pres_levels > 32 ||
      p.n_thickness_levels > IMEXSW_MAX_LEVELS || p.n_dyn_pres_levels > IMEXSW_MAX_LEVELS)
    return fail(nullptr, "hazard table larger than 32 entries");
  for (int i = 0; i < 4 + p.n_solid + 2; i++)
    if (p.limiter[i] < 0 || p.limiter[i] > 7) return fail(nullptr, "limiter out of range");

  imexsw_ctx* c = new imexsw_ctx();
  c->p = p;
  c->ns = p.n_solid;
  c->nv = 4 + p.n_solid;
  DevParams& P = c->P;
  P.nx = p.nx;
  P.ny = p.ny;
  P.pitch = (p.nx + 31) / 32 * 32;
  P.ns = p.n_solid;
  P.nv = c->nv;
  P.nrk = p.n_rk;
  P.energy = p.energy_flag;
  P.alpha = p.alpha_flag;
  P.slope = p.slope_correction_flag;
  P.curv = p.curvature_term_flag;
  P.topo = p.topo_change_flag;
  P.radial = p.radial_source_flag;
  P.rheo = p.rheology_flag;
  P.rheo_model = p.rheology_model;
  P.settling = p.settling_flag;
  P.entrain = p.entrainment_flag;
  P.erod_dep = p.erodible_deposit_flag;
  P.n_thk = p.n_thickness_levels;
  P.n_pdyn = p.n_thickness_levels > 0 ? p.n_dyn_pres_levels : 0;
  P.x0 = p.x0;
  P.y0 = p.y0;
  // init_grid (src/geometry_2d.f90:251-300)
  P.dx = (p.nx > 1) ? p.cell_size : 1.0;
  P.dy = (p.ny > 1) ? p.cell_size : 1.0;
  P.dx2 = P.dx / 2.0;
  P.dy2 = P.dy / 2.0;
  P.one_by_dx = 1.0 / P.dx;
  P.one_by_dy = 1.0 / P.dy;
  P.eps_sing = std::fmin(std::fmin(std::pow(P.dx, 4.0), std::pow(P.dy, 4.0)), 1.0e-10);
  {
    double e2 = P.eps_sing * P.eps_sing;
    P.eps_sing4 = e2 * e2;
  }
  P.cfl = p.cfl;
  P.max_dt = p.max_dt;
  P.reconstr_coeff = p.reconstr_coeff;
  P.theta = p.theta;
  for (int i = 0; i < IMEXSW_MAX_VARS + 2; i++) P.limiter[i] = p.limiter[i];
  P.lim_c = 0;
  for (int i = 0; i < IMEXSW_MAX_VARS + 2; i++)
    P.lim_c |= (p.limiter[i] >= 3 && p.limiter[i] <= 5) ? 1 : 0;
  for (int s = 0; s < 4; s++)
    for (int i = 0; i < IMEXSW_MAX_VARS; i++) {
      P.bc_flag[s][i] = p.bc_flag[s][i];
      P.bc_value[s][i] = p.bc_value[s][i];
    }
  for (int s = 0; s < 4; s++)
    P.bc_quiet[s] = (p.bc_flag[s][0] == 2 || p.bc_value[s][0] == 0.0) ? 1 : 0;
  P.grav = p.grav;
  P.pres = p.pres;
  P.inv_pres = 1.0 / p.pres;  // src/inpout_2d.f90:1330
  P.T_ambient = p.T_ambient;
  P.sp_heat_a = p.sp_heat_a;
  P.sp_gas_const_a = p.sp_gas_const_a;
  P.rho_a_amb = p.pres / (p.sp_gas_const_a * p.T_ambient);  // :1353
  P.kin_visc_c = p.kin_visc_c;
  for (int i = 0; i < IMEXSW_MAX_SOLID; i++) {
    P.rho_s[i] = p.rho_s[i];
    P.inv_rho_s[i] = (i < p.n_solid) ? 1.0 / p.rho_s[i] : 0.0;  // :1882
    P.diam_s[i] = p.diam_s[i];
    P.sp_heat_s[i] = p.sp_heat_s[i];
    P.erodible_fract[i] = p.erodible_fract[i];
    P.alphas_source[i] = p.alphas_source[i];
  }
  P.mu = p.mu;
  P.xi = p.xi;
  P.erosion_coeff = p.erosion_coeff;
  P.erodible_porosity = p.erodible_porosity;
  P.T_erodible = p.T_erodible;
  P.rho_c_sub = p.pres / (p.sp_gas_const_a * p.T_erodible);  // :1733
  P.coeff_porosity = (p.erosion_coeff > 0.0) ? p.erodible_porosity / (1.0 - p.erodible_porosity) : 0.0;
  P.alphastot_min = p.alphastot_min;
  P.max_packing = p.maximum_solid_packing;
  P.x_source = p.x_source;
  P.y_source = p.y_source;
  P.r_source = p.r_source;
  P.h_source = p.h_source;
  P.vel_source = p.vel_source;
  P.T_source = p.T_source;
  for (int i = 0; i < 4; i++) P.time_param[i] = p.time_param[i];
  P.pi_g = 4.0 * std::atan(1.0);
  for (int i = 0; i < IMEXSW_MAX_LEVELS; i++) {
    P.thickness_levels[i] = p.thickness_levels[i];
    P.dyn_pres_levels[i] = p.dyn_pres_levels[i];
  }
  P.h_cs = (4 + p.n_solid) * kEps;  // src/solver_2d.f90:273-274
  P.one_by_h = 1.0 / P.h_cs;
  P.ff = p.friction_factor;
  P.sqrt_ff = std::sqrt(p.friction_factor);
  P.vp = p.vertical_profiles_flag ? 1 : 0;
  P.bottom_conc = p.bottom_conc_flag ? 1 : 0;
  P.nq = P.vp ? p.n_quad : 0;
  P.vonK = p.vonK;
  P.k_s = p.k_s;
  P.Sc = p.Sc;
  P.h_crit_rel = 0.0;
  if (P.vp) gauss_legendre(P.nq, P.zq, P.wq);

  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    fail(nullptr, "hipStreamCreate failed");
    delete c;
    return -1;
  }
  c->plane = (size_t)P.pitch * P.ny;
  const size_t pl = c->plane;
  DevPtrs& D = c->D;
  int rc = 0;
  rc |= dalloc(c, &D.q, c->nv * pl);
  rc |= dalloc(c, &D.D, c->nv * pl);
  rc |= dalloc(c, &D.mask, pl);
  rc |= dalloc(c, &D.sflag, pl);
  rc |= dalloc(c, &D.B, pl);
  rc |= dalloc(c, &D.Bx, pl);
  rc |= dalloc(c, &D.By, pl);
  if (P.curv || P.slope) {
    rc |= dalloc(c, &D.Bxx, pl);
    rc |= dalloc(c, &D.Bxy, pl);
    rc |= dalloc(c, &D.Byy, pl);
  }
  if (P.slope) {
    rc |= dalloc(c, &D.gcoef, pl);
    rc |= dalloc(c, &D.dgdx, pl);
    rc |= dalloc(c, &D.gsx, (size_t)(P.pitch + 1) * P.ny);
    rc |= dalloc(c, &D.gsy, (size_t)P.pitch * (P.ny + 1));
  }
  if (c->ns > 0) {
    rc |= dalloc(c, &D.erodible, c->ns * pl);
    rc |= dalloc(c, &D.deposit, c->ns * pl);
    rc |= dalloc(c, &D.erosion, c->ns * pl);
  }
  rc |= dalloc(c, &D.hmax, pl);
  rc |= dalloc(c, &D.pdynmax, pl);
  rc |= dalloc(c, &D.modvel, pl);
  if (P.n_thk * P.n_pdyn > 0) rc |= dalloc(c, &D.vuln, pl);
  rc |= dalloc(c, &D.dt_bits, 1);
  rc |= dalloc(c, &D.scal, S_NSLOTS);
  rc |= dalloc(c, &D.counters, C_NCOUNTERS * kSlots);
  c->grid_tiles = dim3((P.nx + kTileX - 1) / kTileX, (P.ny + kTileY - 1) / kTileY);
  c->grid_mask = dim3((P.nx + kStripW - 1) / kStripW, (P.ny + kTileH - 1) / kTileH);
  c->grid_stag = dim3((P.nx + 1 + kTileX - 1) / kTileX, (P.ny + 1 + kTileY - 1) / kTileY);
  c->n_tiles = (int)(c->grid_mask.x * c->grid_mask.y);
  if (p.n_rk != 2 || P.vp || (P.rheo && P.rheo_model != 1)) {
    // the tableau-driven stage pipeline (N_RK != 2, and every N_RK with
    // the profile closures or rheology models 6 and 8): stage states,
    // residuals and implicit terms of every stage, two byte planes for the
    // mask dilation
    c->general = true;
    fill_tableau(p.n_rk, c->tab);
    const size_t sp = (size_t)c->nv * pl * p.n_rk;
    rc |= dalloc(c, &c->G.qrk, sp);
    rc |= dalloc(c, &c->G.Dx, sp);
    rc |= dalloc(c, &c->G.Im, sp);
    rc |= dalloc(c, &c->G.ma, pl);
    rc |= dalloc(c, &c->G.mb, pl);
  }
  rc |= dalloc(c, &D.tile_sums, 4 * (size_t)c->n_tiles * kTileH);
  rc |= dalloc(c, &D.tile_flag, (size_t)c->n_tiles);
  rc |= dalloc(c, &D.tile_dirty, (size_t)c->n_tiles);
  rc |= dalloc(c, &D.tile_nact, (size_t)c->n_tiles);
  rc |= dalloc(c, &D.tile_list, (size_t)c->n_tiles);
  rc |= dalloc(c, &D.row_list, (size_t)c->n_tiles * kTileH);
  rc |= dalloc(c, &D.cell_list, (size_t)P.nx * P.ny);
  // the stored physical variables of the state (vertical profiles only)
  if (c->P.vp) rc |= dalloc(c, &D.qpp, (c->nv + 2) * pl);
  // the face sides of the 2-D profile stencil
  if (c->P.vp && P.ny > 1) rc |= dalloc(c, &D.sides, 4 * (2 * c->nv + 3) * pl);
  c->strip_h = (P.ny > 1) ? kTileH : 1;
  c->row_unit = 1;
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
      c->device = dev;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
          ncu > 0)
        c->n_cu = ncu;
      if (!rc && resolve_scratch(c, dev)) rc = 1;
    }
  }
  if (!rc && P.vp && vp_h_crit(c)) rc = 1;
  if (rc) {
    g_create_error = c->err;
    imexsw_destroy(c);
    return -1;
  }
  // no tile has been scanned yet: the first mask scans all
  if (hipMemsetAsync(D.tile_dirty, 1, (size_t)c->n_tiles, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    fail(nullptr, "stream synchronise failed after allocation");
    imexsw_destroy(c);
    return -1;
  }
  *out = c;
  return 0;
}

void imexsw_destroy(imexsw_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  collect_timing(c);
  drop_ctx_graph(c);
  for (hipEvent_t e : c->cnt_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  if (c->h_done) (void)hipHostFree(c->h_done);
  if (c->D.harvest) (void)hipFree(c->D.harvest);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stage) (void)hipFree(c->stage);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int imexsw_set_topography(imexsw_ctx* c, const double* B_cent, const uint8_t* nodata) {
  if (!c || !B_cent) return fail(c, "null argument");
  const DevParams& P = c->P;
  if (upload_plane(c, B_cent, c->D.B, P.nx, P.ny, P.pitch)) return -1;
  if (nodata) {
    uint8_t* nd = nullptr;
    if (!c->D.nodata) {
      if (dalloc(c, &nd, c->plane)) return -1;
      c->D.nodata = nd;
    }
    if (upload_plane(c, nodata, const_cast<uint8_t*>(c->D.nodata), P.nx, P.ny, P.pitch)) return -1;
  }
  if (run_topo(c, 0)) return -1;
  c->d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int imexsw_set_geometry(imexsw_ctx* c, const double* B_cent, const double* Bx,
                        const double* By, const double* Bxx, const double* Bxy,
                        const double* Byy, const double* gcoef, const double* gsx,
                        const double* gsy, const double* dgdx, const uint8_t* nodata) {
  if (!c || !B_cent) return fail(c, "null argument");
  const DevParams& P = c->P;
  DevPtrs& D = c->D;
  if (upload_plane(c, B_cent, D.B, P.nx, P.ny, P.pitch)) return -1;
  if (Bx && upload_plane(c, Bx, D.Bx, P.nx, P.ny, P.pitch)) return -1;
  if (By && upload_plane(c, By, D.By, P.nx, P.ny, P.pitch)) return -1;
  if (D.Bxx) {
    if (Bxx && upload_plane(c, Bxx, D.Bxx, P.nx, P.ny, P.pitch)) return -1;
    if (Bxy && upload_plane(c, Bxy, D.Bxy, P.nx, P.ny, P.pitch)) return -1;
    if (Byy && upload_plane(c, Byy, D.Byy, P.nx, P.ny, P.pitch)) return -1;
  }
  if (D.gcoef) {
    if (gcoef && upload_plane(c, gcoef, D.gcoef, P.nx, P.ny, P.pitch)) return -1;
    if (dgdx && upload_plane(c, dgdx, D.dgdx, P.nx, P.ny, P.pitch)) return -1;
    if (gsx && upload_plane(c, gsx, D.gsx, P.nx + 1, P.ny, P.pitch + 1)) return -1;
    if (gsy && upload_plane(c, gsy, D.gsy, P.nx, P.ny + 1, P.pitch)) return -1;
  }
  if (nodata) {
    uint8_t* nd = nullptr;
    if (!D.nodata) {
      if (dalloc(c, &nd, c->plane)) return -1;
      D.nodata = nd;
    }
    if (upload_plane(c, nodata, const_cast<uint8_t*>(D.nodata), P.nx, P.ny, P.pitch)) return -1;
  }
  c->d_valid = false;
  return 0;
}

int imexsw_set_source(imexsw_ctx* c, const int32_t* source_cell, const uint8_t* dir_bits,
                      const double* vect) {
  if (!c || !source_cell || !dir_bits || !vect) return fail(c, "null argument");
  const DevParams& P = c->P;
  DevPtrs& D = c->D;
  // one source byte per cell (source_cell, and for ring cells the boundary
  // directions) and the direction vector of the first direction a ring
  // cell has (E, W, S, N order, the order the stencil tests them in)
  const size_t n = (size_t)P.nx * P.ny;
  std::vector<uint8_t> info(n);
  std::vector<double> ring(2 * n, 0.0);
  for (size_t i = 0; i < n; i++) {
    if (source_cell[i] < 0 || source_cell[i] > 2) return fail(c, "source_cell values must be 0, 1 or 2");
    uint8_t b = source_cell[i] == 1 ? SI_SRC : source_cell[i] == 2 ? SI_RING : 0;
    if (source_cell[i] == 2) {
      const uint8_t d = dir_bits[i];
      const int o = (d & IMEXSW_SRC_E) ? 0 : (d & IMEXSW_SRC_W) ? 2 : (d & IMEXSW_SRC_S) ? 4
                    : (d & IMEXSW_SRC_N) ? 6 : -1;
      b |= ((d & IMEXSW_SRC_E) ? SI_E : 0) | ((d & IMEXSW_SRC_W) ? SI_W : 0) |
           ((d & IMEXSW_SRC_S) ? SI_S : 0) | ((d & IMEXSW_SRC_N) ? SI_N : 0);
      if (o >= 0) {
        b |= SI_DIR;
        ring[2 * i] = vect[8 * i + o];
        ring[2 * i + 1] = vect[8 * i + o + 1];
      }
    }
    info[i] = b;
  }
  if (!D.sinfo) {
    uint8_t* si = nullptr;
    double* sr = nullptr;
    if (dalloc(c, &si, c->plane) || dalloc(c, &sr, 2 * c->plane)) return -1;
    D.sinfo = si;
    D.sring = sr;
  }
  if (upload_plane(c, info.data(), const_cast<uint8_t*>(D.sinfo), P.nx, P.ny, P.pitch)) return -1;
  if (upload_aos(c, ring.data(), 2, const_cast<double*>(D.sring))) return -1;
  HIPCHK(c, hipMemsetAsync(D.tile_dirty, 1, (size_t)c->n_tiles, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->d_valid = false;
  return 0;
}

int imexsw_upload_q(imexsw_ctx* c, const double* q) {
  if (!c || !q) return fail(c, "null argument");
  if (upload_aos(c, q, c->nv, c->D.q)) return -1;
  hipLaunchKernelGGL(k_sflag<>, c->grid_tiles, dim3(256), 0, c->stream, c->P, c->D);
  if (check_launch(c, "k_sflag")) return -1;
  if (c->D.qpp) {
#define CALL(N) launch_qp_planes<N>(c)
    NS_SWITCH(c->ns, CALL)
#undef CALL
    if (check_launch(c, "k_qp_planes")) return -1;
  }
  // every tile's neighbourhood may have changed: the next mask scans all
  HIPCHK(c, hipMemsetAsync(c->D.tile_dirty, 1, (size_t)c->n_tiles, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->d_valid = false;
  // a new state starts a new run: the next imexsw_advance takes its step
  // history from dt0 again
  c->dt_hist = false;
  return 0;
}

int imexsw_upload_erosion(imexsw_ctx* c, const double* erodible, const double* deposit,
                          const double* erosion) {
  if (!c) return fail(c, "null argument");
  if (c->ns == 0) return 0;
  const DevParams& P = c->P;
  if (erodible && upload_aos(c, erodible, c->ns, c->D.erodible)) return -1;
  for (int i = 0; i < c->ns; i++) {
    const size_t off = (size_t)i * P.nx * P.ny;
    if (deposit && upload_plane(c, deposit + off, c->D.deposit + i * c->plane, P.nx, P.ny, P.pitch)) return -1;
    if (erosion && upload_plane(c, erosion + off, c->D.erosion + i * c->plane, P.nx, P.ny, P.pitch)) return -1;
  }
  return 0;
}

int imexsw_check_solve(imexsw_ctx* c, int solve_all, double t, int32_t* solve_cells) {
  if (!c) return fail(c, "null argument");
  if (set_scalar(c, S_T, t)) return -1;
  if (set_scalar(c, S_DONE, 0.0)) return -1;
  HIPCHK(c, hipMemsetAsync(c->D.counters + C_COMPUTE_CELLS * kSlots, 0,
                           kSlots * sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipMemsetAsync(c->D.counters + C_TILES * kSlots, 0,
                           2 * kSlots * sizeof(unsigned long long), c->stream));  // tiles, rows
  HIPCHK(c, hipMemsetAsync(c->D.counters + C_CELLS * kSlots, 0,
                           kSlots * sizeof(unsigned long long), c->stream));
  if (run_mask(c, solve_all)) return -1;
  // the mask now reflects this call's solve_all: the next step's mask
  // scans every tile again
  HIPCHK(c, hipMemsetAsync(c->D.tile_dirty, 1, (size_t)c->n_tiles, c->stream));
  c->d_valid = false;
  if (solve_cells) {
    unsigned long long v[kSlots], n = 0;
    HIPCHK(c, hipMemcpyAsync(v, c->D.counters + C_SOLVE_CELLS * kSlots, sizeof v,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < kSlots; i++) n += v[i];
    *solve_cells = (int32_t)n;
  }
  return 0;
}

int imexsw_timestep(imexsw_ctx* c, double* dt) {
  if (!c || !dt) return fail(c, "null argument");
  // the procedures called one at a time run whatever state imexsw_advance
  // stopped in
  if (set_scalar(c, S_DONE, 0.0)) return -1;
  const unsigned long long init = (unsigned long long)(*(const long long*)&c->P.max_dt);
  HIPCHK(c, hipMemcpyAsync(c->D.dt_bits, &init, sizeof init, hipMemcpyHostToDevice, c->stream));
  if (run_stencil(c, c->P.cfl != -1.0)) return -1;
  unsigned long long bits = 0;
  HIPCHK(c, hipMemcpyAsync(&bits, c->D.dt_bits, sizeof bits, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double v;
  std::memcpy(&v, &bits, sizeof v);
  *dt = (c->P.cfl != -1.0) ? v : c->P.max_dt;
  return 0;
}

int imexsw_rk_step(imexsw_ctx* c, double t, double dt) {
  if (!c) return fail(c, "null argument");
  if (!(dt > 0.0)) return fail(c, "dt must be positive");
  if (set_scalar(c, S_T, t) || set_scalar(c, S_DT, dt) || set_scalar(c, S_DONE, 0.0)) return -1;
  if (c->general) {
    if (init_stages(c)) return -1;
    if (!c->d_valid && run_stencil(c, 0)) return -1;
    if (run_stages(c, 0)) return -1;
  } else {
    if (!c->d_valid && run_stencil(c, 0)) return -1;
    if (run_local(c, OP_RK, K_LOCAL)) return -1;
  }
  c->d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int imexsw_erosion_deposition(imexsw_ctx* c, double dt) {
  if (!c) return fail(c, "null argument");
  if (set_scalar(c, S_DT, dt) || set_scalar(c, S_DONE, 0.0)) return -1;
  if (run_local(c, OP_EROSION, K_EROSION)) return -1;
  c->d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int imexsw_topography_reconstruction(imexsw_ctx* c) {
  if (!c) return fail(c, "null argument");
  if (run_topo(c, 0)) return -1;
  c->d_valid = false;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int imexsw_post_step(imexsw_ctx* c, double sums[4]) {
  if (!c) return fail(c, "null argument");
  if (set_scalar(c, S_DONE, 0.0)) return -1;
  if (sums)
    HIPCHK(c, hipMemsetAsync(c->D.tile_sums, 0, 4 * (size_t)c->n_tiles * kTileH * sizeof(double),
                             c->stream));
  if (run_local(c, OP_POST | (sums ? OP_SUMS : 0), K_POST)) return -1;
  if (sums) {
    const size_t nt = (size_t)c->n_tiles * kTileH;
    std::vector<double> ts(4 * nt);
    HIPCHK(c, hipMemcpyAsync(ts.data(), c->D.tile_sums, ts.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int m = 0; m < 4; m++) sums[m] = 0.0;
    for (size_t t = 0; t < nt; t++)
      for (int m = 0; m < 4; m++) sums[m] += ts[4 * t + m];
  } else {
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

int imexsw_advance(imexsw_ctx* c, int32_t nsteps, double dt0, double* t_io, double t_end,
                   double t_output, int32_t* steps_done) {
  if (!c || !t_io) return fail(c, "null argument");
  double s[S_NSLOTS];
  if (get_scalars(c, s)) return -1;
  s[S_T] = *t_io;
  s[S_T_END] = t_end;
  s[S_T_OUT] = t_output;
  s[S_DONE] = 0.0;
  if (!c->dt_hist) {
    // src/IMEX_SfloW2D.f90:250-251
    s[S_DT_OLD] = dt0;
    s[S_DT_OLD_OLD] = dt0;
    c->dt_hist = true;
  }
  HIPCHK(c, hipMemcpyAsync(c->D.scal, s, sizeof s, hipMemcpyHostToDevice, c->stream));
  {
    // already at the output time (or the end): no step, nothing launched
    double t_out = t_output;
    if (t_end - t_out < 1.0e-7) t_out = t_end;
    if (!(*t_io < t_end) || !(*t_io < t_out) || nsteps <= 0) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      if (steps_done) *steps_done = 0;
      return 0;
    }
  }
  if (!c->h_cnt) {
    HIPCHK(c, hipHostMalloc((void**)&c->h_cnt, 2 * kSlots * sizeof(unsigned long long)));
    HIPCHK(c, hipHostMalloc((void**)&c->h_done, 2 * sizeof(double)));
    for (hipEvent_t& e : c->cnt_ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  const unsigned long long* d_cnt = c->D.counters + C_COMPUTE_CELLS * kSlots;
  unsigned long long steps0 = 0;
  HIPCHK(c, hipMemcpyAsync(&steps0, c->D.counters + C_STEPS * kSlots, sizeof steps0, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->h_cnt, d_cnt, kSlots * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->sparse = want_sparse(c, sum_slots(c->h_cnt));
  // Without per-kernel timing the steps run as replays of a captured graph
  // of kGraphSteps steps (one host call per replay instead of eight kernel
  // launches per step); the rest, and every step when timing, is launched
  // directly.  The graphs are recaptured when the context's constants or
  // pointers changed since the capture.  Each replay is followed by a copy of
  // its last step's computed-cell count and of the done flag: the count
  // decides between sparse and dense steps two replays later (so the device
  // always has a replay queued), and the flag ends the call once the run has
  // reached t_output or t_end (the steps after that are no-ops on the
  // device).
  int n = 0;
  bool stop = false;
  if (!c->timing && nsteps >= kGraphSteps) {
    if (std::memcmp(&c->graph_P, &c->P, sizeof(DevParams)) ||
        std::memcmp(&c->graph_D, &c->D, sizeof(DevPtrs)) ||
        c->graph_shape != c->strip_h * 1000003 + c->target_strips || c->graph_warm != c->warm ||
        c->graph_closed != c->closed_start ||
        std::memcmp(&c->graph_mlp, &c->mlp, sizeof(MlpDev))) {
      drop_ctx_graph(c);
      c->graph_P = c->P;
      c->graph_D = c->D;
      c->graph_shape = c->strip_h * 1000003 + c->target_strips;
      c->graph_warm = c->warm;
      c->graph_closed = c->closed_start;
      c->graph_mlp = c->mlp;
    }
    for (int r = 0; n + kGraphSteps <= nsteps; n += kGraphSteps, r++) {
      if (r >= 2) {
        HIPCHK(c, hipEventSynchronize(c->cnt_ev[r & 1]));
        if (c->h_done[r & 1] != 0.0) {
          stop = true;
          break;
        }
        c->sparse = want_sparse(c, sum_slots(c->h_cnt + (r & 1) * kSlots));
      }
      const int m = c->sparse ? 1 : 0;
      if (!c->exec[m] && capture_steps(c, m)) return -1;
      HIPCHK(c, hipGraphLaunch(c->exec[m], c->stream));
      if (c->sparse) c->sparse_steps += kGraphSteps;
      HIPCHK(c, hipMemcpyAsync(c->h_cnt + (r & 1) * kSlots, d_cnt, kSlots * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->h_done + (r & 1), c->D.scal + S_DONE, sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipEventRecord(c->cnt_ev[r & 1], c->stream));
    }
  }
  for (int i = 0; !stop && n < nsteps; n++, i++) {
    if (i > 0 && i % kGraphSteps == 0) {
      double done = 0.0;
      HIPCHK(c, hipMemcpyAsync(&done, c->D.scal + S_DONE, sizeof done, hipMemcpyDeviceToHost,
                               c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      if (done != 0.0) break;
    }
    if (enqueue_step(c, false) || end_steps(c)) return -1;
    if (c->sparse) c->sparse_steps++;
  }
  c->d_valid = false;
  if (get_scalars(c, s)) return -1;
  *t_io = s[S_T];
  unsigned long long steps1 = 0;
  HIPCHK(c, hipMemcpyAsync(&steps1, c->D.counters + C_STEPS * kSlots, sizeof steps1, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (steps_done) *steps_done = (int32_t)(steps1 - steps0);
  return 0;
}

int imexsw_download(imexsw_ctx* c, double* q, double* qp, double* hmax, double* pdynmax,
                    double* mod_vel_max, uint8_t* vuln) {
  if (!c) return fail(c, "null argument");
  const DevParams& P = c->P;
  if (q && download_aos(c, c->D.q, c->nv, q)) return -1;
  if (qp) {
    const size_t n = (size_t)(c->nv + 2) * P.nx * P.ny;
    if (ensure_stage(c, n * sizeof(double))) return -1;
#define CALL(N) launch_qp<N>(c, c->stage, 0, P.ny)
    NS_SWITCH(c->ns, CALL)
#undef CALL
    if (check_launch(c, "k_qp_out")) return -1;
    HIPCHK(c, hipMemcpyAsync(qp, c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  if (hmax && download_plane(c, c->D.hmax, hmax, P.nx, P.ny, P.pitch)) return -1;
  if (pdynmax && download_plane(c, c->D.pdynmax, pdynmax, P.nx, P.ny, P.pitch)) return -1;
  if (mod_vel_max && download_plane(c, c->D.modvel, mod_vel_max, P.nx, P.ny, P.pitch)) return -1;
  const int nl = P.n_thk * P.n_pdyn;
  if (vuln && nl > 0) {
    std::vector<uint32_t> bits((size_t)P.nx * P.ny);
    if (download_plane(c, c->D.vuln, bits.data(), P.nx, P.ny, P.pitch)) return -1;
    for (size_t cc = 0; cc < bits.size(); cc++)
      for (int i = 0; i < nl; i++) vuln[cc * nl + i] = (bits[cc] >> i) & 1u;
  }
  return 0;
}

int imexsw_download_noalloc(imexsw_ctx* c, double* q, double* qp) {
  if (!c) return fail(c, "null argument");
  const DevParams& P = c->P;
  double* ptr = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  if (!ptr) return fail(c, "the allocation-free download has no staging memory on device %d",
                        c->device);
  for (int which = 0; which < 2; which++) {
    double* host = which == 0 ? q : qp;
    if (!host) continue;
    const int nvar = which == 0 ? c->nv : c->nv + 2;
    const size_t row = (size_t)nvar * P.nx;
    if (row > kScratchDoubles)
      return fail(c, "imexsw_download_noalloc: a grid row holds %zu values, more than the %zu "
                     "of its staging memory (nx <= %zu for %d variables)",
                  row, kScratchDoubles, kScratchDoubles / nvar, nvar);
    const int rows = (int)std::min<size_t>((size_t)P.ny, kScratchDoubles / row);
    for (int k0 = 0; k0 < P.ny; k0 += rows) {
      const int k1 = std::min(P.ny, k0 + rows);
      if (which == 0) {
        const dim3 grid(c->grid_tiles.x, (k1 - k0 + kTileY - 1) / kTileY);
        hipLaunchKernelGGL(k_planes_to_aos<>, grid, dim3(256), 0, c->stream, P, c->D.q, nvar, ptr,
                           k0, k1);
      } else {
#define CALL(N) launch_qp<N>(c, ptr, k0, k1)
        NS_SWITCH(c->ns, CALL)
#undef CALL
      }
      if (check_launch(c, "imexsw_download_noalloc")) return -1;
      HIPCHK(c, hipMemcpyAsync(host + (size_t)k0 * row, ptr, (size_t)(k1 - k0) * row * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

int imexsw_download_erosion(imexsw_ctx* c, double* erodible, double* deposit, double* erosion,
                            double* B_cent) {
  if (!c) return fail(c, "null argument");
  const DevParams& P = c->P;
  if (c->ns > 0) {
    if (erodible && download_aos(c, c->D.erodible, c->ns, erodible)) return -1;
    for (int i = 0; i < c->ns; i++) {
      const size_t off = (size_t)i * P.nx * P.ny;
      if (deposit && download_plane(c, c->D.deposit + i * c->plane, deposit + off, P.nx, P.ny, P.pitch)) return -1;
      if (erosion && download_plane(c, c->D.erosion + i * c->plane, erosion + off, P.nx, P.ny, P.pitch)) return -1;
    }
  }
  if (B_cent && download_plane(c, c->D.B, B_cent, P.nx, P.ny, P.pitch)) return -1;
  return 0;
}

int imexsw_download_mask(imexsw_ctx* c, uint8_t* mask) {
  if (!c || !mask) return fail(c, "null argument");
  if (download_plane(c, c->D.mask, mask, c->P.nx, c->P.ny, c->P.pitch)) return -1;
  const size_t n = (size_t)c->P.nx * c->P.ny;
  for (size_t i = 0; i < n; i++) mask[i] &= M_ACTIVE;  // solve_mask only
  return 0;
}

int imexsw_get_dt(imexsw_ctx* c, double* dt) {
  if (!c || !dt) return fail(c, "null argument");
  double s[S_NSLOTS];
  if (get_scalars(c, s)) return -1;
  *dt = s[S_DT];
  return 0;
}

// n packed states through k_convert on the context's stream, with device
// buffers that live for the call (host set-up helpers, not step-path code)
static int convert_points(imexsw_ctx* c, int64_t n, const double* in, const double* bx,
                          const double* by, double* out, int mode) {
  if (!c) return fail(c, "null argument");
  if (n < 0) return fail(c, "n must be >= 0 (got %lld)", (long long)n);
  if (n == 0) return 0;
  if (!in || !out || (mode == 2 && (!bx || !by))) return fail(c, "null argument");
  const size_t nin = (size_t)n * (mode == 1 ? c->nv : c->nv + 2);
  const size_t nout = (size_t)n * (mode == 1 ? c->nv + 2 : c->nv);
  const size_t nb = mode == 2 ? (size_t)n : 0;
  double* buf = nullptr;
  HIPCHK(c, hipMalloc((void**)&buf, (nin + nout + 2 * nb) * sizeof(double)));
  double *din = buf, *dout = buf + nin, *dbx = buf + nin + nout, *dby = dbx + nb;
  int rc = 0;
  if (hipMemcpyAsync(din, in, nin * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      (nb && hipMemcpyAsync(dbx, bx, nb * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
                 hipSuccess) ||
      (nb && hipMemcpyAsync(dby, by, nb * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
                 hipSuccess)) {
    rc = fail(c, "upload of the states failed");
  } else {
#define CALL(N) launch_convert<N>(c->P, c->stream, c->n_cu, n, din, dbx, dby, dout, mode)
    NS_SWITCH(c->ns, CALL)
#undef CALL
    if (check_launch(c, "k_convert") ||
        hipMemcpyAsync(out, dout, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = fail(c, "point conversion failed");
  }
  (void)hipFree(buf);
  return rc;
}

int imexsw_qp_to_qc(imexsw_ctx* c, int64_t n, const double* qp, double* qc) {
  return convert_points(c, n, qp, nullptr, nullptr, qc, 0);
}

int imexsw_qc_to_qp(imexsw_ctx* c, int64_t n, const double* qc, double* qp) {
  return convert_points(c, n, qc, nullptr, nullptr, qp, 1);
}

int imexsw_flux_coeffs(imexsw_ctx* c, int64_t n, const double* qp, const double* Bx,
                       const double* By, double* shape) {
  return convert_points(c, n, qp, Bx, By, shape, 2);
}

int imexsw_lambertw(int32_t branch, int64_t n, const double* x, double* w) {
  if (branch != 0 && branch != -1) return fail(nullptr, "branch must be 0 or -1");
  if (n < 0 || (n > 0 && (!x || !w))) return fail(nullptr, "bad argument");
  if (n == 0) return 0;
  double* buf = nullptr;
  if (hipMalloc((void**)&buf, 2 * (size_t)n * sizeof(double)) != hipSuccess)
    return fail(nullptr, "hipMalloc failed");
  int rc = 0;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 1024);
  if (hipMemcpy(buf, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(nullptr, "upload failed");
  } else {
    hipLaunchKernelGGL(k_lambertw<>, dim3(grid), dim3(256), 0, nullptr, (int)branch,
                       (long long)n, buf, buf + n);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(w, buf + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(nullptr, "Lambert W evaluation failed");
  }
  (void)hipFree(buf);
  return rc;
}

int imexsw_gauss_legendre(int32_t n, double* x, double* w) {
  if (n < 1 || n > IMEXSW_MAX_QUAD || !x || !w)
    return fail(nullptr, "n must be 1..%d and x, w non-null", IMEXSW_MAX_QUAD);
  gauss_legendre(n, x, w);
  return 0;
}

int imexsw_stats_get(imexsw_ctx* c, imexsw_stats* st, int reset) {
  if (!c || !st) return fail(c, "null argument");
  std::vector<unsigned long long> raw((size_t)C_NCOUNTERS * kSlots);
  HIPCHK(c, hipMemcpyAsync(raw.data(), c->D.counters, raw.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  unsigned long long v[C_NCOUNTERS];
  for (int m = 0; m < C_NCOUNTERS; m++) {
    v[m] = 0;
    for (int i = 0; i < kSlots; i++) {
      const unsigned long long x = raw[(size_t)m * kSlots + i];
      v[m] = (m == C_ITMAX) ? std::max(v[m], x) : v[m] + x;
    }
  }
  std::memset(st, 0, sizeof *st);
  st->steps = (int64_t)v[C_STEPS];
  st->cell_updates_active = (int64_t)v[C_ACTIVE_UPD];
  st->cell_updates_computed = (int64_t)v[C_COMPUTE_UPD];
  st->cell_updates_grid = (int64_t)v[C_STEPS] * c->P.nx * c->P.ny;
  st->newton_cells = (int64_t)v[C_NEWTON_CELLS];
  st->newton_iters = (int64_t)v[C_NEWTON_ITERS];
  st->newton_nonconverged = (int64_t)v[C_NONCONV];
  st->nan_events = (int64_t)v[C_NAN];
  st->neg_h_events = (int64_t)v[C_NEGH];
  st->neg_solid_events = (int64_t)v[C_NEGS];
  st->sum_solid_events = (int64_t)v[C_SUMS];
  st->low_T_events = (int64_t)v[C_LOWT];
  st->line_search_evals = (int64_t)v[C_LS_EVALS];
  for (int b = 0; b < 8; b++) st->newton_iter_hist[b] = (int64_t)v[C_HIST0 + b];
  st->newton_iter_max = (int64_t)v[C_ITMAX];
  st->warm_cells = (int64_t)v[C_WARM_USED];
  st->warm_fallbacks = (int64_t)v[C_WARM_FALLBACK];
  st->sparse_steps = c->sparse_steps;
  st->solve_cells = (int32_t)v[C_SOLVE_CELLS];
  st->compute_cells = (int32_t)v[C_COMPUTE_CELLS];
  if (reset) {
    c->sparse_steps = 0;
    // keep the current solve_cells / compute_cells slots, clear the rest
    for (size_t i = 0; i < raw.size(); i++)
      if (i / kSlots != C_SOLVE_CELLS && i / kSlots != C_COMPUTE_CELLS) raw[i] = 0;
    HIPCHK(c, hipMemcpyAsync(c->D.counters, raw.data(), raw.size() * sizeof(unsigned long long),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

int imexsw_set_timing(imexsw_ctx* c, int enable) {
  if (!c) return fail(c, "null argument");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  c->timing = enable != 0;
  for (int i = 0; i < K_NK; i++) {
    c->ms[i] = 0.0;
    c->calls[i] = 0;
  }
  return 0;
}

int imexsw_kernel_times(imexsw_ctx* c, double* ms, int64_t* calls, int cap) {
  if (!c) return fail(c, "null argument");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  const int n = std::min(cap, (int)K_NK);
  for (int i = 0; i < n; i++) {
    if (ms) ms[i] = c->ms[i];
    if (calls) calls[i] = c->calls[i];
  }
  return n;
}

const char* imexsw_kernel_name(int index) {
  return (index >= 0 && index < K_NK) ? kKernelNames[index] : "";
}

int imexsw_set_tuning(imexsw_ctx* c, int max_strip_rows, int target_strips) {
  if (!c) return fail(c, "null argument");
  const int m = max_strip_rows;
  if (!(m == 1 || m == 2 || m == 4 || m == 8 || m == 16))
    return fail(c, "max_strip_rows must be 1, 2, 4, 8 or 16 (got %d)", m);
  if (target_strips < 1) return fail(c, "target_strips must be >= 1 (got %d)", target_strips);
  if (c->P.ny > 1) c->strip_h = m;
  c->target_strips = target_strips;
  return 0;
}

int imexsw_sync(imexsw_ctx* c) {
  if (!c) return fail(c, "null argument");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int imexsw_q2d_write(const char* path, int32_t nv, int32_t nx, int32_t ny,
                     const double* x_comp, const double* y_comp, const double* q) {
  if (!path || !x_comp || !y_comp || !q || nv < 1 || nx < 1 || ny < 1) return -1;
  FILE* fp = std::fopen(path, "w");
  if (!fp) return -1;
  std::string rec;
  char fld[21];
  for (int k = 0; k < ny; k++) {
    for (int j = 0; j < nx; j++) {
      rec.clear();
      fortran_e20_12(x_comp[j], fld);
      rec += fld;
      fortran_e20_12(y_comp[k], fld);
      rec += fld;
      const double* qc = q + ((size_t)k * nx + j) * nv;
      for (int i = 0; i < nv; i++) {
        fortran_e20_12(std::fabs(qc[i]) < 1.0e-20 ? 0.0 : qc[i], fld);
        rec += fld;
      }
      rec += '\n';
      std::fputs(rec.c_str(), fp);
    }
    std::fputs("  \n", fp);
  }
  std::fputs("  \n  \n", fp);
  return std::fclose(fp) == 0 ? 0 : -1;
}

int imexsw_q2d_read(const char* path, int32_t nv, int32_t n_solid, int32_t nx, int32_t ny,
                    double* q) {
  if (!path || !q || nv < 1 || nx < 1 || ny < 1 || n_solid < 0 || 4 + n_solid > nv) return -1;
  FILE* fp = std::fopen(path, "r");
  if (!fp) return -1;
  std::string rec;
  int rc = 0;
  for (int k = 0; k < ny && !rc; k++) {
    for (int j = 0; j < nx && !rc; j++) {
      if (!read_record(fp, rec)) {
        rc = -1;
        break;
      }
      rec.resize((size_t)20 * (nv + 2), ' ');  // a short record pads with blanks
      double* qc = q + ((size_t)k * nx + j) * nv;
      double xy;
      if (!fortran_e_field(rec.c_str(), 20, &xy) || !fortran_e_field(rec.c_str() + 20, 20, &xy)) {
        rc = -1;
        break;
      }
      for (int i = 0; i < nv; i++)
        if (!fortran_e_field(rec.c_str() + 20 * (i + 2), 20, &qc[i])) rc = -1;
      if (rc) break;
      if (qc[0] <= 0.0)
        for (int i = 0; i < nv; i++) qc[i] = 0.0;
      for (int i = 4; i < 4 + n_solid; i++)
        if (qc[i] <= 0.0) qc[i] = 0.0;
    }
    if (!rc && !read_record(fp, rec) && k < ny - 1) rc = -1;  // the row's closing record
  }
  std::fclose(fp);
  return rc;
}

}  // extern "C"

// ===========================================================================
// ensembles: many member grids advanced by one set of launches per step,
// the steps captured in a HIP graph
// ===========================================================================

namespace {

int efail(imexsw_ensemble* e, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  else g_create_error = buf;
  return -1;
}

#define EHIPCHK(e, call)                                                            \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) return efail((e), "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)



template <int NS>
void launch_batch_step(imexsw_ensemble* e) {
  const BatchLists L = batch_lists(e);
  const int M = L.members;
  const dim3 mgrid((M + 255) / 256);
  hipLaunchKernelGGL(kb_reset<>, mgrid, dim3(256), 0, e->stream, e->d_params, e->d_ptrs, L);
  hipLaunchKernelGGL(kb_mask<>, dim3((e->total_tiles + kMaskGroup - 1) / kMaskGroup),
                     dim3(64 * kMaskGroup), 0, e->stream, e->d_params, e->d_ptrs, L);
  launch_batch_stencil<NS>(e);
  hipLaunchKernelGGL(kb_dtpolicy<>, mgrid, dim3(256), 0, e->stream, e->d_params, e->d_ptrs, M);
  launch_batch_local<NS>(e);
  hipLaunchKernelGGL(kb_advance<>, mgrid, dim3(256), 0, e->stream, e->d_ptrs, M);
}

int batch_step(imexsw_ensemble* e) {
  switch (e->ns) {
    case 1: launch_batch_step<1>(e); break;
    case 2: launch_batch_step<2>(e); break;
    case 3: launch_batch_step<3>(e); break;
    default: return efail(e, "ensembles support 1 to 3 solid classes");
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return efail(e, "batch launch: %s", hipGetErrorString(err));
  return 0;
}

void drop_graph(imexsw_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->graph_strip_h = 0;
}

// graph of graph_steps batched steps (kernels only: no memset nodes)
int build_graph(imexsw_ensemble* e) {
  drop_graph(e);
  EHIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < e->graph_steps && !rc; s++) rc = batch_step(e);
  hipGraph_t g = nullptr;
  hipError_t err = hipStreamEndCapture(e->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (err != hipSuccess) return efail(e, "hipStreamEndCapture: %s", hipGetErrorString(err));
  e->graph = g;
  EHIPCHK(e, hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0));
  e->graph_strip_h = e->strip_h;
  return 0;
}

// member scalars for a run: t, t_end, t_output = t_end, not done, dt history
int seed_members(imexsw_ensemble* e, double dt0, const double* t, double t_end) {
  for (size_t m = 0; m < e->members.size(); m++) {
    imexsw_ctx* c = e->members[m];
    double s[S_NSLOTS];
    if (get_scalars(c, s)) return efail(e, "member %zu: %s", m, c->err.c_str());
    s[S_T] = t ? t[m] : s[S_T];
    s[S_T_END] = t_end;
    s[S_T_OUT] = t_end;
    s[S_DONE] = 0.0;
    if (!c->dt_hist) {
      s[S_DT_OLD] = dt0;
      s[S_DT_OLD_OLD] = dt0;
      c->dt_hist = true;
    }
    EHIPCHK(e, hipMemcpy(c->D.scal, s, sizeof s, hipMemcpyHostToDevice));
    c->d_valid = false;
  }
  return 0;
}

int run_steps(imexsw_ensemble* e, int nsteps, bool use_graph) {
  if (use_graph && nsteps >= e->graph_steps &&
      (!e->exec || e->graph_strip_h != e->strip_h) && build_graph(e))
    return -1;
  int done = 0;
  while (use_graph && nsteps - done >= e->graph_steps) {
    EHIPCHK(e, hipGraphLaunch(e->exec, e->stream));
    done += e->graph_steps;
  }
  for (; done < nsteps; done++)
    if (batch_step(e)) return -1;
  return 0;
}

}  // namespace

extern "C" {

int imexsw_ensemble_create(int32_t n_members, const imexsw_params* params,
                           imexsw_ensemble** out) {
  if (!params || !out || n_members < 1) return efail(nullptr, "null argument or no members");
  if (n_members > (1 << (31 - kMemberShift)))
    return efail(nullptr, "at most %d members", 1 << (31 - kMemberShift));
  *out = nullptr;
  const int ns = params[0].n_solid;
  if (ns < 1 || ns > 3) return efail(nullptr, "ensembles support 1 to 3 solid classes");
  for (int m = 0; m < n_members; m++) {
    const imexsw_params& p = params[m];
    if (p.n_solid != ns) return efail(nullptr, "member %d: all members need the same n_solid", m);
    if (p.ny < 2) return efail(nullptr, "member %d: ensembles need 2-D grids", m);
    if (p.n_rk != 2) return efail(nullptr, "member %d: ensembles run N_RK = 2", m);
    if (p.vertical_profiles_flag || (p.rheology_flag && p.rheology_model != 1))
      return efail(nullptr, "member %d: ensembles run without vertical profiles and with "
                            "rheology model 1 or none", m);
    if (p.topo_change_flag || p.radial_source_flag || p.slope_correction_flag || p.curvature_term_flag)
      return efail(nullptr, "member %d: ensembles run the flat-source-free model (no topography "
                            "change, radial source, slope correction or curvature)", m);
    const long tiles = (long)((p.nx + kStripW - 1) / kStripW) * ((p.ny + kTileH - 1) / kTileH);
    if (tiles * kTileH > kIdMask) return efail(nullptr, "member %d: grid too large for an ensemble", m);
  }
  imexsw_ensemble* e = new imexsw_ensemble();
  e->ns = ns;
  std::vector<DevParams> hp(n_members);
  std::vector<DevPtrs> hd(n_members);
  std::vector<int> off(n_members + 1, 0);
  for (int m = 0; m < n_members; m++) {
    imexsw_ctx* c = nullptr;
    if (imexsw_create(&params[m], &c)) {
      std::string msg = g_create_error;
      imexsw_ensemble_destroy(e);
      return efail(nullptr, "member %d: %s", m, msg.c_str());
    }
    e->members.push_back(c);
    hp[m] = c->P;
    hd[m] = c->D;
    off[m + 1] = off[m] + c->n_tiles;
    e->n_cu = c->n_cu;
  }
  e->total_tiles = off[n_members];
  int rc = 0;
  auto alloc = [&](void** p, size_t bytes) {
    if (!rc && hipMalloc(p, std::max<size_t>(bytes, 1)) != hipSuccess) rc = 1;
  };
  alloc((void**)&e->d_params, sizeof(DevParams) * n_members);
  alloc((void**)&e->d_ptrs, sizeof(DevPtrs) * n_members);
  alloc((void**)&e->d_tile_off, sizeof(int) * (n_members + 1));
  alloc((void**)&e->d_tiles, sizeof(int) * e->total_tiles);
  alloc((void**)&e->d_rows, sizeof(int) * (size_t)e->total_tiles * kTileH);
  alloc((void**)&e->d_counts, sizeof(unsigned long long) * 2);
  if (!rc && hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) rc = 1;
  if (!rc && (hipMemcpy(e->d_params, hp.data(), sizeof(DevParams) * n_members, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(e->d_ptrs, hd.data(), sizeof(DevPtrs) * n_members, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(e->d_tile_off, off.data(), sizeof(int) * (n_members + 1), hipMemcpyHostToDevice) != hipSuccess))
    rc = 1;
  if (rc) {
    imexsw_ensemble_destroy(e);
    return efail(nullptr, "ensemble allocation failed");
  }
  *out = e;
  return 0;
}

void imexsw_ensemble_destroy(imexsw_ensemble* e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graph(e);
  for (void* p : {(void*)e->d_params, (void*)e->d_ptrs, (void*)e->d_tile_off, (void*)e->d_tiles,
                  (void*)e->d_rows, (void*)e->d_counts})
    if (p) (void)hipFree(p);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  for (imexsw_ctx* c : e->members) imexsw_destroy(c);
  delete e;
}

const char* imexsw_ensemble_last_error(const imexsw_ensemble* e) {
  return e ? e->err.c_str() : g_create_error.c_str();
}

imexsw_ctx* imexsw_ensemble_member(imexsw_ensemble* e, int32_t i) {
  if (!e || i < 0 || i >= (int32_t)e->members.size()) return nullptr;
  return e->members[i];
}

int imexsw_ensemble_set_tuning(imexsw_ensemble* e, int strip_rows, int graph_steps) {
  if (!e) return efail(e, "null argument");
  if (!(strip_rows == 4 || strip_rows == 8 || strip_rows == 16))
    return efail(e, "strip_rows must be 4, 8 or 16 (got %d)", strip_rows);
  if (graph_steps < 1 || graph_steps > 1000) return efail(e, "graph_steps must be 1..1000");
  e->strip_h = strip_rows;
  if (graph_steps != e->graph_steps) drop_graph(e);
  e->graph_steps = graph_steps;
  return 0;
}

int imexsw_ensemble_advance(imexsw_ensemble* e, int32_t nsteps, double dt0, double* t,
                            double t_end, int32_t* steps_done) {
  if (!e || nsteps < 0) return efail(e, "bad argument");
  if (seed_members(e, dt0, t, t_end)) return -1;
  // member 0's step count before the call (steps_done is this call's)
  unsigned long long before = 0;
  if (!e->members.empty()) {
    EHIPCHK(e, hipMemcpyAsync(&before, e->members[0]->D.counters + C_STEPS * kSlots,
                              sizeof before, hipMemcpyDeviceToHost, e->stream));
    EHIPCHK(e, hipStreamSynchronize(e->stream));
  }
  hipEvent_t a, b;
  EHIPCHK(e, hipEventCreate(&a));
  EHIPCHK(e, hipEventCreate(&b));
  (void)hipEventRecord(a, e->stream);
  if (run_steps(e, nsteps, true)) return -1;
  (void)hipEventRecord(b, e->stream);
  EHIPCHK(e, hipStreamSynchronize(e->stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  e->last_ms = ms;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  int64_t steps0 = -1;
  for (size_t m = 0; m < e->members.size(); m++) {
    imexsw_ctx* c = e->members[m];
    double s[S_NSLOTS];
    if (get_scalars(c, s)) return efail(e, "member %zu: %s", m, c->err.c_str());
    if (t) t[m] = s[S_T];
    if (m == 0) {
      unsigned long long v = 0;
      EHIPCHK(e, hipMemcpy(&v, c->D.counters + C_STEPS * kSlots, sizeof v, hipMemcpyDeviceToHost));
      steps0 = (int64_t)(v - before);
    }
  }
  if (steps_done) *steps_done = (int32_t)steps0;
  return 0;
}

// Try each strip height for `steps_per_trial` real steps (results do not
// depend on the shape, so the trials advance the run) and keep the fastest.
int imexsw_ensemble_autotune(imexsw_ensemble* e, int32_t steps_per_trial, double dt0,
                             double* t, double t_end, int* best_strip_rows) {
  if (!e || steps_per_trial < 1) return efail(e, "bad argument");
  if (seed_members(e, dt0, t, t_end)) return -1;
  const int shapes[3] = {4, 8, 16};
  float best = 1e30f;
  int best_h = e->strip_h;
  hipEvent_t a, b;
  EHIPCHK(e, hipEventCreate(&a));
  EHIPCHK(e, hipEventCreate(&b));
  for (int h : shapes) {
    e->strip_h = h;
    (void)hipEventRecord(a, e->stream);
    if (run_steps(e, steps_per_trial, false)) return -1;
    (void)hipEventRecord(b, e->stream);
    EHIPCHK(e, hipStreamSynchronize(e->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    if (ms < best) {
      best = ms;
      best_h = h;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  e->strip_h = best_h;
  if (best_strip_rows) *best_strip_rows = best_h;
  if (t)
    for (size_t m = 0; m < e->members.size(); m++) {
      double s[S_NSLOTS];
      if (get_scalars(e->members[m], s)) return efail(e, "member %zu", m);
      t[m] = s[S_T];
    }
  return 0;
}

double imexsw_ensemble_last_ms(const imexsw_ensemble* e) { return e ? e->last_ms : 0.0; }

}  // extern "C"

// ===========================================================================
// learned Newton warm start: the MLP, its training on the device, and the
// harvest of (state entering Newton, converged solution) pairs
// ===========================================================================
struct imexsw_mlp {
  std::string err;
  hipStream_t stream = nullptr;
  float* w = nullptr;     // fp32 master parameters
  short* wb = nullptr;    // bf16 copies
  float* norm = nullptr;  // kMlpNorm
  float* m1 = nullptr;
  float* m2 = nullptr;
  float* grad = nullptr;
  float* loss = nullptr;
  int cap = 0;            // training buffer capacity (samples, multiple of 64)
  float* tz = nullptr;    // 3 * cap * kMlpH
  short* ta = nullptr;    // cap * 16 + 3 * cap * kMlpH
  short* tdz = nullptr;   // 3 * cap * kMlpH + cap * 16
  float* x = nullptr;     // data buffers (standardised): n * 16, n * 2
  float* t = nullptr;
  size_t data_cap = 0;
  long steps = 0;         // Adam steps taken
  float host_norm[kMlpNorm];
};

namespace {

int mfail(imexsw_mlp* m, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (m) m->err = buf;
  else g_create_error = buf;
  return -1;
}

#define MHIPCHK(m, call)                                                            \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) return mfail((m), "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

MlpDev mlp_dev(imexsw_mlp* m) { return MlpDev{m->w, m->wb, m->norm}; }

int mlp_reserve(imexsw_mlp* m, int n) {
  const int cap = (n + 63) / 64 * 64;
  if (cap <= m->cap) return 0;
  for (void* p : {(void*)m->tz, (void*)m->ta, (void*)m->tdz})
    if (p) (void)hipFree(p);
  m->tz = nullptr;
  m->ta = nullptr;
  m->tdz = nullptr;
  MHIPCHK(m, hipMalloc(&m->tz, sizeof(float) * 3 * cap * kMlpH));
  MHIPCHK(m, hipMalloc(&m->ta, sizeof(short) * ((size_t)cap * kMlpInPad + 3 * (size_t)cap * kMlpH)));
  MHIPCHK(m, hipMalloc(&m->tdz, sizeof(short) * (3 * (size_t)cap * kMlpH + (size_t)cap * kMlpOutPad)));
  m->cap = cap;
  return 0;
}

MlpTrainBuf mlp_buf(imexsw_mlp* m) {
  MlpTrainBuf T;
  for (int k = 0; k < 3; k++) T.z[k] = m->tz + (size_t)k * m->cap * kMlpH;
  T.a[0] = m->ta;
  for (int k = 1; k < 4; k++) T.a[k] = m->ta + (size_t)m->cap * kMlpInPad + (size_t)(k - 1) * m->cap * kMlpH;
  for (int k = 0; k < 3; k++) T.dz[k] = m->tdz + (size_t)k * m->cap * kMlpH;
  T.dz[3] = m->tdz + 3 * (size_t)m->cap * kMlpH;
  T.grad = m->grad;
  T.loss = m->loss;
  return T;
}

int mlp_data(imexsw_mlp* m, size_t n) {
  if (n <= m->data_cap) return 0;
  if (m->x) (void)hipFree(m->x);
  if (m->t) (void)hipFree(m->t);
  m->x = nullptr;
  m->t = nullptr;
  MHIPCHK(m, hipMalloc(&m->x, sizeof(float) * kMlpInPad * std::max<size_t>(n, 1)));
  MHIPCHK(m, hipMalloc(&m->t, sizeof(float) * kMlpOut * std::max<size_t>(n, 1)));
  m->data_cap = n;
  return 0;
}

// raw features / targets -> standardised, inputs padded to 16
void standardise(const float* norm, const float* xr, const float* tr, size_t n,
                 std::vector<float>& xs, std::vector<float>& ts, size_t xstride) {
  xs.assign(n * kMlpInPad, 0.f);
  if (tr) ts.assign(n * kMlpOut, 0.f);
  for (size_t s = 0; s < n; s++) {
    for (int f = 0; f < kMlpIn; f++)
      xs[s * kMlpInPad + f] = (xr[s * xstride + f] - norm[f]) / norm[kMlpIn + f];
    if (tr)
      for (int o = 0; o < kMlpOut; o++)
        ts[s * kMlpOut + o] =
            (tr[s * xstride + o] - norm[2 * kMlpIn + o]) / norm[2 * kMlpIn + kMlpOut + o];
  }
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double urand(uint64_t& s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
double nrand(uint64_t& s) {
  const double u = urand(s), v = urand(s);
  return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
}

}  // namespace

extern "C" {

int imexsw_mlp_num_params(void) { return kMlpParams; }

int imexsw_mlp_create(imexsw_mlp** out, uint64_t seed) {
  if (!out) return mfail(nullptr, "null argument");
  *out = nullptr;
  imexsw_mlp* m = new imexsw_mlp();
  int rc = 0;
  auto al = [&](void** p, size_t bytes) {
    if (!rc && hipMalloc(p, bytes) != hipSuccess) rc = 1;
  };
  al((void**)&m->w, sizeof(float) * kMlpParams);
  al((void**)&m->wb, sizeof(short) * kMlpParams);
  al((void**)&m->norm, sizeof(float) * kMlpNorm);
  al((void**)&m->m1, sizeof(float) * kMlpParams);
  al((void**)&m->m2, sizeof(float) * kMlpParams);
  al((void**)&m->grad, sizeof(float) * kMlpParams);
  al((void**)&m->loss, sizeof(float));
  if (!rc && hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) rc = 1;
  if (rc) {
    imexsw_mlp_destroy(m);
    return mfail(nullptr, "MLP allocation failed");
  }
  // He-normal initial weights (deterministic in seed), zero biases; padded
  // input columns and output rows stay zero
  std::vector<float> w(kMlpParams, 0.f);
  uint64_t st = seed ^ 0x5DEECE66Dull;
  auto fill = [&](int off, int rows, int cols, int use_rows, int use_cols, int fan_in) {
    const double sd = std::sqrt(2.0 / fan_in);
    for (int r = 0; r < use_rows; r++)
      for (int cc = 0; cc < use_cols; cc++) w[off + r * cols + cc] = (float)(sd * nrand(st));
  };
  fill(kOffW1, kMlpH, kMlpInPad, kMlpH, kMlpIn, kMlpIn);
  fill(kOffW2, kMlpH, kMlpH, kMlpH, kMlpH, kMlpH);
  fill(kOffW3, kMlpH, kMlpH, kMlpH, kMlpH, kMlpH);
  fill(kOffW4, kMlpOutPad, kMlpH, kMlpOut, kMlpH, 4 * kMlpH);
  for (int f = 0; f < kMlpNorm; f++) m->host_norm[f] = (f >= kMlpIn && f < 2 * kMlpIn) ||
                                                       f >= 2 * kMlpIn + kMlpOut ? 1.f : 0.f;
  if (imexsw_mlp_set_params(m, w.data(), m->host_norm)) {
    std::string e = m->err;
    imexsw_mlp_destroy(m);
    return mfail(nullptr, "%s", e.c_str());
  }
  *out = m;
  return 0;
}

void imexsw_mlp_destroy(imexsw_mlp* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : {(void*)m->w, (void*)m->wb, (void*)m->norm, (void*)m->m1, (void*)m->m2,
                  (void*)m->grad, (void*)m->loss, (void*)m->tz, (void*)m->ta, (void*)m->tdz,
                  (void*)m->x, (void*)m->t})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char* imexsw_mlp_last_error(const imexsw_mlp* m) {
  return m ? m->err.c_str() : g_create_error.c_str();
}

int imexsw_mlp_set_params(imexsw_mlp* m, const float* params, const float* norm) {
  if (!m) return mfail(m, "null argument");
  if (params) {
    MHIPCHK(m, hipMemcpyAsync(m->w, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_mlp_cast<>, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream, m->w, m->wb);
    MHIPCHK(m, hipMemsetAsync(m->m1, 0, sizeof(float) * kMlpParams, m->stream));
    MHIPCHK(m, hipMemsetAsync(m->m2, 0, sizeof(float) * kMlpParams, m->stream));
    m->steps = 0;
  }
  if (norm) {
    for (int f = 0; f < kMlpNorm; f++) m->host_norm[f] = norm[f];
    MHIPCHK(m, hipMemcpyAsync(m->norm, m->host_norm, sizeof(float) * kMlpNorm, hipMemcpyHostToDevice, m->stream));
  }
  MHIPCHK(m, hipStreamSynchronize(m->stream));
  return 0;
}

int imexsw_mlp_get_params(imexsw_mlp* m, float* params, float* norm) {
  if (!m) return mfail(m, "null argument");
  if (params) MHIPCHK(m, hipMemcpy(params, m->w, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  if (norm)
    for (int f = 0; f < kMlpNorm; f++) norm[f] = m->host_norm[f];
  return 0;
}

// inference on raw features x[n][9] -> raw targets y[n][2]
int imexsw_mlp_forward(imexsw_mlp* m, const float* x, int64_t n, float* y) {
  if (!m || !x || !y || n < 0) return mfail(m, "bad argument");
  if (n == 0) return 0;
  std::vector<float> xs, ts;
  standardise(m->host_norm, x, nullptr, (size_t)n, xs, ts, kMlpIn);
  if (mlp_data(m, (size_t)n)) return -1;
  MHIPCHK(m, hipMemcpyAsync(m->x, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_mlp_forward<>, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, m->stream,
                     mlp_dev(m), m->x, (int)n, m->t);
  std::vector<float> ys((size_t)n * kMlpOut);
  MHIPCHK(m, hipMemcpyAsync(ys.data(), m->t, ys.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MHIPCHK(m, hipStreamSynchronize(m->stream));
  const float* on = m->host_norm + 2 * kMlpIn;
  for (int64_t s = 0; s < n; s++)
    for (int o = 0; o < kMlpOut; o++)
      y[s * kMlpOut + o] = ys[s * kMlpOut + o] * on[kMlpOut + o] + on[o];
  return 0;
}

// loss and parameter gradient of one minibatch, in the standardised space
// (x[n][9], t[n][2] already standardised), for gradient checks
int imexsw_mlp_grad(imexsw_mlp* m, const float* x, const float* t, int32_t n, float* grad,
                    float* loss) {
  if (!m || !x || !t || n < 1) return mfail(m, "bad argument");
  if (mlp_reserve(m, n) || mlp_data(m, (size_t)n)) return -1;
  std::vector<float> xs((size_t)n * kMlpInPad, 0.f);
  for (int s = 0; s < n; s++)
    for (int f = 0; f < kMlpIn; f++) xs[(size_t)s * kMlpInPad + f] = x[(size_t)s * kMlpIn + f];
  MHIPCHK(m, hipMemcpyAsync(m->x, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MHIPCHK(m, hipMemcpyAsync(m->t, t, sizeof(float) * kMlpOut * n, hipMemcpyHostToDevice, m->stream));
  MHIPCHK(m, hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
  const MlpTrainBuf T = mlp_buf(m);
  hipLaunchKernelGGL(k_mlp_fwd_bwd<>, dim3((n + 63) / 64), dim3(64), 0, m->stream, mlp_dev(m), T,
                     m->x, m->t, n);
  hipLaunchKernelGGL(k_mlp_wgrad<>, dim3(40), dim3(64), 0, m->stream, T, n);
  if (grad) MHIPCHK(m, hipMemcpyAsync(grad, m->grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
  float l = 0.f;
  MHIPCHK(m, hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MHIPCHK(m, hipStreamSynchronize(m->stream));
  if (loss) *loss = l;
  return 0;
}

// Train on raw pairs[n][11] (9 features, 2 targets): the standardisation is
// taken from the data, then `epochs` passes of Adam over minibatches of
// `batch` samples in a seeded shuffled order.  loss_hist[e] = mean minibatch
// loss of epoch e (standardised units).
int imexsw_mlp_train(imexsw_mlp* m, const float* pairs, int64_t n, int32_t epochs, int32_t batch,
                     float lr, uint64_t seed, float* loss_hist) {
  if (!m || !pairs || n < 1 || epochs < 1 || batch < 64) return mfail(m, "bad argument");
  // rows with a non-finite value are left out
  std::vector<float> kept;
  {
    int64_t good = 0;
    for (int64_t s = 0; s < n; s++) {
      bool ok = true;
      for (int f = 0; f < kHarvestWidth; f++) ok = ok && std::isfinite(pairs[s * kHarvestWidth + f]);
      good += ok;
    }
    if (good < n) {
      kept.reserve((size_t)good * kHarvestWidth);
      for (int64_t s = 0; s < n; s++) {
        bool ok = true;
        for (int f = 0; f < kHarvestWidth; f++) ok = ok && std::isfinite(pairs[s * kHarvestWidth + f]);
        if (ok) kept.insert(kept.end(), pairs + s * kHarvestWidth, pairs + (s + 1) * kHarvestWidth);
      }
      pairs = kept.data();
      n = good;
      if (n < 1) return mfail(m, "no finite training pairs");
    }
  }
  // standardisation from the data
  double mean[kHarvestWidth] = {0}, var[kHarvestWidth] = {0};
  for (int64_t s = 0; s < n; s++)
    for (int f = 0; f < kHarvestWidth; f++) mean[f] += pairs[s * kHarvestWidth + f];
  for (int f = 0; f < kHarvestWidth; f++) mean[f] /= (double)n;
  for (int64_t s = 0; s < n; s++)
    for (int f = 0; f < kHarvestWidth; f++) {
      const double d = pairs[s * kHarvestWidth + f] - mean[f];
      var[f] += d * d;
    }
  float norm[kMlpNorm];
  for (int f = 0; f < kMlpIn; f++) {
    norm[f] = (float)mean[f];
    norm[kMlpIn + f] = (float)std::max(std::sqrt(var[f] / (double)n), 1e-6);
  }
  for (int o = 0; o < kMlpOut; o++) {
    norm[2 * kMlpIn + o] = (float)mean[kMlpIn + o];
    norm[2 * kMlpIn + kMlpOut + o] = (float)std::max(std::sqrt(var[kMlpIn + o] / (double)n), 1e-6);
  }
  if (imexsw_mlp_set_params(m, nullptr, norm)) return -1;
  std::vector<float> xs, ts;
  standardise(norm, pairs, pairs + kMlpIn, (size_t)n, xs, ts, kHarvestWidth);
  if (mlp_reserve(m, batch) || mlp_data(m, (size_t)n)) return -1;
  std::vector<int64_t> perm((size_t)n);
  for (int64_t s = 0; s < n; s++) perm[s] = s;
  uint64_t st = seed;
  std::vector<float> xe((size_t)n * kMlpInPad), te((size_t)n * kMlpOut);
  const MlpTrainBuf T = mlp_buf(m);
  const int nb = (int)((n + batch - 1) / batch);
  for (int ep = 0; ep < epochs; ep++) {
    for (int64_t s = n - 1; s > 0; s--) std::swap(perm[s], perm[splitmix(st) % (uint64_t)(s + 1)]);
    for (int64_t s = 0; s < n; s++) {
      std::memcpy(&xe[s * kMlpInPad], &xs[perm[s] * kMlpInPad], sizeof(float) * kMlpInPad);
      std::memcpy(&te[s * kMlpOut], &ts[perm[s] * kMlpOut], sizeof(float) * kMlpOut);
    }
    MHIPCHK(m, hipMemcpyAsync(m->x, xe.data(), xe.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipMemcpyAsync(m->t, te.data(), te.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIPCHK(m, hipMemsetAsync(m->loss, 0, sizeof(float), m->stream));
    for (int b = 0; b < nb; b++) {
      const int64_t off = (int64_t)b * batch;
      const int cnt = (int)std::min<int64_t>(batch, n - off);
      hipLaunchKernelGGL(k_mlp_fwd_bwd<>, dim3((cnt + 63) / 64), dim3(64), 0, m->stream, mlp_dev(m),
                         T, m->x + off * kMlpInPad, m->t + off * kMlpOut, cnt);
      hipLaunchKernelGGL(k_mlp_wgrad<>, dim3(40), dim3(64), 0, m->stream, T, cnt);
      m->steps++;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.f - std::pow(b1, (float)m->steps), c2 = 1.f - std::pow(b2, (float)m->steps);
      hipLaunchKernelGGL(k_mlp_adam<>, dim3((kMlpParams + 255) / 256), dim3(256), 0, m->stream, m->w,
                         m->wb, m->grad, m->m1, m->m2, lr, b1, b2, 1e-8f, c1, c2);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return mfail(m, "training launch: %s", hipGetErrorString(e));
    float l = 0.f;
    MHIPCHK(m, hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIPCHK(m, hipStreamSynchronize(m->stream));
    if (loss_hist) loss_hist[ep] = l / (float)nb;
  }
  return 0;
}

// layout probe of the matrix-core operands: D = A . B for 16x16 bf16 inputs
// (row-major bit patterns), fp32 D
int imexsw_mfma_probe(const uint16_t* A, const uint16_t* B, float* D) {
  short *a = nullptr, *b = nullptr;
  float* d = nullptr;
  if (hipMalloc(&a, 512) != hipSuccess || hipMalloc(&b, 512) != hipSuccess ||
      hipMalloc(&d, 1024) != hipSuccess)
    return mfail(nullptr, "probe allocation failed");
  (void)hipMemcpy(a, A, 512, hipMemcpyHostToDevice);
  (void)hipMemcpy(b, B, 512, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_mfma_probe<>, dim3(1), dim3(64), 0, 0, a, b, d);
  hipError_t e = hipMemcpy(D, d, 1024, hipMemcpyDeviceToHost);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(d);
  return e == hipSuccess ? 0 : mfail(nullptr, "probe: %s", hipGetErrorString(e));
}

// harvest: converged Newton pairs of every following stage-2 update are
// appended to a device buffer of `capacity` pairs (0 stops and frees)
int imexsw_harvest(imexsw_ctx* c, int64_t capacity) {
  if (!c || capacity < 0) return fail(c, "bad argument");
  if (c->general && capacity > 0) return fail(c, "the harvest runs with N_RK = 2 only");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->D.harvest) {
    (void)hipFree(c->D.harvest);
    c->D.harvest = nullptr;
    c->D.harvest_cap = 0;
  }
  if (capacity > 0) {
    HIPCHK(c, hipMalloc(&c->D.harvest, sizeof(float) * kHarvestWidth * (size_t)capacity));
    c->D.harvest_cap = capacity;
  }
  HIPCHK(c, hipMemsetAsync(c->D.counters + C_HARVEST * kSlots, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// copy out the harvested pairs (at most cap); n = pairs produced so far
// (may exceed the capacity: the rest were dropped)
int imexsw_harvest_get(imexsw_ctx* c, float* pairs, int64_t cap, int64_t* n) {
  if (!c) return fail(c, "null argument");
  unsigned long long cnt = 0;
  HIPCHK(c, hipMemcpyAsync(&cnt, c->D.counters + C_HARVEST * kSlots, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (n) *n = (int64_t)cnt;
  const int64_t k = std::min<int64_t>(std::min<int64_t>((int64_t)cnt, c->D.harvest_cap), cap);
  if (pairs && k > 0 && c->D.harvest)
    HIPCHK(c, hipMemcpy(pairs, c->D.harvest, sizeof(float) * kHarvestWidth * (size_t)k, hipMemcpyDeviceToHost));
  return 0;
}

// closed-form Newton start for the following stage-2 solves
int imexsw_set_closed_form_start(imexsw_ctx* c, int enable) {
  if (!c) return fail(c, "null argument");
  if (c->general && enable) return fail(c, "the closed-form start runs with N_RK = 2 only");
  if (enable && c->warm) return fail(c, "switch the learned warm start off first");
  c->closed_start = enable != 0;
  return 0;
}

// learned warm start for the following stage-2 Newton solves (NULL: off)
int imexsw_set_warm_start(imexsw_ctx* c, imexsw_mlp* m) {
  if (!c) return fail(c, "null argument");
  if (!m) {
    c->warm = false;
    return 0;
  }
  if (c->general) return fail(c, "the learned warm start runs with N_RK = 2 only");
  if (c->closed_start) return fail(c, "switch the closed-form start off first");
  if (!c->D.guess && dalloc(c, &c->D.guess, 2 * c->plane)) return -1;
  c->mlp = mlp_dev(m);
  c->warm = true;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
