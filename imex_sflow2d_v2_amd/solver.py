"""Host mirror of the reference's ``solver_2d`` module over the C-ABI.

``Solver`` exposes the public procedures the reference main program calls
(src/IMEX_SfloW2D.f90:193-533): ``check_solve``, ``timestep``,
``imex_RK_solver``, ``update_erosion_deposition_cell``,
``topography_reconstruction`` and the post-step diagnostics loop, plus
``advance`` which runs the whole time loop on the device.  Every call goes
through ``libimexsw.so`` (include/imexsw.h); there is no CPU fallback: a
missing library raises.
"""
import ctypes as C
import os

import numpy as np

from .params import ImexswParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimexsw.so")

_D = C.POINTER(C.c_double)
_U8 = C.POINTER(C.c_uint8)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_F = C.POINTER(C.c_float)


class ImexswStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "steps", "cell_updates_active", "cell_updates_grid", "cell_updates_computed",
        "newton_cells",
        "newton_iters", "newton_nonconverged", "nan_events", "neg_h_events",
        "neg_solid_events", "sum_solid_events", "low_T_events",
        "line_search_evals")] + [
        ("newton_iter_hist", C.c_int64 * 8), ("newton_iter_max", C.c_int64),
        ("warm_cells", C.c_int64), ("warm_fallbacks", C.c_int64),
        ("sparse_steps", C.c_int64),
        ("solve_cells", C.c_int32), ("compute_cells", C.c_int32)]


# name -> (restype, argtypes); the full set of symbols declared in
# include/imexsw.h
ABI = {
    "imexsw_version": (C.c_char_p, []),
    "imexsw_last_error": (C.c_char_p, [C.c_void_p]),
    "imexsw_create": (C.c_int, [C.POINTER(ImexswParams), C.POINTER(C.c_void_p)]),
    "imexsw_destroy": (None, [C.c_void_p]),
    "imexsw_set_topography": (C.c_int, [C.c_void_p, _D, _U8]),
    "imexsw_set_geometry": (C.c_int, [C.c_void_p] + [_D] * 10 + [_U8]),
    "imexsw_set_source": (C.c_int, [C.c_void_p, _I32, _U8, _D]),
    "imexsw_upload_q": (C.c_int, [C.c_void_p, _D]),
    "imexsw_upload_erosion": (C.c_int, [C.c_void_p, _D, _D, _D]),
    "imexsw_check_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_double, _I32]),
    "imexsw_timestep": (C.c_int, [C.c_void_p, _D]),
    "imexsw_rk_step": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "imexsw_erosion_deposition": (C.c_int, [C.c_void_p, C.c_double]),
    "imexsw_topography_reconstruction": (C.c_int, [C.c_void_p]),
    "imexsw_post_step": (C.c_int, [C.c_void_p, _D]),
    "imexsw_advance": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, _D, C.c_double,
                                 C.c_double, _I32]),
    "imexsw_download": (C.c_int, [C.c_void_p, _D, _D, _D, _D, _D, _U8]),
    "imexsw_download_noalloc": (C.c_int, [C.c_void_p, _D, _D]),
    "imexsw_download_erosion": (C.c_int, [C.c_void_p, _D, _D, _D, _D]),
    "imexsw_download_mask": (C.c_int, [C.c_void_p, _U8]),
    "imexsw_get_dt": (C.c_int, [C.c_void_p, _D]),
    "imexsw_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ImexswStats), C.c_int]),
    "imexsw_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "imexsw_kernel_times": (C.c_int, [C.c_void_p, _D, _I64, C.c_int]),
    "imexsw_kernel_name": (C.c_char_p, [C.c_int]),
    "imexsw_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "imexsw_ensemble_create": (C.c_int, [C.c_int32, C.POINTER(ImexswParams),
                                         C.POINTER(C.c_void_p)]),
    "imexsw_ensemble_destroy": (None, [C.c_void_p]),
    "imexsw_ensemble_last_error": (C.c_char_p, [C.c_void_p]),
    "imexsw_ensemble_member": (C.c_void_p, [C.c_void_p, C.c_int32]),
    "imexsw_ensemble_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "imexsw_ensemble_advance": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, _D,
                                          C.c_double, C.POINTER(C.c_int32)]),
    "imexsw_ensemble_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, _D,
                                           C.c_double, C.POINTER(C.c_int)]),
    "imexsw_ensemble_last_ms": (C.c_double, [C.c_void_p]),
    "imexsw_mlp_num_params": (C.c_int, []),
    "imexsw_mlp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint64]),
    "imexsw_mlp_destroy": (None, [C.c_void_p]),
    "imexsw_mlp_last_error": (C.c_char_p, [C.c_void_p]),
    "imexsw_mlp_set_params": (C.c_int, [C.c_void_p, _F, _F]),
    "imexsw_mlp_get_params": (C.c_int, [C.c_void_p, _F, _F]),
    "imexsw_mlp_forward": (C.c_int, [C.c_void_p, _F, C.c_int64, _F]),
    "imexsw_mlp_grad": (C.c_int, [C.c_void_p, _F, _F, C.c_int32, _F, _F]),
    "imexsw_mlp_train": (C.c_int, [C.c_void_p, _F, C.c_int64, C.c_int32, C.c_int32,
                                   C.c_float, C.c_uint64, _F]),
    "imexsw_mfma_probe": (C.c_int, [C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), _F]),
    "imexsw_harvest": (C.c_int, [C.c_void_p, C.c_int64]),
    "imexsw_harvest_get": (C.c_int, [C.c_void_p, _F, C.c_int64, C.POINTER(C.c_int64)]),
    "imexsw_set_warm_start": (C.c_int, [C.c_void_p, C.c_void_p]),
    "imexsw_set_closed_form_start": (C.c_int, [C.c_void_p, C.c_int]),
    "imexsw_sync": (C.c_int, [C.c_void_p]),
    "imexsw_qp_to_qc": (C.c_int, [C.c_void_p, C.c_int64, _D, _D]),
    "imexsw_qc_to_qp": (C.c_int, [C.c_void_p, C.c_int64, _D, _D]),
    "imexsw_flux_coeffs": (C.c_int, [C.c_void_p, C.c_int64, _D, _D, _D, _D]),
    "imexsw_lambertw": (C.c_int, [C.c_int32, C.c_int64, _D, _D]),
    "imexsw_gauss_legendre": (C.c_int, [C.c_int32, _D, _D]),
    "imexsw_q2d_write": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _D, _D, _D]),
    "imexsw_q2d_read": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _D]),
}

_lib = None


class ImexswError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load libimexsw.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImexswError("libimexsw.so is not built (run __graft_entry__.build() or "
                          "python -m imex_sflow2d_v2_amd.build): %s" % path)
    lib = C.CDLL(path)
    for name, (res, args) in ABI.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_D)


def _f(a, dtype=np.float64):
    return np.asfortranarray(a, dtype=dtype)


def lambertw(x, branch=0):
    """Lambert W on branch 0 or -1 as the device evaluates it."""
    lib = load_library()
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
    w = np.zeros_like(x)
    if lib.imexsw_lambertw(int(branch), x.size, _dp(x), _dp(w)) != 0:
        raise ImexswError(lib.imexsw_last_error(None).decode())
    return w


def gauss_legendre(n):
    """Gauss-Legendre points and weights on [-1, 1] (the library's rule)."""
    lib = load_library()
    x, w = np.zeros(n), np.zeros(n)
    if lib.imexsw_gauss_legendre(int(n), _dp(x), _dp(w)) != 0:
        raise ImexswError(lib.imexsw_last_error(None).decode())
    return x, w


def q2d_write(path, q, x_comp, y_comp):
    """output_solution's .q_2d file (src/inpout_2d.f90:5766-5891) for the
    state q[nv, nx, ny]; x_comp, y_comp: cell-centre coordinates."""
    lib = load_library()
    q = _f(q)
    nv, nx, ny = q.shape
    x = np.ascontiguousarray(x_comp, dtype=np.float64)
    y = np.ascontiguousarray(y_comp, dtype=np.float64)
    if x.shape != (nx,) or y.shape != (ny,):
        raise ValueError("x_comp / y_comp do not match q's shape")
    if lib.imexsw_q2d_write(os.fsencode(path), nv, nx, ny, _dp(x), _dp(y), _dp(q)) != 0:
        raise ImexswError("cannot write %s" % path)


def q2d_read(path, nv, n_solid, nx, ny):
    """read_solution's .q_2d branch (src/inpout_2d.f90:5432-5453): the state
    q[nv, nx, ny] (Fortran order)."""
    lib = load_library()
    q = np.zeros((nv, nx, ny), order="F")
    if lib.imexsw_q2d_read(os.fsencode(path), nv, n_solid, nx, ny, _dp(q)) != 0:
        raise ImexswError("cannot read %s as a %d x %d .q_2d state with %d variables"
                          % (path, nx, ny, nv))
    return q


class Solver:
    """One device context: the state of one grid."""

    def __init__(self, params):
        self.lib = load_library()
        self.params = params.copy()
        h = C.c_void_p()
        rc = self.lib.imexsw_create(C.byref(self.params), C.byref(h))
        if rc != 0:
            raise ImexswError(self.lib.imexsw_last_error(None).decode())
        self.h = h
        self.nv = 4 + self.params.n_solid
        self.ns = self.params.n_solid
        self.nx, self.ny = self.params.nx, self.params.ny
        self.t = 0.0

    # ---------------------------------------------------------------- utils
    def _chk(self, rc):
        if rc < 0:
            raise ImexswError(self.lib.imexsw_last_error(self.h).decode())
        return rc

    @classmethod
    def _member(cls, lib, handle, params, owner):
        """A non-owning view of an ensemble member's context."""
        self = cls.__new__(cls)
        self.lib = lib
        self.params = params.copy()
        self.h = C.c_void_p(handle)
        self.nv = 4 + self.params.n_solid
        self.ns = self.params.n_solid
        self.nx, self.ny = self.params.nx, self.params.ny
        self.t = 0.0
        self._owner = owner
        return self

    def close(self):
        if getattr(self, "_owner", None) is not None:
            self.h = None
            return
        if getattr(self, "h", None):
            self.lib.imexsw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- set-up
    def set_topography(self, B, nodata=None):
        B = _f(B)
        nd = None if nodata is None else _f(nodata, np.uint8).ctypes.data_as(_U8)
        self._chk(self.lib.imexsw_set_topography(self.h, _dp(B), nd))

    def set_source(self, source_cell, bits, vect):
        sc = _f(source_cell, np.int32)
        bi = _f(bits, np.uint8)
        ve = _f(vect)
        self._chk(self.lib.imexsw_set_source(self.h, sc.ctypes.data_as(_I32),
                                             bi.ctypes.data_as(_U8), _dp(ve)))

    def upload_q(self, q):
        q = _f(q)
        assert q.shape == (self.nv, self.nx, self.ny)
        self._chk(self.lib.imexsw_upload_q(self.h, _dp(q)))

    def upload_erosion(self, erodible=None, deposit=None, erosion=None):
        args = []
        keep = []
        for a in (erodible, deposit, erosion):
            if a is None:
                args.append(None)
            else:
                a = _f(a)
                keep.append(a)
                args.append(_dp(a))
        self._chk(self.lib.imexsw_upload_erosion(self.h, *args))

    def load_deck(self, deck, init=True):
        """Set-up order of the reference main program (src/IMEX_SfloW2D.f90:
        189-311): topography, initial state, source ring, first check_solve,
        initial diagnostics."""
        from .decks import init_source_np
        self.set_topography(deck.B_cent)
        if self.params.radial_source_flag:
            self.set_source(*init_source_np(self.params))
        self.upload_q(deck.q)
        if self.ns:
            z = np.zeros((self.nx, self.ny, self.ns), order="F")
            self.upload_erosion(deck.erodible, z, z)
        self.t = 0.0
        if init:
            self.check_solve(True)
            self.post_step(sums=False)
        return self

    # ---------------------------------------------------------------- solver_2d
    def check_solve(self, solve_all=False, t=None):
        n = C.c_int32(0)
        self._chk(self.lib.imexsw_check_solve(self.h, int(solve_all),
                                              self.t if t is None else t, C.byref(n)))
        return n.value

    def timestep(self):
        dt = C.c_double(0.0)
        self._chk(self.lib.imexsw_timestep(self.h, C.byref(dt)))
        return dt.value

    def imex_RK_solver(self, dt, t=None):
        self._chk(self.lib.imexsw_rk_step(self.h, self.t if t is None else t, dt))

    def update_erosion_deposition_cell(self, dt):
        self._chk(self.lib.imexsw_erosion_deposition(self.h, dt))

    def topography_reconstruction(self):
        self._chk(self.lib.imexsw_topography_reconstruction(self.h))

    def post_step(self, sums=True):
        s = np.zeros(4)
        self._chk(self.lib.imexsw_post_step(self.h, _dp(s) if sums else None))
        return s

    def advance(self, nsteps, dt0=1e-4, t_end=1e30, t_output=None):
        t = np.array([self.t])
        n = C.c_int32(0)
        self._chk(self.lib.imexsw_advance(self.h, int(nsteps), dt0, _dp(t), t_end,
                                          t_end if t_output is None else t_output,
                                          C.byref(n)))
        self.t = float(t[0])
        return n.value

    def get_dt(self):
        """The last step's dt (after the host loop's clipping)."""
        v = np.zeros(1)
        self._chk(self.lib.imexsw_get_dt(self.h, _dp(v)))
        return float(v[0])

    def set_closed_form_start(self, enable=True):
        """Newton starts from the closed-form Voellmy solution (see
        imexsw_set_closed_form_start)."""
        self._chk(self.lib.imexsw_set_closed_form_start(self.h, int(bool(enable))))

    def set_tuning(self, max_strip_rows=16, target_strips=2048):
        self._chk(self.lib.imexsw_set_tuning(self.h, int(max_strip_rows), int(target_strips)))

    # ------------------------------------------------- point conversions
    def qp_to_qc(self, qp):
        """qp_to_qc of states qp(n_vars+2, ...) with the context's model
        (vertical profiles included), on the device."""
        qp = np.asarray(qp, dtype=np.float64)
        nv = self.params.n_vars
        flat = np.ascontiguousarray(qp.reshape(nv + 2, -1).T)
        out = np.zeros((flat.shape[0], nv))
        self._chk(self.lib.imexsw_qp_to_qc(self.h, flat.shape[0], _dp(flat), _dp(out)))
        return out.T.reshape((nv,) + qp.shape[1:])

    def qc_to_qp(self, qc):
        """qc_to_qp of states qc(n_vars, ...) on the device."""
        qc = np.asarray(qc, dtype=np.float64)
        nv = self.params.n_vars
        flat = np.ascontiguousarray(qc.reshape(nv, -1).T)
        out = np.zeros((flat.shape[0], nv + 2))
        self._chk(self.lib.imexsw_qc_to_qp(self.h, flat.shape[0], _dp(flat), _dp(out)))
        return out.T.reshape((nv + 2,) + qc.shape[1:])

    def flux_coeffs(self, qp, Bx=None, By=None):
        """eval_flux_coeffs of states qp(n_vars+2, ...) (1 without profiles)."""
        qp = np.asarray(qp, dtype=np.float64)
        nv = self.params.n_vars
        flat = np.ascontiguousarray(qp.reshape(nv + 2, -1).T)
        n = flat.shape[0]
        bx = np.zeros(n) if Bx is None else np.ascontiguousarray(np.asarray(Bx, float).ravel())
        by = np.zeros(n) if By is None else np.ascontiguousarray(np.asarray(By, float).ravel())
        out = np.zeros((n, nv))
        self._chk(self.lib.imexsw_flux_coeffs(self.h, n, _dp(flat), _dp(bx), _dp(by), _dp(out)))
        return out.T.reshape((nv,) + qp.shape[1:])

    def sync(self):
        self._chk(self.lib.imexsw_sync(self.h))

    # ---------------------------------------------------------------- outputs
    def q(self):
        a = np.zeros((self.nv, self.nx, self.ny), order="F")
        self._chk(self.lib.imexsw_download(self.h, _dp(a), None, None, None, None, None))
        return a

    def qp(self):
        a = np.zeros((self.nv + 2, self.nx, self.ny), order="F")
        self._chk(self.lib.imexsw_download(self.h, None, _dp(a), None, None, None, None))
        return a

    def post(self):
        hm, pd, mv = (np.zeros((self.nx, self.ny), order="F") for _ in range(3))
        nl = self.params.n_thickness_levels * self.params.n_dyn_pres_levels
        vu = np.zeros((max(nl, 1), self.nx, self.ny), dtype=np.uint8, order="F")
        self._chk(self.lib.imexsw_download(self.h, None, None, _dp(hm), _dp(pd), _dp(mv),
                                           vu.ctypes.data_as(_U8)))
        return hm, pd, mv, vu[:nl]

    def erosion_state(self):
        er = np.zeros((self.ns, self.nx, self.ny), order="F")
        dep = np.zeros((self.nx, self.ny, self.ns), order="F")
        ers = np.zeros((self.nx, self.ny, self.ns), order="F")
        B = np.zeros((self.nx, self.ny), order="F")
        self._chk(self.lib.imexsw_download_erosion(self.h, _dp(er), _dp(dep), _dp(ers), _dp(B)))
        return er, dep, ers, B

    def mask(self):
        m = np.zeros((self.nx, self.ny), dtype=np.uint8, order="F")
        self._chk(self.lib.imexsw_download_mask(self.h, m.ctypes.data_as(_U8)))
        return m

    def stats(self, reset=False):
        s = ImexswStats()
        self._chk(self.lib.imexsw_stats_get(self.h, C.byref(s), int(reset)))
        out = {n: getattr(s, n) for n, _ in ImexswStats._fields_}
        out["newton_iter_hist"] = list(s.newton_iter_hist)
        return out

    def set_timing(self, on=True):
        self._chk(self.lib.imexsw_set_timing(self.h, int(on)))

    def kernel_times(self):
        ms = np.zeros(16)
        calls = np.zeros(16, dtype=np.int64)
        n = self._chk(self.lib.imexsw_kernel_times(self.h, _dp(ms), calls.ctypes.data_as(_I64), 16))
        return {self.lib.imexsw_kernel_name(i).decode(): (float(ms[i]), int(calls[i]))
                for i in range(n)}


class Ensemble:
    """Many member grids advanced by batched launches (imexsw_ensemble_*).
    ``members[i]`` is a Solver view of member i for set-up and output."""

    def __init__(self, params_list):
        self.lib = load_library()
        n = len(params_list)
        self._params = (ImexswParams * n)(*[p.copy() for p in params_list])
        h = C.c_void_p()
        rc = self.lib.imexsw_ensemble_create(n, self._params, C.byref(h))
        if rc != 0:
            raise ImexswError(self.lib.imexsw_ensemble_last_error(None).decode())
        self.h = h
        self.members = [Solver._member(self.lib, self.lib.imexsw_ensemble_member(h, i),
                                       params_list[i], self) for i in range(n)]
        self.t = np.zeros(n)

    def _chk(self, rc):
        if rc < 0:
            raise ImexswError(self.lib.imexsw_ensemble_last_error(self.h).decode())
        return rc

    def load(self, decks):
        for m, d in zip(self.members, decks):
            m.load_deck(d)
        self.t[:] = 0.0
        return self

    def set_tuning(self, strip_rows=8, graph_steps=25):
        self._chk(self.lib.imexsw_ensemble_set_tuning(self.h, int(strip_rows), int(graph_steps)))

    def autotune(self, steps_per_trial, dt0, t_end=1e30):
        best = C.c_int(0)
        self._chk(self.lib.imexsw_ensemble_autotune(self.h, int(steps_per_trial), dt0,
                                                    _dp(self.t), t_end, C.byref(best)))
        for m, t in zip(self.members, self.t):
            m.t = float(t)
        return best.value

    def advance(self, nsteps, dt0, t_end=1e30):
        n = C.c_int32(0)
        self._chk(self.lib.imexsw_ensemble_advance(self.h, int(nsteps), dt0, _dp(self.t),
                                                   t_end, C.byref(n)))
        for m, t in zip(self.members, self.t):
            m.t = float(t)
        return n.value

    def last_ms(self):
        return self.lib.imexsw_ensemble_last_ms(self.h)

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.h = None
            self.lib.imexsw_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
