"""Synthetic input decks for the benchmark and test configurations.

The reference reads its inputs from NAMELIST decks, ESRI DEMs and ``.q_2d``
restart files (src/inpout_2d.f90); these builders produce the same module
state directly: an ``ImexswParams``, the cell-centred topography ``B_cent``,
the conservative state ``q(n_vars, nx, ny)`` (Fortran order) and, for a
radial source, the ring arrays that ``init_source`` derives
(src/geometry_2d.f90:872-1102).

Host-side set-up only (runs once per run); nothing here is on the per-step
path.
"""
import math

import numpy as np

from .params import make_params, E, N, S, W


# --------------------------------------------------------------------------
# host restatements of the reference's one-time set-up arithmetic
# --------------------------------------------------------------------------

def qp_to_qc_np(p, qp):
    """Vectorised physical -> conservative conversion for initial states.

    Follows ``qp_to_qc`` (src/constitutive_2d.f90:1709-2140, non-profile,
    gas-carrier branch) with the reference's evaluation order.  ``qp`` has
    shape (n_vars+2, ...) with rows h, hu, hv, T, (h*)alphas..., u, v.
    """
    nv = p.n_vars
    ns = p.n_solid
    qp = np.asarray(qp, dtype=np.float64)
    h = qp[0]
    qc = np.zeros((nv,) + h.shape)
    wet = h > np.finfo(np.float64).eps
    with np.errstate(divide="ignore", invalid="ignore"):
        T = qp[3]
        al = [qp[4 + i] if p.alpha_flag else qp[4 + i] / qp[0] for i in range(ns)]
        at = np.zeros_like(h)
        for a in al:
            at = at + a
        ars = [al[i] * p.rho_s[i] for i in range(ns)]
        rho_a = p.pres / (p.sp_gas_const_a * T)
        rho_c = ((1.0 - 0.0 - at - 0.0) * rho_a + 0.0) / (1.0 - 0.0 - at)
        alphac = 1.0 - at
        sar = np.zeros_like(h)
        for a in ars:
            sar = sar + a
        rho_m = alphac * rho_c + sar
        inv_rhom = 1.0 / rho_m
        xs = [a * inv_rhom for a in ars]
        xc = alphac * rho_c * inv_rhom
        shc = np.where(xc > np.finfo(np.float64).eps,
                       ((xc - 0.0) * p.sp_heat_a + 0.0) / xc, p.sp_heat_a)
        dsh = np.zeros_like(h)
        for i in range(ns):
            dsh = dsh + xs[i] * p.sp_heat_s[i]
        shm = dsh + xc * shc
        q1 = rho_m * h
        u, v = qp[nv], qp[nv + 1]
        qc[0] = q1
        qc[1] = rho_m * qp[1]
        qc[2] = rho_m * qp[2]
        if p.energy_flag:
            qc[3] = np.where(h > 0.0, h * rho_m * (shm * T + 0.5 * (u * u + v * v)), 0.0)
        else:
            qc[3] = h * rho_m * shm * T
        for i in range(ns):
            qc[4 + i] = xs[i] * q1
    qc[:, ~wet] = 0.0
    return qc


def mixt_var_np(p, qp):
    """(rho_m, rho_c, red_grav) of one physical state (src/constitutive_2d.f90:1333-1483)."""
    h = qp[0]
    if h <= np.finfo(np.float64).eps:
        rho_a_amb = p.pres / (p.sp_gas_const_a * p.T_ambient)
        return rho_a_amb, rho_a_amb, 0.0
    T = qp[3]
    al = [qp[4 + i] if p.alpha_flag else qp[4 + i] / qp[0] for i in range(p.n_solid)]
    at = 0.0
    for a in al:
        at += a
    rho_a = p.pres / (p.sp_gas_const_a * T)
    rho_c = ((1.0 - 0.0 - at - 0.0) * rho_a + 0.0) / (1.0 - 0.0 - at)
    d = 0.0
    for i in range(p.n_solid):
        d += al[i] * p.rho_s[i]
    rho_m = (1.0 - at) * rho_c + d
    rho_a_amb = p.pres / (p.sp_gas_const_a * p.T_ambient)
    rg = (rho_m - rho_a_amb) / rho_m * p.grav
    return rho_m, rho_c, rg


def h_source_from_mfr(p, mfr, vel, T_source, alphas_source, r_source):
    """MFR_SOURCE + VEL_SOURCE -> H_SOURCE (src/inpout_2d.f90:3930-3945, 4466-4480)."""
    source_length = math.pi * (r_source + r_source) * 1.0
    nv = p.n_vars
    qp = np.zeros(nv + 2)
    qp[0] = 1.0
    qp[1] = 1.0 * vel
    qp[3] = T_source
    for i in range(p.n_solid):
        qp[4 + i] = alphas_source[i] if p.alpha_flag else alphas_source[i] * 1.0
    qp[nv] = vel
    rho_m, _, _ = mixt_var_np(p, qp)
    return mfr / (source_length * rho_m * vel * 1.0)


def cell_centres(p):
    dx, dy = p.dx, p.dy
    x_stag = p.x0 + np.arange(p.nx + 1) * dx
    y_stag = p.y0 + np.arange(p.ny + 1) * dy
    x_comp = 0.5 * (x_stag[:-1] + x_stag[1:])
    y_comp = 0.5 * (y_stag[:-1] + y_stag[1:])
    return x_comp, y_comp, x_stag, y_stag


def init_source_np(p):
    """Radial source ring (init_source, src/geometry_2d.f90:1005-1100).

    Returns (source_cell[int32], dir_bits[uint8], vect[8, nx, ny]) in
    Fortran (j, k) order.
    """
    nx, ny = p.nx, p.ny
    sc = np.zeros((nx, ny), dtype=np.int32)
    bits = np.zeros((nx, ny), dtype=np.uint8)
    vect = np.zeros((8, nx, ny))
    if not p.radial_source_flag:
        return sc, bits, vect
    xc, yc, xs_, ys_ = cell_centres(p)
    xs, ys, rs = p.x_source, p.y_source, p.r_source
    r2 = rs * rs

    def sq(a):
        return a * a

    for k in range(1, ny - 1):
        for j in range(1, nx - 1):
            if sq(xc[j] - xs) + sq(yc[k] - ys) <= r2:
                sc[j, k] = 1
                if sq(xc[j - 1] - xs) + sq(yc[k] - ys) >= r2:
                    sc[j - 1, k] = 2
                    bits[j - 1, k] |= 1
                    d = math.sqrt(sq(xs_[j] - xs) + sq(yc[k] - ys))
                    vect[0, j - 1, k] = (xs_[j] - xs) * rs / sq(d)
                    vect[1, j - 1, k] = (yc[k] - ys) * rs / sq(d)
                elif sq(xc[j + 1] - xs) + sq(yc[k] - ys) >= r2:
                    sc[j + 1, k] = 2
                    bits[j + 1, k] |= 2
                    d = math.sqrt(sq(xs_[j + 1] - xs) + sq(yc[k] - ys))
                    vect[2, j + 1, k] = (xs_[j + 1] - xs) * rs / sq(d)
                    vect[3, j + 1, k] = (yc[k] - ys) * rs / sq(d)
                if sq(xc[j] - xs) + sq(yc[k - 1] - ys) >= r2:
                    sc[j, k - 1] = 2
                    bits[j, k - 1] |= 8
                    d = math.sqrt(sq(xc[j] - xs) + sq(ys_[k] - ys))
                    vect[6, j, k - 1] = (xc[j] - xs) * rs / sq(d)
                    vect[7, j, k - 1] = (ys_[k] - ys) * rs / sq(d)
                elif sq(xc[j] - xs) + sq(yc[k + 1] - ys) >= r2:
                    sc[j, k + 1] = 2
                    bits[j, k + 1] |= 4
                    d = math.sqrt(sq(xc[j] - xs) + sq(ys_[k + 1] - ys))
                    vect[4, j, k + 1] = (xc[j] - xs) * rs / sq(d)
                    vect[5, j, k + 1] = (ys_[k + 1] - ys) * rs / sq(d)
    return sc, bits, vect


class Deck:
    """One run's inputs: params, topography, initial q, erodible layer."""

    def __init__(self, name, params, B_cent, q, erodible=None, t_end=None,
                 steps=None, dt0=1e-4, description=""):
        self.name = name
        self.params = params
        self.B_cent = np.asfortranarray(B_cent, dtype=np.float64)
        self.q = np.asfortranarray(q, dtype=np.float64)
        nx, ny = params.nx, params.ny
        if erodible is None:
            erodible = np.zeros((params.n_solid, nx, ny))
        self.erodible = np.asfortranarray(erodible, dtype=np.float64)
        self.t_end = t_end
        self.steps = steps
        self.dt0 = dt0
        self.description = description

    @property
    def n_cells(self):
        return self.params.nx * self.params.ny


def _state_from_h(p, h, T, alphas, u=None, v=None):
    nv = p.n_vars
    qp = np.zeros((nv + 2,) + h.shape)
    qp[0] = h
    u = np.zeros_like(h) if u is None else u
    v = np.zeros_like(h) if v is None else v
    qp[1] = h * u
    qp[2] = h * v
    qp[3] = T
    for i, a in enumerate(alphas):
        qp[4 + i] = a if p.alpha_flag else a * h
    qp[nv] = u
    qp[nv + 1] = v
    return qp_to_qc_np(p, qp)


# --------------------------------------------------------------------------
# the five benchmark configurations
# --------------------------------------------------------------------------

def dambreak_1d(nx=400, energy=True, t_end=200.0, n_rk=2, cfl=0.4):
    """1D dam break, 400 x 1 cells, dx = 1 m, flat bed: h = 2 m for x < 200 m
    and 1 m elsewhere, at rest; one solid class (2500 kg/m3, 1 mm, 0.01);
    ideal-gas carrier at 300 K; Voellmy mu = 0.1, xi = 1000 m/s2; N_RK = 2,
    CFL 0.4, 200 s."""
    p = make_params(nx, 1, 1.0, 1, n_rk=n_rk, energy_flag=energy, cfl=cfl,
                    max_dt=1.0, mu=0.1, xi=1000.0, rho_s=(2500.0,),
                    diam_s=(1.0e-3,), sp_heat_s=(1617.0,), T_ambient=300.0)
    xc, _, _, _ = cell_centres(p)
    h = np.where(xc < 0.5 * nx * p.dx, 2.0, 1.0)[:, None]
    q = _state_from_h(p, h, 300.0, [0.01])
    return Deck("dambreak_1d", p, np.zeros((nx, 1)), q, t_end=t_end,
                dt0=1e-3, description="1D dam break 400x1 dx=1")


def incline_2d(n=512, dx=10.0, steps=2000, energy=True, settling=True):
    """2D collapse on an incline: n x n cells (512 x 512), dx = 10 m, planar
    10 degree slope; Gaussian column h = 20 exp(-r^2/(2 200^2)) m at rest
    centred at (L/4, L/2); two solid classes (0.5 mm and 2 mm, volume
    fractions 0.02 and 0.01); carrier at 600 K; sedimentation on, erosion off;
    2000 time steps; CFL 0.24; Voellmy 0.1 / 1000."""
    p = make_params(n, n, dx, 2, energy_flag=energy, cfl=0.24, max_dt=1.0,
                    rho_s=(2500.0, 2500.0), diam_s=(0.5e-3, 2.0e-3),
                    sp_heat_s=(1617.0, 1617.0), settling=settling,
                    T_ambient=300.0)
    xc, yc, _, _ = cell_centres(p)
    L = n * dx
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    B = (L - X) * math.tan(math.radians(10.0))
    r2 = (X - 0.25 * L) ** 2 + (Y - 0.5 * L) ** 2
    h = 20.0 * np.exp(-r2 / (2.0 * 200.0 ** 2))
    q = _state_from_h(p, h, 600.0, [0.02, 0.01])
    return Deck("incline_2d", p, B, q, steps=steps, dt0=1e-3,
                description="2D Gaussian column on 10 deg incline %dx%d dx=%g" % (n, n, dx))


def volcano_dem(n, dx, seed=3, cone_h=1500.0, cone_r=3000.0, n_bumps=32,
                noise_rms=20.0):
    """Cone + seeded Gaussian bumps (amplitude U[10,80] m, sigma U[50,300] m)
    + fractal (1/f^2 spectrum) noise rescaled to noise_rms, clamped at zero as
    the reference clamps its DEM (src/geometry_2d.f90:322)."""
    rng = np.random.default_rng(seed)
    L = n * dx
    xc = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    r = np.hypot(X - 0.5 * L, Y - 0.5 * L)
    B = cone_h * np.clip(1.0 - r / cone_r, 0.0, None)
    amp = rng.uniform(10.0, 80.0, n_bumps)
    sig = rng.uniform(50.0, 300.0, n_bumps)
    cx = rng.uniform(0.0, L, n_bumps)
    cy = rng.uniform(0.0, L, n_bumps)
    for a, s, x0, y0 in zip(amp, sig, cx, cy):
        B += a * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / (2.0 * s * s))
    kx = np.fft.fftfreq(n)
    K2 = kx[:, None] ** 2 + kx[None, :] ** 2
    K2[0, 0] = 1.0
    spec = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / K2
    spec[0, 0] = 0.0
    noise = np.real(np.fft.ifft2(spec))
    noise *= noise_rms / max(noise.std(), 1e-30)
    B = np.maximum(B + noise, 0.0)
    return B


def volcano_2048(n=2048, dx=5.0, steps=1000, energy=True, mfr=1.0e6,
                 vel_source=10.0, erosion_coeff=1.0e-5, erodible=1.0,
                 topo_change=True, seed=3):
    """Large synthetic volcanic DEM: n x n (2048 x 2048) cells, dx = 5 m;
    cone 1500 m high, 3 km radius, 32 bumps, 20 m RMS fractal noise; radial
    source of 1e6 kg/s over a 50 m radius at the summit (VEL_SOURCE 10 m/s,
    T_SOURCE 800 K); three solid classes (0.1, 1, 4 mm); erosion (coefficient
    1e-5, 1 m erodible layer, porosity 0.3) and deposition on, topography
    change on; 1000 time steps, CFL 0.24."""
    L = n * dx
    alphas_src = [0.02, 0.02, 0.01]
    p = make_params(n, n, dx, 3, energy_flag=energy, cfl=0.24, max_dt=1.0,
                    rho_s=(2500.0, 2500.0, 2500.0),
                    diam_s=(0.1e-3, 1.0e-3, 4.0e-3),
                    sp_heat_s=(1617.0, 1617.0, 1617.0), settling=True,
                    erosion_coeff=erosion_coeff, erodible_porosity=0.3,
                    erodible_fract=(0.4, 0.4, 0.2), T_erodible=300.0,
                    topo_change=topo_change, radial_source=True,
                    T_ambient=300.0)
    r_src = 50.0
    hs = h_source_from_mfr(p, mfr, vel_source, 800.0, alphas_src, r_src)
    p2 = make_params(n, n, dx, 3, energy_flag=energy, cfl=0.24, max_dt=1.0,
                     rho_s=(2500.0, 2500.0, 2500.0),
                     diam_s=(0.1e-3, 1.0e-3, 4.0e-3),
                     sp_heat_s=(1617.0, 1617.0, 1617.0), settling=True,
                     erosion_coeff=erosion_coeff, erodible_porosity=0.3,
                     erodible_fract=(0.4, 0.4, 0.2), T_erodible=300.0,
                     topo_change=topo_change, radial_source=True,
                     T_ambient=300.0,
                     source=dict(x=0.5 * L, y=0.5 * L, r=r_src, h=hs,
                                 vel=vel_source, T=800.0, alphas=alphas_src,
                                 time_param=(1e9, 1e9, 0.0, 1e9)))
    B = volcano_dem(n, dx, seed=seed)
    q = np.zeros((p2.n_vars, n, n))
    er = np.full((3, n, n), erodible)
    return Deck("volcano_%d" % n, p2, B, q, erodible=er, steps=steps, dt0=1e-3,
                description="volcanic DEM %dx%d dx=%g, radial source" % (n, n, dx))


def wet_2048(n=2048, dx=5.0, steps=1000, energy=True):
    """Fully wet 2048^2 variant (the incline geometry scaled) used for the
    bandwidth roofline: every cell is active from the first step."""
    p = make_params(n, n, dx, 3, energy_flag=energy, cfl=0.24, max_dt=1.0,
                    rho_s=(2500.0, 2500.0, 2500.0),
                    diam_s=(0.1e-3, 1.0e-3, 4.0e-3),
                    sp_heat_s=(1617.0, 1617.0, 1617.0), settling=True,
                    T_ambient=300.0)
    xc, yc, _, _ = cell_centres(p)
    L = n * dx
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    B = (L - X) * math.tan(math.radians(10.0))
    r2 = (X - 0.25 * L) ** 2 + (Y - 0.5 * L) ** 2
    h = 1.0 + 20.0 * np.exp(-r2 / (2.0 * (0.05 * L) ** 2))
    q = _state_from_h(p, h, 600.0, [0.02, 0.01, 0.005])
    return Deck("wet_%d" % n, p, B, q, steps=steps, dt0=1e-3,
                description="fully wet %dx%d dx=%g incline" % (n, n, dx))


def ensemble_members(count=256, seed=4, n_min=128, n_max=1024, dx=10.0,
                     steps=500, energy=True, sizes=None):
    """Batched ensemble: `count` (256) members, square grids with
    n = round(exp(U[ln 128, ln 1024])) (seed 4), dx = 10 m.  Each member has
    its own seeded DEM (a 5 degree plane plus 12 Gaussian bumps of amplitude
    U[5, 40] m and width U[3, 12] % of L) and a Gaussian column
    h = 15 exp(-r^2 / (2 (0.08 L)^2)) m at rest at (0.3 L, 0.5 L), dry
    (h < 1e-3 m cut to 0) elsewhere.  Sampled per member: Voellmy mu ~
    U[0.05, 0.3] (xi = 1000 m/s^2), carrier temperature ~ U[300, 900] K, total
    solid volume fraction ~ U[0.005, 0.05] split 2:1 over two classes (0.5 mm,
    2 mm); settling on.  500 steps each, CFL 0.24.  `sizes` overrides the
    sampled grid sizes (tests)."""
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), count))).astype(int)
    out = []
    for m, n in enumerate(sizes):
        n = int(n)
        mu = float(rng.uniform(0.05, 0.3))
        T = float(rng.uniform(300.0, 900.0))
        at = float(rng.uniform(0.005, 0.05))
        p = make_params(n, n, dx, 2, energy_flag=energy, cfl=0.24, max_dt=1.0,
                        mu=mu, xi=1000.0, rho_s=(2500.0, 2500.0),
                        diam_s=(0.5e-3, 2.0e-3), sp_heat_s=(1617.0, 1617.0),
                        settling=True, T_ambient=300.0)
        xc, yc, _, _ = cell_centres(p)
        L = n * dx
        X, Y = np.meshgrid(xc, yc, indexing="ij")
        B = (L - X) * math.tan(math.radians(5.0))
        mrng = np.random.default_rng([seed, m])
        for _ in range(12):
            a = mrng.uniform(5.0, 40.0)
            sg = mrng.uniform(0.03, 0.12) * L
            x0, y0 = mrng.uniform(0.0, L, 2)
            # separable Gaussian: outer product of the x and y factors
            B += a * np.outer(np.exp(-(xc - x0) ** 2 / (2.0 * sg * sg)),
                              np.exp(-(yc - y0) ** 2 / (2.0 * sg * sg)))
        r2 = (X - 0.3 * L) ** 2 + (Y - 0.5 * L) ** 2
        h = 15.0 * np.exp(-r2 / (2.0 * (0.08 * L) ** 2))
        wet = h > 1e-3
        q = np.zeros((p.n_vars, n, n))
        q[:, wet] = _state_from_h(p, h[wet], T, [at * 2.0 / 3.0, at / 3.0])
        out.append(Deck("member_%d" % m, p, B, q, steps=steps, dt0=1e-3,
                        description="ensemble member %d: %dx%d mu=%.3f T=%.0f alpha=%.4f"
                        % (m, n, n, mu, T, at)))
    return out


def small_2d(n=48, ny=None, dx=10.0, seed=0, energy=True, alpha=False,
             settling=True, erosion=False, source=False, topo_change=False,
             slope_correction=False, curvature=False, n_solid=2, rheology=True,
             bc=None, limiter=1, n_rk=2, wet_floor=0.0, entrainment=True,
             source_time=(1e9, 1e9, 0.0, 1e9), rheology_model=1, friction_factor=-1.0):
    """Small seeded 2D decks for the parity tests (seconds on the oracle).
    ``source_time`` is the radial source's time_param (period, duration,
    ramp time, end time; src/constitutive_2d.f90:5041-5088)."""
    ny = n if ny is None else ny
    rng = np.random.default_rng(seed)
    rho = (2500.0, 2200.0, 2700.0)[:n_solid]
    dia = (0.5e-3, 2.0e-3, 0.1e-3)[:n_solid]
    shs = (1617.0, 1500.0, 1400.0)[:n_solid]
    L = n * dx
    src = None
    if source:
        src = dict(x=0.5 * L + 0.3 * dx, y=0.5 * ny * dx - 0.2 * dx, r=2.6 * dx,
                   h=2.0, vel=5.0, T=700.0, alphas=[0.02, 0.01, 0.01][:n_solid],
                   time_param=tuple(source_time))
    p = make_params(n, ny, dx, n_solid, energy_flag=energy, alpha_flag=alpha,
                    cfl=0.24, max_dt=1.0, rho_s=rho, diam_s=dia, sp_heat_s=shs,
                    settling=settling, erosion_coeff=(1e-3 if erosion else 0.0),
                    erodible_porosity=0.3, T_erodible=300.0,
                    topo_change=topo_change, radial_source=source,
                    slope_correction=slope_correction, curvature=curvature,
                    rheology=rheology, bc=bc, limiter=limiter, n_rk=n_rk,
                    rheology_model=rheology_model, friction_factor=friction_factor,
                    source=src, entrainment=entrainment,
                    thickness_levels=(0.1, 1.0), dyn_pres_levels=(10.0, 1000.0))
    xc, yc, _, _ = cell_centres(p)
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    B = 0.15 * (L - X) + 3.0 * np.sin(X / (7 * dx)) * np.cos(Y / (5 * dx)) + 5.0
    B = np.maximum(B, 0.0)
    r2 = (X - 0.35 * L) ** 2 + (Y - 0.45 * ny * dx) ** 2
    h = 4.0 * np.exp(-r2 / (2.0 * (0.12 * L) ** 2))
    h = np.where(h > 1e-3, h, wet_floor)
    u = 0.5 * rng.standard_normal(h.shape) * (h > 0)
    v = 0.5 * rng.standard_normal(h.shape) * (h > 0)
    T = 400.0 + 50.0 * rng.random(h.shape)
    al = [0.02 * (1 + 0.2 * rng.random(h.shape)) for _ in range(n_solid)]
    if alpha:
        pass
    q = _state_from_h(p, h, T, al, u, v)
    er = np.full((n_solid, n, ny), 0.5) if erosion else None
    return Deck("small2d_%d" % seed, p, B, q, erodible=er, steps=20, dt0=1e-3,
                description="small 2D test deck")


def profile_channel(nx=64, ny=48, dx=5.0, n_rk=2, n_solid=2, rheology_model=8,
                    vertical_profiles=True, bottom_conc=True, n_quad=20,
                    slope_correction=True, settling=True, seed=11, convert=None):
    """A channel of moving flow for the vertical-profile closures (and
    rheology model 8), which need every wet cell to move: the profile
    momentum, speed and coefficients are NaN for a layer at rest, and so is
    model 8's friction step, as in the reference.  A 3 degree incline along
    x with seeded bumps; h ~ 1.5 m, u ~ 3 m/s, v ~ +-0.4 m/s everywhere;
    Dirichlet inflow on the west (h = 2 m, u = 4 m/s), zero-gradient
    outflow elsewhere; two classes (20 and 40 um, 2500 and 1800 kg/m3)
    settling with the bottom-concentration coefficient; friction factor
    0.01, vonK 0.4, k_s 0.05 m, Sc 0.9.  ``convert(params, qp) -> qc``
    converts the initial physical state (qp_to_qc with the profiles, which
    the host restatement qp_to_qc_np does not implement); by default the
    closures must be off."""
    rng = np.random.default_rng(seed)
    rho = (2500.0, 1800.0, 2200.0)[:n_solid]
    dia = (2e-5, 4e-5, 3e-5)[:n_solid]
    shs = (1617.0, 1200.0, 1300.0)[:n_solid]
    ab = (0.03, 0.01, 0.005)[:n_solid]
    nv = 4 + n_solid
    hb, ub, Tb = 2.0, 4.0, 400.0
    bcw = [(0, hb), (0, hb * ub), (0, 0.0), (0, Tb)] + [(0, a * hb) for a in ab]
    bc = {0: bcw}
    for side in (1, 2, 3):
        bc[side] = [(1, 0.0)] * nv
    vp = None
    if vertical_profiles:
        vp = dict(vonK=0.4, k_s=0.05, Sc=0.9, n_quad=n_quad, bottom_conc=bottom_conc)
    p = make_params(nx, ny, dx, n_solid, n_rk=n_rk, energy_flag=False, cfl=0.24, max_dt=1.0,
                    rho_s=rho, diam_s=dia, sp_heat_s=shs, settling=settling,
                    slope_correction=slope_correction, bc=bc,
                    rheology=rheology_model is not None,
                    rheology_model=rheology_model if rheology_model is not None else 1,
                    friction_factor=0.01, vertical_profiles=vp)
    xc, yc, _, _ = cell_centres(p)
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    L = nx * dx
    B = (L - X) * math.tan(math.radians(3.0)) + 20.0
    for _ in range(6):
        x0, y0 = rng.uniform(0, L), rng.uniform(0, ny * dx)
        B = B + rng.uniform(0.2, 0.8) * np.exp(-((X - x0) ** 2 + (Y - y0) ** 2)
                                               / (2.0 * (rng.uniform(3, 8) * dx) ** 2))
    h = 1.5 + 0.3 * np.sin(X / (9 * dx)) * np.cos(Y / (7 * dx))
    u = 3.0 + 0.5 * np.cos(Y / (6 * dx))
    v = 0.4 * np.sin(X / (8 * dx))
    T = 400.0 + 20.0 * rng.random(h.shape)
    al = [a * (0.8 + 0.4 * rng.random(h.shape)) for a in (0.02, 0.01, 0.005)[:n_solid]]
    qp = np.zeros((nv + 2,) + h.shape)
    qp[0], qp[1], qp[2], qp[3] = h, h * u, h * v, T
    for i, a in enumerate(al):
        qp[4 + i] = a * h
    qp[nv], qp[nv + 1] = u, v
    if convert is None:
        if vertical_profiles:
            raise ValueError("the profile closures need a qp_to_qc with profiles (convert=)")
        q = qp_to_qc_np(p, qp)
    else:
        q = convert(p, qp)
    name = "profile_channel" if vertical_profiles else "friction8_channel"
    return Deck(name, p, B, q, steps=30, dt0=1e-3,
                description="%s %dx%d dx=%g N_RK=%d" % (name, nx, ny, dx, n_rk))


def reference_box(q_path):
    """The reference's own TESTS/TEST_BOX run (its IMEX_SfloW2D.inp): a
    400 x 1 grid of 125 m cells on a flat bed, one solid class (2000 kg/m3,
    0.1 mm) with settling and air entrainment, the energy equation,
    rheology model 6 with FRICTION_FACTOR 1e-3, N_RK = 2, the KT scheme with
    limiter 1 on every variable, THETA 1.3, CFL 0.24 as the single-precision
    literal 0.23999999463558197, MAX_DT 0.2, DT0 1e-4, T_END 200 s; a wall
    (hu = 0) on the west and zero-gradient outflow on the east.  The state
    is its restart file INIT_BOX_0000.q_2d (``q_path``), read as
    read_solution reads it."""
    from .solver import q2d_read
    nx, ny, ns = 400, 1, 1
    nv = 4 + ns
    # (flag, value) per variable: h, hu, hv, T, h alpha
    bcw = [(1, 0.0), (0, 0.0), (0, 0.0), (1, 0.0), (1, 0.0)]
    bce = [(1, 0.0), (1, 0.0), (0, 0.0), (1, 0.0), (1, 0.0)]
    p = make_params(nx, ny, 125.0, ns, n_rk=2, energy_flag=True, cfl=0.23999999463558197,
                    max_dt=0.2, theta=1.3, reconstr_coeff=1.0, limiter=1,
                    bc={0: bcw, 1: bce}, grav=9.81, pres=101300.0, T_ambient=300.0,
                    sp_heat_a=998.0, sp_gas_const_a=287.051, kin_visc_a=1.48e-5,
                    rho_s=(2000.0,), diam_s=(1e-4,), sp_heat_s=(1617.0,), settling=True,
                    entrainment=True, erosion_coeff=0.0, rheology=True, rheology_model=6,
                    friction_factor=1e-3)
    q = q2d_read(q_path, nv, ns, nx, ny)
    return Deck("reference_box", p, np.zeros((nx, ny)), q, t_end=200.0, dt0=1e-4,
                description="TESTS/TEST_BOX 400x1 dx=125 model 6")


def _open_text(path):
    import gzip
    if str(path).endswith(".gz"):
        return gzip.open(path, "rt")
    return open(path)


def read_esri_dem(path):
    """An ESRI ASCII grid as the reference reads a DEM
    (src/inpout_2d.f90:3480-3603): the header values and the node heights
    f[ncols, nrows], row k = 1 at the bottom (the file starts at the top
    row).  ``path`` may be gzip-compressed."""
    with _open_text(path) as fh:
        lines = fh.read().split("\n")
    hdr = {}
    for ln in lines[:6]:
        key, val = ln.split()[:2]
        hdr[key.lower()] = float(val)
    ncols, nrows = int(hdr["ncols"]), int(hdr["nrows"])
    vals = np.array(" ".join(lines[6:]).split(), dtype=np.float64)
    if vals.size != ncols * nrows:
        raise ValueError("%s: %d heights for a %d x %d grid" % (path, vals.size, ncols, nrows))
    rows = vals.reshape(nrows, ncols)
    f = np.empty((ncols, nrows))
    for r in range(nrows):
        f[:, nrows - 1 - r] = rows[r]
    return hdr, f


def _interp_index(x1, x2):
    """Cell index and weight of interp_2d_scalar along one axis
    (src/geometry_2d.f90:1174-1226), 0-based."""
    n = x1.size
    if n == 1:
        return np.zeros(x2.shape, dtype=np.int64), np.zeros(x2.shape)
    ix = np.floor((x2 - x1[0]) / (x1[1] - x1[0])).astype(np.int64)
    ix = np.minimum(ix, n - 2)
    alfa = (x1[ix + 1] - x2) / (x1[ix + 1] - x1[ix])
    return ix, alfa


def dem_to_b_cent(hdr, f, nx, ny, x0, y0, cell_size):
    """B_cent(nx, ny) and B_nodata(nx, ny) from a DEM as init_grid makes them
    (src/geometry_2d.f90:173-365): cell centres from the staggered nodes,
    the nodata test on the four surrounding heights (interp_2d_nodata,
    :1126-1172), then bilinear interpolation (interp_2d_scalar, :1174-1226)
    of the heights clipped at zero."""
    ncols, nrows = f.shape
    cs, xll, yll = hdr["cellsize"], hdr["xllcorner"], hdr["yllcorner"]
    x1 = xll + (np.arange(1, ncols + 1) - 0.5) * cs
    y1 = yll + (np.arange(1, nrows + 1) - 0.5) * cs
    dx = cell_size if nx > 1 else 1.0
    dy = cell_size if ny > 1 else 1.0
    xs = x0 + np.arange(nx + 1) * dx
    ys = y0 + np.arange(ny + 1) * dy
    xc = 0.5 * (xs[:-1] + xs[1:])
    yc = 0.5 * (ys[:-1] + ys[1:])
    ix, ax = _interp_index(x1, xc)
    iy, ay = _interp_index(y1, yc)
    I, J = ix[:, None], iy[None, :]
    AX, AY = ax[:, None], ay[None, :]
    nod = hdr.get("nodata_value", -9999.0)
    if ncols == 1:
        nodata = (f[I, J] == nod) | (f[I, J + 1] == nod)
    elif nrows == 1:
        nodata = (f[I, J] == nod) | (f[I + 1, J] == nod)
    else:
        nodata = ((f[I, J] == nod) | (f[I, J + 1] == nod) | (f[I + 1, J] == nod)
                  | (f[I + 1, J + 1] == nod))
    g = np.maximum(0.0, f)
    if ncols == 1:
        b = AY * g[I, J] + (1.0 - AY) * g[I, J + 1]
    elif nrows == 1:
        b = AX * g[I, J] + (1.0 - AX) * g[I + 1, J]
    else:
        b = (AX * (AY * g[I, J] + (1.0 - AY) * g[I, J + 1])
             + (1.0 - AX) * (AY * g[I + 1, J] + (1.0 - AY) * g[I + 1, J + 1]))
    return np.asfortranarray(b), np.asfortranarray(nodata.astype(np.uint8))


def _read_q2d_any(path, nv, ns, nx, ny):
    from .solver import q2d_read
    if not str(path).endswith(".gz"):
        return q2d_read(path, nv, ns, nx, ny)
    import gzip
    import tempfile
    with gzip.open(path, "rb") as g, tempfile.NamedTemporaryFile(suffix=".q_2d") as t:
        t.write(g.read())
        t.flush()
        return q2d_read(t.name, nv, ns, nx, ny)


def reference_bw(q_path, dem_path):
    """The reference's own TESTS/TEST_BW run: TEST_BOX's deck (400 x 1 cells
    of 125 m, model 6 with FRICTION_FACTOR 1e-3, settling of one 2000 kg/m3
    class and air entrainment, the energy equation, N_RK = 2, limiter 1,
    THETA 1.3, CFL 0.23999999463558197, MAX_DT 0.2, DT0 1e-4) with 1 mm
    particles on its DEM (a 3-degree incline, ``dem_path``), a Dirichlet
    inflow on the west (h = 400 m, hu = 10^4 m2/s, h alpha = 0.1, T = 900 K)
    and zero-gradient outflow on the east, to T_END = 400 s from its restart
    file INIT_BW_0000.q_2d (``q_path``)."""
    nx, ny, ns = 400, 1, 1
    nv = 4 + ns
    bcw = [(0, 400.0), (0, 10000.0), (0, 0.0), (0, 900.0), (0, 0.1)]
    bce = [(1, 0.0), (1, 0.0), (0, 0.0), (1, 0.0), (1, 0.0)]
    p = make_params(nx, ny, 125.0, ns, n_rk=2, energy_flag=True, cfl=0.23999999463558197,
                    max_dt=0.2, theta=1.3, reconstr_coeff=1.0, limiter=1,
                    bc={0: bcw, 1: bce}, grav=9.81, pres=101300.0, T_ambient=300.0,
                    sp_heat_a=998.0, sp_gas_const_a=287.051, kin_visc_a=1.48e-5,
                    rho_s=(2000.0,), diam_s=(1e-3,), sp_heat_s=(1617.0,), settling=True,
                    entrainment=True, erosion_coeff=0.0, rheology=True, rheology_model=6,
                    friction_factor=1e-3)
    hdr, f = read_esri_dem(dem_path)
    B, _ = dem_to_b_cent(hdr, f, nx, ny, 0.0, 0.0, 125.0)
    q = _read_q2d_any(q_path, nv, ns, nx, ny)
    return Deck("reference_bw", p, B, q, t_end=400.0, dt0=1e-4,
                description="TESTS/TEST_BW 400x1 dx=125 incline, inflow, model 6")


def reference_2d(q_path, dem_path):
    """The reference's own TESTS/TEST_2D run: 100 x 68 cells of 0.3 m from
    (0, -10.2) on its DEM (``dem_path``), one solid class (2000 kg/m3,
    0.1 mm) with settling and air entrainment, the energy equation, rheology
    model 6 with FRICTION_FACTOR 1e-3, N_RK = 2, limiter 1, THETA 1.3,
    CFL 0.24, MAX_DT 1, DT0 1e-4, zero-gradient boundaries on every side,
    to T_END = 1.5 s from its restart file INIT_2D_0000.q_2d (``q_path``):
    a pile at rest released on the bed."""
    nx, ny, ns = 100, 68, 1
    nv = 4 + ns
    p = make_params(nx, ny, 0.3, ns, x0=0.0, y0=-10.2, n_rk=2, energy_flag=True, cfl=0.24,
                    max_dt=1.0, theta=1.3, reconstr_coeff=1.0, limiter=1,
                    grav=9.81, pres=101300.0, T_ambient=300.0,
                    sp_heat_a=998.0, sp_gas_const_a=287.051, kin_visc_a=1.48e-5,
                    rho_s=(2000.0,), diam_s=(1e-4,), sp_heat_s=(1617.0,), settling=True,
                    entrainment=True, erosion_coeff=0.0, rheology=True, rheology_model=6,
                    friction_factor=1e-3)
    hdr, f = read_esri_dem(dem_path)
    B, _ = dem_to_b_cent(hdr, f, nx, ny, 0.0, -10.2, 0.3)
    q = _read_q2d_any(q_path, nv, ns, nx, ny)
    return Deck("reference_2d", p, B, q, t_end=1.5, dt0=1e-4,
                description="TESTS/TEST_2D 100x68 dx=0.3 DEM, model 6")


CONFIGS = {
    "dambreak_1d": dambreak_1d,
    "incline_2d": incline_2d,
    "volcano_2048": volcano_2048,
    "wet_2048": wet_2048,
}
