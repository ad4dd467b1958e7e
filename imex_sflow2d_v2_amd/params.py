"""ctypes mirror of ``imexsw_params`` (include/imexsw.h).

The fields correspond one to one to the reference's namelist-driven module
variables (src/parameters_2d.f90, src/constitutive_2d.f90:20-260,
src/geometry_2d.f90:20-170).  ``make_params`` fills defaults the way
``init_param`` does (src/inpout_2d.f90:393-941) for the supported model subset.
"""
import ctypes as C

MAX_SOLID = 6
MAX_VARS = 4 + MAX_SOLID
MAX_LEVELS = 8
MAX_QUAD = 64

W, E, S, N = 0, 1, 2, 3


class ImexswParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("x0", C.c_double), ("y0", C.c_double), ("cell_size", C.c_double),
        ("n_solid", C.c_int32), ("n_rk", C.c_int32),
        ("energy_flag", C.c_int32), ("alpha_flag", C.c_int32),
        ("slope_correction_flag", C.c_int32), ("curvature_term_flag", C.c_int32),
        ("topo_change_flag", C.c_int32), ("radial_source_flag", C.c_int32),
        ("rheology_flag", C.c_int32), ("rheology_model", C.c_int32),
        ("settling_flag", C.c_int32), ("entrainment_flag", C.c_int32),
        ("erodible_deposit_flag", C.c_int32),
        ("cfl", C.c_double), ("max_dt", C.c_double),
        ("reconstr_coeff", C.c_double), ("theta", C.c_double),
        ("limiter", C.c_int32 * (MAX_VARS + 2)),
        ("bc_flag", (C.c_int32 * MAX_VARS) * 4),
        ("bc_value", (C.c_double * MAX_VARS) * 4),
        ("grav", C.c_double), ("pres", C.c_double), ("T_ambient", C.c_double),
        ("sp_heat_a", C.c_double), ("sp_gas_const_a", C.c_double),
        ("kin_visc_c", C.c_double),
        ("rho_s", C.c_double * MAX_SOLID), ("diam_s", C.c_double * MAX_SOLID),
        ("sp_heat_s", C.c_double * MAX_SOLID),
        ("mu", C.c_double), ("xi", C.c_double),
        ("erosion_coeff", C.c_double), ("erodible_porosity", C.c_double),
        ("erodible_fract", C.c_double * MAX_SOLID),
        ("T_erodible", C.c_double), ("alphastot_min", C.c_double),
        ("maximum_solid_packing", C.c_double),
        ("x_source", C.c_double), ("y_source", C.c_double), ("r_source", C.c_double),
        ("h_source", C.c_double), ("vel_source", C.c_double), ("T_source", C.c_double),
        ("alphas_source", C.c_double * MAX_SOLID),
        ("time_param", C.c_double * 4),
        ("n_thickness_levels", C.c_int32), ("n_dyn_pres_levels", C.c_int32),
        ("thickness_levels", C.c_double * MAX_LEVELS),
        ("dyn_pres_levels", C.c_double * MAX_LEVELS),
        ("vertical_profiles_flag", C.c_int32), ("bottom_conc_flag", C.c_int32),
        ("n_quad", C.c_int32),
        ("vonK", C.c_double), ("k_s", C.c_double), ("Sc", C.c_double),
        ("z_dyn", C.c_double), ("friction_factor", C.c_double),
    ]

    @property
    def n_vars(self):
        return 4 + self.n_solid

    @property
    def dx(self):
        return self.cell_size if self.nx > 1 else 1.0

    @property
    def dy(self):
        return self.cell_size if self.ny > 1 else 1.0

    def copy(self):
        p = ImexswParams()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(ImexswParams))
        return p


def make_params(nx, ny, cell_size, n_solid, *, x0=0.0, y0=0.0, n_rk=2,
                energy_flag=False, alpha_flag=False, slope_correction=False,
                curvature=False, topo_change=False, radial_source=False,
                rheology=True, mu=0.1, xi=1000.0, settling=False,
                entrainment=False, erodible_deposit=False, cfl=0.24,
                max_dt=1.0, reconstr_coeff=1.0, theta=1.0, limiter=1,
                bc=None, grav=9.81, pres=101300.0, T_ambient=300.0,
                sp_heat_a=998.0, sp_gas_const_a=287.051, kin_visc_a=1.48e-5,
                rho_s=(2500.0,), diam_s=(1.0e-3,), sp_heat_s=(1617.0,),
                erosion_coeff=0.0, erodible_porosity=0.0, erodible_fract=None,
                T_erodible=300.0, alphastot_min=0.0, maximum_solid_packing=0.64,
                source=None, thickness_levels=(), dyn_pres_levels=(),
                rheology_model=1, friction_factor=-1.0, vertical_profiles=None):
    """Build an ``ImexswParams``.

    ``bc`` maps side (W/E/S/N index) to a list of ``(flag, value)`` per
    conservative variable; default is zero-gradient (flag 1, value 0) for all,
    the choice of the shipped decks (e.g. TESTS/TEST_DAMBREAK/*/IMEX_SfloW2D.inp).
    ``source`` is a dict with x, y, r, h, vel, T, alphas, time_param.
    ``rheology_model`` 8 integrates the turbulent friction ``friction_factor``
    in closed form.  ``vertical_profiles`` is a dict with vonK, k_s, Sc,
    n_quad and optionally bottom_conc, z_dyn (VERTICAL_PROFILES_PARAMETERS).
    """
    p = ImexswParams()
    nv = 4 + n_solid
    if not (0 <= n_solid <= MAX_SOLID):
        raise ValueError("n_solid out of range")
    p.nx, p.ny, p.x0, p.y0, p.cell_size = nx, ny, x0, y0, cell_size
    p.n_solid, p.n_rk = n_solid, n_rk
    p.energy_flag = int(energy_flag)
    p.alpha_flag = int(alpha_flag)
    p.slope_correction_flag = int(slope_correction)
    p.curvature_term_flag = int(curvature)
    p.topo_change_flag = int(topo_change)
    p.radial_source_flag = int(radial_source)
    p.rheology_flag = int(rheology)
    p.rheology_model = rheology_model if rheology else 0
    p.friction_factor = friction_factor
    if vertical_profiles is not None:
        vp = vertical_profiles
        p.vertical_profiles_flag = 1
        p.bottom_conc_flag = int(vp.get("bottom_conc", False))
        p.n_quad = vp["n_quad"]
        p.vonK, p.k_s, p.Sc = vp["vonK"], vp["k_s"], vp["Sc"]
        p.z_dyn = vp.get("z_dyn", -1.0)
    p.settling_flag = int(settling)
    p.entrainment_flag = int(entrainment)
    p.erodible_deposit_flag = int(erodible_deposit)
    p.cfl, p.max_dt = cfl, max_dt
    p.reconstr_coeff, p.theta = reconstr_coeff, theta
    lim = [limiter] * (nv) if isinstance(limiter, int) else list(limiter)
    # limiter(n_vars+1:n_vars+2) = limiter(2:3)  (src/inpout_2d.f90:2079-2080)
    lim = lim[:nv] + [lim[1], lim[2]]
    for i, v in enumerate(lim):
        p.limiter[i] = v
    for side in range(4):
        for i in range(nv):
            f, v = (1, 0.0)
            if bc is not None and side in bc:
                f, v = bc[side][i]
            p.bc_flag[side][i] = f
            p.bc_value[side][i] = v
    p.grav, p.pres, p.T_ambient = grav, pres, T_ambient
    p.sp_heat_a, p.sp_gas_const_a, p.kin_visc_c = sp_heat_a, sp_gas_const_a, kin_visc_a
    for i in range(n_solid):
        p.rho_s[i] = rho_s[i]
        p.diam_s[i] = diam_s[i]
        p.sp_heat_s[i] = sp_heat_s[i]
    p.mu, p.xi = mu, xi
    p.erosion_coeff = erosion_coeff
    if erosion_coeff == 0.0:
        # src/inpout_2d.f90:1664-1668
        p.erodible_porosity = 0.0
        ef = [1.0 / n_solid] * n_solid if n_solid else []
        T_erodible = 300.0
    else:
        p.erodible_porosity = erodible_porosity
        ef = list(erodible_fract) if erodible_fract is not None else (
            [1.0] if n_solid == 1 else [1.0 / n_solid] * n_solid)
    for i in range(n_solid):
        p.erodible_fract[i] = ef[i]
    p.T_erodible = T_erodible
    p.alphastot_min = alphastot_min
    p.maximum_solid_packing = maximum_solid_packing
    if source is not None:
        p.x_source, p.y_source, p.r_source = source["x"], source["y"], source["r"]
        p.h_source, p.vel_source, p.T_source = source["h"], source["vel"], source["T"]
        for i in range(n_solid):
            p.alphas_source[i] = source["alphas"][i]
        for i in range(4):
            p.time_param[i] = source["time_param"][i]
    p.n_thickness_levels = len(thickness_levels)
    p.n_dyn_pres_levels = len(dyn_pres_levels) if thickness_levels else 0
    if thickness_levels and not dyn_pres_levels:
        # src/inpout_2d.f90:3650-3653: one dynamic-pressure level of 0
        p.n_dyn_pres_levels = 1
        dyn_pres_levels = (0.0,)
    for i, v in enumerate(thickness_levels):
        p.thickness_levels[i] = v
    for i, v in enumerate(dyn_pres_levels):
        p.dyn_pres_levels[i] = v
    return p
