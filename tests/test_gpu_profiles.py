"""Device versus oracle for the vertical-profile closures and rheology models
6 and 8 (src/constitutive_2d.f90:797-1123, 1976-2107, 2532-2986, 3252-3306,
3639-3650, src/geometry_2d.f90:1776-2038, 2407-2449), including the
reference's own TESTS/TEST_BOX run.

The device evaluates the closures with the same operation sequence as the
oracle (deterministic exp/log, the same Lambert W iteration, the same
quadrature sums in the same order), so the bar is bit-equality for the
special functions and point conversions, and the parity tests' 1e-12 per
conserved variable for whole runs.  The runs use a channel of moving flow:
the reference's closures are NaN for a layer at rest.
"""
import numpy as np
import pytest

from imex_sflow2d_v2_amd import decks
from imex_sflow2d_v2_amd.solver import Solver, lambertw
from oracle import Oracle
from oracle.pyoracle import load_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _conv_oracle(p, qp):
    return Oracle(p).qp_to_qc(qp)


def test_lambertw_device_matches_oracle():
    L = load_oracle()
    x = np.concatenate([-np.exp(-1.0) + np.logspace(-9, np.log10(np.exp(-1.0)), 500),
                        -np.logspace(-25, -0.44, 300), [-0.4, 0.0, 0.3, 2.0, 40.0]])
    print("case lambertw", x.size, flush=True)
    w0 = lambertw(x, 0)
    wm = lambertw(x, -1)
    for i, xi in enumerate(x):
        assert w0[i] == L.orc_lambertw0(float(xi)), (xi, w0[i], L.orc_lambertw0(float(xi)))
        assert wm[i] == L.orc_lambertwm1(float(xi)), (xi, wm[i], L.orc_lambertwm1(float(xi)))


@pytest.mark.parametrize("ns", [1, 2, 3])
def test_point_conversions_match_oracle(ns):
    """qp_to_qc, qc_to_qp (the Aitken mean speed) and the flux shape
    coefficients of random moving states, device against oracle, to the
    bit."""
    d = decks.profile_channel(nx=8, ny=8, n_solid=ns, convert=_conv_oracle)
    p = d.params
    print("case conversions ns", ns, flush=True)
    rng = np.random.default_rng(ns)
    n = 400
    nv = p.n_vars
    h = 10 ** rng.uniform(-1.5, 1.3, n)
    U = 10 ** rng.uniform(-0.5, 1.3, n)
    ang = rng.uniform(0, 2 * np.pi, n)
    qp = np.zeros((nv + 2, n))
    qp[0] = h
    qp[nv] = U * np.cos(ang)
    qp[nv + 1] = U * np.sin(ang)
    qp[1], qp[2] = h * qp[nv], h * qp[nv + 1]
    qp[3] = rng.uniform(300, 700, n)
    for i in range(ns):
        qp[4 + i] = rng.uniform(0.001, 0.04, n) * h
    bx = rng.uniform(-0.3, 0.3, n)
    by = rng.uniform(-0.3, 0.3, n)
    orc = Oracle(p)
    dev = Solver(p)
    try:
        qc_d, qc_o = dev.qp_to_qc(qp), orc.qp_to_qc(qp)
        assert np.array_equal(qc_d, qc_o), np.abs(qc_d - qc_o).max()
        back_d, back_o = dev.qc_to_qp(qc_o), orc.qc_to_qp(qc_o)
        assert np.array_equal(back_d, back_o), np.abs(back_d - back_o).max()
        sc_d, sc_o = dev.flux_coeffs(qp, bx, by), orc.flux_coeffs(qp, bx, by)
        assert np.array_equal(sc_d, sc_o), np.abs(sc_d - sc_o).max()
        # the closures are on: shape coefficients differ from 1
        assert np.abs(sc_d[0] - 1.0).min() > 1e-6
    finally:
        dev.close()


def _run_pair(deck, nsteps):
    print("case", deck.name, deck.description, flush=True)
    orc = Oracle(deck.params, fresh=True).load_deck(deck)
    dev = Solver(deck.params).load_deck(deck)
    n1 = orc.run(nsteps, dt0=deck.dt0)
    n2 = dev.advance(nsteps, dt0=deck.dt0)
    assert n1 == n2 == nsteps
    assert abs(orc.t - dev.t) <= 1e-12 * max(1.0, abs(orc.t)), (orc.t, dev.t)
    return orc, dev


def _check(orc, dev, p):
    qo, qd = orc.q(), dev.q()
    assert not np.isnan(qo).any()
    for v in range(qo.shape[0]):
        e = np.abs(qd[v] - qo[v]).max() / max(np.abs(qo[v]).max(), 1e-300)
        print("var", v, "rel err", e, flush=True)
        assert e <= TOL, (v, e)
    qpo, qpd = orc.qp(), dev.qp()
    nv = p.n_vars
    for v in (nv, nv + 1):
        assert np.abs(qpd[v] - qpo[v]).max() <= TOL * np.abs(qpo[v]).max()
    _, dep_o, _, _ = orc.erosion_state()
    _, dep_d, _, _ = dev.erosion_state()
    assert np.abs(dep_d - dep_o).max() <= TOL * max(np.abs(dep_o).max(), 1e-300)
    hm_o, pd_o, mv_o, _ = orc.post()
    hm_d, pd_d, mv_d, _ = dev.post()
    for a, b in ((hm_d, hm_o), (pd_d, pd_o), (mv_d, mv_o)):
        assert np.abs(a - b).max() <= TOL * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("kw", [
    dict(),                                   # model 8, N_RK = 2, bottom concentration
    dict(n_rk=3),                             # IMEX-SSP(3,3,2)
    dict(rheology_model=None),                # no rheology: Newton with profiles
    dict(bottom_conc=False, n_quad=7, n_solid=1),
    dict(slope_correction=False, n_solid=3, n_rk=1),
], ids=["model8", "nrk3", "norheo", "nobottom", "noslope_ns3_nrk1"])
def test_profile_channel_matches_oracle(kw):
    d = decks.profile_channel(convert=_conv_oracle, **kw)
    orc, dev = _run_pair(d, 30)
    try:
        _check(orc, dev, d.params)
    finally:
        dev.close()


@pytest.mark.parametrize("tuning", [(16, 1), (4, 1), (2, 1)])
def test_profile_channel_strips_match_oracle(tuning):
    """The stencil's multi-row strips (a 2048^2 grid runs 16-row strips; the
    small decks above run one-row strips): forced on the 64 x 48 channel."""
    d = decks.profile_channel(convert=_conv_oracle)
    print("case strips", tuning, flush=True)
    orc = Oracle(d.params, fresh=True).load_deck(d)
    dev = Solver(d.params).load_deck(d)
    try:
        dev.set_tuning(*tuning)
        n1 = orc.run(10, dt0=d.dt0)
        n2 = dev.advance(10, dt0=d.dt0)
        assert n1 == n2 == 10
        assert dev.stats()["nan_events"] == 0
        _check(orc, dev, d.params)
    finally:
        dev.close()


@pytest.mark.parametrize("kind", ["nrk3", "model8"])
def test_tableau_path_strips_match_oracle(kind):
    """The tableau-driven path without the profiles on multi-row strips."""
    if kind == "nrk3":
        d = decks.small_2d(n=40, ny=36, seed=5, n_rk=3)
    else:
        d = decks.profile_channel(vertical_profiles=False)
    print("case tableau strips", kind, flush=True)
    orc = Oracle(d.params, fresh=True).load_deck(d)
    dev = Solver(d.params).load_deck(d)
    try:
        dev.set_tuning(16, 1)
        n1 = orc.run(10, dt0=d.dt0)
        n2 = dev.advance(10, dt0=d.dt0)
        assert n1 == n2 == 10
        _check(orc, dev, d.params)
    finally:
        dev.close()


def test_profile_entry_points_match_oracle():
    """The solver_2d entry points one at a time on the profile channel
    (check_solve, timestep, imex_RK_solver, update_erosion_deposition_cell,
    post-step), with the initial state uploaded again half-way: the stored physical
    variables that the profile stencil and erosion pass read follow every
    kernel that changes the state and every upload."""
    d = decks.profile_channel(convert=_conv_oracle)
    print("case entry points", flush=True)
    orc = Oracle(d.params, fresh=True).load_deck(d)
    dev = Solver(d.params).load_deck(d)
    try:
        t = 0.0
        for it in range(6):
            if it == 3:
                # back to the initial state: stale stored variables would
                # still hold the state of step 3
                dev.upload_q(d.q)
                orc.set_q(d.q)
            assert orc.check_solve(False) == dev.check_solve(False, t=t)
            dt_o, dt_d = orc.timestep(), dev.timestep()
            assert dt_o == dt_d, (dt_o, dt_d)
            orc.imex_rk(t, dt_o)
            dev.imex_RK_solver(dt_d, t=t)
            orc.erosion_deposition(dt_o)
            dev.update_erosion_deposition_cell(dt_d)
            s_o, s_d = orc.post_step(), dev.post_step()
            assert np.allclose(s_o, s_d, rtol=1e-12, atol=0.0), (s_o, s_d)
            t += dt_o
        _check(orc, dev, d.params)
    finally:
        dev.close()


def test_profile_channel_1d_matches_oracle():
    d = decks.profile_channel(nx=200, ny=1, convert=_conv_oracle)
    orc, dev = _run_pair(d, 60)
    try:
        _check(orc, dev, d.params)
    finally:
        dev.close()


def test_device_initial_state_conversion():
    """The library's own qp_to_qc of the initial physical state gives the
    oracle's conservative state, so a host without the oracle starts the
    same run."""
    d = decks.profile_channel(convert=_conv_oracle)
    dev = Solver(d.params)
    try:
        d2 = decks.profile_channel(convert=lambda p, qp: dev.qp_to_qc(qp))
        assert np.array_equal(d2.q, d.q)
    finally:
        dev.close()


@pytest.mark.parametrize("n_rk", [2, 3])
def test_model8_without_profiles_matches_oracle(n_rk):
    d = decks.profile_channel(vertical_profiles=False, n_rk=n_rk)
    orc, dev = _run_pair(d, 40)
    try:
        _check(orc, dev, d.params)
    finally:
        dev.close()


# --------------------------------------------------------------------------
# rheology model 6 (src/constitutive_2d.f90:3639-3650)
# --------------------------------------------------------------------------
def test_reference_box_run_matches_oracle():
    """The reference's own TESTS/TEST_BOX run (model 6, settling and
    entrainment, energy on) from its restart file to T_END = 200 s: 1103
    steps, device against oracle."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "INIT_BOX_0000.q_2d")
    d = decks.reference_box(path)
    print("case", d.name, flush=True)
    orc = Oracle(d.params, fresh=True).load_deck(d)
    dev = Solver(d.params).load_deck(d)
    try:
        n1 = orc.run(100000, dt0=d.dt0, t_end=d.t_end)
        n2 = dev.advance(100000, dt0=d.dt0, t_end=d.t_end)
        assert n1 == n2 and orc.t == dev.t == d.t_end, (n1, n2, orc.t, dev.t)
        _check(orc, dev, d.params)
        so, sd = orc.counters(), dev.stats()
        assert sd["newton_iters"] == so["newton_iters"]
        assert sd["newton_cells"] == so["newton_cells"]
    finally:
        dev.close()


def _golden(name):
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)


@pytest.mark.parametrize("which", ["2d", "bw"])
def test_reference_decks_run_to_t_end_match_oracle(which):
    """The reference's own TESTS/TEST_2D (a pile released on its 100 x 68
    DEM, model 6, settling and entrainment, to T_END = 1.5 s) and TESTS/TEST_BW
    (a hot dilute Dirichlet inflow on a 1-D 3-degree incline, to T_END =
    400 s: the inflow is lighter than the ambient air and lifts off in the
    first cell every step, as the reference's erosion-deposition update
    prescribes), device against oracle from their restart files."""
    if which == "2d":
        d = decks.reference_2d(_golden("INIT_2D_0000.q_2d.gz"),
                               _golden("TEST_2D_topography_dem.asc.gz"))
    else:
        d = decks.reference_bw(_golden("INIT_BW_0000.q_2d.gz"),
                               _golden("TEST_BW_topography_dem.asc.gz"))
    print("case", d.name, flush=True)
    orc = Oracle(d.params, fresh=True).load_deck(d)
    dev = Solver(d.params).load_deck(d)
    try:
        n1 = orc.run(100000, dt0=d.dt0, t_end=d.t_end)
        n2 = dev.advance(100000, dt0=d.dt0, t_end=d.t_end)
        assert n1 == n2 and orc.t == dev.t == d.t_end, (n1, n2, orc.t, dev.t)
        so, sd = orc.counters(), dev.stats()
        assert sd["newton_iters"] == so["newton_iters"]
        assert sd["newton_cells"] == so["newton_cells"]
        assert sd["nan_events"] == 0
        if which == "2d":
            _check(orc, dev, d.params)
            assert np.array_equal(dev.mask(), orc.mask())
            assert (orc.q()[0] > 0).sum() > 1000          # the pile has spread
        else:
            assert np.array_equal(dev.q(), orc.q())
    finally:
        dev.close()


@pytest.mark.parametrize("n_rk", [2, 3])
@pytest.mark.parametrize("slope", [False, True])
def test_model6_small_2d_matches_oracle(n_rk, slope):
    d = decks.small_2d(n=40, ny=36, seed=5, n_rk=n_rk, rheology_model=6, friction_factor=2e-3,
                       slope_correction=slope, curvature=slope)
    orc, dev = _run_pair(d, 30)
    try:
        _check(orc, dev, d.params)
        assert np.array_equal(dev.mask(), orc.mask())
    finally:
        dev.close()
