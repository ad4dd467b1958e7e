"""Ensembles (imexsw_ensemble_*): members of different sizes and parameters
advanced by batched launches and captured graphs give exactly the bits of
each member advanced alone, which the parity tests tie to the oracle."""
import numpy as np
import pytest

from imex_sflow2d_v2_amd import decks
from imex_sflow2d_v2_amd.solver import Ensemble, Solver
from oracle import Oracle

pytestmark = pytest.mark.gpu

SIZES = [40, 73, 61, 128, 97, 50]


def _members():
    return decks.ensemble_members(count=len(SIZES), sizes=SIZES, seed=21)


@pytest.mark.parametrize("strip_rows,graph_steps", [(8, 7), (4, 25), (16, 1)])
def test_ensemble_bitwise_equals_members_alone(strip_rows, graph_steps):
    ds = _members()
    print("ensemble sizes", SIZES, "strip", strip_rows, "graph", graph_steps, flush=True)
    ens = Ensemble([d.params for d in ds]).load(ds)
    ens.set_tuning(strip_rows, graph_steps)
    n = ens.advance(30, dt0=ds[0].dt0)
    assert n == 30
    for i, d in enumerate(ds):
        alone = Solver(d.params).load_deck(d)
        alone.advance(30, dt0=d.dt0)
        assert alone.t == ens.t[i], (i, alone.t, ens.t[i])
        assert np.array_equal(alone.q(), ens.members[i].q()), i
        a, b = alone.stats(), ens.members[i].stats()
        assert a["newton_iters"] == b["newton_iters"], i
    ens.close()


def test_ensemble_autotune_then_advance():
    ds = _members()
    ens = Ensemble([d.params for d in ds]).load(ds)
    best = ens.autotune(3, dt0=ds[0].dt0)
    assert best in (4, 8, 16)
    # the count is this call's steps, not the run's (the trials took 9)
    assert ens.advance(11, dt0=ds[0].dt0) == 11
    for i in (0, 3):
        alone = Solver(ds[i].params).load_deck(ds[i])
        alone.advance(20, dt0=ds[i].dt0)
        assert alone.t == ens.t[i]
        assert np.array_equal(alone.q(), ens.members[i].q())
    ens.close()


def test_ensemble_member_matches_oracle():
    ds = _members()[:2]
    ens = Ensemble([d.params for d in ds]).load(ds)
    ens.advance(15, dt0=ds[0].dt0)
    orc = Oracle(ds[1].params, fresh=True).load_deck(ds[1])
    orc.run(15, dt0=ds[1].dt0)
    assert orc.t == ens.t[1]
    qo, qd = orc.q(), ens.members[1].q()
    scale = np.abs(qo).max(axis=(1, 2))[:, None, None]
    assert np.all(np.abs(qo - qd) <= 1e-12 * scale)
    ens.close()


def test_ensemble_full_config_members_bitwise():
    """The benchmark's ensemble at full size (256 members of 130^2 to 1021^2
    cells, 7.4e7 cells): one 25-step graph replay of the batch gives the
    bits of the smallest, the largest, the first and the last member
    advanced alone."""
    ds = decks.ensemble_members(count=256)
    sizes = [d.params.nx for d in ds]
    picks = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 0, len(ds) - 1})
    print("ensemble 256 members, cells", sum(d.params.nx * d.params.ny for d in ds),
          "checking", picks, flush=True)
    ens = Ensemble([d.params for d in ds]).load(ds)
    ens.set_tuning(16, 25)
    assert ens.advance(25, dt0=ds[0].dt0) == 25
    for i in picks:
        alone = Solver(ds[i].params).load_deck(ds[i])
        alone.advance(25, dt0=ds[i].dt0)
        assert alone.t == ens.t[i], (i, alone.t, ens.t[i])
        assert np.array_equal(alone.q(), ens.members[i].q()), i
        del alone
    ens.close()
