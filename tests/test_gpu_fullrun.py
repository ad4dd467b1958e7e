"""The whole benchmark run at full size: the 2048 x 2048 volcano workload
that bench.py times, 1000 timesteps from t = 0, on the device against the
oracle in the reference's stale-face interface semantics (module arrays that
keep an inactive neighbour's last interface values), bit for bit: every
conserved variable, the running maxima and vulnerability table, the erosion
fields and bed, the mask, the active cell-update and Newton iteration
counts.  About two minutes, most of it the oracle on 16 host threads."""
import time

import numpy as np
import pytest

from imex_sflow2d_v2_amd import decks
from imex_sflow2d_v2_amd.solver import Solver
from oracle import Oracle

pytestmark = pytest.mark.gpu


def compare_full_run(steps, fresh):
    t0 = time.time()
    d = decks.volcano_2048()
    dev = Solver(d.params).load_deck(d)
    n = dev.advance(steps, dt0=d.dt0)
    print("device steps", n, "t", dev.t, "%.1f s" % (time.time() - t0), flush=True)
    orc = Oracle(d.params, fresh=fresh, threads=16).load_deck(d)
    done = 0
    while done < steps:
        k = orc.run(min(100, steps - done), dt0=d.dt0)
        if k == 0:
            break
        done += k
        print("oracle steps", done, "t", orc.t, "%.1f s" % (time.time() - t0), flush=True)
    assert done == n and orc.t == dev.t, (done, n, orc.t, dev.t)
    qo, qd = orc.q(), dev.q()
    for v in range(qo.shape[0]):
        assert np.array_equal(qd[v], qo[v]), v
    for name, a, b in zip(("hmax", "pdynmax", "modvel", "vuln"), dev.post(), orc.post()):
        assert np.array_equal(a, b), name
    for name, a, b in zip(("erodible", "deposit", "erosion", "B"), dev.erosion_state(),
                          orc.erosion_state()):
        assert np.array_equal(a, b), name
    assert np.array_equal(dev.mask(), orc.mask())
    s, c = dev.stats(), orc.counters()
    print("active cell-updates", s["cell_updates_active"], c["active_cell_updates"],
          "newton iterations", s["newton_iters"], c["newton_iters"], flush=True)
    assert s["cell_updates_active"] == c["active_cell_updates"]
    assert s["newton_iters"] == c["newton_iters"]


def test_volcano_2048_full_run_bitwise_stale_faces():
    compare_full_run(1000, fresh=False)
