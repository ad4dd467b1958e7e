"""Whole-benchmark parity check (not collected by pytest: about two minutes,
most of it the oracle on 16 host threads).

Runs the 2048 x 2048 volcano workload that bench.py times for N steps
(default 1000, the benchmark's length) on the device, and the oracle from
the same deck in fresh-face semantics (the device's) or, with `stale`, in
the reference's stale-face semantics, then compares every conserved
variable, the running maxima and vulnerability table, the erosion fields and
bed, the mask, the active cell-update count and the Newton iteration count.

    python3 -u tests/full_run_check.py [steps] [stale]

Exit status 0 when everything is bit-identical.
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from imex_sflow2d_v2_amd import decks  # noqa: E402
from imex_sflow2d_v2_amd.solver import Solver  # noqa: E402
from oracle import Oracle  # noqa: E402


def main(steps, fresh=True):
    t0 = time.time()
    d = decks.volcano_2048()
    print("deck %.1f s" % (time.time() - t0), flush=True)
    dev = Solver(d.params).load_deck(d)
    n = dev.advance(steps, dt0=d.dt0)
    print("device steps", n, "t", dev.t, flush=True)
    orc = Oracle(d.params, fresh=fresh, threads=16).load_deck(d)
    print("oracle faces:", "fresh" if fresh else "stale", flush=True)
    done = 0
    while done < steps:
        k = orc.run(min(100, steps - done), dt0=d.dt0)
        if k == 0:
            break
        done += k
        print("oracle steps", done, "t", orc.t, "%.1f s" % (time.time() - t0), flush=True)
    ok = done == n and orc.t == dev.t
    print("steps", done, n, "t", orc.t, dev.t)
    qo, qd = orc.q(), dev.q()
    for v in range(qo.shape[0]):
        sc = max(np.abs(qo[v]).max(), 1e-300)
        same = np.array_equal(qd[v], qo[v])
        ok &= same
        print("var", v, "max rel err %.3e" % (np.abs(qd[v] - qo[v]).max() / sc), "bitwise", same)
    for name, a, b in zip(("hmax", "pdynmax", "modvel", "vuln"), dev.post(), orc.post()):
        ok &= np.array_equal(a, b)
        print(name, "equal", np.array_equal(a, b))
    for name, a, b in zip(("erodible", "deposit", "erosion", "B"), dev.erosion_state(),
                          orc.erosion_state()):
        ok &= np.array_equal(a, b)
        print(name, "equal", np.array_equal(a, b))
    ok &= np.array_equal(dev.mask(), orc.mask())
    s, c = dev.stats(), orc.counters()
    ok &= s["cell_updates_active"] == c["active_cell_updates"] and s["newton_iters"] == c["newton_iters"]
    print("mask equal", np.array_equal(dev.mask(), orc.mask()))
    print("active cell-updates", s["cell_updates_active"], c["active_cell_updates"])
    print("newton iterations", s["newton_iters"], c["newton_iters"])
    print("ALL BIT-IDENTICAL" if ok else "DIFFERENCES FOUND", flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main(int(sys.argv[1]) if len(sys.argv) > 1 else 1000,
                  fresh=not (len(sys.argv) > 2 and sys.argv[2] == "stale")))
