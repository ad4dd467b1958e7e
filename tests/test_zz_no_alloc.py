"""imexsw_download_noalloc (the allocation-free variant of imexsw_download
for q and qp, staging through the library's static per-GPU array) against
imexsw_download, through the C ABI: two calls in a row, the second with a
smaller grid; the first grid's rows need several staging chunks."""
import ctypes as C

import numpy as np
import pytest

from imex_sflow2d_v2_amd import decks
from imex_sflow2d_v2_amd.params import make_params
from imex_sflow2d_v2_amd.solver import ImexswError, Solver, _dp

pytestmark = pytest.mark.gpu


def _noalloc(s, want_q=True, want_qp=True):
    q = np.zeros((s.nv, s.nx, s.ny), order="F") if want_q else None
    qp = np.zeros((s.nv + 2, s.nx, s.ny), order="F") if want_qp else None
    rc = s.lib.imexsw_download_noalloc(s.h, None if q is None else _dp(q),
                                       None if qp is None else _dp(qp))
    if rc != 0:
        raise ImexswError(s.lib.imexsw_last_error(s.h).decode())
    return q, qp


def test_download_noalloc_matches_download():
    # 2048 x 300 cells, n_vars = 6: a qp row is 16384 values, so the 2^21
    # values of staging hold 128 rows and the download takes three chunks
    big = decks.small_2d(n=2048, ny=300, seed=21, dx=5.0)
    small = decks.small_2d(n=40, ny=36, seed=22)
    runs = []
    for d in (big, small):
        print("case", d.params.nx, d.params.ny, flush=True)
        s = Solver(d.params).load_deck(d)
        s.advance(25, dt0=d.dt0)
        runs.append(s)
    for s in runs:
        q, qp = _noalloc(s)
        assert np.array_equal(q, s.q())
        assert np.array_equal(qp, s.qp())
        assert np.abs(q).max() > 0
    # only one of the two
    q, qp = _noalloc(runs[1], want_qp=False)
    assert qp is None and np.array_equal(q, runs[1].q())


def test_download_noalloc_rejects_rows_beyond_its_staging():
    # 300000 x 1 cells, n_vars = 5: a qp row of 2.1e6 values exceeds 2^21
    p = make_params(300000, 1, 1.0, 1, rho_s=(2500.0,), diam_s=(1e-3,), sp_heat_s=(1617.0,))
    s = Solver(p)
    with pytest.raises(ImexswError, match="staging memory"):
        _noalloc(s, want_q=False)
    q, _ = _noalloc(s, want_qp=False)   # q rows of 1.5e6 values fit
    assert q.shape == (5, 300000, 1)
