"""The warm-start MLP's matrix-core kernels against the CPU reference
(oracle/mlp_ref.py: the same bfloat16 rounding points, float64
accumulation), and the warm-started Newton against the plain one."""
import numpy as np
import pytest

from imex_sflow2d_v2_amd import decks, learned
from imex_sflow2d_v2_amd.solver import Solver, load_library
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).view(np.uint32) >> 16,
                                dtype=np.uint16)


def test_mfma_operand_layout():
    import ctypes as C
    lib = load_library()
    rng = np.random.default_rng(1)
    A = mlp_ref.bf16(rng.standard_normal((16, 16)).astype(np.float32))
    B = mlp_ref.bf16(rng.standard_normal((16, 16)).astype(np.float32))
    D = np.zeros((16, 16), dtype=np.float32)
    a, b = _bits(A), _bits(B)
    assert lib.imexsw_mfma_probe(a.ctypes.data_as(C.POINTER(C.c_uint16)),
                                 b.ctypes.data_as(C.POINTER(C.c_uint16)),
                                 D.ctypes.data_as(C.POINTER(C.c_float))) == 0
    np.testing.assert_allclose(D, A @ B, rtol=1e-5, atol=1e-5)


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp_ref.N_PARAMS, dtype=np.float32)
    for k, (o, n) in mlp_ref.OFF.items():
        p[o:o + n] = (0.25 * rng.standard_normal(n)).astype(np.float32)
    W1 = p[mlp_ref.OFF["W1"][0]:mlp_ref.OFF["W1"][0] + mlp_ref.OFF["W1"][1]].reshape(64, 16)
    W1[:, 9:] = 0.0
    W4 = p[mlp_ref.OFF["W4"][0]:mlp_ref.OFF["W4"][0] + mlp_ref.OFF["W4"][1]].reshape(16, 64)
    W4[2:] = 0.0
    p[mlp_ref.OFF["b4"][0] + 2:mlp_ref.OFF["b4"][0] + 16] = 0.0
    return p


IDENTITY_NORM = np.array([0.0] * 9 + [1.0] * 9 + [0.0, 0.0, 1.0, 1.0], dtype=np.float32)


def test_forward_matches_cpu_reference():
    m = learned.Mlp(seed=3)
    p = _random_params(3)
    m.set_params(p, IDENTITY_NORM)
    x = np.random.default_rng(4).standard_normal((203, 9)).astype(np.float32)
    y = m.forward(x)
    ref = mlp_ref.forward(p, x.astype(np.float64), quantise=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    print("forward max rel err", err, flush=True)
    assert err < 2e-2
    m.close()


def test_gradient_matches_cpu_reference():
    m = learned.Mlp(seed=5)
    p = _random_params(5)
    m.set_params(p, IDENTITY_NORM)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((300, 9)).astype(np.float32)
    t = rng.standard_normal((300, 2)).astype(np.float32)
    g, loss = m.grad(x, t)
    lref, gref = mlp_ref.loss_and_grad(p, x.astype(np.float64), t.astype(np.float64),
                                       quantise=True)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    cos = float(g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref)))
    print("loss", loss, lref, "grad rel err", rel, "cos", cos, flush=True)
    assert abs(loss - lref) / lref < 1e-3
    assert rel < 2e-2 and cos > 0.9998
    for k, (o, n) in mlp_ref.OFF.items():   # every block, not just the norm
        blk = np.linalg.norm(g[o:o + n] - gref[o:o + n])
        assert blk <= 5e-2 * max(np.linalg.norm(gref[o:o + n]), 1e-12), k
    m.close()


def test_training_reduces_loss():
    rng = np.random.default_rng(7)
    x = rng.standard_normal((20000, 9)).astype(np.float32)
    t = np.stack([np.tanh(x[:, 0] - 0.5 * x[:, 1]), 0.3 * x[:, 2] * x[:, 3]], axis=1)
    pairs = np.concatenate([x, t.astype(np.float32)], axis=1)
    m = learned.Mlp(seed=8)
    hist = m.train(pairs, epochs=6, batch=1024, lr=2e-3, seed=9)
    print("loss history", hist, flush=True)
    assert hist[-1] < 0.3 * hist[0]
    m.close()


def test_warm_start_harvest_train_and_parity():
    d = decks.small_2d(n=64, ny=60, seed=10, energy=True, wet_floor=0.05)
    plain = Solver(d.params).load_deck(d)
    learned.start_harvest(plain, 1 << 20)
    plain.advance(40, dt0=d.dt0)
    pairs, produced = learned.harvested(plain)
    s = plain.stats()
    print("harvested", pairs.shape, produced, "newton cells", s["newton_cells"],
          "nonconverged", s["newton_nonconverged"], flush=True)
    # every converged solve is a pair, except those whose predictor has no
    # momentum (no ratio target) or a non-finite value
    conv = s["newton_cells"] - s["newton_nonconverged"]
    assert 0.95 * conv <= produced <= conv
    assert np.all(np.isfinite(pairs))
    learned.stop_harvest(plain)
    m = learned.Mlp(seed=11)
    hist = m.train(pairs, epochs=4, batch=512, lr=2e-3, seed=12)
    print("loss history", hist, flush=True)
    # warm-started run from the same deck
    warm = Solver(d.params).load_deck(d)
    learned.set_warm_start(warm, m)
    ref = Solver(d.params).load_deck(d)
    warm.advance(40, dt0=d.dt0)
    ref.advance(40, dt0=d.dt0)
    sw, sr = warm.stats(), ref.stats()
    print("iters/cell warm %.3f plain %.3f warm cells %d fallbacks %d" % (
        sw["newton_iters"] / sw["newton_cells"], sr["newton_iters"] / sr["newton_cells"],
        sw["warm_cells"], sw["warm_fallbacks"]), flush=True)
    assert sw["warm_cells"] + sw["warm_fallbacks"] == sw["newton_cells"]
    assert sw["newton_nonconverged"] == 0
    qw, qr = warm.q(), ref.q()
    for v in range(qr.shape[0]):
        scale = max(np.abs(qr[v]).max(), 1e-300)
        err = np.abs(qw[v] - qr[v]).max() / scale
        print("var", v, "warm vs plain rel diff", err, flush=True)
        assert err <= 1e-8
    assert abs(warm.t - ref.t) <= 1e-9 * ref.t
    m.close()
