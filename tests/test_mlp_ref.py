"""The CPU reference of the warm-start MLP (oracle/mlp_ref.py) pinned by what
does not depend on it: bfloat16 rounding of known values, and the exact
float64 gradient against central finite differences of the loss."""
import numpy as np

from oracle import mlp_ref


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3.0 * 2.0 ** -9, 1.0 + 2.0 ** -7, -2.5, 0.0,
                  3.0e38], dtype=np.float32)
    want = [1.0, 1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -7, -2.5, 0.0]
    got = mlp_ref.bf16(x)
    assert np.array_equal(got[:6], want)       # ties to even, exact values kept
    assert abs(got[6] - 3.0e38) / 3.0e38 < 2.0 ** -8


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = np.zeros(mlp_ref.N_PARAMS)
    for k, (o, n) in mlp_ref.OFF.items():
        p[o:o + n] = 0.3 * rng.standard_normal(n)
    # padded input columns / output rows are zero in the device layout
    W1 = p[mlp_ref.OFF["W1"][0]:mlp_ref.OFF["W1"][0] + mlp_ref.OFF["W1"][1]].reshape(64, 16)
    W1[:, 9:] = 0.0
    x = rng.standard_normal((7, 9))
    t = rng.standard_normal((7, 2))
    loss, g = mlp_ref.loss_and_grad(p, x, t, quantise=False)
    idx = list(range(0, mlp_ref.N_PARAMS, 97))
    for k in ("b1", "b2", "b3"):
        idx.append(mlp_ref.OFF[k][0] + 3)
    idx += [mlp_ref.OFF["b4"][0], mlp_ref.OFF["b4"][0] + 1, mlp_ref.OFF["W4"][0] + 5]
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss_and_grad(pp, x, t, quantise=False)[0]
              - mlp_ref.loss_and_grad(pm, x, t, quantise=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # the quantised variant stays close to the exact one
    lq, gq = mlp_ref.loss_and_grad(p, x, t, quantise=True)
    assert abs(lq - loss) / loss < 2e-2
    assert np.linalg.norm(gq - g) / np.linalg.norm(g) < 5e-2
