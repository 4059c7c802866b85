"""The CPU restatement of the learned MAC's network (oracle/mlp.py), pinned
by central finite differences of its loss and by its parameter layout."""
import numpy as np

from oracle import mlp


def test_parameter_count_matches_the_abi():
    import barytreek_amd as btk
    assert mlp.N_PARAMS == btk.MAC_WEIGHTS == 5644
    assert (mlp.N_IN, mlp.N_OUT) == (btk.MAC_FEATURES, btk.MAC_OUTPUTS)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, mlp.N_PARAMS)
    X = rng.normal(0, 1, (7, mlp.N_IN))
    Y = rng.normal(0, 1, (7, mlp.N_OUT))
    loss, g = mlp.loss_and_grad(w, X, Y)
    h = 1e-6
    idx = rng.choice(mlp.N_PARAMS, 200, replace=False)
    for i in idx:
        e = np.zeros_like(w)
        e[i] = h
        fd = (mlp.loss_and_grad(w + e, X, Y)[0] - mlp.loss_and_grad(w - e, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # a plausible slip -- a dropped ReLU mask -- changes the gradient
    assert np.abs(g).sum() > 0 and loss > 0


def test_forward_is_relu_mlp():
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.5, mlp.N_PARAMS)
    x = rng.normal(0, 1, (1, mlp.N_IN))
    p = mlp.unpack(w)
    h1 = np.maximum(p["W1"] @ x[0] + p["b1"], 0)
    h2 = np.maximum(p["W2"] @ h1 + p["b2"], 0)
    assert np.allclose(mlp.forward(w, x)[0][0], p["W3"] @ h2 + p["b3"])
