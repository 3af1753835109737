"""The learned-alpha oracle (oracle/o_mlp.py) pinned by mathematics: its
gradient against central finite differences of its own loss (FP64), GELU's
derivative against the finite difference of GELU, Adam's first step
(|update| = lr for every parameter with a non-zero gradient), and the
features against their definitions on an equilibrium and a hand-built state."""
import numpy as np

import o_mlp
import oracle as o


def test_parameter_count():
    assert o_mlp.N_PARAMS == 9217


def test_gelu_derivative():
    x = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (o_mlp.gelu(x + h) - o_mlp.gelu(x - h)) / (2 * h)
    assert np.abs(fd - o_mlp.gelu_grad(x)).max() < 1e-8


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = o_mlp.init_params(5) + rng.normal(0, 0.05, o_mlp.N_PARAMS)
    X = rng.uniform(-1, 1, (32, 12))
    y = rng.normal(0, 1, 32)
    loss, g = o_mlp.loss_and_grad(p, X, y)
    idx = rng.choice(o_mlp.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [0, 767, 768, 9151, 9152, 9216]])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (o_mlp.loss_and_grad(p + e, X, y)[0] - o_mlp.loss_and_grad(p - e, X, y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_adam_first_step_is_lr_sign():
    rng = np.random.default_rng(4)
    p = rng.normal(0, 1, 100)
    g = rng.normal(0, 1, 100)
    p1, m, v = o_mlp.adam(p, g, np.zeros(100), np.zeros(100), 1, 1e-3)
    assert np.allclose(p - p1, 1e-3 * np.sign(g), rtol=1e-6)


def test_features_at_equilibrium_and_by_definition():
    cx, cy, cz, opp, w = o.lattice27()
    feq = o.feq27(1.02, 0.03, -0.01, 0.02)
    rng = np.random.default_rng(8)
    f = feq * (1 + 1e-3 * rng.uniform(-1, 1, 27))
    phi, m = o_mlp.features27(np.stack([f, feq], 1))
    # second column: equilibrium, m = 0 up to round-off
    assert m[1] < 1e-12
    rho = f.sum()
    ux = (cx * f).sum() / rho
    D = o.feq27(rho, ux, (cy * f).sum() / rho, (cz * f).sum() / rho) - f
    x = D / f
    S2, S3 = (f * x ** 2).sum(), (f * x ** 3).sum()
    assert abs(m[0] - np.abs(x).max()) < 1e-15
    assert abs(phi[0, 1] - S3 / S2 / m[0]) < 1e-10
    assert abs(phi[0, 7] - 5 * ux) < 1e-14
    assert abs(phi[0, 11] - (np.log10(S2) / 8 + 1)) < 1e-12
