"""Learned-alpha model, FP64 numpy restatement.  TEST INFRASTRUCTURE ONLY:
the checker for the library's feature, forward, backward and Adam kernels.

The reference has no learned component; its closest precedent is the
hand-fitted Lagrange-multiplier guesses of amr/modules/amr_processes.f95:
1871-2039, whose use is commented out (calculate_lagrangian_mults.f95:729-741),
and "Implement machine learning to shorten Jacobian solution time" (Future
Work.txt:34).  The model here predicts the entropic alpha of a D3Q27 node as
the initial guess of the Newton root-find of H(f + alpha D) = H(f).

Features (12) of a node's pre-collision state, with x_i = D_i / f_i,
S_n = sum_i f_i x_i^n = sum_i D_i x_i^(n-1), m = max|x_i|:
    0  log10(m) / 4 + 1
    1-4  (S_n / S2) / m^(n-2), n = 3..6
    5  min_i x_i / m
    6  S2 / (rho m^2)
    7-9  5 u
    10 25 |u|^2
    11 log10(S2) / 8 + 1
Target y = (alpha - 2) / m; alpha0 = 2 + m (y_mean + y_std * net(phi)).
Network: 12 -> 64 -> 64 -> 64 -> 1, GELU (tanh form, Hendrycks & Gimpel
2016) after the three hidden layers, 9,217 parameters, packed as
W1 (12x64, row-major in x out), b1, W2 (64x64), b2, W3 (64x64), b3, W4 (64), b4.
"""
from __future__ import annotations

import numpy as np

SIZES = [(12, 64), (64, 64), (64, 64), (64, 1)]
N_PARAMS = sum(i * o + o for i, o in SIZES)  # 9217
_C = np.sqrt(2.0 / np.pi)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(_C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(_C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * _C * (1.0 + 3 * 0.044715 * x * x)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, k = [], 0
    for i, o in SIZES:
        W = p[k:k + i * o].reshape(i, o)
        k += i * o
        b = p[k:k + o]
        k += o
        out.append((W, b))
    return out


def init_params(seed):
    """He-style uniform init (the library draws the same numbers from
    Philox4x32-10; here any draw is fine for the checks)."""
    rng = np.random.default_rng(seed)
    p = []
    for i, o in SIZES:
        lim = np.sqrt(6.0 / i)
        p.append(rng.uniform(-lim, lim, i * o))
        p.append(np.zeros(o))
    return np.concatenate(p)


def forward(p, X):
    """X: (B, 12).  Returns (yhat (B,), cache)."""
    layers = unpack(p)
    a = np.asarray(X, np.float64)
    cache = []
    for li, (W, b) in enumerate(layers):
        z = a @ W + b
        cache.append((a, z))
        a = gelu(z) if li < 3 else z
    return a[:, 0], cache


def loss_and_grad(p, X, y):
    """Mean squared error 0.5 mean (yhat - y)^2 and its gradient (packed)."""
    yhat, cache = forward(p, X)
    B = len(y)
    r = yhat - np.asarray(y, np.float64)
    loss = 0.5 * np.mean(r * r)
    layers = unpack(p)
    grads = [None] * 4
    d = (r / B)[:, None]  # dL/dz of the output layer
    for li in range(3, -1, -1):
        a, z = cache[li]
        W, b = layers[li]
        if li < 3:
            d = d * gelu_grad(z)
        grads[li] = (a.T @ d, d.sum(0))
        d = d @ W.T
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015), bias-corrected; returns p, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


W27 = None


def features27(f_pre, rho_u=None):
    """Features of pre-collision D3Q27 populations f_pre (27, N) (full values),
    in the lattice order of linkwise.f95:30-37.  Returns (phi (N, 12), m (N,))."""
    import oracle as _o
    cx, cy, cz, opp, w = _o.lattice27()
    f = np.asarray(f_pre, np.float64)
    rho = f.sum(0)
    ux = (cx[:, None] * f).sum(0) / rho
    uy = (cy[:, None] * f).sum(0) / rho
    uz = (cz[:, None] * f).sum(0) / rho
    N = f.shape[1]
    feq = np.empty_like(f)
    for k in range(N):
        feq[:, k] = _o.feq27(rho[k], ux[k], uy[k], uz[k])
    D = feq - f
    x = D / f
    m = np.abs(x).max(0)
    S = {n: (D * x ** (n - 1)).sum(0) for n in range(2, 7)}
    phi = np.empty((N, 12))
    mm = np.maximum(m, 1e-30)
    phi[:, 0] = np.log10(mm) / 4 + 1
    for j, n in enumerate(range(3, 7)):
        phi[:, 1 + j] = S[n] / S[2] / mm ** (n - 2)
    phi[:, 5] = x.min(0) / mm
    phi[:, 6] = S[2] / (rho * mm * mm)
    phi[:, 7], phi[:, 8], phi[:, 9] = 5 * ux, 5 * uy, 5 * uz
    phi[:, 10] = 25 * (ux * ux + uy * uy + uz * uz)
    phi[:, 11] = np.log10(np.maximum(S[2], 1e-300)) / 8 + 1
    return phi, m
