"""Learned alpha on the GPU through the C ABI.
  gradient (FP32 kernels) vs FP64 oracle: relative L2 <= 1e-5, loss rel <= 1e-6
  forward FP32 path vs oracle: max abs <= 1e-5 * (1 + max|y|)
  forward BF16-MFMA path vs oracle: max abs <= 0.03 max|y| + 0.01 (BF16 inputs
    and weights, FP32 accumulation)
  trained network as Newton start on the TGV box: fewer Newton updates than the
    plain start, fields equal to the plain run within 1e-7 relative L2.
  logged features (FP32 on the device) vs the oracle's FP64 definition at the
    same nodes: max abs <= 1e-4 per feature (the features are O(1); measured
    <= 4e-6)."""
import math

import numpy as np
import pytest

import o_mlp

pytestmark = pytest.mark.gpu


def _solver(pkg, **kw):
    s = pkg.Solver(**kw)
    s.define_level(0, (16, 16, 16))
    return s


def test_mlp_gradient_matches_oracle(pkg):
    s = _solver(pkg)
    s.mlp_init(11)
    p, _, _ = s.mlp_get_params()
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (300, 12)).astype(np.float32)
    y = rng.normal(0, 1, 300).astype(np.float32)
    loss_g, g_g = s.mlp_loss_grad(X, y)
    loss_o, g_o = o_mlp.loss_and_grad(p.astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    e = float(np.linalg.norm(g_g - g_o) / np.linalg.norm(g_o))
    print(f"MLP gradient GPU vs oracle: relL2 {e:.2e}, loss {loss_g:.6e} vs {loss_o:.6e}")
    assert e < 1e-5 and abs(loss_g - loss_o) <= 1e-6 * loss_o


def test_mlp_forward_paths_match_oracle(pkg):
    s = _solver(pkg)
    s.mlp_init(12)
    p, _, _ = s.mlp_get_params()
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (1000, 12)).astype(np.float32)
    ref, _ = o_mlp.forward(p.astype(np.float64), X.astype(np.float64))
    y32 = s.mlp_forward(X, mfma=False)
    y16 = s.mlp_forward(X, mfma=True)
    e32 = float(np.abs(y32 - ref).max())
    e16 = float(np.abs(y16 - ref).max())
    print(f"MLP forward: FP32 max abs {e32:.2e}, BF16-MFMA max abs {e16:.2e}, max|y| {np.abs(ref).max():.3f}")
    assert e32 <= 1e-5 * (1 + np.abs(ref).max())
    assert e16 <= 0.03 * np.abs(ref).max() + 0.01


def _tgv(pkg, mode, L=48):
    u0 = 0.1 / math.sqrt(3.0)
    nu = u0 * (L / (2 * math.pi)) / 1600.0
    s = pkg.Solver(precision="fp32s", nu=nu, alpha_mode=mode, alpha_tol=1e-10)
    s.define_level(0, (L, L, L))
    s.init_tgv(0, u0, 1e-3, 20260101)
    return s


def test_learned_start_reduces_newton_updates(pkg, orc):
    """Train on pairs logged from the plain solver, then run the same state
    with the plain start, an untrained network and the trained network."""
    L = 48
    plain = _tgv(pkg, "newton", L)
    plain.step(0, 10)
    plain.mlp_reserve(4_000_000)
    for _ in range(8):
        n = plain.mlp_collect(0, 1)
    assert n > 200_000
    params0 = None
    plain.mlp_init(5)
    params0 = plain.mlp_get_params()
    l0, l1 = plain.mlp_train(3000, 4096, 2e-3, 9)
    print(f"training: {n} pairs, loss {l0:.3e} -> {l1:.3e}")
    assert l1 < 0.2 * l0
    trained = plain.mlp_get_params()
    f_state = plain.download_f(0)
    # FP64 oracle from the same state: both starts must meet the same field error
    u0 = 0.1 / math.sqrt(3.0)
    nu = u0 * (L / (2 * math.pi)) / 1600.0
    f_ref, _, _ = orc.step_periodic27((L, L, L), f_state, 1 / (6 * nu + 1), nsteps=10, tol=1e-12)
    m_ref = orc.macro27(f_ref)
    results = {}
    for name, mode, params in (("plain", "newton", None), ("untrained", "learned", params0),
                               ("trained", "learned", trained)):
        s = _tgv(pkg, mode, L)
        if params is not None:
            s.mlp_set_params(*params)
        s.upload_f(0, f_state)
        s.reset_stats()
        s.step(0, 10)
        st = s.stats()
        results[name] = (st["newton_iters"] / st["node_updates"], st["n_fallback"] / st["node_updates"],
                         s.macro(0))
    ip, _, mp = results["plain"]
    it, ft, mt = results["trained"]
    iu, fu, _ = results["untrained"]

    def err(m):
        u = np.stack(m[1:])
        ur = np.stack(m_ref[1:])
        return float(np.linalg.norm(u - ur) / np.linalg.norm(ur))
    ep, et = err(mp), err(mt)
    print(f"Newton updates/node: plain {ip:.3f}, untrained {iu:.3f} (fallback {fu:.3f}), "
          f"trained {it:.3f} (fallback {ft:.4f}); ratio plain/trained {ip / it:.2f}; "
          f"u relL2 vs FP64 oracle: plain {ep:.2e}, trained {et:.2e}")
    assert it < 0.8 * ip and it < iu
    # equal field error against the FP64 oracle
    assert et < 1e-5 and et < 2 * ep + 1e-7


def test_logged_features_match_oracle(pkg):
    L = 24
    s = _tgv(pkg, "newton", L)
    s.step(0, 5)
    f_post = s.download_f(0).reshape(27, L, L, L)  # (q, z, y, x)
    s.mlp_reserve(L ** 3)
    n = s.mlp_collect(0, 1)
    assert n > L ** 3 // 2
    X, _ = s.mlp_samples(n)
    nodes = s.mlp_sample_nodes(n)
    assert len(np.unique(nodes)) == n and nodes.max() < L ** 3
    import oracle as orc_mod
    cx, cy, cz, _, _ = orc_mod.lattice27()
    # pre-collision populations: periodic pull f_i(x) <- f_i(x - c_i)
    f_pre = np.stack([np.roll(f_post[i], (cz[i], cy[i], cx[i]), axis=(0, 1, 2)) for i in range(27)])
    f_pre = f_pre.reshape(27, -1)[:, nodes]
    phi, m = o_mlp.features27(f_pre)
    err = np.abs(X.astype(np.float64) - phi).max(0)
    print("logged features vs oracle, max abs per feature:", " ".join(f"{e:.1e}" for e in err))
    assert np.all(err <= 1e-4)


def test_learned_alpha_at_bench_schedule_duct(pkg):
    """Regression: pairs logged in the impulsive start of the walled duct (the
    benchmark's 5 + 20 step schedule, scaled to 96^3) are finite, include
    near-equilibrium nodes (m down to 1e-8, where the ratio features once
    overflowed FP32), train to a finite, falling loss, and the trained start
    needs fewer Newton updates than the plain start."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "bench", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    L = 96
    s, _, _, _ = bench.make_c2(pkg, L, 0.002, "newton")
    s.step(0, 25)
    s.mlp_reserve(3 * L ** 3)
    for _ in range(3):
        n = s.mlp_collect(0, 1)
    X, y = s.mlp_samples(n)
    print(f"duct {L}^3 after 25 steps: {n} pairs, |phi| max {np.abs(X).max():.3f}, "
          f"min m 10^{4 * (X[:, 0].min() - 1):.2f}")
    assert n > L ** 3
    assert np.isfinite(X).all() and np.isfinite(y).all()
    assert X[:, 0].min() < 1 - 1.5  # some m below 10^-6 (log10(m) / 4 + 1 < -0.5)
    assert np.abs(X[:, 1:5]).max() <= 1.0 + 1e-5  # |S_n / S_2| / m^(n-2) <= 1
    s.mlp_init(20260105)
    l0, l1 = s.mlp_train(800, 4096, 2e-3, 20260105)
    print(f"training loss {l0:.3e} -> {l1:.3e}")
    assert math.isfinite(l0) and math.isfinite(l1) and l1 < 0.5 * l0
    params = s.mlp_get_params()
    f_state = s.download_f(0)
    sl, _, _, _ = bench.make_c2(pkg, L, 0.002, "learned")
    sl.mlp_set_params(*params)
    sl.upload_f(0, f_state)
    its = {}
    for name, solver in (("plain", s), ("learned", sl)):
        solver.reset_stats()
        solver.step(0, 10)
        st = solver.stats()
        its[name] = st["newton_iters"] / st["node_updates"]
    print(f"Newton updates/node: plain {its['plain']:.3f}, learned {its['learned']:.3f}")
    assert its["learned"] < its["plain"]


def test_collected_pairs_are_ordered_and_reproducible(pkg):
    """Each collection's pairs sit in node order, whatever order the
    wavefronts claimed their slots in, so two identical runs log bit-equal
    training sets and train to bit-equal parameters."""
    L = 24
    runs = []
    for _ in range(2):
        s = _tgv(pkg, "newton", L)
        s.step(0, 5)
        s.mlp_reserve(3 * L ** 3)
        n1 = s.mlp_collect(0, 1)
        n2 = s.mlp_collect(0, 2)
        X, y = s.mlp_samples(n2)
        nodes = s.mlp_sample_nodes(n2)
        for a, b in ((0, n1), (n1, n2)):
            assert b - a > 1 and np.all(np.diff(nodes[a:b].astype(np.int64)) > 0)
        s.mlp_init(7)
        l0, l1 = s.mlp_train(50, 1024, 2e-3, 7)
        runs.append((X, y, nodes, l0, l1, s.mlp_get_params()))
    (X0, y0, n0, a0, b0, p0), (X1, y1, n1_, a1, b1, p1) = runs
    assert np.array_equal(n0, n1_) and np.array_equal(X0, X1) and np.array_equal(y0, y1)
    assert a0 == a1 and b0 == b1
    for u, v in zip(p0, p1):
        assert np.array_equal(np.asarray(u), np.asarray(v))
