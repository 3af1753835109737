"""Oracle D3Q39 two-population step (restating calculate_flow_variables,
calculate_lagrangian_mults + lu_decomp, fieq_comp_39, calculate_alpha's
estimate, collision + find_q_bar + find_little_q_bar, quick_calcs, damp_it).

Pinned by: the reference's own lu_decomp and quick_calcs compiled from its
sources (bit-exact where built), the moment constraints the multipliers must
meet (closed form), the symmetry of the rest state, Newton's quadratic
convergence, exact conservation of mass by the collision for the un-shifted
lattice, and invariance of a uniform state."""
import ctypes
import math

import numpy as np
import pytest

import oracle as o

G = 1.4
T0 = 1 / G


def _mrows(shifted):
    cx, cy, cz = o.velocities39(shifted)
    c2 = cx ** 2 + cy ** 2 + cz ** 2
    return np.stack([np.ones(39), cx, cy, cz, cx ** 2, cy ** 2, cz ** 2, cx * cy, cx * cz, cy * cz,
                     cx * c2, cy * c2, cz * c2], 1).astype(float)


def _targets(rho, u, T):
    u = np.asarray(u, float)
    P = np.array([rho * (T + u[0] ** 2), rho * (T + u[1] ** 2), rho * (T + u[2] ** 2),
                  rho * u[0] * u[1], rho * u[0] * u[2], rho * u[1] * u[2]])
    q = (rho * (u @ u) + 5 * rho * T) * u
    return P, np.concatenate([[rho], rho * u, P, q])


def _maxwell_lm(u, T):
    u = np.asarray(u, float)
    lm = np.zeros(13)
    lm[0] = -1.5 * math.log(2 * math.pi * T) - (u @ u) / (2 * T)
    lm[1:4] = u / T
    lm[4:7] = -1 / (2 * T)
    return lm


need_ref = pytest.mark.skipif(not o.ref_available(),
                              reason="reference routines not built (no reference tree)")


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@need_ref
@pytest.mark.parametrize("shifted", [False, True])
def test_lu_step_bit_exact_with_reference(shifted):
    """ref_mode reproduces lu_decomp.f95 (including its two index slips) bit for bit."""
    M = _mrows(shifted)
    ref = o.ref_lib()
    rng = np.random.default_rng(11)
    for t in range(300):
        expo = rng.uniform(0.005, 0.3, 39)
        J = (M.T * expo) @ M
        mom = J[0] + rng.normal(0, 1e-3, 13)
        mom[1:4] /= mom[0]
        lm0 = rng.normal(0, 0.2, 13)
        lm_o, mx_o = o.lu_step39(J, mom, lm0, ref_mode=True)
        lm_r = lm0.copy()
        Jf = np.asfortranarray(J)
        mx_r = ref.ref_lu_decomp(_p(Jf), _p(mom), _p(lm_r))
        assert np.array_equal(lm_o, lm_r) and mx_o == mx_r


@pytest.mark.parametrize("shifted", [False, True])
def test_exact_lu_step_is_newton_step(shifted):
    M = _mrows(shifted)
    rng = np.random.default_rng(12)
    for t in range(100):
        expo = rng.uniform(0.005, 0.3, 39)
        J = (M.T * expo) @ M
        mom = J[0] + rng.normal(0, 1e-3, 13)
        tgt = mom.copy()
        tgt[1:4] *= 1.0
        mom[1:4] /= mom[0]
        tgt[1:4] = mom[0] * mom[1:4]
        lm0 = rng.normal(0, 0.2, 13)
        lm_e, mx = o.lu_step39(J, mom, lm0, ref_mode=False)
        d = np.linalg.solve(J, J[0] - tgt)
        assert np.allclose(lm_e, lm0 - d, rtol=0, atol=1e-11 * max(1.0, np.abs(d).max()))
        assert abs(mx - np.abs(d).max()) <= 1e-11 * np.abs(d).max()


@need_ref
def test_sutherland_and_conductivity_match_reference():
    """quick_sutherlands / quick_therm_cond bit-exact over 0.2..12 x T0 and five
    densities, in a child process (see ref_checks.py); inputs on the
    reference's error path are checked to be rejected by the oracle too."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, PYTHONPATH=os.pathsep.join(
        [here, os.path.join(here, "..", "oracle"), os.environ.get("PYTHONPATH", "")]))
    r = subprocess.run([sys.executable, os.path.join(here, "ref_checks.py"), "therm_cond"],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "REF-CHECK-COMPLETE" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("shifted,u", [(False, (0.0, 0.0, 0.0)), (False, (0.3, -0.1, 0.05)),
                                       (True, (1.5, 0.1, -0.2)), (True, (1.2, 0.0, 0.0))])
@pytest.mark.parametrize("rho,T", [(1.0, T0), (0.8, 0.6 * T0), (1.3, 1.4 * T0)])
def test_multipliers_meet_moment_constraints(shifted, u, rho, T):
    P, tgt = _targets(rho, u, T)
    lm, e, it = o.lm_solve39(rho, u, T, P, _maxwell_lm(u, T), shifted=shifted, maxit=100)
    assert it > 0
    feq = o.feq39(lm, rho, shifted)
    mom = _mrows(shifted).T @ feq
    assert np.allclose(mom, tgt, rtol=0, atol=1e-9 * np.abs(tgt).max())
    assert np.all(feq > 0)


def test_rest_state_symmetry():
    P, tgt = _targets(1.0, (0, 0, 0), T0)
    lm, e, it = o.lm_solve39(1.0, (0, 0, 0), T0, P, _maxwell_lm((0, 0, 0), T0), maxit=100)
    assert np.abs(lm[[1, 2, 3, 7, 8, 9, 10, 11, 12]]).max() < 1e-12
    assert abs(lm[4] - lm[5]) < 1e-12 and abs(lm[4] - lm[6]) < 1e-12


def test_newton_converges_quadratically():
    u, T, rho = np.array([0.25, 0.1, -0.05]), 0.9 * T0, 1.1
    P, tgt = _targets(rho, u, T)
    M = _mrows(False)
    lm = _maxwell_lm(u, T) + 0.02
    steps = []
    for k in range(6):
        feq = o.feq39(lm, rho)
        J = (M.T * feq) @ M
        mom = np.concatenate([[rho], u, P, tgt[10:]])
        lm, mx = o.lu_step39(J, mom, lm)
        steps.append(mx)
        if mx < 1e-13:
            break
    # |delta_{k+1}| <= C |delta_k|^2 once in the basin
    for a, b in zip(steps[1:-1], steps[2:]):
        assert b <= 50 * a * a + 1e-14


def test_alpha_estimate_at_equilibrium_and_near_it():
    c = o.consts39(mach=0.3)
    u, T = np.array([0.3, 0.0, 0.0]), T0
    P, _ = _targets(1.0, u, T)
    lm, feq, it = o.lm_solve39(1.0, u, T, P, _maxwell_lm(u, T), maxit=100)
    feq = o.feq39(lm, 1.0)
    a, st = o.alpha_est39(feq, np.zeros(39), T, 100, c)
    assert a == 2.0 and st == 1   # no non-equilibrium: alpha = 2 (calculate_alpha.f95:171-175)
    rng = np.random.default_rng(5)
    M = _mrows(False)
    for amp in (1e-3, 1e-2):
        # a non-equilibrium part with none of the 13 constrained moments, so
        # that feq is the equilibrium of f
        d = amp * feq * rng.uniform(-1, 1, 39)
        d -= M @ np.linalg.solve(M.T @ M, M.T @ d)
        f = feq + d
        a, st = o.alpha_est39(f, feq - f, T, 100, c)
        # the estimate targets the non-trivial root of H(f + a D) = H(f), H = sum f ln f
        D = feq - f
        H = lambda x: float(np.sum(x * np.log(x)))
        lo, hi = 1.0, 3.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if H(f + mid * D) - H(f) < 0:
                lo = mid
            else:
                hi = mid
        assert abs(a - lo) < 20 * amp, (amp, a, lo)


def test_collision_conserves_mass_unshifted_and_keeps_equilibrium():
    c = o.consts39(mach=0.3)
    rng = np.random.default_rng(6)
    u, T, rho = np.array([0.2, -0.1, 0.05]), 1.1 * T0, 0.95
    P, _ = _targets(rho, u, T)
    lm, e, it = o.lm_solve39(rho, u, T, P, _maxwell_lm(u, T), maxit=100)
    feq = o.feq39(lm, rho)
    geq = c.lead_g * T * feq
    fo, go = o.collide39(feq, geq, lm, rho, u, T, 2.0, 100, c)
    assert np.abs(fo - feq).max() < 1e-15 and np.abs(go - geq).max() < 1e-15
    f = feq * (1 + 1e-3 * rng.uniform(-1, 1, 39))
    g = geq * (1 + 1e-3 * rng.uniform(-1, 1, 39))
    r, uu, TT, PP = o.moments39(f, g, c)
    lm2, e2, it2 = o.lm_solve39(r, uu, TT, PP, lm, maxit=100)
    fo, go = o.collide39(f, g, lm2, r, uu, TT, 1.9, 100, c)
    assert abs(fo.sum() - f.sum()) < 1e-12


@pytest.mark.parametrize("shifted,mach", [(False, 0.3), (True, 1.5)])
def test_uniform_state_is_a_fixed_point(shifted, mach):
    c = o.consts39(mach=mach, shifted=shifted)
    shape = (6, 5, 4)
    f, g, lm = o.init39(shape, 1.0, mach, 0.0, 0.0, T0, c)
    f2, g2, lm2, al, st = o.step_periodic39(shape, f, g, lm, c, step0=30, nsteps=5, keep_alpha=True)
    assert np.abs(f2 - f).max() < 1e-12 and np.abs(g2 - g).max() < 1e-12
    assert np.all(al == 2.0) and st["n_repair"] == 0


def test_damp_it_rules_and_order():
    """A negative population takes the mean of its positive neighbours along
    +-c_i (damp_it.f95:64-92).  Isolated negatives: the sweep order cannot
    matter, so the reference order and the order-free form agree exactly."""
    shape = (9, 8, 7)
    n = 9 * 8 * 7
    rng = np.random.default_rng(3)
    fo = rng.uniform(0.01, 0.1, (39, n))
    neg = rng.choice(n, 12, replace=False)
    for k in neg:
        fo[rng.integers(1, 39), k] = -0.005
    a = fo.copy()
    b = fo.copy()
    na = o.lib().o39_damp_periodic(9, 8, 7, o._dp(a), 0, 1)
    nb = o.lib().o39_damp_periodic(9, 8, 7, o._dp(b), 0, 0)
    assert na == nb == 12 and np.array_equal(a, b) and np.all(a > 0)
    # a chain of negatives along one link: orders differ, both stay positive
    cx, cy, cz = o.velocities39(False)
    fo2 = rng.uniform(0.01, 0.1, (39, n))
    for x in range(2, 5):
        fo2[1, x + 9 * (3 + 8 * 3)] = -0.01
    a = fo2.copy()
    b = fo2.copy()
    o.lib().o39_damp_periodic(9, 8, 7, o._dp(a), 0, 1)
    o.lib().o39_damp_periodic(9, 8, 7, o._dp(b), 0, 0)
    assert np.all(a > 0) and np.all(b > 0) and not np.array_equal(a, b)
