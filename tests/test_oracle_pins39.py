"""Pins of the D3Q39 oracle's heuristic pieces by mathematics that does not
restate them.

alpha estimate (calculate_alpha.f95:118-289).  For a small non-equilibrium
part of amplitude e (x_i = D_i / f_i = O(e), D moment-free) the exact
non-trivial root of H(f + a D) = H(f), H = sum f ln f, is
    a* = 2 + S3 / (3 S2) + O(e^2),   S_n = sum_i f_i x_i^n
(expand (1 + a x) ln(1 + a x) to third order).  The quadratic of the
estimate, a A^2 + b A + c = 0, expanded to the same order gives
    A = 2 - (S3 - (4/3) B^2 S3m - (2/3) a_k B^2 S3p) / S2 + O(e^2),
B = beta_mfp, S3m / S3p the sums over x in [-1, 0] / x >= 0, a_k the
start-up value of the lower bound a_sm (a_sm itself past the start-up, so at
B = 1 the estimate is the exact root to first order).  The tests check
  * A - a* = O(e^2) at B ~ 1 (ratio ~4 per doubling of e), against a root
    found by bisection on the log1p form of H;
  * A - 2 - (prediction) = O(e^2) for B ~ 1 and B = 0.02 (tiny grid
    spacing), at steps 0, 10 and past the start-up, with B from the kinetic
    theory mean free time tau = lambda / sqrt(3 k T / m),
    lambda = mu / p sqrt(pi k T / (2 m)) (CODATA k) and mu the Sutherland
    viscosity (bit-exact against the reference elsewhere);
  * the early exits (no cubic deviation; |a| < 1e-14) give 2 and the lower
    clamp returns the lower bound a_sm, which lies below the exact root.
A wrong 1/6, 2/15, the (2, 1, 2) / (4, 4, 4) weights of second_term_2, the
60/60/90 first-order part of third_total, B vs B^2 or the ramp form all
change the first-order coefficient.  Not pinned here: the second-order
coefficients (the 11 x^4 term, 36 and 3 in third_total, the alpha_hi
weighted first_term_2..4) and the upper clamp alpha_hi = 2.125, which no
moment-free state reached in a seeded search.

q-bar and g-bar (find_q_bar.f95:52-220, find_little_q_bar.f95:23-52).
  * a deviation with a single central third moment Q_xxx (x axis only,
    u along x) gives qbar_i = eps Q_xxx (c_x^3 - 3 T c_x) / (6 T^3), likewise
    along y and z; one with only Q_xyz (the eight (+-1, +-1, +-1) links)
    gives eps Q_xyz c_x c_y c_z / T^3;
  * any deviation: qbar_i = eps / (6 T^3) sum_abc Q_abc H_abc(c_i) with the
    symmetric Hermite tensor H_abc = c_a c_b c_c - T (c_a d_bc + c_b d_ac +
    c_c d_ab) and Q from a tensor contraction (equal to the reference's
    27-entry table and kroen_val for a symmetric Q);
  * g-bar for a deviation along x only is eps c_x (sum c_x dg) / T;
  * the collision of a non-equilibrium state is f + a b1 (D - qbar) + b2 qbar
    with b1, b2 from the (pinned) viscosity and conductivity and the
    start-up epsilon ramp (step + 1) / startup eps."""
import math

import numpy as np
import pytest

import oracle as o

T0 = 1 / 1.4
K_CODATA = 1.380649e-23


def _mrows(shifted=False):
    cx, cy, cz = o.velocities39(shifted)
    c2 = cx * cx + cy * cy + cz * cz
    return np.stack([np.ones(39), cx, cy, cz, cx * cx, cy * cy, cz * cz, cx * cy, cx * cz, cy * cz,
                     cx * c2, cy * c2, cz * c2], 1).astype(float)


def _eq(u, T=T0, rho=1.0):
    u = np.asarray(u, float)
    P = np.array([rho * (T + u[0] ** 2), rho * (T + u[1] ** 2), rho * (T + u[2] ** 2),
                  rho * u[0] * u[1], rho * u[0] * u[2], rho * u[1] * u[2]])
    lm0 = np.zeros(13)
    lm0[0] = -1.5 * math.log(2 * math.pi * T) - (u @ u) / (2 * T)
    lm0[1:4] = u / T
    lm0[4:7] = -1 / (2 * T)
    lm, _, it = o.lm_solve39(rho, u, T, P, lm0, maxit=100)
    assert it > 0
    return lm, o.feq39(lm, rho)


def _shape(feq, seed):
    """A moment-free deviation shape (none of the 13 constrained moments)."""
    M = _mrows()
    d = feq * np.random.default_rng(seed).uniform(-1, 1, 39)
    return d - M @ np.linalg.solve(M.T @ M, M.T @ d)


def _sums(f, D):
    x = D / np.abs(f)
    S2, S3 = np.sum(f * x ** 2), np.sum(f * x ** 3)
    m = (x <= 0) & (x >= -1)
    return x, S2, S3, np.sum((f * x ** 3)[m]), np.sum((f * x ** 3)[x >= 0])


def _beta_mfp(c, T):
    _, dmu, dT = o.sutherland39(T, c)
    lam = dmu / c.p_ref * math.sqrt(math.pi * K_CODATA * dT / (2 * c.molec_weight))
    tau = lam / math.sqrt(3 * K_CODATA * dT / c.molec_weight)
    return c.dt / (2 * tau + c.dt)


def _entropic_root(f, D, lnfeq):
    """Non-trivial root of H(f + a D) - H(f) = 0 in (1, 3) by bisection on
    sum f ((1 + a x) log1p(a x) - a x log1p(x)) + a sum D ln feq."""
    x = D / f
    def g(a):
        return float(np.sum(f * ((1 + a * x) * np.log1p(a * x) - a * x * np.log1p(x)))
                     + a * np.sum(D * lnfeq))
    lo, hi = 1.0, 3.0
    assert g(lo) < 0 < g(hi)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if g(mid) < 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_alpha_estimate_is_the_entropic_root_to_first_order():
    c = o.consts39(mach=0.3)  # dx 1 mm: beta_mfp = 1 - 9e-5
    assert _beta_mfp(c, T0) > 0.9998
    lm, feq = _eq((0.3, 0.05, -0.02))
    d = _shape(feq, 5)
    r = []
    for e in (2.5e-4, 5e-4, 1e-3):
        f = feq + e * d
        D = feq - f
        A, st = o.alpha_est39(f, D, T0, 100, c)
        ast = _entropic_root(f, D, np.log(feq))
        _, S2, S3, _, _ = _sums(f, D)
        assert st == 0
        r.append(A - ast)
        # first order of the exact root
        assert abs((ast - 2) - S3 / (3 * S2)) < 0.03 * abs(S3 / (3 * S2))
    print("estimate - exact root:", " ".join(f"{v:.3e}" for v in r))
    assert abs(r[0]) < 0.02 * abs(S3 / (3 * S2)) * 0.25
    for k in range(2):
        assert 3.3 < r[k + 1] / r[k] < 4.7, r


@pytest.mark.parametrize("dx", [1e-3, 2e-9])
@pytest.mark.parametrize("step", [0, 10, 100])
def test_alpha_estimate_first_order_coefficients(dx, step):
    c = o.consts39(mach=0.3, dx=dx, startup=25)
    B = _beta_mfp(c, T0)
    lm, feq = _eq((0.3, 0.05, -0.02))
    d = _shape(feq, 5)
    r, pred = [], []
    # with beta_mfp = 0.02 the quadratic's leading coefficient B^2 S3m / 6
    # falls below the 1e-14 exit at amplitudes under ~1e-3
    amps = (5e-4, 1e-3, 2e-3) if B > 0.5 else (1e-3, 2e-3, 4e-3)
    for e in amps:
        f = feq + e * d
        D = feq - f
        A, st = o.alpha_est39(f, D, T0, step, c)
        assert st == 0
        x, S2, S3, S3m, S3p = _sums(f, D)
        # Atif et al.'s lower bound (the chi-sum root) and its start-up ramp
        cs1 = np.sum(f * 2 * x ** 2 / (2 + x))
        a_sm = (S2 - math.sqrt(S2 ** 2 - 8 * S3m * cs1)) / (2 * S3m)
        a_k = (a_sm - 1e-6) / 25 * step if step < 25 else a_sm
        p = -(S3 - (4 / 3) * B ** 2 * S3m - (2 / 3) * a_k * B ** 2 * S3p) / S2
        pred.append(p)
        r.append(A - 2 - p)
    print(f"dx {dx:g} beta_mfp {B:.5f} step {step}: first-order prediction {pred[0]:+.4e}, "
          f"remainders {r[0]:.2e} {r[1]:.2e} {r[2]:.2e}")
    assert abs(r[0]) < 0.03 * abs(pred[0])
    for k in range(2):
        assert 3.3 < r[k + 1] / r[k] < 4.7, r


def test_alpha_estimate_early_exits_and_lower_clamp():
    c = o.consts39(mach=0.3)
    lm, feq = _eq((0.3, 0.0, 0.0))
    assert o.alpha_est39(feq, np.zeros(39), T0, 100, c) == (2.0, 1)
    # grid spacing 0.5 nm: beta_mfp = 0.005, |a| of the quadratic < 1e-14
    c_small = o.consts39(mach=0.3, dx=5e-10)
    f = feq + 1e-3 * _shape(feq, 5)
    assert o.alpha_est39(f, feq - f, T0, 100, c_small) == (2.0, 1)
    # sparse over-populations (seeded search): the lower clamp returns the
    # lower bound, which lies below the exact root where that exists
    M = _mrows()
    rng = np.random.default_rng(1)
    hits = 0
    for t in range(3000):
        d = np.zeros(39)
        ks = rng.choice(np.arange(1, 39), rng.integers(1, 4), replace=False)
        d[ks] = feq[ks] * 10 ** rng.uniform(-1, 1.5, len(ks))
        d -= M @ np.linalg.solve(M.T @ M, M.T @ d)
        f = feq + d
        if np.any(f <= 0):
            continue
        D = feq - f
        A, st = o.alpha_est39(f, D, T0, 100, c)
        if st != 2:
            continue
        hits += 1
        x, S2, S3, S3m, _ = _sums(f, D)
        cs1 = np.sum(f * 2 * x ** 2 / (2 + x))
        a_sm = (S2 - math.sqrt(S2 ** 2 - 8 * S3m * cs1)) / (2 * S3m)
        assert abs(A - min(a_sm, 2.125 - 1e-9)) < 1e-13
        neg = D < 0
        amax = np.min(-f[neg] / D[neg])
        x2 = D / f
        g = lambda a: float(np.sum(f * ((1 + a * x2) * np.log1p(a * x2) - a * x2 * np.log1p(x2)))
                            + a * np.sum(D * np.log(feq)))
        if amax > 3 or g(0.999999 * amax) > 0:
            assert g(A) < 0  # A below the root: H(f + A D) < H(f)
    assert hits >= 2


def _cvec(shifted=False):
    cx, cy, cz = o.velocities39(shifted)
    return np.stack([cx, cy, cz], 1).astype(float)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_qbar_single_axis_moment(axis):
    c = _cvec()
    T, eps = 0.8, 3e-3
    on_axis = np.all(np.delete(c, axis, 1) == 0, 1) & (c[:, axis] != 0)
    D = np.where(on_axis, c[:, axis] * np.array([0.7 if abs(v) == 1 else (-0.2 if abs(v) == 2 else 0.05)
                                                   for v in c[:, axis]]), 0.0) * 1e-3
    u = np.zeros(3)
    u[axis] = 0.37
    Q = np.sum((c[:, axis] - u[axis]) ** 3 * D)
    assert abs(Q) > 1e-6
    want = eps * Q * (c[:, axis] ** 3 - 3 * T * c[:, axis]) / (6 * T ** 3)
    got = o.qbar39(D, u, T, eps)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-18)


def test_qbar_xyz_moment_and_general_hermite_form():
    c = _cvec()
    T, eps = 0.75, 2e-3
    corner = np.all(np.abs(c) == 1, 1)
    D = np.where(corner, c[:, 0] * c[:, 1] * c[:, 2], 0.0) * 4e-4
    Qxyz = np.sum(c[:, 0] * c[:, 1] * c[:, 2] * D)
    got = o.qbar39(D, np.zeros(3), T, eps)
    assert np.allclose(got, eps * Qxyz * c[:, 0] * c[:, 1] * c[:, 2] / T ** 3, rtol=1e-13, atol=1e-18)
    rng = np.random.default_rng(4)
    I3 = np.eye(3)
    for t in range(20):
        D = rng.normal(0, 1e-3, 39)
        u = rng.uniform(-0.4, 0.4, 3)
        T = rng.uniform(0.5, 1.2)
        xi = c - u
        Q = np.einsum("i,ia,ib,ic->abc", D, xi, xi, xi)
        H = (np.einsum("ia,ib,ic->iabc", c, c, c)
             - T * (np.einsum("ia,bc->iabc", c, I3) + np.einsum("ib,ac->iabc", c, I3)
                    + np.einsum("ic,ab->iabc", c, I3)))
        want = eps / (6 * T ** 3) * np.einsum("abc,iabc->i", Q, H)
        got = o.qbar39(D, u, T, eps)
        assert np.allclose(got, want, rtol=1e-11, atol=1e-15 * np.abs(want).max())


def test_little_qbar_single_axis():
    c = _cvec()
    T, eps = 0.9, 1e-3
    on_x = (c[:, 1] == 0) & (c[:, 2] == 0) & (c[:, 0] != 0)
    dg = np.where(on_x, c[:, 0] * 1e-3, 0.0)
    u = np.array([0.2, 0.0, 0.0])
    want = eps * c[:, 0] * np.sum(c[:, 0] * dg) / T
    assert np.allclose(o.little_qbar39(dg, u, T, eps), want, rtol=1e-13, atol=1e-18)


@pytest.mark.parametrize("step", [3, 100])
def test_collision_of_a_nonequilibrium_state(step):
    c = o.consts39(mach=0.3, startup=25)
    rho, u, T = 0.97, np.array([0.25, -0.05, 0.08]), 1.05 * T0
    lm, feq = _eq(u, T, rho)
    geq = c.lead_g * T * feq
    rng = np.random.default_rng(8)
    f = feq * (1 + 2e-3 * rng.uniform(-1, 1, 39))
    g = geq * (1 + 2e-3 * rng.uniform(-1, 1, 39))
    alpha = 1.93
    fo, go = o.collide39(f, g, lm, rho, u, T, alpha, step, c)
    D, Dg = feq - f, geq - g
    eps = c.eps * ((step + 1) / 25 if step < 25 else 1.0)
    cv = _cvec()
    xi = cv - u
    Q = np.einsum("i,ia,ib,ic->abc", D, xi, xi, xi)
    I3 = np.eye(3)
    H = (np.einsum("ia,ib,ic->iabc", cv, cv, cv)
         - T * (np.einsum("ia,bc->iabc", cv, I3) + np.einsum("ib,ac->iabc", cv, I3)
                + np.einsum("ic,ab->iabc", cv, I3)))
    qb = eps / (6 * T ** 3) * np.einsum("abc,iabc->i", Q, H)
    hq = eps / T * cv @ (xi.T @ Dg)
    if step < 25:
        b1 = c.init_beta1 + c.beta1_scale * step
        b2 = c.init_beta2 + c.beta2_scale * step
    else:
        mu, _, _ = o.sutherland39(T, c)
        kappa = o.therm_cond39(rho, T, c)
        b1 = 1 / (2 * mu / (rho * T) + 1)
        b2 = 1 / (2 * kappa / (rho * c.cp * T) + 1)
    # eps is the level time step dt (timing_calcs.f95:24-26): qbar is small but
    # far above the round-off of the comparison
    assert np.abs(qb).max() > 1e-11 and np.abs(hq).max() > 1e-11
    assert np.allclose(fo, f + alpha * b1 * (D - qb) + b2 * qb, rtol=0, atol=1e-15)
    assert np.allclose(go, g + alpha * b1 * (Dg - hq) + b2 * hq, rtol=0, atol=1e-14)
