"""Pins of the oracle's shifted-lattice wall rule (o_walls39.c), whose
reference (shifted_ibb.f95, derivatives.f95, streaming.f95's shifted branch)
needs AMReX and cannot be compiled here:

  stencil kinds: for every direction and axis the wall derivative follows the
    case table of derivatives.f95:816-1235 (velocity) and :1236- (temperature)
    as extracted into tests/golden/wall_kinds39.json (the kind is the sign of
    the link's unshifted component, except for the four links (2,0,-2),
    (-2,0,2), (0,2,-2), (0,-2,2), whose cases carry the reversed kinds): + kind (0 - u_t) /
    (1 + w), - kind (u_t - 0) / (1 + w), w = q |c_i| with the shifted
    lattice's link length (directions_and_weights.f95:477-519), temperature
    zero for the wall kinds, the local difference otherwise;
  local differences are exact for a linear field (central, and one-sided
    when a neighbour is not fluid);
  the 81 transcribed P, Q, R entries equal the textbook Grad-Hermite tensors
    P_ab = (S_ab - d_ab div/c_v) / beta_1,
    Q_abc = d_bc t_a + d_ac t_b + d_ab t_c + u_a P_bc + u_b P_ac + u_c P_ab,
    r_ab = T S_ab + u_a dT_b + u_b dT_a - div/c_v (2 T d_ab + u_a u_b)
    (S_ab = d_a u_b + d_b u_a, t = grad T / beta_2) summed over the 27 index
    triples against H2 = c c - T I and H3 = c c c - T (c I)_sym, with the
    reference's three term slips and nothing else: P_zzx counted twice
    (shifted_ibb.f95:170), Q_zxx taking dT/dx for dT/dz (:154; its basis
    term is c_x^2 c_z, :180), the zy entries of R repeating yy (:231-233).  The entries differ in which
    diagonal velocity derivative stands for the divergence; with the three
    equal every version coincides, which is the case pinned here;
  zero gradient: a fluid at rest at uniform T around a body rebuilds every
    wall link as the target equilibrium feq_i(rho_t, 0, T) with rho_t the
    sum with wall links replaced by their real opposites, g = (2 c_v - 3) T
    feq_i, from the node's multipliers and from the start-up guesses alike."""
import json
import os

import numpy as np
import pytest

from conftest import load_pkg

KINDS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wall_kinds39.json")
FACES = [-1000, -1000, -11, -6, -1000, -11]
T0 = 1 / 1.4


def _shifted_len(orc):
    cx, cy, cz = orc.velocities39(True)
    return np.sqrt(cx * cx + cy * cy + cz * cz)


def test_stencil_kinds_follow_reference_table(orc):
    tab = json.load(open(KINDS))
    link = _shifted_len(orc)
    rng = np.random.default_rng(3)
    for j in range(39):
        ut = rng.uniform(-1, 1, 3) + np.array([1.0, 0.0, 0.0])
        su = np.zeros((7, 3))
        G_T = rng.uniform(-0.1, 0.1, 3)
        ex = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]])
        sT = 0.7 + ex @ G_T
        q = rng.uniform(0.05, 0.95, 3)
        scode = np.zeros(7, np.int32)
        scode[1:4] = -1001 - np.floor(q * 1e9).astype(np.int64)   # walls at x + e_a ...
        scode[4:7] = 0                                            # ... fluid at x - e_a
        du, dT = orc.wall_derivs39(j, ut, su, sT, scode)
        for ax in range(3):
            wd = abs(scode[1 + ax] + 1001.0) / 1e9 * link[j]
            kv, kt = tab["velocity_kinds"][j][ax], tab["temperature_kinds"][j][ax]
            if kv == 1:
                assert np.allclose(du[ax], (0.0 - ut) / (1 + wd), rtol=0, atol=1e-15), (j, ax)
            elif kv == -1:
                assert np.allclose(du[ax], (ut - 0.0) / (1 + wd), rtol=0, atol=1e-15), (j, ax)
            else:   # local, one-sided (x + e_a is not fluid): rest - up on a zero field
                assert np.all(du[ax] == 0.0), (j, ax)
            if kt != 0:
                assert tab["temperature_ibb_zero"][j][ax] and dT[ax] == 0.0, (j, ax)
            else:
                assert abs(dT[ax] - G_T[ax]) < 1e-15, (j, ax)


def test_local_differences_exact_for_linear_fields(orc):
    rng = np.random.default_rng(5)
    ex = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]])
    j = 1   # unshifted (1, 0, 0): y and z are local axes
    for missing in (None, 2, 5):
        G = rng.uniform(-0.2, 0.2, (3, 3))       # u_b = u0_b + sum_a G[a, b] x_a
        u0 = rng.uniform(-1, 1, 3)
        su = u0 + ex @ G
        sT = 0.7 + ex @ rng.uniform(-0.1, 0.1, 3)
        scode = np.zeros(7, np.int32)
        if missing is not None:
            scode[missing] = -1
        du, _ = orc.wall_derivs39(j, u0 + 0.1, su, sT, scode)
        assert np.abs(du[1:] - G[1:]).max() < 1e-15


def _textbook(i, rho, u, T, feq, du, dT, eps, b1, b2, cv, lead_g, cvec):
    I = np.eye(3)
    S = du + du.T
    div = du[0, 0]                                  # = du[1, 1] = du[2, 2] here
    P = (S - I * div / cv) / b1
    H2 = np.outer(cvec, cvec) - T * I
    p_comp = (3 * np.sum(P * H2) + P[2, 2] * H2[2, 2]) / (2 * T * T)   # P_zzx twice
    t = dT / b2
    Q = np.zeros((3, 3, 3))
    H3 = np.zeros((3, 3, 3))
    for a in range(3):
        for b in range(3):
            for c in range(3):
                Q[a, b, c] = (I[b, c] * t[a] + I[a, c] * t[b] + I[a, b] * t[c] +
                              u[a] * P[b, c] + u[b] * P[a, c] + u[c] * P[a, b])
                H3[a, b, c] = (cvec[a] * cvec[b] * cvec[c] -
                               T * (cvec[a] * I[b, c] + cvec[b] * I[a, c] + cvec[c] * I[a, b]))
    # Q_zxx takes dT/dx; its basis term in the listing is c_x^2 c_z (the
    # 3 T c_z of the (x, x, z) class sits on Q_xxz), so the slip adds
    # (t_x - t_z) c_x^2 c_z to the symmetric sum
    q_comp = (np.sum(Q * H3) + (t[0] - t[2]) * cvec[0] ** 2 * cvec[2]) / (6 * T ** 3)
    f = feq + eps * (-rho * T / 2) * (p_comp + q_comp)
    r = T * S + np.outer(u, dT) + np.outer(dT, u) - div / cv * (2 * T * I + np.outer(u, u))
    R = 3 * (np.sum(r * H2) - r[2, 1] * cvec[2] * cvec[1] + r[1, 1] * cvec[2] * cvec[1])
    k1 = 3 * dT.sum() / (b1 * cv)
    g_alpha = np.sum(cvec / T * (9 * t + u * k1))
    g = lead_g * T * feq + eps * (-rho * T * (2 * cv - 3) / 2) * (3 * k1 + g_alpha + R / (2 * b1 * T))
    return f, g


def test_grad_expansion_equals_textbook_tensors(orc):
    c = orc.consts39(mach=1.5, shifted=True)
    cx, cy, cz = orc.velocities39(True)
    rng = np.random.default_rng(11)
    worst = 0.0
    for trial in range(200):
        i = int(rng.integers(1, 39))
        rho, T = rng.uniform(0.7, 1.4), rng.uniform(0.55, 0.95)
        u = rng.uniform(-0.4, 0.4, 3) + np.array([1.0, 0, 0])
        du = rng.uniform(-0.05, 0.05, (3, 3))
        np.fill_diagonal(du, rng.uniform(-0.05, 0.05))
        dT = rng.uniform(-0.03, 0.03, 3)
        feq = rng.uniform(0.01, 0.2)
        eps, b1, b2 = rng.uniform(1e-4, 1e-2), rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0)
        f, g = orc.wall_grad39(i, rho, u, T, feq, du, dT, eps, b1, b2, c)
        fe, ge = _textbook(i, rho, u, T, feq, du, dT, eps, b1, b2, c.cv, c.lead_g,
                           np.array([cx[i], cy[i], cz[i]], float))
        worst = max(worst, abs(f - fe) / abs(f - feq), abs(g - ge) / abs(g - c.lead_g * T * feq))
    assert worst < 1e-11


def test_zero_gradient_rebuilds_target_equilibrium(orc):
    pkg = load_pkg()
    shape = (16, 14, 12)
    c = orc.consts39(mach=1.5, shifted=True)
    tri, nrm = pkg.icosphere(3, 3.3, (7.5, 6.75, 5.5), 7)
    st, ns, _ = orc.build_links39s(shape, FACES, tri, nrm)
    f, g, lm = orc.init39(shape, 1.0, 0.0, 0.0, 0.0, T0, c)
    cxu, cyu, czu = orc.velocities39(False)
    erp = [int(np.nonzero((cxu == -cxu[i]) & (cyu == -cyu[i]) & (czu == -czu[i]))[0][0])
           for i in range(39)]
    wall_nodes = np.nonzero((st[0] >= 0) & (st[1:] <= -1001).any(0))[0]
    assert ns > 50 and wall_nodes.size > 100
    n_adj = 0
    for lvl_step in (0, 10):
        for k in wall_nodes[::3]:
            fi = f[:, k]
            wl = st[:, k] <= -1001
            wl[0] = False
            fw, gw, nrep, nc, ke = orc.wall_rebuild39(shape, FACES, st, f, g, lm, k, fi, c, 50,
                                                      lvl_step)
            assert nrep == 0 and nc == 0
            if (st[1:15, k] < -1000).any():
                n_adj += 1
                rho_t = sum(fi[erp[i]] if st[i, k] < -1000 else fi[i] for i in range(39))
                lm0 = lm[:, k] if lvl_step > 4 else \
                    [-2.2, 0.4, 0, 0, -0.45, -0.45, -0.45, 0, 0, 0, 0.015, 0, 0]
                P = rho_t * T0 * np.array([1, 1, 1, 0, 0, 0.0])
                lmt, _, it = orc.lm_solve39(rho_t, [0.0, 0.0, 0.0], T0, P, lm0, shifted=True,
                                            maxit=c.maxit)
                assert it >= 0
                e = orc.feq39(lmt, rho_t, True)
            else:
                e = orc.feq39(lm[:, k], fi.sum(), True)
            assert np.abs(fw[wl] - e[wl]).max() < 1e-12 * e.max(), k
            assert np.abs(gw[wl] - c.lead_g * T0 * e[wl]).max() < 1e-12 * c.lead_g * T0 * e.max()
    assert n_adj > 50
