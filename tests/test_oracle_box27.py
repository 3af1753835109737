"""Oracle D3Q27 step with boundary links, pinned by invariants: halfway
bounce-back conserves mass exactly in a closed box, a fluid at rest around a
curved wall stays at rest, and equilibrium inlet/outlet tables feed exactly
their equilibria."""
import numpy as np
import pytest

from conftest import load_pkg

WALLS = [-10] * 6


def test_closed_box_conserves_mass(orc):
    shape = (12, 10, 9)
    st, _, _ = orc.build_links27(shape, WALLS, np.zeros((0, 9)), np.zeros((0, 3)))
    f = orc.init_tgv27(shape, 0.05, 1e-3, 7)
    m0 = f.sum()
    g, alpha, ost, bst = orc.step_box27(shape, WALLS, st, f, beta=0.8, nsteps=15)
    assert abs(g.sum() - m0) < 1e-11 * m0
    assert ost["nodes"] == 15 * 12 * 10 * 9 and ost["n_noconv"] == 0


def test_rest_state_around_sphere_stays_at_rest(orc):
    pkg = load_pkg()
    shape = (24, 20, 20)
    tri, nrm = pkg.icosphere(3, 5.2, (11.5, 9.5, 10.5), 3)
    st, ns, _ = orc.build_links27(shape, WALLS, tri, nrm)
    assert ns > 300
    f = orc.uniform27(st, 1.0, (0.0, 0.0, 0.0))
    g, alpha, ost, bst = orc.step_box27(shape, WALLS, st, f, beta=0.9, nsteps=10)
    w = orc.lattice27()[4]
    assert np.abs(g - w[:, None]).max() < 1e-15
    assert bst["n_ibb_links"] > 1000 and bst["n_repair"] == 0


def test_equilibrium_inlet_outlet_channel(orc):
    """Uniform flow equal to the inlet / outlet state with slip-free faces
    replaced by the same equilibrium is a fixed point."""
    shape = (16, 8, 8)
    u = (0.05, 0.0, 0.0)
    faces = [-101, -1000, -1000, -201, -1000, -1000]
    st, _, _ = orc.build_links27(shape, faces, np.zeros((0, 9)), np.zeros((0, 3)))
    f = orc.uniform27(st, 1.0, u)
    g, alpha, ost, bst = orc.step_box27(shape, faces, st, f, beta=0.7, nsteps=5,
                                        inlets=[(1.0,) + u], outlets=[(1.0,) + u],
                                        freestream=(1.0,) + u)
    assert np.abs(g - f).max() < 1e-15
    assert np.all(alpha == 2.0)


def _channel_table(orc, L, H, qb, qt):
    """Link table of a channel along x: inlet 1 at x-, outlet 1 at x+, flat
    walls at y = -qb and y = H - 1 + qt (every link leaving the fluid rows
    through a wall has fraction q along the link; first crossed face in the
    order x-, y-, x+, y+), one periodic z layer."""
    import math
    cx, cy, cz, _, _ = orc.lattice27()
    n = L * H
    st = np.zeros((27, n), np.int32)
    cb, ct = -1001 - int(math.floor(qb * 1e9)), -1001 - int(math.floor(qt * 1e9))
    x = np.arange(n) % L
    y = np.arange(n) // L
    for i in range(1, 27):
        xn, yn = x + cx[i], y + cy[i]
        code = np.zeros(n, np.int32)
        code = np.where(yn > H - 1, ct, code)
        code = np.where(xn > L - 1, -201, code)
        code = np.where(yn < 0, cb, code)
        code = np.where(xn < 0, -101, code)
        st[i] = code
    return st


def _poiseuille_error(orc, H, q, U=0.01, nu=1 / 6):
    """Fully developed channel flow between walls at fraction q: relative L2
    distance of the mid-channel u_x profile from the best-fitting parabola
    that vanishes on the walls y = -q and y = H - 1 + q."""
    L = 4 * H
    st = _channel_table(orc, L, H, q, q)
    f = orc.uniform27(st, 1.0, (U, 0.0, 0.0))
    faces = [-101, -10, 2147483647, -201, -10, 2147483647]
    f, _, _, bst = orc.step_box27((L, H, 1), faces, st, f, 1 / (6 * nu + 1),
                                  nsteps=int(3 * H * H / nu), inlets=[(1.0, U, 0.0, 0.0)],
                                  outlets=[(1.0, U, 0.0, 0.0)], fixed=True, want_alpha=False)
    assert bst["n_repair"] == 0 and bst["n_ibb_links"] > 0
    ux = orc.macro27(f)[1].reshape(H, L)[:, L // 2]
    y = np.arange(H)
    p = (y + q) * (H - 1 + q - y)
    A = ux @ p / (p @ p)
    return float(np.linalg.norm(ux - A * p) / np.linalg.norm(ux))


@pytest.mark.parametrize("q", [0.2, 0.5, 0.8])
def test_wall_links_poiseuille_second_order(orc, q):
    """Pressure-driven channel between flat walls cut at fraction q of the
    links: the wall rule places the no-slip wall where the link codes put it,
    to second order (error ratio ~4 per halving of the grid spacing).  A
    rule that mislocates the wall by O(dx) converges at ratio ~2."""
    errs = [_poiseuille_error(orc, H, q) for H in (8, 16, 32)]
    print(f"q {q}: Poiseuille profile error {errs[0]:.2e} {errs[1]:.2e} {errs[2]:.2e}")
    assert errs[2] < 2e-3
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0
