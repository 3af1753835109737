"""Oracle D3Q27 step with boundary links, pinned by invariants: halfway
bounce-back conserves mass exactly in a closed box, a fluid at rest around a
curved wall stays at rest, and equilibrium inlet/outlet tables feed exactly
their equilibria."""
import numpy as np

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
