"""The coarse-fine operators of oracle/o_amr.py pinned by exactness: nodal
interpolation reproduces fields linear in space at every fine node (coincident,
edge, face and cell-centre positions), linear time interpolation reproduces a
field linear in time, and injection returns the coincident fine value."""
import numpy as np

import o_amr


def _linear(shape, origin, coef, scale=1.0):
    loc = o_amr._coords(shape) + np.asarray(origin)[:, None]
    X = loc * scale
    return coef[:, :1] + coef[:, 1:2] * X[0] + coef[:, 2:3] * X[1] + coef[:, 3:4] * X[2]


def test_interpolation_exact_for_linear_fields():
    rng = np.random.default_rng(1)
    coef_a = rng.normal(size=(27, 4))
    coef_b = rng.normal(size=(27, 4))
    cshape, corigin = (12, 10, 9), (0, 0, 0)
    fshape, forigin = (9, 8, 7), (5, 4, 3)  # stored fine nodes incl. ghosts
    ca = _linear(cshape, corigin, coef_a)
    cb = _linear(cshape, corigin, coef_b)
    for theta in (0.0, 0.5, 0.25, 1.0):
        got = o_amr.coarse_to_fine(ca, cb, cshape, corigin, fshape, forigin, theta)
        exp = (1 - theta) * _linear(fshape, forigin, coef_a, 0.5) + theta * _linear(fshape, forigin, coef_b, 0.5)
        assert np.abs(got - exp).max() < 1e-12


def test_ghost_nodes_are_the_shell():
    g = o_amr.ghost_nodes((5, 4, 6))
    assert len(g) == 5 * 4 * 6 - 3 * 2 * 4


def test_injection_takes_coincident_values():
    rng = np.random.default_rng(2)
    cshape, corigin = (10, 10, 10), (0, 0, 0)
    flo, fhi = (6, 5, 4), (13, 12, 11)
    fshape = tuple(fhi[d] - flo[d] + 3 for d in range(3))
    forigin = tuple(flo[d] - 1 for d in range(3))
    ff = rng.normal(size=(27, np.prod(fshape)))
    fc = rng.normal(size=(27, np.prod(cshape)))
    out = o_amr.fine_to_coarse(ff, fshape, forigin, flo, fhi, fc, cshape, corigin)
    changed = np.nonzero(np.any(out != fc, axis=0))[0]
    # covered coarse nodes: x in [3, 6], y in [3, 6], z in [2, 5]
    assert len(changed) == 4 * 4 * 4
    x, y, z = 4, 5, 3
    kc = x + 10 * (y + 10 * z)
    kf = (2 * x - forigin[0]) + fshape[0] * ((2 * y - forigin[1]) + fshape[1] * (2 * z - forigin[2]))
    assert np.array_equal(out[:, kc], ff[:, kf])
