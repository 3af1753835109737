"""The oracle's geometry helpers against the reference's own routines, compiled
from the reference sources (tri_volume.f95, box_test.f95, linkwise.f95:45-57)
into oracle/_ref by oracle/ref/Makefile.  Bit-exact: same integer result on
every input, including the degenerate ones (ray through an edge or vertex,
segment ends on a bounding-box face, crossings at q = 0 and near q = 1).
Skipped where the reference tree was not there to compile from."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.skipif(
    not __import__("oracle").ref_available(), reason="reference routines not built (no reference tree)")


def _p(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ref_tri_volume(R, xc, e, t):
    a, pa = _p(xc); b, pb = _p(e); c, pc = _p(t)
    return R.ref_tri_volume(pa, pb, pc)


def _ref_box_test(R, p1, p2, bb):
    a, pa = _p(p1); b, pb = _p(p2); c, pc = _p(bb)
    return R.ref_box_test(pa, pb, pc)


def test_tri_volume_matches_reference(orc):
    R = orc.ref_lib()
    rng = np.random.default_rng(7)
    n = 0
    for k in range(20000):
        t = rng.uniform(-1, 1, 9)
        e = rng.uniform(-2, 2, 3)
        if k % 4 == 0:      # crossing exactly at a vertex or an edge midpoint
            v = t.reshape(3, 3)
            xc = v[k % 3] if k % 8 == 0 else 0.5 * (v[0] + v[1])
        elif k % 4 == 1:    # small integer lattice coordinates: exact zeros
            t = rng.integers(-2, 3, 9).astype(float)
            e = rng.integers(-2, 3, 3).astype(float)
            xc = rng.integers(-2, 3, 3).astype(float)
        else:
            w = rng.dirichlet([1, 1, 1]) if k % 2 else rng.uniform(-0.5, 1.5, 3)
            xc = w @ t.reshape(3, 3)
        r_o = orc.tri_volume(xc, e, t)
        r_r = _ref_tri_volume(R, xc, e, t)
        assert r_o == r_r, (k, xc, e, t, r_o, r_r)
        n += r_o != 0
    assert n > 1000


def test_box_test_matches_reference(orc):
    R = orc.ref_lib()
    rng = np.random.default_rng(8)
    for k in range(20000):
        lo = rng.integers(-3, 3, 3).astype(float)
        hi = lo + rng.integers(0, 3, 3)
        bb = np.array([hi[0], lo[0], hi[1], lo[1], hi[2], lo[2]])
        if k % 2:
            p1 = rng.integers(-4, 5, 3).astype(float)
            p2 = rng.integers(-4, 5, 3).astype(float)
        else:
            p1 = rng.uniform(-4, 4, 3)
            p2 = rng.uniform(-4, 4, 3)
        assert orc.box_test(p1, p2, bb) == _ref_box_test(R, p1, p2, bb), (k, p1, p2, bb)


def test_link_code_matches_reference(orc):
    R = orc.ref_lib()
    rng = np.random.default_rng(9)
    dirs = [(1, 0, 0), (1, 1, 0), (1, 1, 1), (-1, 0, 1), (0, -1, -1)]
    for k in range(20000):
        c = np.array(dirs[k % len(dirs)], float)
        dx = (1.0, 0.5, 0.037)[k % 3]
        e1 = rng.uniform(-50, 50, 3)
        if k % 5 == 0:
            q = 0.0
        elif k % 5 == 1:
            q = 1.0 - 1e-12 * rng.uniform()
        else:
            q = rng.uniform()
        xc = e1 + q * c * dx
        a, pa = _p(xc); b, pb = _p(e1)
        r_r = R.ref_link_code(pa, pb, int(c @ c), dx)
        r_o = orc.link_code(e1, xc, math.sqrt(c @ c) * dx)
        assert r_o == r_r, (k, q, r_o, r_r)
