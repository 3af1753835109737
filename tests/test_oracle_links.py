"""Oracle link tables (restatement of fluid_or_solid / inter_bb_setup /
state_directions), pinned against an analytic sphere: the triangulated sphere
converges to the analytic one, so node classes and wall fractions q must match
the closed-form ray-sphere answers away from the chord error of the facets."""
import math

import numpy as np
import pytest

from conftest import load_pkg

FACES_DUCT = [-101, -10, -10, -201, -10, -10]   # inlet x-, walls y/z, outlet x+


@pytest.fixture(scope="module")
def sphere(orc):
    pkg = load_pkg()
    R, C = 9.3, np.array([20.5, 15.25, 16.5])
    tri, nrm = pkg.icosphere(5, R, C, 20260102)
    st, ns, nt = orc.build_links27((44, 32, 34), FACES_DUCT, tri, nrm)
    return R, C, tri, nrm, st, ns, nt


def test_icosphere_is_closed_and_outward(sphere):
    R, C, tri, nrm, *_ = sphere
    assert tri.shape == (20480, 9)
    v = tri.reshape(-1, 3, 3)
    # outward normals, vertices on the sphere
    assert np.all(np.einsum("tk,tk->t", nrm, v.mean(1) - C) > 0)
    assert np.allclose(np.linalg.norm(v - C, axis=2), R, rtol=1e-12)
    # closed surface: every edge appears exactly twice
    edges = {}
    for t in v.round(12):
        for a, b in ((0, 1), (1, 2), (2, 0)):
            key = tuple(sorted((tuple(t[a]), tuple(t[b]))))
            edges[key] = edges.get(key, 0) + 1
    assert set(edges.values()) == {2}


def test_node_classes_match_analytic_sphere(sphere):
    R, C, tri, nrm, st, ns, nt = sphere
    nx, ny, nz = 44, 32, 34
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = np.sqrt((x - C[0]) ** 2 + (y - C[1]) ** 2 + (z - C[2]) ** 2).ravel()
    sag = R * (1 - math.cos(math.radians(2.5)))  # > facet chord error at level 5
    solid = st[0] == -1001
    assert nt == 0
    assert np.all(solid[d < R - sag]) and not np.any(solid[d > R])
    assert ns == int(solid.sum())
    assert abs(ns - 4 / 3 * math.pi * R ** 3) / ns < 0.02


def test_wall_fractions_match_analytic_sphere(orc, sphere):
    R, C, tri, nrm, st, ns, nt = sphere
    cx, cy, cz, opp, w = orc.lattice27()
    nx, ny, nz = 44, 32, 34
    n = nx * ny * nz
    # largest radial gap between a level-5 facet and the sphere (facet angular
    # circumradius < 1.8 deg); along a link it grows as 1 / cos(incidence)
    gap = R * (1 - math.cos(math.radians(1.8)))
    checked = 0
    for i in range(1, 27):
        c = np.array([cx[i], cy[i], cz[i]], float)
        L = np.linalg.norm(c)
        ks = np.nonzero(st[i] <= -1001)[0]
        for k in ks:
            a, b, cc = k % nx, (k // nx) % ny, k // (nx * ny)
            # the neighbour along c_i is solid (inter_bb_setup is only asked then)
            nb = (a + cx[i]) + nx * ((b + cy[i]) + ny * (cc + cz[i]))
            assert st[0, nb] == -1001
            q = abs(int(st[i, k]) + 1001) / 1e9
            # analytic ray-sphere intersection along the link
            p = np.array([a, b, cc], float) - C
            dd = c / L
            bq = p @ dd
            disc = bq * bq - (p @ p - R * R)
            t = -bq - math.sqrt(disc)
            cosi = abs(dd @ ((p + t * dd) / R))
            if cosi < 0.1:
                continue  # grazing link
            assert abs(q * L - t) < 1.5 * gap / cosi, (i, k, q * L, t, cosi)
            checked += 1
    assert checked > 2000


def test_link_code_encoding_round_trip():
    # -1001 - floor(q * 1e9): decode q = |s + 1001| / 1e9 (target_finder.f95:40)
    for q in (0.0, 1e-9, 0.123456789, 0.5, 0.999999999):
        s = -1001 - math.floor(q * 1e9)
        assert s <= -1001
        assert abs(abs(s + 1001) / 1e9 - q) < 1e-9


def test_face_and_interior_codes(sphere, orc):
    R, C, tri, nrm, st, ns, nt = sphere
    cx, cy, cz, opp, w = orc.lattice27()
    nx, ny, nz = 44, 32, 34
    n = nx * ny * nz
    solid = st[0] == -1001
    # solid nodes: -666 on every link (state_directions.f95:87)
    assert np.all(st[1:, solid] == -666)
    k = np.arange(n)
    a, b, c = k % nx, (k // nx) % ny, k // (nx * ny)
    for i in range(1, 27):
        na, nb_, nc = a + cx[i], b + cy[i], c + cz[i]
        codes = st[i, ~solid]
        fa, fb, fc = na[~solid], nb_[~solid], nc[~solid]
        # faces in reference order x-, y-, z-, x+, y+, z+
        exp = np.full(codes.shape, 12345, dtype=np.int64)
        for face, m in ((FACES_DUCT[0], fa < 0), (FACES_DUCT[1], fb < 0), (FACES_DUCT[2], fc < 0),
                        (FACES_DUCT[3], fa > nx - 1), (FACES_DUCT[4], fb > ny - 1),
                        (FACES_DUCT[5], fc > nz - 1)):
            sel = m & (exp == 12345)
            exp[sel] = face
        inside = exp == 12345
        assert np.array_equal(codes[~inside], exp[~inside])
        # interior: 0 for fluid neighbours, a wall code for solid neighbours
        nbk = (fa + nx * (fb + ny * fc))[inside]
        nb_solid = solid[nbk]
        ci = codes[inside]
        assert np.all(ci[~nb_solid] == 0)
        assert np.all(ci[nb_solid] <= -1001)


def test_periodic_faces_wrap(orc):
    st, ns, nt = orc.build_links27((6, 5, 4), [orc.FACE_PERIODIC] * 6, np.zeros((0, 9)),
                                   np.zeros((0, 3)))
    assert ns == 0 and np.all(st == 0)
