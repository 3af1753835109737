"""Oracle D3Q39 link table on the shifted lattice (shifted_state_directions.f95:
every link keyed on its upstream node x - c_i, c_x shifted by +1), pinned by:
  * every shifted link whose reversed vector -c is a D3Q27 velocity tests the
    segment x -> x - c, the same segment as that D3Q27 link, so where the
    upstream node is solid its code equals the (pinned, bit-exact) D3Q27
    table's code; node codes equal;
  * every wall code decodes to the analytic ray-sphere fraction along x -> x - c
    to the facet chord error, and appears exactly where the upstream node is
    solid;
  * a link whose upstream is fluid carries the downstream node's code (0),
    links across faces the face code (an outflow face the downstream code),
    a solid node -666 everywhere, and the rest direction keeps the node code."""
import math

import numpy as np
import pytest

from conftest import load_pkg

PER = 2147483647
FACES = [-1000, -11, PER, -6, -11, PER]


@pytest.fixture(scope="module")
def sphere39s(orc):
    pkg = load_pkg()
    R, C = 6.3, np.array([14.5, 10.25, 9.75])
    tri, nrm = pkg.icosphere(5, R, C, 20260102)
    shape = (30, 22, 20)
    st, ns, nt = orc.build_links39s(shape, FACES, tri, nrm)
    st27, ns27, _ = orc.build_links27(shape, FACES, tri, nrm)
    return R, C, shape, st, ns, nt, st27, ns27


def _shifted_dirs(orc):
    cx, cy, cz, *_ = orc.lattice39()
    return np.stack([cx + 1, cy, cz], 1)


def _xyz(shape, k):
    nx, ny, _ = shape
    return np.stack([k % nx, (k // nx) % ny, k // (nx * ny)], -1)


def _up_index(shape, faces, x, c):
    """Upstream node index of positions x for vector c, -1 beyond a
    non-periodic face (periodic axes wrap)."""
    u = x - c
    ok = np.ones(len(x), bool)
    for d in range(3):
        n = shape[d]
        if faces[d] == PER:
            u[:, d] %= n
        else:
            ok &= (u[:, d] >= 0) & (u[:, d] < n)
    idx = u[:, 0] + shape[0] * (u[:, 1] + shape[1] * u[:, 2])
    return np.where(ok, idx, -1)


def test_shifted_links_equal_d3q27_segments(orc, sphere39s):
    R, C, shape, st, ns, nt, st27, ns27 = sphere39s
    assert nt == 0 and ns == ns27 > 500
    assert np.array_equal(st[0], st27[0])
    cs = _shifted_dirs(orc)
    cx27, cy27, cz27, *_ = orc.lattice27()
    c27 = np.stack([cx27, cy27, cz27], 1)
    n = st.shape[1]
    x = _xyz(shape, np.arange(n))
    fluid = st[0] >= 0
    shared = walls = 0
    for i in range(1, 39):
        j = np.nonzero((c27 == -cs[i]).all(1))[0]
        if len(j) == 0:
            continue
        shared += 1
        up = _up_index(shape, FACES, x.copy(), cs[i])
        solid_up = fluid & (up >= 0) & (st[0][np.maximum(up, 0)] < 0)
        assert np.array_equal(st[i][solid_up], st27[j[0]][solid_up]), i
        walls += int(solid_up.sum())
    assert shared >= 10 and walls > 500


def test_shifted_wall_codes_match_sphere(orc, sphere39s):
    R, C, shape, st, ns, nt, _, _ = sphere39s
    cs = _shifted_dirs(orc)
    n = st.shape[1]
    x = _xyz(shape, np.arange(n)).astype(float)
    fluid = st[0] >= 0
    checked = 0
    for i in range(1, 39):
        up = _up_index(shape, FACES, _xyz(shape, np.arange(n)), cs[i])
        solid_up = fluid & (up >= 0) & (st[0][np.maximum(up, 0)] < 0)
        wall = fluid & (st[i] <= -1001)
        assert np.array_equal(wall, solid_up), i
        if not wall.any():
            continue
        q = np.abs(st[i][wall].astype(np.int64) + 1001) / 1e9
        # analytic crossing of the segment p - t c, t in (0, 1], with the sphere
        p = x[wall] - C
        c = -cs[i].astype(float)
        a2 = c @ c
        b2 = p @ c
        cc = np.einsum("ij,ij->i", p, p) - R * R
        t = (-b2 - np.sqrt(b2 * b2 - a2 * cc)) / a2
        assert np.all(np.abs(q - t) < 0.01 * R / math.sqrt(a2) + 2e-9), (i, np.abs(q - t).max())
        checked += int(wall.sum())
    assert checked > 2000


def test_shifted_face_interior_and_solid_codes(orc, sphere39s):
    R, C, shape, st, ns, nt, _, _ = sphere39s
    cs = _shifted_dirs(orc)
    n = st.shape[1]
    x = _xyz(shape, np.arange(n))
    fluid = st[0] >= 0
    assert np.all(st[1:, ~fluid] == -666)
    for i in range(1, 39):
        u = x - cs[i]
        # first non-periodic face crossed by the upstream node, order x-, y-, x+, y+ (z periodic)
        face = np.full(n, -1)
        for f, (d, lo) in reversed(list(enumerate([(0, True), (1, True), (2, True),
                                                   (0, False), (1, False), (2, False)]))):
            if FACES[f] == PER:
                continue
            out = u[:, d] < 0 if lo else u[:, d] > shape[d] - 1
            face = np.where(out, f, face)
        up = _up_index(shape, FACES, x.copy(), cs[i])
        solid_up = (up >= 0) & (st[0][np.maximum(up, 0)] < 0)
        sel = fluid & ~solid_up
        dn = x + cs[i]
        dn[:, 2] %= shape[2]
        dn_in = np.all((dn >= 0) & (dn < np.array(shape)), 1)
        didx = np.where(dn_in, dn[:, 0] + shape[0] * (dn[:, 1] + shape[1] * dn[:, 2]), 0)
        down = np.where(dn_in, st[0][didx], 0)
        down = np.where(down == -1001, 0, down)
        fc = np.array(FACES)[np.maximum(face, 0)]
        want = np.where(face >= 0, np.where(fc == -6, down, fc), down)
        # an upstream beyond a face is solid only inside the sphere's extension (none here)
        assert np.array_equal(st[i][sel], want[sel]), i
