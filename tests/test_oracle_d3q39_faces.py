"""Oracle D3Q39 box faces: freestream (-1000), pure outflow (-6), bounce-back
(-10) and inviscid (-11) edge links, in pull form.

Pinned by:
  * a particle tracer: the inviscid source is where a particle launched from a
    node with a lattice velocity ends after one step with specular reflection
    off the half-way planes of the walls (event-driven geometry, not the
    oracle's index formula);
  * mirror symmetry: a box with inviscid walls evolves as the lower part of a
    periodic box of twice the extent holding the mirrored state (every node
    update in the walled box sees the same inputs as its twin, up to the
    summation order of the mirrored node), on the plain and the shifted lattice;
  * a uniform freestream with the supersonic face set (freestream in,
    outflow out, inviscid sides) is a fixed point;
  * a box closed by bounce-back or inviscid faces conserves mass and energy
    (every lost population is replaced by exactly one missing one);
  * the repair's one-sided rule at a face (damp_it.f95:109-123);
  * a refined level's interior step (ghost layer of depth 4, not updated)
    equals the periodic step when the layer holds the periodic continuation."""
import numpy as np
import pytest

import oracle as o

G = 1.4
T0 = 1 / G
P = o.FACE_PERIODIC


def _fields(shape, mach, amp=1.0):
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    rho = 1.0 + amp * 0.01 * np.sin(2 * np.pi * (x / nx + 0.7 * y / ny) + 0.3)
    ux = mach + amp * 0.02 * np.sin(2 * np.pi * x / nx + 0.1 * z)
    uy = amp * 0.015 * np.sin(2 * np.pi * y / (2 * ny) + 0.4 * x)
    uz = amp * 0.01 * np.cos(2 * np.pi * (x / nx + z / nz))
    T = T0 * (1 + amp * 0.01 * np.cos(2 * np.pi * z / nz + 0.2 * y))
    return rho, ux, uy, uz, T


def _mirror_dirs(axis, shifted):
    cx, cy, cz = o.velocities39(shifted)
    c = np.stack([cx, cy, cz], 1)
    out = np.zeros(39, int)
    for k in range(39):
        m = c[k].copy()
        m[axis] = -m[axis]
        out[k] = int(np.nonzero((c == m).all(1))[0][0])
    return out


# multiplier components odd under a reflection of y / z:
# chi, zeta(3), pi_xx, pi_yy, pi_zz, pi_xy, pi_xz, pi_yz, lambda(3)
_LM_ODD = {1: [2, 7, 9, 11], 2: [3, 8, 9, 12]}


def _double(shape, f, g, lm, axis, shifted):
    """State on twice the extent along axis: the original below, its mirror
    image about the upper wall above."""
    nx, ny, nz = shape
    sh = (nz, ny, nx)
    ax = 2 - axis  # numpy axis of the (z, y, x) array
    mir = _mirror_dirs(axis, shifted)

    def dbl(a, q, perm=None, odd=()):
        a = a.reshape((q,) + sh)
        b = np.flip(a, axis=ax + 1)
        if perm is not None:
            b = b[perm]
        b = b.copy()
        for j in odd:
            b[j] = -b[j]
        return np.concatenate([a, b], axis=ax + 1).reshape(q, -1)

    dshape = list(shape)
    dshape[axis] *= 2
    return (tuple(dshape), dbl(f, 39, mir), dbl(g, 39, mir), dbl(lm, 13, None, _LM_ODD[axis]))


def _lower(dshape, a, q, axis, n):
    nx, ny, nz = dshape
    a = a.reshape((q, nz, ny, nx))
    sl = [slice(None)] * 4
    sl[3 - axis] = slice(0, n)
    return a[tuple(sl)].reshape(q, -1)


def _trace(node, c, shape, walls):
    """End node and velocity of a particle leaving node with velocity c for
    one time unit, reflecting specularly off the planes -1/2 and n - 1/2 of the
    axes in walls (event-driven: the reflection times are found exactly)."""
    p = np.array(node, float)
    v = np.array(c, float)
    t = 0.0
    while True:
        times = {}
        for d in walls:
            if v[d] < 0:
                th = (-0.5 - p[d]) / v[d]
            elif v[d] > 0:
                th = (shape[d] - 0.5 - p[d]) / v[d]
            else:
                continue
            if th > 1e-12 and t + th <= 1.0:
                times[d] = t + th
        if not times:
            p = p + v * (1.0 - t)
            return tuple(int(round(x)) for x in p), tuple(int(x) for x in v)
        hit = min(times.values())
        p = p + v * (hit - t)
        for d, th in times.items():
            if abs(th - hit) < 1e-12:  # corners: every plane reached at that instant
                v[d] = -v[d]
        t = hit


@pytest.mark.parametrize("shifted", [False, True])
def test_inviscid_source_is_the_specular_particle_path(shifted):
    shape = (7, 6, 5)
    faces = [P, -11, -11, P, -11, -11]
    cx, cy, cz = o.velocities39(shifted)
    n_checked = 0
    for z in range(shape[2]):
        for y in range(shape[1]):
            for x in (0, 3):
                for k in range(39):
                    (sx, sy, sz), d = o.source39((x, y, z), k, shape, faces, shifted)
                    # the particle leaving the source with velocity c_d ends here with c_k
                    end, v = _trace((sx, sy, sz), (cx[d], cy[d], cz[d]), shape, (1, 2))
                    assert (end[1], end[2]) == (y, z) and v == (cx[k], cy[k], cz[k]), (x, y, z, k)
                    assert (sx - (x - cx[k])) % shape[0] == 0
                    n_checked += int(d != k)
    assert n_checked > 500


@pytest.mark.parametrize("shifted,mach", [(False, 0.3), (True, 1.5)])
def test_inviscid_walls_equal_mirrored_periodic_box(shifted, mach):
    shape = (6, 5, 5)
    c = o.consts39(mach=mach, shifted=shifted)
    f0, g0, lm0 = o.init39(shape, *_fields(shape, mach), c)
    steps = 6
    # walls on y only
    fw, gw, lw, _, stw = o.step_faces39(shape, [P, -11, P, P, -11, P], f0, g0, lm0, c, 0, steps)
    dshape, fd, gd, ld = _double(shape, f0, g0, lm0, 1, shifted)
    fp, gp, lp, _, stp = o.step_periodic39(dshape, fd, gd, ld, c, 0, steps)
    fl, gl = _lower(dshape, fp, 39, 1, shape[1]), _lower(dshape, gp, 39, 1, shape[1])
    e = max(np.abs(fl - fw).max() / np.abs(fw).max(), np.abs(gl - gw).max() / np.abs(gw).max())
    assert e < 1e-12, e
    assert np.abs(fw - f0).max() > 1e-4  # the state moved
    assert stw["lm_noconv"] == 0 and stp["lm_noconv"] == 0
    # walls on y and z: four mirror images
    fw, gw, _, _, _ = o.step_faces39(shape, [P, -11, -11, P, -11, -11], f0, g0, lm0, c, 0, steps)
    d1, f1, g1, l1 = _double(shape, f0, g0, lm0, 1, shifted)
    d2, f2, g2, l2 = _double(d1, f1, g1, l1, 2, shifted)
    fp, gp, _, _, _ = o.step_periodic39(d2, f2, g2, l2, c, 0, steps)
    fl = _lower(d1, _lower(d2, fp, 39, 2, shape[2]), 39, 1, shape[1])
    assert np.abs(fl - fw).max() / np.abs(fw).max() < 1e-12


@pytest.mark.parametrize("shifted,mach", [(False, 0.3), (True, 1.5)])
def test_uniform_freestream_is_a_fixed_point_of_the_supersonic_faces(shifted, mach):
    shape = (8, 6, 5)
    c = o.consts39(mach=mach, shifted=shifted)
    n = shape[0] * shape[1] * shape[2]
    f0, g0, lm0 = o.init39(shape, np.ones(n), mach, 0.0, 0.0, T0, c)
    fsf, fsg = o.freestream39(c)
    assert np.abs(f0 - fsf[:, None]).max() < 1e-14
    faces = [-1000, -11, -11, -6, -11, -11]
    f1, g1, _, _, st = o.step_faces39(shape, faces, f0, g0, lm0, c, 30, 3)
    assert np.abs(f1 - f0).max() < 1e-13 and np.abs(g1 - g0).max() < 1e-13
    assert st["n_repair"] == 0 and st["lm_noconv"] == 0


@pytest.mark.parametrize("code", [-10, -11])
def test_closed_box_conserves_mass_and_energy(code):
    shape = (6, 5, 5)
    c = o.consts39(mach=0.05)
    f0, g0, lm0 = o.init39(shape, *_fields(shape, 0.05), c)
    cx, cy, cz = o.velocities39(False)
    c2 = (cx ** 2 + cy ** 2 + cz ** 2)[:, None]
    f1, g1, _, _, _ = o.step_faces39(shape, [code] * 6, f0, g0, lm0, c, 0, 4)
    m0, m1 = f0.sum(), f1.sum()
    e0, e1 = (c2 * f0).sum() + g0.sum(), (c2 * f1).sum() + g1.sum()
    assert abs(m1 - m0) / m0 < 1e-13 and abs(e1 - e0) / e0 < 1e-13
    assert np.abs(f1 - f0).max() > 1e-5


def test_outflow_and_freestream_sources():
    shape = (7, 6, 5)
    faces = [-1000, -10, P, -6, -10, P]
    cx, cy, cz = o.velocities39(False)
    opp = o.lattice39()[3]
    for k in range(39):
        # upstream node across the inflow face: freestream value
        src, d = o.source39((0, 3, 2), k, shape, faces)
        assert (src is None) == (cx[k] > 0)
        # across the outflow face: the node's own population of the same direction
        src, d = o.source39((6, 3, 2), k, shape, faces)
        if cx[k] < 0:
            assert src == (6, 3, 2) and d == k
        # across the y- wall: the node's own opposite population
        src, d = o.source39((3, 0, 4), k, shape, faces)
        if cy[k] > 0:
            assert src == (3, 0, 4) and d == opp[k]
        else:
            assert src == (3 - cx[k], -cy[k], (4 - cz[k]) % 5) and d == k
    # chain order at an edge: -x before -y
    src, _ = o.source39((0, 0, 0), 7, shape, faces)  # c = (1, 1, 1)
    assert src is None


def test_damp_one_sided_at_faces():
    shape = (6, 5, 5)
    n = 150
    fo = np.full((39, n), 0.5)
    fo[1, 0] = -0.2            # direction +x at x = 0: upstream neighbour across the -x face
    fo[1, 1] = 0.3             # its downstream neighbour
    fo[4, 5] = -0.1            # direction -x at x = 5: x - c = 6 lies across the +x face
    out, nrep = o.damp_faces39(shape, [-1000, -11, -11, -6, -11, -11], fo)
    assert nrep == 2
    assert out[1, 0] == (0.3 + 0.2) / 2
    assert out[4, 5] == (0.5 + 0.1) / 2           # one side: x + c = 4
    per, _ = o.damp_faces39(shape, [P] * 6, fo)
    assert per[1, 0] == (0.3 + 0.5) / 2            # periodic: both sides positive


@pytest.mark.parametrize("shifted,mach", [(False, 0.3), (True, 1.5)])
def test_ghost_layer_step_equals_periodic_step(shifted, mach):
    """A refined level's interior step (ghost layer of depth 4, not updated):
    with the layer holding the periodic continuation of the box, the valid
    nodes step exactly as the periodic box does."""
    shape = (7, 6, 5)
    G = 4
    c = o.consts39(mach=mach, shifted=shifted)
    f0, g0, lm0 = o.init39(shape, *_fields(shape, mach), c)
    nx, ny, nz = shape
    ext = (nx + 2 * G, ny + 2 * G, nz + 2 * G)

    def pad(a, q):
        a = a.reshape(q, nz, ny, nx)
        return np.pad(a, ((0, 0), (G, G), (G, G), (G, G)), mode="wrap").reshape(q, -1)

    fp, gp, _, _, stp = o.step_periodic39(shape, f0, g0, lm0, c, 40, 1)
    fg, gg, _, _, stg = o.step_ghost39(ext, G, pad(f0, 39), pad(g0, 39), pad(lm0, 13), c, 40, 1)
    inner = (slice(None), slice(G, G + nz), slice(G, G + ny), slice(G, G + nx))
    assert np.array_equal(fg.reshape(39, *ext[::-1])[inner].reshape(39, -1), fp)
    assert np.array_equal(gg.reshape(39, *ext[::-1])[inner].reshape(39, -1), gp)
    assert stg["nodes"] == stp["nodes"] == nx * ny * nz
    # the layer is left as it was
    gh = o.ghost_nodes39(ext, G)
    assert np.array_equal(fg[:, gh], pad(f0, 39)[:, gh])
