"""The GGM refinement indicator of the oracle (oracle/o_ggm.py), pinned on the
CPU:
  - the vectorised box form equals the scalar windage functions node by node
    (random fields, random link codes with -666 nodes, every scheme,
    periodic and not)
  - closed forms: a linear field has |grad| exactly on nodes two or more from
    the faces and a zero GGM four or more from them, in every scheme; q = x^2
    gives gm = 2|x| and ggm = 2 (central and upwind)
  - the reference's backward formula under the forward test
    (derivatives.f95:672-673, 689-690): upwind with a negative wind on q = x^4 gives
    |3 q(x) - 4 q(x-1) + q(x-2)| / 2, not the forward difference
  - a cut link falls back to the one-sided difference (q = x^2: 2 x0 - 1),
    and the x + 2 test reads link 15 + a, on D3Q27 a diagonal link
  - gm is computed on -666 nodes and ggm is 0 there; both are 0 on solid
    nodes."""
import itertools

import numpy as np
import pytest

import o_ggm

SHAPE = (9, 7, 6)


def coords(shape):
    nx, ny, nz = shape
    k = np.arange(nx * ny * nz)
    return k % nx, (k // nx) % ny, k // (nx * ny)


def scalar_gradmag(src, vel, shape, periodic, state, scheme, skip_null):
    nx, ny, nz = shape
    ext = (nx, ny, nz)
    out = np.zeros(nx * ny * nz)
    for k in range(nx * ny * nz):
        p = [k % nx, (k // nx) % ny, k // (nx * ny)]
        if state is not None and state[0][k] < 0 and not (skip_null and state[0][k] == -666):
            continue
        s2 = 0.0
        for ax in range(3):
            q, ok = [src[k]], [True]
            for off, comp in ((1, 1 + ax), (2, 15 + ax), (-1, 4 + ax), (-2, 18 + ax)):
                c = list(p)
                c[ax] += off
                inb = 0 <= c[ax] < ext[ax]
                if periodic[ax]:
                    c[ax] %= ext[ax]
                    inb = True
                j = c[0] + nx * (c[1] + ny * c[2])
                q.append(src[j] if inb else 0.0)
                ok.append(inb and (state is None or state[comp][k] >= 0))
            d = o_ggm.windage2(q, ok, vel[ax][k], scheme)
            s2 += d * d
        out[k] = np.sqrt(s2)
    return out


@pytest.mark.parametrize("scheme,periodic,table", list(itertools.product(
    (0, 1, 2), ((0, 0, 0), (1, 0, 1)), (False, True))))
def test_box_form_equals_scalar_windage(scheme, periodic, table):
    rng = np.random.default_rng(7 + scheme)
    n = int(np.prod(SHAPE))
    src = rng.standard_normal(n)
    vel = rng.standard_normal((3, n))
    vel[:, ::11] = 0.0
    state = None
    if table:
        state = np.where(rng.random((27, n)) < 0.2, -1001, 0).astype(np.int64)
        state[0] = np.where(rng.random(n) < 0.1, -666, state[0])
    for skip in (True, False):
        a = o_ggm.gradmag(src, vel, SHAPE, periodic, state, scheme, skip_null=skip)
        b = scalar_gradmag(src, vel, SHAPE, periodic, state, scheme, skip)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_linear_field(scheme):
    x, y, z = coords(SHAPE)
    g = np.array([0.3, -0.7, 0.45])
    q = g[0] * x + g[1] * y + g[2] * z + 1.0
    vel = np.random.default_rng(3).standard_normal((3, q.size))
    gm = o_ggm.gradmag(q, vel, SHAPE, scheme=scheme)
    ggm = o_ggm.gradmag(gm, vel, SHAPE, scheme=scheme, skip_null=False)
    nx, ny, nz = SHAPE

    def inner(m):
        return (x >= m) & (x < nx - m) & (y >= m) & (y < ny - m) & (z >= m) & (z < nz - m)

    assert np.allclose(gm[inner(2)], np.linalg.norm(g), rtol=0, atol=1e-13)
    assert inner(2).sum() > 10
    assert np.abs(ggm[inner(4)]).max() < 1e-13 if inner(4).any() else True


@pytest.mark.parametrize("scheme", [0, 1])
def test_quadratic_field(scheme):
    shape = (16, 9, 9)
    x, y, z = coords(shape)
    q = (x - 3.0) ** 2
    vel = np.full((3, q.size), 0.1)
    gm = o_ggm.gradmag(q, vel, shape, scheme=scheme)
    ggm = o_ggm.gradmag(gm, vel, shape, scheme=scheme, skip_null=False)
    mid = (x >= 2) & (x <= 13)
    assert np.allclose(gm[mid], 2 * np.abs(x[mid] - 3.0), rtol=0, atol=1e-12)
    # gm = 2|x - 3| has a kink at x = 3; away from it and the faces, ggm = 2
    far = (x >= 6) & (x <= 11)
    assert np.allclose(ggm[far], 2.0, rtol=0, atol=1e-12)


def test_backward_formula_under_forward_test():
    shape = (12, 5, 5)
    x, _, _ = coords(shape)
    q = x.astype(float) ** 4
    vel = np.zeros((3, q.size))
    vel[0] = -0.2  # upwind with a negative wind: the forward branch
    gm = o_ggm.gradmag(q, vel, shape, scheme=o_ggm.UPWIND)
    for xv in (4, 6, 8):
        k = xv + shape[0] * (2 + shape[1] * 2)
        want = abs(3 * xv ** 4 - 4 * (xv - 1) ** 4 + (xv - 2) ** 4) / 2
        forward = abs(-3 * xv ** 4 + 4 * (xv + 1) ** 4 - (xv + 2) ** 4) / 2
        assert gm[k] == pytest.approx(want, rel=1e-14) and abs(want - forward) > 1.0


def test_cut_links_and_the_second_node_test():
    shape = (12, 5, 5)
    x, y, z = coords(shape)
    q = x.astype(float) ** 2
    vel = np.zeros((3, q.size))
    vel[0] = -0.2
    n = q.size
    state = np.zeros((27, n), np.int64)
    x0 = 6
    node = x0 + shape[0] * (2 + shape[1] * 2)
    # central, link +x cut: the backward one-sided difference
    state[1, node] = -1001
    gm = o_ggm.gradmag(q, vel, shape, state=state, scheme=o_ggm.CENTRAL)
    assert gm[node] == pytest.approx(2 * x0 - 1, rel=1e-14)
    # upwind, negative wind (forward test on x+1 and link 15), link 15 cut:
    # the first-order difference q(x) - q(x+1)
    state[1, node] = 0
    state[15, node] = -1001
    gm = o_ggm.gradmag(q, vel, shape, state=state, scheme=o_ggm.UPWIND)
    assert gm[node] == pytest.approx(2 * x0 + 1, rel=1e-14)
    state[15, node] = 0
    gm = o_ggm.gradmag(q, vel, shape, state=state, scheme=o_ggm.UPWIND)
    assert gm[node] == pytest.approx(abs(3 * x0 ** 2 - 4 * (x0 - 1) ** 2 + (x0 - 2) ** 2) / 2, rel=1e-14)


def test_null_and_solid_nodes():
    rng = np.random.default_rng(5)
    n = int(np.prod(SHAPE))
    src = rng.standard_normal(n)
    vel = rng.standard_normal((3, n))
    state = np.zeros((27, n), np.int64)
    state[0, 10] = -666
    state[0, 20] = -1
    gm = o_ggm.gradmag(src, vel, SHAPE, state=state)
    ggm = o_ggm.gradmag(gm, vel, SHAPE, state=state, skip_null=False)
    assert gm[10] > 0 and gm[20] == 0.0
    assert ggm[10] == 0.0 and ggm[20] == 0.0
