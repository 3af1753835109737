"""GGM refinement indicator, Python restatement.  TEST INFRASTRUCTURE ONLY.

Restates the reference's gradient-of-gradient-magnitude tagging on one box:
  GGM/ggm_calculator.f95:47-68          populate_q, calculate_grad_mag,
                                        calculate_gradgrad_mag per variable
  GGM/populate_q.f95:23-56              the quantity: 1 density, 2 u, 3 v,
                                        4 w, 7 velocity magnitude
  GGM/derivatives/calculate_grad_mag.f95:33-61
                                        gm = |grad q|; 0 where the node's own
                                        code is < 0 and not -666
  GGM/derivatives/calculate_gradgrad_mag.f95:37-75
                                        ggm = |grad gm|; 0 where the node's own
                                        code is < 0; tag where ggm >= limit
  derivatives/modules/derivatives.f95:589-700
                                        derivative_windage_firstord / _secondord
                                        over (x, x+1, x+2, x-1, x-2) with the
                                        codes of links 1+a, 15+a, 4+a, 18+a
                                        (reference component indices 2+a,
                                        16+a, 5+a, 19+a) deciding use
A stencil node beyond a non-periodic face of the box is never usable and
reads as 0 (the reference reads its AMReX ghost value there).  A refined level
of many boxes is restated on a dense array over its stored nodes' bounding
box (boxes_dense): ghost-shell nodes first (every link usable), valid nodes
over them (the level's link codes); positions no box stores are unfilled,
never computed and never usable.

windage1 / windage2 are the scalar functions as the reference writes them;
gradmag is the vectorised form over a box, pinned against them node by node.
"""
import numpy as np

CENTRAL, UPWIND, DOWNWIND = 0, 1, 2


def windage1(q, ok, wind, scheme):
    """derivative_windage_firstord (derivatives.f95:589-642); q, ok 1-based
    as (x, x+1, x-1) through index 3 (the secondord fallback's packing)."""
    if scheme == UPWIND:
        if wind > 0.0:
            return q[2] - q[0] if ok[2] else 0.0
        return q[0] - q[1] if ok[1] else 0.0
    if scheme == DOWNWIND:
        if wind < 0.0:
            return q[2] - q[0] if ok[2] else 0.0
        return q[0] - q[1] if ok[1] else 0.0
    raise ValueError("the secondord fallback never calls the central first-order form")


def windage2(q, ok, wind, scheme):
    """derivative_windage_secondord (derivatives.f95:644-700): q, ok over
    (x, x+1, x+2, x-1, x-2)."""
    if scheme == CENTRAL:
        if ok[1] and ok[3]:
            return (q[3] - q[1]) / 2.0
        if ok[1]:
            return q[1] - q[0]
        if ok[3]:
            return q[0] - q[3]
        return 0.0
    back = wind > 0.0 if scheme == UPWIND else wind < 0.0
    if back:
        if ok[3] and ok[4]:
            return (3 * q[0] - 4 * q[3] + q[4]) / 2
    elif ok[1] and ok[2]:
        return (-3 * q[0] + 4 * q[3] - q[4]) / 2
    return windage1((q[0], q[1], q[3]), (ok[0], ok[1], ok[3]), wind, scheme)


def _stencil(shape, periodic, state, ax, filled=None):
    """Per node: neighbour index (or -1) and usability of (x+1, x+2, x-1, x-2)."""
    nx, ny, nz = shape
    n = nx * ny * nz
    k = np.arange(n)
    p = [k % nx, (k // nx) % ny, k // (nx * ny)]
    ext = (nx, ny, nz)
    nd, ok = [], []
    for off, comp in ((1, 1 + ax), (2, 15 + ax), (-1, 4 + ax), (-2, 18 + ax)):
        c = p[ax] + off
        inb = (c >= 0) & (c < ext[ax])
        if periodic[ax]:
            c = c % ext[ax]
            inb = np.ones(n, bool)
        q = list(p)
        q[ax] = np.where(inb, c, 0)
        idx = np.where(inb, q[0] + nx * (q[1] + ny * q[2]), -1)
        use = inb.copy()
        if state is not None:
            use &= state[comp] >= 0
        if filled is not None:
            use &= np.where(idx >= 0, filled[np.maximum(idx, 0)] != 0, False)
        nd.append(idx)
        ok.append(use)
    return nd, ok


def gradmag(src, vel, shape, periodic=(0, 0, 0), state=None, scheme=CENTRAL, skip_null=True,
            filled=None):
    """|grad src| per node (calculate_grad_mag.f95:33-61 when skip_null,
    calculate_gradgrad_mag.f95:37-63 otherwise); vel [3][n] the wind;
    filled (optional): only filled positions are computed or usable."""
    src = np.asarray(src, np.float64)
    d2 = np.zeros(src.shape)
    for ax in range(3):
        nd, ok = _stencil(shape, periodic, state, ax, filled)
        q = [src] + [np.where(i >= 0, src[np.maximum(i, 0)], 0.0) for i in nd]
        o1, o2, o3, o4 = ok
        w = vel[ax]
        if scheme == CENTRAL:
            d = np.where(o1 & o3, (q[3] - q[1]) / 2.0,
                         np.where(o1, q[1] - q[0], np.where(o3, q[0] - q[3], 0.0)))
        else:
            back = w > 0.0 if scheme == UPWIND else w < 0.0
            first = np.where(back, np.where(o3, q[3] - q[0], 0.0), np.where(o1, q[0] - q[1], 0.0))
            d = np.where(back, np.where(o3 & o4, (3 * q[0] - 4 * q[3] + q[4]) / 2, first),
                         np.where(o1 & o2, (-3 * q[0] + 4 * q[3] - q[4]) / 2, first))
        d2 = d2 + d * d
    out = np.sqrt(d2)
    if state is not None:
        own = state[0]
        dead = (own < 0) & ~(skip_null & (own == -666))
        out = np.where(dead, 0.0, out)
    if filled is not None:
        out = np.where(filled != 0, out, 0.0)
    return out


def fields27(f, cx, cy, cz):
    """rho, (u, v, w) of full D3Q27 populations f [27][n]."""
    rho = f.sum(0)
    vel = np.stack([np.asarray(c, np.float64) @ f for c in (cx, cy, cz)]) / rho
    return rho, vel


def quantity(rho, vel, which):
    """populate_q.f95:23-56 for the quantities a D3Q27 level has."""
    if which == 1:
        return rho
    if which in (2, 3, 4):
        return vel[which - 2]
    if which == 7:
        return np.sqrt(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    raise ValueError(which)


def ggm27(f, shape, cx, cy, cz, which=1, scheme=CENTRAL, periodic=(0, 0, 0), state=None):
    """(gm, ggm) of a D3Q27 box (ggm_calculator.f95:47-68)."""
    rho, vel = fields27(f, cx, cy, cz)
    q = quantity(rho, vel, which)
    gm = gradmag(q, vel, shape, periodic, state, scheme, skip_null=True)
    return gm, gradmag(gm, vel, shape, periodic, state, scheme, skip_null=False)


def tags(ggm, limit, valid=None):
    """calculate_gradgrad_mag.f95:67-71: tag where ggm >= limit (valid nodes)."""
    t = ggm >= limit
    return t if valid is None else t & valid


def boxes_dense(f, state, boxes, cx, cy, cz, which=1, scheme=CENTRAL):
    """ggm of a refined level of many boxes, in its stored order: f [27][n]
    and state [27][n] (or None) stored box after box, each with its one-node
    ghost shell; boxes [(lo, hi)] their valid regions.  Returns (ggm, valid)
    over the stored nodes (0 on ghost shells)."""
    dlo = np.min([b[0] for b in boxes], axis=0) - 1
    dhi = np.max([b[1] for b in boxes], axis=0) + 1
    ext = dhi - dlo + 1
    nd = int(np.prod(ext))
    rho, vel = fields27(f, cx, cy, cz)
    q = quantity(rho, vel, which)
    qd = np.zeros(nd)
    vd = np.zeros((3, nd))
    filled = np.zeros(nd, np.int8)
    sd = np.zeros((27, nd), np.int64) if state is not None else None
    pos, off = [], 0
    for lo, hi in boxes:
        e = np.array(hi) - np.array(lo) + 3
        z, y, x = np.mgrid[0:e[2], 0:e[1], 0:e[0]]
        x, y, z = x.ravel(), y.ravel(), z.ravel()
        shell = (x == 0) | (y == 0) | (z == 0) | (x == e[0] - 1) | (y == e[1] - 1) | (z == e[2] - 1)
        g = np.stack([x + lo[0] - 1, y + lo[1] - 1, z + lo[2] - 1]) - dlo[:, None]
        kd = g[0] + ext[0] * (g[1] + ext[1] * g[2])
        pos.append((np.arange(off, off + x.size), kd, shell))
        off += x.size
    for pass_ in (0, 1):
        for ks, kd, shell in pos:
            m = shell if pass_ == 0 else ~shell
            qd[kd[m]] = q[ks[m]]
            vd[:, kd[m]] = vel[:, ks[m]]
            filled[kd[m]] = 1 + pass_
            if sd is not None:
                sd[:, kd[m]] = 0 if pass_ == 0 else state[:, ks[m]]
    shape = tuple(int(v) for v in ext)
    gm = gradmag(qd, vd, shape, (0, 0, 0), sd, scheme, True, filled)
    ggm = gradmag(gm, vd, shape, (0, 0, 0), sd, scheme, False, filled)
    out = np.zeros(off)
    valid = np.zeros(off, bool)
    for ks, kd, shell in pos:
        out[ks[~shell]] = ggm[kd[~shell]]
        valid[ks[~shell]] = True
    return out, valid
