"""Coarse-fine exchange operators, numpy restatement.  TEST INFRASTRUCTURE ONLY.

Restates the operators the reference delegates to AMReX (not vendored; "22.06
or higher", README.txt): amrex_fillcoarsepatch with amrex_interp_node_bilinear
(interlevel_interpolation.f95:114-257: nodal interpolation per axis at ratio
2, linear in time between the coarse level's previous and current states) and
amrex_average_down_nodes (fine_to_coarse.f95:30-64: nodal injection).  Parity
with AMReX itself is unpinned (AMReX is absent); these definitions are pinned
by exactness on fields linear in space and time.

Arrays are (Q, n) with node = x + nx (y + ny z) of the stored box; a level's
stored node (a, b, c) has global index origin + (a, b, c).
"""
import numpy as np


def _coords(shape):
    nx, ny, nz = shape
    k = np.arange(nx * ny * nz)
    return np.stack([k % nx, (k // nx) % ny, k // (nx * ny)])


def coarse_to_fine(cprev, cnew, cshape, corigin, fshape, forigin, theta, nodes=None):
    """Fine populations at the stored fine nodes `nodes` (all if None)."""
    cprev = np.asarray(cprev, np.float64)
    cnew = np.asarray(cnew, np.float64)
    c = (1 - theta) * cprev + theta * cnew
    loc = _coords(fshape)
    if nodes is not None:
        loc = loc[:, nodes]
    G = loc + np.asarray(forigin)[:, None]
    lo = np.floor_divide(G, 2) - np.asarray(corigin)[:, None]
    odd = G % 2
    out = np.zeros((c.shape[0], G.shape[1]))
    nx, ny, nz = cshape
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                d = np.array([dx, dy, dz])[:, None]
                w = np.prod(np.where(odd == 1, 0.5, np.where(d == 0, 1.0, 0.0)), axis=0)
                idx = lo + d * odd
                k = idx[0] + nx * (idx[1] + ny * idx[2])
                out += w[None, :] * c[:, k]
    return out


def ghost_nodes(fshape):
    loc = _coords(fshape)
    nx, ny, nz = fshape
    g = (loc[0] == 0) | (loc[1] == 0) | (loc[2] == 0) | (loc[0] == nx - 1) | (loc[1] == ny - 1) | \
        (loc[2] == nz - 1)
    return np.nonzero(g)[0]


def fine_to_coarse(ffine, fshape, forigin, flo, fhi, fcoarse, cshape, corigin):
    """Covered coarse nodes C (2C inside the fine valid box [flo, fhi]) take
    the fine populations at 2C; returns the updated coarse array."""
    out = np.array(fcoarse, np.float64, copy=True)
    clo = [(flo[d] + 1) // 2 for d in range(3)]
    chi = [fhi[d] // 2 for d in range(3)]
    cz, cy, cx = np.meshgrid(*[np.arange(clo[d], chi[d] + 1) for d in (2, 1, 0)], indexing="ij")
    C = np.stack([cx.ravel(), cy.ravel(), cz.ravel()])
    kc = (C[0] - corigin[0]) + cshape[0] * ((C[1] - corigin[1]) + cshape[1] * (C[2] - corigin[2]))
    F = 2 * C - np.asarray(forigin)[:, None]
    kf = F[0] + fshape[0] * (F[1] + fshape[1] * F[2])
    out[:, kc] = np.asarray(ffine)[:, kf]
    return out
