"""Oblique-shock measurements for the supersonic wedge case (the reference's
demonstration case of a 10 degree wedge at Mach 1.5, README.txt:80-92): the
weak-shock solution of the theta-beta-Mach relation with its
Rankine-Hugoniot jumps, and the shock angle and shock-layer density read off
a computed density plane."""
from __future__ import annotations

import math

import numpy as np


def oblique_shock(mach: float, theta_deg: float, gamma: float = 1.4):
    """Weak-branch shock angle (degrees) of the theta-beta-Mach relation
    tan(theta) = 2 cot(beta) (M^2 sin^2 beta - 1) / (M^2 (gamma + cos 2 beta) + 2)
    and the density and pressure ratios across that shock (Rankine-Hugoniot).
    Bisection between the Mach angle and the angle of maximum deflection."""
    th = math.radians(theta_deg)

    def deflection(b):
        return math.atan(2 / math.tan(b) * (mach ** 2 * math.sin(b) ** 2 - 1) /
                         (mach ** 2 * (gamma + math.cos(2 * b)) + 2))
    lo = math.asin(1 / mach)
    # angle of maximum deflection (the weak branch lies below it)
    grid = np.linspace(lo + 1e-9, math.pi / 2 - 1e-9, 20001)
    hi = float(grid[int(np.argmax([deflection(b) for b in grid]))])
    if deflection(hi) < th:
        raise ValueError(f"no attached shock: theta {theta_deg} exceeds the maximum deflection "
                         f"{math.degrees(deflection(hi)):.3f} at Mach {mach}")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if deflection(mid) < th:
            lo = mid
        else:
            hi = mid
    b = 0.5 * (lo + hi)
    mn2 = (mach * math.sin(b)) ** 2
    rho_ratio = (gamma + 1) * mn2 / ((gamma - 1) * mn2 + 2)
    p_ratio = 1 + 2 * gamma / (gamma + 1) * (mn2 - 1)
    return math.degrees(b), rho_ratio, p_ratio


def shock_fit(rho: np.ndarray, solid: np.ndarray, apex, theta_deg: float, span,
              rho_inf: float = 1.0):
    """Oblique shock above a wedge's upper face (the line from apex at
    theta_deg) on one plane rho[y, x].  For each column x in
    apex_x + span[0] .. apex_x + span[1]: a first shock height where rho falls
    through the mean of rho_inf and the column's maximum above the face; the
    post-shock density as the mean over the middle 35-65 % of the layer
    between face and shock (clear of the near-wall layer and of the shock's
    own thickness); the shock height again at the mean of rho_inf and that
    value.  Returns (angle of the least-squares shock line in degrees, mean
    post-shock density, number of columns)."""
    x0, y0 = apex
    ny, nx = rho.shape
    tn = math.tan(math.radians(theta_deg))
    pts = []
    for x in range(int(x0) + span[0], min(nx, int(x0) + span[1])):
        col = rho[:, x]
        ys = y0 + (x - x0) * tn

        def crossing(level):
            yy = None
            for y in range(int(ys) + 2, ny - 1):
                if not solid[y, x] and col[y] >= level > col[y + 1]:
                    yy = y + (col[y] - level) / (col[y] - col[y + 1])
            return yy
        ysh = crossing(0.5 * (col[int(ys) + 2:].max() + rho_inf))
        if ysh is None:
            continue
        band = np.arange(int(math.ceil(ys + 0.35 * (ysh - ys))), int(ys + 0.65 * (ysh - ys)) + 1)
        if band.size == 0:
            continue
        post = float(col[band].mean())
        yy = crossing(0.5 * (post + rho_inf))
        if yy is not None:
            pts.append((x - x0, yy - y0, post))
    if len(pts) < 2:
        return float("nan"), float("nan"), len(pts)
    pts = np.array(pts)
    k = np.polyfit(pts[:, 0], pts[:, 1], 1)
    return math.degrees(math.atan(k[0])), float(pts[:, 2].mean()), len(pts)
