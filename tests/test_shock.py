"""The oblique-shock helpers used by the wedge case: the theta-beta-Mach weak
solution against tabulated values (NACA 1135 charts: Mach 2 / 10 degrees ->
39.31 degrees; Mach 1.5 / 10 degrees -> 56.7 degrees), its limits (zero
deflection -> the Mach angle, beyond the maximum deflection -> no attached
shock), and the shock fit on a synthetic density plane with a known shock."""
import importlib
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def sh(pkg):
    return importlib.import_module("uclbs_amd.shock")


def test_theta_beta_mach_tabulated(sh):
    b, r, p = sh.oblique_shock(2.0, 10.0)
    assert abs(b - 39.314) < 0.01
    assert abs(r - 1.4583) < 1e-3 and abs(p - 1.7065) < 1e-3
    b, r, p = sh.oblique_shock(1.5, 10.0)
    assert abs(b - 56.68) < 0.02
    # normal-shock relations at the normal Mach number
    mn = 1.5 * math.sin(math.radians(b))
    assert abs(p - (2.8 * mn * mn - 0.4) / 2.4) < 1e-12


def test_theta_beta_mach_limits(sh):
    b, r, p = sh.oblique_shock(1.5, 1e-6)
    assert abs(b - math.degrees(math.asin(1 / 1.5))) < 1e-3 and abs(r - 1) < 1e-6
    with pytest.raises(ValueError):
        sh.oblique_shock(1.5, 13.0)   # maximum deflection at Mach 1.5 is about 12.1 degrees


def test_shock_fit_synthetic(sh):
    ny, nx = 140, 220
    apex, theta, beta, post = (50.3, 40.27), 10.0, 50.0, 1.3
    y, x = np.mgrid[0:ny, 0:nx].astype(float)
    dx = x - apex[0]
    face = apex[1] + dx * math.tan(math.radians(theta))
    shock = apex[1] + dx * math.tan(math.radians(beta))
    rho = 1.0 + (post - 1.0) * 0.5 * (1 - np.tanh((y - shock) / 1.5))
    solid = (dx >= 0) & (y <= face) & (y >= apex[1] - dx * math.tan(math.radians(theta)))
    rho[dx < 0] = 1.0
    rho[solid] = 1.0
    ang, pr, n = sh.shock_fit(rho, solid, apex, theta, (8, 60))
    assert n >= 40
    assert abs(ang - beta) < 0.2 and abs(pr - post) < 1e-3
