"""Refined D3Q27 levels on the GPU through the C ABI, against oracle/o_amr.py:
  ghost-shell interpolation (space and time) and injection: FP64 storage,
  max abs difference <= 1e-14 (same arithmetic up to summation order);
  a two-level Taylor-Green run stays finite, conserves mass to 1e-4 and
  tracks the single-level run within 5e-2 relative L2 (interface error)."""
import math

import numpy as np
import pytest

import o_amr

pytestmark = pytest.mark.gpu


def _two_levels(pkg, prec, nu=0.01, coarse=(16, 14, 12), flo=(8, 7, 6), fhi=(21, 18, 15)):
    s = pkg.Solver(precision=prec, nu=nu, alpha_tol=1e-12)
    s.define_level(0, coarse)
    fshape = tuple(fhi[d] - flo[d] + 1 for d in range(3))
    s.define_level(1, fshape, periodic=(0, 0, 0), lo=flo)
    return s


def test_interlevel_fill_matches_oracle(pkg, orc):
    coarse, flo, fhi = (16, 14, 12), (8, 7, 6), (21, 18, 15)
    s = _two_levels(pkg, "fp64", coarse=coarse, flo=flo, fhi=fhi)
    fstored = tuple(fhi[d] - flo[d] + 3 for d in range(3))
    forigin = tuple(flo[d] - 1 for d in range(3))
    f0 = orc.init_tgv27(coarse, 0.05, 1e-3, 3)
    s.upload_f(0, f0)
    s.step(0, 1)              # previous state f0, current state f1
    f1 = s.download_f(0)
    s.init_uniform(1, 1.0, (0.0, 0.0, 0.0))
    ff0 = s.download_f(1)
    g = o_amr.ghost_nodes(fstored)
    for theta in (0.0, 0.5, 1.0):
        s.interlevel(1, theta)
        ff = s.download_f(1)
        exp = o_amr.coarse_to_fine(f0, f1, coarse, (0, 0, 0), fstored, forigin, theta, nodes=g)
        assert np.abs(ff[:, g] - exp).max() < 1e-14
        inner = np.setdiff1d(np.arange(ff.shape[1]), g)
        assert np.array_equal(ff[:, inner], ff0[:, inner])   # only the shell is written
    s.init_from_coarse(1)
    ff = s.download_f(1)
    exp = o_amr.coarse_to_fine(f1, f1, coarse, (0, 0, 0), fstored, forigin, 1.0)
    assert np.abs(ff - exp).max() < 1e-14


def test_fine_to_coarse_matches_oracle(pkg, orc):
    coarse, flo, fhi = (16, 14, 12), (8, 7, 6), (21, 18, 15)
    s = _two_levels(pkg, "fp64", coarse=coarse, flo=flo, fhi=fhi)
    fstored = tuple(fhi[d] - flo[d] + 3 for d in range(3))
    forigin = tuple(flo[d] - 1 for d in range(3))
    fc = orc.init_tgv27(coarse, 0.05, 1e-3, 4)
    ff = orc.init_tgv27(fstored, 0.03, 1e-3, 5)
    s.upload_f(0, fc)
    s.upload_f(1, ff)
    s.fine_to_coarse(1)
    got = s.download_f(0)
    exp = o_amr.fine_to_coarse(ff, fstored, forigin, flo, fhi, fc, coarse, (0, 0, 0))
    assert np.array_equal(got, exp)


def test_rejects_fine_box_outside_coarse(pkg):
    s = pkg.Solver(precision="fp64", nu=0.01)
    s.define_level(0, (16, 16, 16))
    with pytest.raises(pkg.UclbsError):
        s.define_level(1, (10, 10, 10), periodic=(0, 0, 0), lo=(0, 4, 4))
    with pytest.raises(pkg.UclbsError):
        s.define_level(1, (10, 10, 10), periodic=(0, 0, 0), lo=(24, 4, 4))


def test_two_level_tgv_tracks_single_level(pkg):
    L = 32
    u0 = 0.05
    nu = 0.01
    single = pkg.Solver(precision="fp32s", nu=nu)
    single.define_level(0, (L, L, L))
    single.init_tgv(0, u0, 0.0, 1)
    two = pkg.Solver(precision="fp32s", nu=nu)
    two.define_level(0, (L, L, L))
    two.define_level(1, (33, 33, 33), periodic=(0, 0, 0), lo=(16, 16, 16))
    two.init_tgv(0, u0, 0.0, 1)
    two.init_from_coarse(1)
    r0 = two.macro(0)[0].sum()
    steps = 40
    single.step(0, steps)
    two.advance(steps)
    rs, uxs, uys, uzs = single.macro(0)
    rt, uxt, uyt, uzt = two.macro(0)
    assert np.all(np.isfinite(uxt))
    e = float(np.linalg.norm(np.stack([uxt, uyt, uzt]) - np.stack([uxs, uys, uzs])) /
              np.linalg.norm(np.stack([uxs, uys, uzs])))
    dm = abs(rt.sum() - r0) / r0
    st = two.stats()
    print(f"two-level TGV vs single level after {steps} coarse steps: u relL2 {e:.3e}, "
          f"mass drift {dm:.2e}, node updates {st['node_updates']}")
    assert e < 5e-2 and dm < 1e-4
