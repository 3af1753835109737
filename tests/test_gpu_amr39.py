"""Two D3Q39 levels (ratio 2 in space and time) on the GPU through the C ABI
against the FP64 oracle's schedule on the same inputs: level 0 periodic, level
1 a box with a four-node ghost layer filled from level 0 at theta = 0 and 1/2
(f, g, rho, T interpolated in space and time, the multipliers in space) and
injected back after its second substep.
  FP64 storage: f, g relative L2 <= 1e-11 on both levels (valid nodes of the
    fine level), multipliers max abs <= 1e-8
  FP32 storage: rho, u, T of level 0 relative L2 <= 1e-6 against FP64
  a uniform freestream is a fixed point of the two-level advance."""
import sys
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 4
T0 = 1 / 1.4


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def fields(shape, mach):
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    rho = 1.0 + 0.01 * np.sin(2 * np.pi * (x / nx + y / ny))
    ux = mach + 0.02 * np.sin(2 * np.pi * x / nx)
    uy = 0.01 * np.sin(2 * np.pi * y / ny)
    uz = 0.01 * np.cos(2 * np.pi * (x / nx + z / nz))
    T = T0 * (1 + 0.01 * np.cos(2 * np.pi * z / nz))
    return rho, ux, uy, uz, T


CSHAPE = (16, 12, 12)
FLO, FHI = (8, 6, 6), (23, 17, 17)
FVALID = tuple(FHI[d] - FLO[d] + 1 for d in range(3))
FSTORED = tuple(v + 2 * G for v in FVALID)
FORIGIN = tuple(FLO[d] - G for d in range(3))


def oracle_run(orc, c0, c1, f0, g0, lm0, ncoarse, step0=100):
    """The library's two-level schedule on the oracle.  Every step's start-up
    ramps read the finest level's counter at that moment (the reference's
    timestep(amr_max_lvl)): the coarse step sees the fine counter before the
    fine level's two substeps."""
    sys.path.insert(0, os.path.dirname(orc.__file__))
    import o_amr
    corg = (0, 0, 0)
    fc, gc, lmc = f0, g0, lm0
    ff, gf, lmf, _ = orc.coarse_to_fine39(fc, fc, gc, gc, lmc, CSHAPE, corg, FSTORED, FORIGIN, 1.0, c0)
    gh = orc.ghost_nodes39(FSTORED, G)
    cs, fs = step0, step0
    for _ in range(ncoarse):
        fp, gp = fc, gc
        fc, gc, lmc, _, _ = orc.step_periodic39(CSHAPE, fc, gc, lmc, c0, fs, 1)
        cs += 1
        for theta in (0.0, 0.5):
            fi, gi, li, _ = orc.coarse_to_fine39(fp, fc, gp, gc, lmc, CSHAPE, corg, FSTORED, FORIGIN,
                                                 theta, c0, nodes=gh)
            ff[:, gh], gf[:, gh], lmf[:, gh] = fi, gi, li
            ff, gf, lmf, _, _ = orc.step_ghost39(FSTORED, G, ff, gf, lmf, c1, fs, 1)
            fs += 1
        inj = lambda a, b: o_amr.fine_to_coarse(a, FSTORED, FORIGIN, FLO, FHI, b, CSHAPE, corg)
        fc, gc, lmc = inj(ff, fc), inj(gf, gc), inj(lmf, lmc)
    return fc, gc, lmc, ff, gf, lmf


def gpu_run(pkg, prec, mach, shifted, f0, g0, lm0, ncoarse, step0=100):
    s = pkg.Solver(lattice="d3q39", precision=prec)
    s.set_gas(mach=mach, shifted=int(shifted))
    s.define_level(0, CSHAPE)
    s.upload_state(0, f0, g0, lm0)
    s.define_level(1, FVALID, periodic=(0, 0, 0), lo=FLO)
    s.init_from_coarse(1)
    s.set_timestep(0, step0)
    s.set_timestep(1, step0)
    s.advance(ncoarse)
    return s


def _valid(a, q):
    a = a.reshape(q, FSTORED[2], FSTORED[1], FSTORED[0])
    return a[:, G:-G, G:-G, G:-G].reshape(q, -1)


@pytest.mark.parametrize("shifted,mach", [(False, 0.3), (True, 1.5)])
def test_two_level_d3q39_fp64_parity(pkg, orc, shifted, mach):
    c0 = orc.consts39(mach=mach, shifted=shifted, lvl=0)
    c1 = orc.consts39(mach=mach, shifted=shifted, lvl=1)
    f0, g0, lm0 = orc.init39(CSHAPE, *fields(CSHAPE, mach), c0)
    print(f"two-level D3Q39 shifted={shifted}", flush=True)
    fc, gc, lmc, ff, gf, lmf = oracle_run(orc, c0, c1, f0, g0, lm0, 3)
    s = gpu_run(pkg, "fp64", mach, shifted, f0, g0, lm0, 3)
    fcg, gcg, lmcg = s.download_state(0)
    ffg, gfg, lmfg = s.download_state(1)
    e = [rel_l2(fcg, fc), rel_l2(gcg, gc), rel_l2(_valid(ffg, 39), _valid(ff, 39)),
         rel_l2(_valid(gfg, 39), _valid(gf, 39))]
    elm = max(float(np.abs(lmcg - lmc).max()), float(np.abs(_valid(lmfg, 13) - _valid(lmf, 13)).max()))
    print("relL2 coarse f %.2e g %.2e fine f %.2e g %.2e, max|lm| %.2e" % (*e, elm))
    assert rel_l2(_valid(ff, 39), _valid(orc.coarse_to_fine39(f0, f0, g0, g0, lm0, CSHAPE, (0, 0, 0),
                                                              FSTORED, FORIGIN, 1.0, c0)[0], 39)) > 1e-5
    assert max(e) < 1e-11 and elm < 1e-8
    st = s.stats()
    assert st["n_noconv"] == 0 and st["n_nonpos"] == 0


@pytest.mark.parametrize("shifted,mach", [(False, 0.3), (True, 1.5)])
def test_two_level_d3q39_fp64_parity_in_startup(pkg, orc, shifted, mach):
    """Both levels inside the 25-step start-up window (fine counter 2 .. 7):
    the coarse level's beta, epsilon and alpha ramps follow the fine counter."""
    c0 = orc.consts39(mach=mach, shifted=shifted, lvl=0)
    c1 = orc.consts39(mach=mach, shifted=shifted, lvl=1)
    f0, g0, lm0 = orc.init39(CSHAPE, *fields(CSHAPE, mach), c0)
    fc, gc, lmc, ff, gf, lmf = oracle_run(orc, c0, c1, f0, g0, lm0, 3, step0=2)
    s = gpu_run(pkg, "fp64", mach, shifted, f0, g0, lm0, 3, step0=2)
    fcg, gcg, lmcg = s.download_state(0)
    ffg, gfg, lmfg = s.download_state(1)
    e = [rel_l2(fcg, fc), rel_l2(gcg, gc), rel_l2(_valid(ffg, 39), _valid(ff, 39)),
         rel_l2(_valid(gfg, 39), _valid(gf, 39))]
    print("start-up window: relL2 coarse f %.2e g %.2e fine f %.2e g %.2e" % tuple(e))
    assert max(e) < 1e-11
    # the coarse level's own counter would give a different coarse state
    c_own, *_ = orc.step_periodic39(CSHAPE, f0, g0, lm0, c0, 2, 1)
    c_fine, *_ = orc.step_periodic39(CSHAPE, f0, g0, lm0, c0, 6, 1)
    assert rel_l2(c_own, c_fine) > 1e-9


def test_two_level_d3q39_fp32(pkg, orc):
    mach, shifted = 1.5, True
    c0 = orc.consts39(mach=mach, shifted=shifted, lvl=0)
    c1 = orc.consts39(mach=mach, shifted=shifted, lvl=1)
    f0, g0, lm0 = orc.init39(CSHAPE, *fields(CSHAPE, mach), c0)
    fc, gc, *_ = oracle_run(orc, c0, c1, f0, g0, lm0, 3)
    s = gpu_run(pkg, "fp32s", mach, shifted, f0, g0, lm0, 3)
    fcg, gcg, _ = s.download_state(0)
    r_o, ux_o, uy_o, uz_o, T_o = orc.macro39(fc, gc, c0)
    r_g, ux_g, uy_g, uz_g, T_g = orc.macro39(fcg, gcg, c0)
    e = (rel_l2(r_g, r_o), rel_l2(np.stack([ux_g, uy_g, uz_g]), np.stack([ux_o, uy_o, uz_o])),
         rel_l2(T_g, T_o))
    print("fp32 two-level: relL2 rho %.2e u %.2e T %.2e" % e)
    assert max(e) < 1e-6


def test_two_level_d3q39_freestream_fixed_point(pkg):
    s = pkg.Solver(lattice="d3q39", precision="fp32s")
    s.set_gas(mach=1.5, shifted=1)
    s.define_level(0, CSHAPE)
    s.init_fields(0, 1.0, 1.5, 0.0, 0.0, T0)
    s.define_level(1, FVALID, periodic=(0, 0, 0), lo=FLO)
    s.init_from_coarse(1)
    s.set_timestep(0, 100)
    s.set_timestep(1, 100)
    f0, g0, _ = s.download_state(0)
    f1i, _, _ = s.download_state(1)
    s.advance(2)
    f1, g1, _ = s.download_state(0)
    f1f, _, _ = s.download_state(1)
    assert np.abs(f1 - f0).max() < 1e-6 and np.abs(g1 - g0).max() < 1e-6
    assert np.abs(_valid(f1f, 39) - _valid(f1i, 39)).max() < 1e-6


def test_two_level_d3q39_native_advance_and_graph(pkg, orc):
    """uclbs_advance runs the schedule of Solver.advance in the library: the
    same fields bit for bit, eager and as hipGraph replays (which start once
    both levels are past the start-up ramps), replay after replay with host
    reads in between."""
    mach, shifted = 1.5, True
    c0 = orc.consts39(mach=mach, shifted=shifted, lvl=0)
    f0, g0, lm0 = orc.init39(CSHAPE, *fields(CSHAPE, mach), c0)

    def make():
        s = pkg.Solver(lattice="d3q39", precision="fp32s")
        s.set_gas(mach=mach, shifted=int(shifted))
        s.define_level(0, CSHAPE)
        s.upload_state(0, f0, g0, lm0)
        s.define_level(1, FVALID, periodic=(0, 0, 0), lo=FLO)
        s.init_from_coarse(1)
        s.set_timestep(0, 10)
        s.set_timestep(1, 20)
        return s
    py, eager, graph = make(), make(), make()
    py.advance(16)             # crosses the 25-step start-up of level 0
    eager.advance_native(16, graph=False)
    graph.advance_native(16, graph=True)
    for lvl in (0, 1):
        a = py.download_state(lvl)
        for other in (eager, graph):
            b = other.download_state(lvl)
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), lvl
    for rep in range(3):
        eager.advance_native(2, graph=False)
        graph.advance_native(2, graph=True)   # past the start-up: one replay
        for lvl in (0, 1):
            a, b = eager.download_state(lvl), graph.download_state(lvl)
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), (rep, lvl)
    assert eager.stats()["node_updates"] == graph.stats()["node_updates"]
