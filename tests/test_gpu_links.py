"""Device-built link tables and the D3Q27 step with walls, inlet and outlet,
against the oracle on the same inputs.
  link tables:  bit-exact (int32 equality of the whole reference-form table)
  FP64 storage: populations relative L2 <= 1e-12, alpha max abs <= 1e-11
  FP32 shifted: u relative L2 <= 1e-5, rho relative L2 <= 1e-7 (vs FP64 oracle)
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FACES = [-101, -10, -10, -201, -10, -10]
SEED_GEOM = 20260102


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _solver(pkg, shape, prec, nu, tol=1e-10):
    s = pkg.Solver(precision=prec, nu=nu, alpha_tol=tol)
    s.define_level(0, shape, periodic=(0, 0, 0))
    s.set_faces(0, FACES)
    return s


def test_device_links_bit_exact_small(pkg, orc):
    shape = (48, 32, 30)
    tri, nrm = pkg.icosphere(4, 8.7, (19.5, 15.5, 14.5), SEED_GEOM)
    st_o, ns_o, nt_o = orc.build_links27(shape, FACES, tri, nrm)
    s = _solver(pkg, shape, "fp64", 1e-2)
    ns_g, nt_g = s.build_links(0, tri, nrm)
    st_g = s.download_links(0)
    assert (ns_g, nt_g) == (ns_o, nt_o)
    assert np.array_equal(st_g, st_o)
    nb, nsolid = s.link_counts(0)
    assert nsolid == ns_o and nb > 0


def test_device_links_bit_exact_c2(pkg, orc):
    """C2 geometry at full size: 256^3 duct, level-5 icosphere (20,480 triangles),
    D = 48, centre (80.5, 128.5, 128.5), rotation seed 20260102."""
    shape = (256, 256, 256)
    tri, nrm = pkg.icosphere(5, 24.0, (80.5, 128.5, 128.5), SEED_GEOM)
    st_o, ns_o, nt_o = orc.build_links27(shape, FACES, tri, nrm)
    s = _solver(pkg, shape, "fp32s", 1e-3)
    ns_g, nt_g = s.build_links(0, tri, nrm)
    assert (ns_g, nt_g) == (ns_o, nt_o) and nt_o == 0
    st_g = s.download_links(0)
    assert np.array_equal(st_g, st_o)
    wall = st_o[1:][st_o[1:] <= -1001]
    print(f"C2 tables: solid {ns_o}, wall links {wall.size}, ties {nt_o}")


def _duct_case(pkg, orc):
    shape = (40, 24, 26)
    tri, nrm = pkg.icosphere(4, 5.6, (14.5, 11.5, 12.5), SEED_GEOM)
    st, ns, nt = orc.build_links27(shape, FACES, tri, nrm)
    u_in = (0.06, 0.0, 0.0)
    return shape, tri, nrm, st, u_in


def test_fp64_box_step_parity(pkg, orc):
    shape, tri, nrm, st, u_in = _duct_case(pkg, orc)
    nu = 0.02
    beta = 1 / (6 * nu + 1)
    f0 = orc.uniform27(st, 1.0, u_in)
    steps = 30
    f_o, a_o, ost, bst = orc.step_box27(shape, FACES, st, f0, beta, nsteps=steps, tol=1e-12,
                                        inlets=[(1.0,) + u_in], outlets=[(1.0,) + u_in])
    s = _solver(pkg, shape, "fp64", nu, tol=1e-12)
    s.set_bc_states(inlets=[(1.0,) + u_in], outlets=[(1.0,) + u_in])
    s.build_links(0, tri, nrm)
    s.keep_alpha(True)
    s.upload_f(0, f0)
    s.step(0, steps)
    f_g = s.download_f(0)
    w = orc.lattice27()[4]
    fluid = st[0] >= 0
    assert rel_l2((f_g - w[:, None])[:, fluid], (f_o - w[:, None])[:, fluid]) < 1e-12
    assert np.abs(s.alpha(0)[fluid] - a_o[fluid]).max() < 1e-11
    g = s.stats()
    assert g["n_ibb_links"] == bst["n_ibb_links"] and g["n_repair"] == bst["n_repair"]
    assert g["node_updates"] == ost["nodes"]


def test_fp32_shifted_box_step(pkg, orc):
    shape, tri, nrm, st, u_in = _duct_case(pkg, orc)
    nu = 5e-3
    beta = 1 / (6 * nu + 1)
    f0 = orc.uniform27(st, 1.0, u_in)
    steps = 100
    f_o, a_o, ost, bst = orc.step_box27(shape, FACES, st, f0, beta, nsteps=steps,
                                        inlets=[(1.0,) + u_in], outlets=[(1.0,) + u_in])
    s = _solver(pkg, shape, "fp32s", nu)
    s.set_bc_states(inlets=[(1.0,) + u_in], outlets=[(1.0,) + u_in])
    s.build_links(0, tri, nrm)
    s.keep_alpha(True)
    s.init_uniform(0, 1.0, u_in)
    s.step(0, steps)
    fluid = st[0] >= 0
    r_o, ux_o, uy_o, uz_o = orc.macro27(f_o)
    r_g, ux_g, uy_g, uz_g = s.macro(0)
    e_u = rel_l2(np.stack([ux_g, uy_g, uz_g])[:, fluid], np.stack([ux_o, uy_o, uz_o])[:, fluid])
    e_r = rel_l2(r_g[fluid], r_o[fluid])
    e_a = float(np.abs(s.alpha(0)[fluid] - a_o[fluid]).max())
    print(f"duct fp32s vs fp64 oracle ({steps} steps): relL2 u {e_u:.2e} rho {e_r:.2e} "
          f"max|alpha| {e_a:.2e}")
    assert e_u < 1e-5 and e_r < 1e-7 and e_a < 1e-6


def test_c2_full_size_runs(pkg):
    """C2 at full size, FP32 shifted, 100 steps: bounded, no non-convergence, inflow
    carried downstream (size-independent checks)."""
    L = 256
    u_in = (0.2 / math.sqrt(3.0), 0.0, 0.0)
    tri, nrm = pkg.icosphere(5, 24.0, (80.5, 128.5, 128.5), SEED_GEOM)
    s = _solver(pkg, (L, L, L), "fp32s", 0.002)
    s.set_bc_states(inlets=[(1.0,) + u_in], outlets=[(1.0,) + u_in])
    s.build_links(0, tri, nrm)
    s.init_uniform(0, 1.0, u_in)
    s.step(0, 100)
    r, ux, uy, uz = s.macro(0)
    assert np.all(np.isfinite(r)) and np.all(np.isfinite(ux))
    assert r.min() > 0.8 and r.max() < 1.2 and np.abs(ux).max() < 0.3
    st = s.stats()
    assert st["n_noconv"] == 0 and st["n_nonpos"] == 0
    assert st["n_ibb_links"] > 100 * 60000
