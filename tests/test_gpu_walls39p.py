"""D3Q39 wall links on the plain lattice (Mach < 1.2) on the GPU, through the
C ABI: the table built on the device (state_directions.f95 form, long links
derived from the shorter ones) and the population opposite each wall link
rebuilt every step by the reference's plain-lattice wall rule (target state,
wall-form and local derivatives, the Hermite projection of
interpolative_bounceback_bc.f95's 38 cases with their slips).
  FP64 storage against the FP64 oracle on the same inputs, 30 steps from step
    0 through the start-up ramps, Mach 0.3 past a sphere with a freestream
    inlet, an outflow and inviscid sides: f, g relative L2 <= 1e-11,
    multipliers max abs <= 1e-8, alpha on fluid nodes relative L2 <= 1e-9;
    wall-link, repair and no-convergence counts equal
  FP32 storage: rho, u, T relative L2 <= 1e-6 against the FP64 oracle at 20
    steps and against the oracle with FP32 storage rounding at 30 steps
  a fluid at rest around the sphere in a closed box stays at rest to 1e-13
    (FP64, 10 steps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T0 = 1 / 1.4
MACH = 0.3
FACES = [-1000, -11, -11, -6, -11, -11]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def sphere(pkg):
    return pkg.icosphere(4, 4.3, (12.5, 11.75, 7.5), 20260103)


def solver(pkg, shape, prec, faces=FACES):
    s = pkg.Solver(lattice="d3q39", precision=prec)
    s.set_gas(mach=MACH, shifted=0, turbulent=True)
    s.define_level(0, shape, periodic=(0, 0, 0))
    s.set_faces(0, faces)
    s.keep_alpha(True)
    return s


def test_plain_wall_rule_fp64_parity(pkg, orc):
    shape, steps = (32, 24, 16), 30
    c = orc.consts39(mach=MACH, shifted=False)
    tri, nrm = sphere(pkg)
    st_o, ns_o, _ = orc.build_links39(shape, FACES, tri, nrm)
    f0, g0, lm0 = orc.init39(shape, 1.0, MACH, 0.0, 0.0, T0, c)
    print(f"solid {ns_o}, wall links {int(np.count_nonzero(st_o[1:] <= -1001))}", flush=True)
    f_o, g_o, lm_o, a_o, so, nwl, _ = orc.step_walls39(shape, FACES, st_o, f0, g0, lm0, c,
                                                       step0=0, nsteps=steps, keep_alpha=True)
    s = solver(pkg, shape, "fp64")
    ns_g, _ = s.build_links(0, tri, nrm)
    assert ns_g == ns_o and np.array_equal(s.download_links(0), st_o)
    s.upload_state(0, f0, g0, lm0)
    s.set_timestep(0, 0)
    s.step(0, steps)
    f_g, g_g, lm_g = s.download_state(0)
    sg = s.stats()
    fl = st_o[0] >= 0
    e_f, e_g = rel_l2(f_g, f_o), rel_l2(g_g, g_o)
    e_lm = float(np.abs(lm_g - lm_o)[:, fl].max())
    e_a = rel_l2(s.alpha(0)[fl], a_o[fl])
    print(f"plain walls: relL2 f {e_f:.2e} g {e_g:.2e} max|lm| {e_lm:.2e} relL2 alpha {e_a:.2e}; "
          f"wall links {sg['n_ibb_links']} vs {nwl}, repairs {sg['n_repair']} vs {so['n_repair']}, "
          f"noconv {sg['n_noconv']} vs {so['lm_noconv']}")
    assert rel_l2(f_o, f0) > 1e-3
    assert e_f < 1e-11 and e_g < 1e-11 and e_lm < 1e-8 and e_a < 1e-9
    assert sg["node_updates"] == so["nodes"]
    assert sg["n_ibb_links"] == nwl and nwl > 1000
    assert sg["n_repair"] == so["n_repair"] and sg["n_noconv"] == so["lm_noconv"]


def test_plain_wall_rule_fp32_storage(pkg, orc):
    shape = (32, 24, 16)
    c = orc.consts39(mach=MACH, shifted=False)
    tri, nrm = sphere(pkg)
    st_o, _, _ = orc.build_links39(shape, FACES, tri, nrm)
    f0, g0, lm0 = orc.init39(shape, 1.0, MACH, 0.0, 0.0, T0, c)
    fsf, fsg = orc.freestream39(c)
    fl = st_o[0] >= 0

    def fields(f, g):
        r, ux, uy, uz, T = [a[fl] for a in orc.macro39(f, g, c)]
        return r, np.stack([ux, uy, uz]), T

    s = solver(pkg, shape, "fp32s")
    s.build_links(0, tri, nrm)
    s.upload_state(0, f0, g0, lm0)
    s.set_timestep(0, 0)
    s.step(0, 20)
    g20 = fields(*s.download_state(0)[:2])
    s.step(0, 10)
    g30 = fields(*s.download_state(0)[:2])
    f64, g64, _, _, _, _, _ = orc.step_walls39(shape, FACES, st_o, f0, g0, lm0, c, 0, 20)
    o20 = fields(f64, g64)
    fe, ge, lme = f0, g0, lm0
    for k in range(30):
        fe, ge, lme, _, _, _, _ = orc.step_walls39(shape, FACES, st_o, fe, ge, lme, c, k, 1)
        fe = fsf[:, None] + (fe - fsf[:, None]).astype(np.float32)
        ge = fsg[:, None] + (ge - fsg[:, None]).astype(np.float32)
    e30 = fields(fe, ge)
    d20 = [rel_l2(a, b) for a, b in zip(g20, o20)]
    d30 = [rel_l2(a, b) for a, b in zip(g30, e30)]
    print("fp32 plain walls vs fp64 oracle at 20 steps: relL2 rho {:.2e} u {:.2e} T {:.2e}".format(*d20))
    print("fp32 plain walls vs oracle with fp32 storage at 30 steps: relL2 rho {:.2e} u {:.2e} "
          "T {:.2e}".format(*d30))
    assert max(d20) < 1e-6 and max(d30) < 1e-6


def test_plain_wall_rest_state(pkg, orc):
    shape = (32, 24, 16)
    walls = [-10] * 6
    c = orc.consts39(mach=MACH, shifted=False)
    tri, nrm = sphere(pkg)
    s = solver(pkg, shape, "fp64", faces=walls)
    s.build_links(0, tri, nrm)
    f0, g0, lm0 = orc.init39(shape, 1.0, 0.0, 0.0, 0.0, T0, c)
    s.upload_state(0, f0, g0, lm0)
    s.set_timestep(0, 0)
    s.step(0, 10)
    f1, g1, _ = s.download_state(0)
    st = s.stats()
    print(f"rest: max|df| {np.abs(f1 - f0).max():.2e} max|dg| {np.abs(g1 - g0).max():.2e}, "
          f"wall links {st['n_ibb_links']}")
    assert st["n_ibb_links"] > 1000 and st["n_repair"] == 0
    assert np.abs(f1 - f0).max() < 1e-13 and np.abs(g1 - g0).max() < 1e-13
