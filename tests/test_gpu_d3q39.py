"""D3Q39 two-population step on the GPU (through the C ABI) against the FP64
oracle on the same inputs.
  FP64 storage: f, g relative L2 <= 1e-11, multipliers max abs <= 1e-8,
                alpha relative L2 <= 1e-11 and max abs <= 1e-9 (the
                closed-form estimate amplifies last-bit differences of the
                multipliers at nodes near equilibrium; the kernel factorises the Jacobian as
                L D L^T, the oracle as LU: same Newton steps up to round-off)
  FP32 storage: rho, u, T relative L2 <= 1e-6 against the FP64 oracle
The startup ramps (25 steps) are crossed in every parity run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 1.4
T0 = 1 / G


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


def solver(pkg, shape, prec, mach, shifted, turbulent=True):
    s = pkg.Solver(lattice="d3q39", precision=prec)
    s.set_gas(mach=mach, shifted=int(shifted), turbulent=turbulent)
    s.define_level(0, shape)
    s.keep_alpha(True)
    return s


CASES = [(False, 0.3), (True, 1.5)]


@pytest.mark.parametrize("shifted,mach", CASES)
def test_d3q39_init_matches_oracle(pkg, orc, shifted, mach):
    shape = (10, 8, 6)
    c = orc.consts39(mach=mach, shifted=shifted)
    fl = fields(shape, mach)
    f_o, g_o, lm_o = orc.init39(shape, *fl, c)
    s = solver(pkg, shape, "fp64", mach, shifted)
    s.init_fields(0, *fl)
    f_g, g_g, lm_g = s.download_state(0)
    assert rel_l2(f_g, f_o) < 1e-13 and rel_l2(g_g, g_o) < 1e-13
    assert np.abs(lm_g - lm_o).max() < 1e-9


@pytest.mark.parametrize("shifted,mach", CASES)
def test_d3q39_fp64_parity(pkg, orc, shifted, mach):
    shape = (12, 10, 8)
    steps = 30
    c = orc.consts39(mach=mach, shifted=shifted)
    fl = fields(shape, mach)
    f0, g0, lm0 = orc.init39(shape, *fl, c)
    f_o, g_o, lm_o, a_o, st_o = orc.step_periodic39(shape, f0, g0, lm0, c, 0, steps, keep_alpha=True)
    s = solver(pkg, shape, "fp64", mach, shifted)
    s.upload_state(0, f0, g0, lm0)
    s.step(0, steps)
    f_g, g_g, lm_g = s.download_state(0)
    st = s.stats()
    e_f, e_g = rel_l2(f_g, f_o), rel_l2(g_g, g_o)
    e_lm = float(np.abs(lm_g - lm_o).max())
    a_g = s.alpha(0)
    e_a = float(np.abs(a_g - a_o).max())
    e_al2 = float(np.linalg.norm(a_g - a_o) / np.linalg.norm(a_o))
    print(f"D3Q39 fp64 shifted={shifted}: relL2 f {e_f:.2e} g {e_g:.2e} max|lm| {e_lm:.2e} "
          f"max|alpha| {e_a:.2e} relL2 alpha {e_al2:.2e} iters {st['newton_iters']} vs {st_o['lm_iters']}")
    assert e_f < 1e-11 and e_g < 1e-11 and e_lm < 1e-8 and e_al2 < 1e-11 and e_a < 1e-9
    assert st["node_updates"] == st_o["nodes"] and st["n_noconv"] == 0
    assert abs(st["newton_iters"] - st_o["lm_iters"]) <= 0.01 * st_o["lm_iters"]
    assert st["n_repair"] == st_o["n_repair"]
    assert st["n_nonpos"] == 0 and st["n_fallback"] == 0


@pytest.mark.parametrize("shifted,mach", CASES)
def test_d3q39_fp64_parity_far_from_equilibrium(pkg, orc, shifted, mach):
    """A state 50% off equilibrium with 1% of the f populations negative,
    stepped past the start-up ramps: the alpha estimate's low clamp
    (calculate_alpha.f95:118-289) and its |f| form for negative populations,
    and the damp_it repair of hundreds of nodes, against the oracle."""
    shape = (12, 10, 8)
    steps, step0 = 3, 30
    c = orc.consts39(mach=mach, shifted=shifted)
    f0, g0, lm0 = orc.init39(shape, *fields(shape, mach), c)
    rng = np.random.default_rng(20261202)
    f0 = f0 * (1 + 0.5 * rng.uniform(-1, 1, f0.shape))
    g0 = g0 * (1 + 0.5 * rng.uniform(-1, 1, g0.shape))
    neg = rng.random(f0.shape) < 0.01
    f0[neg] *= -0.1
    f_o, g_o, lm_o, a_o, st_o = orc.step_periodic39(shape, f0, g0, lm0, c, step0, steps,
                                                    keep_alpha=True)
    assert st_o["n_alpha_lo"] > 0 and st_o["n_repair"] > 100  # the branches are reached
    s = solver(pkg, shape, "fp64", mach, shifted)
    s.upload_state(0, f0, g0, lm0)
    s.set_timestep(0, step0)
    s.step(0, steps)
    f_g, g_g, lm_g = s.download_state(0)
    st = s.stats()
    a_g = s.alpha(0)
    e_f, e_g = rel_l2(f_g, f_o), rel_l2(g_g, g_o)
    e_lm = float(np.abs(lm_g - lm_o).max())
    e_a = float(np.abs(a_g - a_o).max())
    e_al2 = float(np.linalg.norm(a_g - a_o) / np.linalg.norm(a_o))
    print(f"D3Q39 fp64 far shifted={shifted}: relL2 f {e_f:.2e} g {e_g:.2e} max|lm| {e_lm:.2e} "
          f"max|alpha| {e_a:.2e} relL2 alpha {e_al2:.2e} bound {st['n_bound']} vs "
          f"{st_o['n_alpha_lo'] + st_o['n_alpha_hi']} repairs {st['n_repair']} vs {st_o['n_repair']}")
    assert e_f < 1e-11 and e_g < 1e-11 and e_lm < 1e-8 and e_al2 < 1e-11 and e_a < 1e-9
    assert st["node_updates"] == st_o["nodes"] and st["n_noconv"] == 0
    assert st["n_repair"] == st_o["n_repair"]
    assert st["n_trivial"] == st_o["n_alpha2"]
    assert st["n_bound"] == st_o["n_alpha_lo"] + st_o["n_alpha_hi"]
    # the D3Q39 step counts an estimate with no real root as n_nonpos and a
    # conductivity outside its fit as n_fallback
    assert st["n_nonpos"] == st_o["n_nan"] and st["n_fallback"] == st_o["n_kappa_err"]


@pytest.mark.parametrize("shifted,mach", CASES)
def test_d3q39_fp32_against_fp64_oracle(pkg, orc, shifted, mach):
    shape = (12, 10, 8)
    steps = 30
    c = orc.consts39(mach=mach, shifted=shifted)
    fl = fields(shape, mach)
    f0, g0, lm0 = orc.init39(shape, *fl, c)
    f_o, g_o, lm_o, a_o, st_o = orc.step_periodic39(shape, f0, g0, lm0, c, 0, steps)
    s = solver(pkg, shape, "fp32s", mach, shifted)
    s.upload_state(0, f0, g0, lm0)
    s.step(0, steps)
    f_g, g_g, lm_g = s.download_state(0)
    r_o, ux_o, uy_o, uz_o, T_o = orc.macro39(f_o, g_o, c)
    r_g, ux_g, uy_g, uz_g, T_g = orc.macro39(f_g, g_g, c)
    e_r = rel_l2(r_g, r_o)
    e_u = rel_l2(np.stack([ux_g, uy_g, uz_g]), np.stack([ux_o, uy_o, uz_o]))
    e_T = rel_l2(T_g, T_o)
    print(f"D3Q39 fp32 shifted={shifted}: relL2 rho {e_r:.2e} u {e_u:.2e} T {e_T:.2e}")
    assert e_r < 1e-6 and e_u < 1e-6 and e_T < 1e-6


def test_d3q39_c3_size_uniform_and_conservation(pkg):
    """C3 shape 384 x 192 x 96, shifted lattice at Mach 1.5, FP32 storage:
    a uniform freestream is a fixed point; a perturbed one keeps its mass."""
    shape = (384, 192, 96)
    n = shape[0] * shape[1] * shape[2]
    s = solver(pkg, shape, "fp32s", 1.5, True)
    s.init_fields(0, 1.0, 1.5, 0.0, 0.0, T0)
    s.set_timestep(0, 100)
    f0, g0, _ = s.download_state(0)
    s.step(0, 3)
    f1, g1, _ = s.download_state(0)
    assert np.abs(f1 - f0).max() < 1e-6 and np.abs(g1 - g0).max() < 1e-6
    st = s.stats()
    assert st["node_updates"] == 3 * n and st["n_noconv"] == 0 and st["n_repair"] == 0
    del f0, g0, f1, g1
    fl = fields(shape, 1.5)
    s.init_fields(0, *fl)
    f0, g0, _ = s.download_state(0)
    m0 = f0.sum()
    del f0, g0
    s.step(0, 5)
    f1, g1, _ = s.download_state(0)
    assert abs(f1.sum() - m0) / m0 < 1e-7
    assert np.all(np.isfinite(f1)) and np.all(np.isfinite(g1))
