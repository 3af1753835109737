"""D3Q39 inlets and outlets with stored states on the GPU (through the C ABI,
uclbs_set_bc_states39) against the FP64 oracle on the same inputs, 30 steps
from the start-up: f, g relative L2 <= 1e-11, multipliers max abs <= 1e-8,
equal repair counts; FP32 storage: rho, u, T relative L2 <= 1e-6.  Cases: rho-v
and v-temperature inlets with rho-v and pure-outflow outlets (plain lattice,
Mach 0.3); a freestream (type 3) inlet with a pressure-temperature outlet that
reads inlet 1's table row; an inlet and a pure-outflow outlet on the shifted
lattice (Mach 1.5)."""
import numpy as np
import pytest

from test_gpu_d3q39_faces import fields, rel_l2

pytestmark = pytest.mark.gpu
T0 = 1 / 1.4

CASES = [
    ("rho-v/v-T inlets", False, 0.3, [-102, -10, -11, -201, -10, -11],
     [(1.0, 0.3, 0.0, 0.0, T0, 1), (1.02, 0.25, 0.01, 0.0, T0 * 1.05, 4)],
     [(0.97, 0.3, 0.0, 0.0, T0 * 0.98, 1), (1.0, 0.0, 0.0, 0.0, T0, 6)]),
    ("freestream inlet, p-T outlet", False, 0.3, [-101, -10, -11, -201, -10, -11],
     [(1.0, 0.0, 0.0, 0.0, 1.0, 3)], [(1.0, 0.3, 0.0, 0.0, T0, 2)]),
    ("shifted inlet, outflow", True, 1.5, [-101, -1000, -11, -202, -1000, -11],
     [(1.01, 1.45, 0.02, 0.0, T0 * 1.02, 1)],
     [(1.0, 1.5, 0.0, 0.0, T0, 1), (1.0, 0.0, 0.0, 0.0, T0, 6)]),
]


def solver(pkg, shape, prec, mach, shifted, faces, inlets, outlets):
    s = pkg.Solver(lattice="d3q39", precision=prec)
    s.set_gas(mach=mach, shifted=int(shifted), turbulent=True)
    s.define_level(0, shape, periodic=(0, 0, 0))
    s.set_faces(0, faces)
    s.set_bc_states39(inlets=inlets, outlets=outlets)
    return s


@pytest.mark.parametrize("name,shifted,mach,faces,inlets,outlets", CASES)
def test_inlets_fp64_parity(pkg, orc, name, shifted, mach, faces, inlets, outlets):
    shape = (12, 10, 8)
    c = orc.consts39(mach=mach, shifted=shifted)
    f0, g0, lm0 = orc.init39(shape, *fields(shape, mach), c)
    print(name, flush=True)
    f_o, g_o, lm_o, _, st_o = orc.step_faces39(shape, faces, f0, g0, lm0, c, 0, 30,
                                               inlets=inlets, outlets=outlets)
    s = solver(pkg, shape, "fp64", mach, shifted, faces, inlets, outlets)
    s.upload_state(0, f0, g0, lm0)
    s.step(0, 30)
    f_g, g_g, lm_g = s.download_state(0)
    st = s.stats()
    e_f, e_g = rel_l2(f_g, f_o), rel_l2(g_g, g_o)
    e_lm = float(np.abs(lm_g - lm_o).max())
    print(f"{name}: relL2 f {e_f:.2e} g {e_g:.2e} max|lm| {e_lm:.2e} repairs {st['n_repair']} "
          f"vs {st_o['n_repair']}")
    assert rel_l2(f_o, f0) > 1e-4
    assert e_f < 1e-11 and e_g < 1e-11 and e_lm < 1e-8
    assert st["n_repair"] == st_o["n_repair"] and st["n_noconv"] == 0


def test_inlets_fp32_against_fp64_oracle(pkg, orc):
    name, shifted, mach, faces, inlets, outlets = CASES[0]
    shape = (12, 10, 8)
    c = orc.consts39(mach=mach, shifted=shifted)
    f0, g0, lm0 = orc.init39(shape, *fields(shape, mach), c)
    f_o, g_o, _, _, _ = orc.step_faces39(shape, faces, f0, g0, lm0, c, 0, 30,
                                         inlets=inlets, outlets=outlets)
    s = solver(pkg, shape, "fp32s", mach, shifted, faces, inlets, outlets)
    s.upload_state(0, f0, g0, lm0)
    s.step(0, 30)
    f_g, g_g, _ = s.download_state(0)
    r_o, ux_o, uy_o, uz_o, T_o = orc.macro39(f_o, g_o, c)
    r_g, ux_g, uy_g, uz_g, T_g = orc.macro39(f_g, g_g, c)
    e = [rel_l2(r_g, r_o), rel_l2(np.stack([ux_g, uy_g, uz_g]), np.stack([ux_o, uy_o, uz_o])),
         rel_l2(T_g, T_o)]
    print("fp32 inlets: relL2 rho {:.2e} u {:.2e} T {:.2e}".format(*e))
    assert max(e) < 1e-6


def test_inlet_without_state_is_refused(pkg):
    s = solver(pkg, (12, 10, 8), "fp64", 0.3, False, [-103, -10, -11, -201, -10, -11],
               [(1.0, 0.3, 0.0, 0.0, T0, 1)], [(1.0, 0.3, 0.0, 0.0, T0, 1)])
    s.init_fields(0, 1.0, 0.3, 0.0, 0.0, T0)
    with pytest.raises(pkg.UclbsError, match="inlet 3"):
        s.step(0, 1)
