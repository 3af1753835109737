"""D3Q39 inlets and outlets with stored states in the oracle
(inlet_bc.f95:21-53, outlet_bc.f95:23-58, inlet_outlet_basis_calcs.f90):
  each table row is the equilibrium of its stored state: its moments are the
    state's rho, rho u, P_ab = rho (T d_ab + u_a u_b) and q (the closed forms
    the multiplier solve targets), g = (2 c_v - 3) T f for a single state;
  a uniform state equal to the inlet and outlet states is a fixed point of
    the step (plain lattice with rho-v inlet and outlet; shifted lattice with
    an inlet and a pure-outflow outlet);
  a type-3 inlet supplies the freestream pair."""
import numpy as np

T0 = 1 / 1.4


def moments(orc, f, shifted):
    cx, cy, cz = [np.asarray(a, float) for a in orc.velocities39(shifted)]
    rho = f.sum()
    u = np.array([cx @ f, cy @ f, cz @ f]) / rho
    P = np.array([[(a * b) @ f for b in (cx, cy, cz)] for a in (cx, cy, cz)])
    c2 = cx * cx + cy * cy + cz * cz
    q = np.array([(a * c2) @ f for a in (cx, cy, cz)])
    return rho, u, P, q


def test_table_rows_hold_their_states(orc):
    for shifted, mach in ((False, 0.3), (True, 1.5)):
        c = orc.consts39(mach=mach, shifted=shifted)
        inlets = [(1.02, mach * 0.9, 0.03, -0.01, T0 * 1.04, 1)]
        outlets = [(0.97, mach * 1.05, 0.0, 0.02, T0 * 0.97, 4)]
        tab = orc.bc_tables39(c, inlets, outlets)
        for row, (rho, u, v, w, T, _) in zip(tab, inlets + outlets):
            r, uu, P, q = moments(orc, row[:39], shifted)
            U = np.array([u, v, w])
            assert abs(r - rho) < 1e-12 and np.abs(uu - U).max() < 1e-12
            assert np.abs(P - rho * (T * np.eye(3) + np.outer(U, U))).max() < 1e-11
            assert np.abs(q - (rho * U @ U + 5 * rho * T) * U).max() < 1e-10
        # one state per family: g is (2 cv - 3) T f of that state
        assert np.allclose(tab[0, 39:], c.lead_g * inlets[0][4] * tab[0, :39], rtol=1e-15)
        assert np.allclose(tab[1, 39:], c.lead_g * outlets[0][4] * tab[1, :39], rtol=1e-15)
        fs = orc.bc_tables39(c, [(1, 0, 0, 0, 1, 3)])
        fsf, fsg = orc.freestream39(c)
        assert np.array_equal(fs[0, :39], fsf) and np.array_equal(fs[0, 39:], fsg)


def test_uniform_inlet_state_is_a_fixed_point(orc):
    shape = (12, 10, 8)
    cases = [(False, 0.3, [-101, -11, -11, -201, -11, -11], [(1.0, 0.3, 0.0, 0.0, T0, 1)],
              [(1.0, 0.3, 0.0, 0.0, T0, 1)]),
             (True, 1.5, [-101, -1000, -11, -201, -1000, -11], [(1.0, 1.5, 0.0, 0.0, T0, 1)],
              [(1.0, 0.0, 0.0, 0.0, T0, 6)])]
    for shifted, mach, faces, inlets, outlets in cases:
        c = orc.consts39(mach=mach, shifted=shifted)
        f0, g0, lm0 = orc.init39(shape, 1.0, mach, 0.0, 0.0, T0, c)
        f, g, lm, _, st = orc.step_faces39(shape, faces, f0, g0, lm0, c, 100, 10,
                                           inlets=inlets, outlets=outlets)
        ef = np.abs(f - f0).max() / f0.max()
        eg = np.abs(g - g0).max() / g0.max()
        print(f"shifted={shifted}: uniform inlet state after 10 steps: max rel change f {ef:.1e} g {eg:.1e}")
        assert ef < 1e-12 and eg < 1e-12 and st["n_repair"] == 0


def test_type3_inlet_is_the_freestream_face(orc):
    shape = (12, 10, 8)
    c = orc.consts39(mach=0.3, shifted=False)
    rng = np.random.default_rng(2)
    rho = 1 + 0.01 * rng.standard_normal(np.prod(shape))
    f0, g0, lm0 = orc.init39(shape, rho, 0.3, 0.0, 0.0, T0, c)
    a = orc.step_faces39(shape, [-101, -10, -11, -6, -10, -11], f0, g0, lm0, c, 0, 5,
                         inlets=[(1, 0, 0, 0, 1, 3)])
    b = orc.step_faces39(shape, [-1000, -10, -11, -6, -10, -11], f0, g0, lm0, c, 0, 5)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
