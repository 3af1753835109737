"""D3Q39 boxes with faces on the GPU (through the C ABI): freestream (-1000),
pure outflow (-6), bounce-back (-10) and inviscid (-11) edge links.
  FP64 storage against the FP64 oracle on the same inputs: f, g relative
    L2 <= 1e-11, multipliers max abs <= 1e-8, alpha relative L2 <= 1e-11 and
    max abs <= 1e-9 (as the
    periodic parity runs: L D L^T against LU factorisation)
  FP32 storage: rho, u, T relative L2 <= 1e-6 against the FP64 oracle
  inviscid walls == the mirrored periodic box of twice the extent, GPU against
    GPU, max relative difference <= 1e-12 (summation order of mirrored nodes)
  the supersonic face set at the 384 x 192 x 96 size: a uniform freestream
    stays put to FP32 round-off; a perturbed one stays finite and repairs
    nothing it should not."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 1.4
T0 = 1 / G
P = 2147483647
C3_FACES = [-1000, -11, -11, -6, -11, -11]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def fields(shape, mach):
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    rho = 1.0 + 0.01 * np.sin(2 * np.pi * (x / nx + 0.7 * y / ny) + 0.3)
    ux = mach + 0.02 * np.sin(2 * np.pi * x / nx + 0.1 * z)
    uy = 0.015 * np.sin(2 * np.pi * y / (2 * ny) + 0.4 * x)
    uz = 0.01 * np.cos(2 * np.pi * (x / nx + z / nz))
    T = T0 * (1 + 0.01 * np.cos(2 * np.pi * z / nz + 0.2 * y))
    return rho, ux, uy, uz, T


def solver(pkg, shape, prec, mach, shifted, faces):
    s = pkg.Solver(lattice="d3q39", precision=prec)
    s.set_gas(mach=mach, shifted=int(shifted), turbulent=True)
    per = [int(faces[d] == P) for d in range(3)]
    s.define_level(0, shape, periodic=per)
    s.set_faces(0, faces)
    s.keep_alpha(True)
    return s


CASES = [(False, 0.3, [-1000, -10, -11, -6, -10, -11]),
         (True, 1.5, C3_FACES),
         (True, 1.5, [-1000, -10, P, -6, -10, P])]


@pytest.mark.parametrize("shifted,mach,faces", CASES)
def test_faces_fp64_parity(pkg, orc, shifted, mach, faces):
    shape = (12, 10, 8)
    steps = 30
    c = orc.consts39(mach=mach, shifted=shifted)
    fl = fields(shape, mach)
    f0, g0, lm0 = orc.init39(shape, *fl, c)
    print(f"faces {faces} shifted={shifted}", flush=True)
    f_o, g_o, lm_o, a_o, st_o = orc.step_faces39(shape, faces, f0, g0, lm0, c, 0, steps,
                                                 keep_alpha=True)
    s = solver(pkg, shape, "fp64", mach, shifted, faces)
    s.upload_state(0, f0, g0, lm0)
    s.step(0, steps)
    f_g, g_g, lm_g = s.download_state(0)
    st = s.stats()
    e_f, e_g = rel_l2(f_g, f_o), rel_l2(g_g, g_o)
    e_lm = float(np.abs(lm_g - lm_o).max())
    a_g = s.alpha(0)
    e_a = float(np.abs(a_g - a_o).max())
    e_al2 = float(np.linalg.norm(a_g - a_o) / np.linalg.norm(a_o))
    print(f"relL2 f {e_f:.2e} g {e_g:.2e} max|lm| {e_lm:.2e} max|alpha| {e_a:.2e} relL2 alpha {e_al2:.2e} "
          f"repairs {st['n_repair']} vs {st_o['n_repair']}")
    assert rel_l2(f_o, f0) > 1e-4  # the faces moved the state
    assert e_f < 1e-11 and e_g < 1e-11 and e_lm < 1e-8 and e_al2 < 1e-11 and e_a < 1e-9
    assert st["node_updates"] == st_o["nodes"] and st["n_noconv"] == 0
    assert st["n_repair"] == st_o["n_repair"]


@pytest.mark.parametrize("shifted,mach,faces", CASES[:2])
def test_faces_fp32_against_fp64_oracle(pkg, orc, shifted, mach, faces):
    shape = (12, 10, 8)
    steps = 30
    c = orc.consts39(mach=mach, shifted=shifted)
    f0, g0, lm0 = orc.init39(shape, *fields(shape, mach), c)
    f_o, g_o, _, _, _ = orc.step_faces39(shape, faces, f0, g0, lm0, c, 0, steps)
    s = solver(pkg, shape, "fp32s", mach, shifted, faces)
    s.upload_state(0, f0, g0, lm0)
    s.step(0, steps)
    f_g, g_g, _ = s.download_state(0)
    r_o, ux_o, uy_o, uz_o, T_o = orc.macro39(f_o, g_o, c)
    r_g, ux_g, uy_g, uz_g, T_g = orc.macro39(f_g, g_g, c)
    e_r = rel_l2(r_g, r_o)
    e_u = rel_l2(np.stack([ux_g, uy_g, uz_g]), np.stack([ux_o, uy_o, uz_o]))
    e_T = rel_l2(T_g, T_o)
    print(f"fp32 faces shifted={shifted}: relL2 rho {e_r:.2e} u {e_u:.2e} T {e_T:.2e}")
    assert e_r < 1e-6 and e_u < 1e-6 and e_T < 1e-6


def _mirror_dirs(orc, shifted):
    cx, cy, cz = orc.velocities39(shifted)
    c = np.stack([cx, cy, cz], 1)
    m = c.copy()
    m[:, 1] = -m[:, 1]
    return np.array([int(np.nonzero((c == v).all(1))[0][0]) for v in m])


def test_inviscid_walls_equal_mirrored_periodic_box_on_gpu(pkg, orc):
    shape = (8, 6, 5)
    mach, steps = 1.5, 12
    c = orc.consts39(mach=mach, shifted=True)
    f0, g0, lm0 = orc.init39(shape, *fields(shape, mach), c)
    nx, ny, nz = shape
    mir = _mirror_dirs(orc, True)

    def dbl(a, q, perm=None, odd=()):
        a = a.reshape(q, nz, ny, nx)
        b = np.flip(a, axis=2)
        b = (b[perm] if perm is not None else b).copy()
        for j in odd:
            b[j] = -b[j]
        return np.concatenate([a, b], axis=2).reshape(q, -1)

    walled = solver(pkg, shape, "fp64", mach, True, [P, -11, P, P, -11, P])
    walled.upload_state(0, f0, g0, lm0)
    doubled = solver(pkg, (nx, 2 * ny, nz), "fp64", mach, True, [P] * 6)
    doubled.upload_state(0, dbl(f0, 39, mir), dbl(g0, 39, mir), dbl(lm0, 13, None, (2, 7, 9, 11)))
    walled.step(0, steps)
    doubled.step(0, steps)
    fw, gw, _ = walled.download_state(0)
    fd, gd, _ = doubled.download_state(0)
    fl = fd.reshape(39, nz, 2 * ny, nx)[:, :, :ny].reshape(39, -1)
    gl = gd.reshape(39, nz, 2 * ny, nx)[:, :, :ny].reshape(39, -1)
    e = max(np.abs(fl - fw).max() / np.abs(fw).max(), np.abs(gl - gw).max() / np.abs(gw).max())
    print(f"walled vs mirrored periodic: {e:.2e}")
    assert e < 1e-12 and rel_l2(fw, f0) > 1e-4


def test_face_code_errors(pkg):
    s = pkg.Solver(lattice="d3q39", precision="fp64")
    s.set_gas(mach=1.5, shifted=1)
    s.define_level(0, (8, 6, 6), periodic=(0, 0, 0))
    with pytest.raises(pkg.UclbsError):
        s.set_faces(0, [-11, -11, -11, -6, -11, -11])   # inviscid x face on the shifted lattice
    with pytest.raises(pkg.UclbsError):
        s.set_faces(0, [-300, -11, -11, -6, -11, -11])  # not a face code
    with pytest.raises(pkg.UclbsError):
        s.set_faces(0, [P, -11, -11, -6, -11, -11])     # periodic code on a walled axis
    s.set_faces(0, C3_FACES)


def test_c3_size_supersonic_faces(pkg):
    """384 x 192 x 96, shifted lattice at Mach 1.5, FP32 storage, freestream
    inflow, outflow, inviscid sides: a uniform freestream is a fixed point; a
    perturbed one stays finite through the faces."""
    shape = (384, 192, 96)
    n = shape[0] * shape[1] * shape[2]
    s = solver(pkg, shape, "fp32s", 1.5, True, C3_FACES)
    s.init_fields(0, 1.0, 1.5, 0.0, 0.0, T0)
    s.set_timestep(0, 100)
    f0, g0, _ = s.download_state(0)
    s.step(0, 3)
    f1, g1, _ = s.download_state(0)
    assert np.abs(f1 - f0).max() < 1e-6 and np.abs(g1 - g0).max() < 1e-6
    st = s.stats()
    assert st["node_updates"] == 3 * n and st["n_noconv"] == 0 and st["n_repair"] == 0
    del f0, g0, f1, g1
    s.init_fields(0, *fields(shape, 1.5))
    s.step(0, 5)
    f1, g1, _ = s.download_state(0)
    assert np.all(np.isfinite(f1)) and np.all(np.isfinite(g1))
