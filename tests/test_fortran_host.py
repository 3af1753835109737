"""Fortran host path: the iso_c_binding module and the drop-in time loop.
CPU: the drivers are built and import the C ABI entry points they bind.
GPU: the Taylor-Green C1 case and a D3Q39 box with the supersonic face set,
advanced by the Fortran loop, match the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG_DIR

DRIVER = os.path.join(PKG_DIR, "fortran", "uclbs_driver")
DRIVER39 = os.path.join(PKG_DIR, "fortran", "uclbs_driver39")


def test_fortran_driver_binds_c_abi():
    assert os.path.exists(DRIVER), "build the package (make -C <pkg>) first"
    out = subprocess.run(["nm", "-D", "--undefined-only", DRIVER], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("uclbs_create", "uclbs_define_level", "uclbs_init_tgv", "uclbs_step_level",
                "uclbs_interlevel", "uclbs_fine_to_coarse", "uclbs_init_from_coarse",
                "uclbs_download_macro", "uclbs_get_stats", "uclbs_last_error", "uclbs_destroy"):
        assert sym in out, sym


def test_fortran_d3q39_driver_binds_c_abi():
    assert os.path.exists(DRIVER39), "build the package (make -C <pkg>) first"
    out = subprocess.run(["nm", "-D", "--undefined-only", DRIVER39], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("uclbs_create", "uclbs_set_gas", "uclbs_define_level", "uclbs_set_faces",
                "uclbs_init_fields", "uclbs_download_state", "uclbs_step_level", "uclbs_destroy"):
        assert sym in out, sym


@pytest.mark.gpu
def test_fortran_loop_d3q39_supersonic_faces_matches_oracle(orc, tmp_path):
    """The D3Q39 box with the supersonic face set (Mach 1.5, shifted lattice)
    advanced by the Fortran time loop, FP64 storage, against the oracle from
    the same initial state: f, g relative L2 <= 1e-11."""
    n3, steps = (24, 16, 12), 20
    out = tmp_path / "d3q39.bin"
    r = subprocess.run([DRIVER39, str(out), *map(str, n3), str(steps), "1"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = n3[0] * n3[1] * n3[2]
    data = np.fromfile(out, dtype=np.float64)
    per = (39 + 39 + 13) * n
    assert data.size == 2 * per
    f0, g0, lm0 = data[:39 * n].reshape(39, n), data[39 * n:78 * n].reshape(39, n), \
        data[78 * n:per].reshape(13, n)
    f1, g1 = data[per:per + 39 * n].reshape(39, n), data[per + 39 * n:per + 78 * n].reshape(39, n)
    c = orc.consts39(mach=1.5, shifted=True)
    f_o, g_o, *_ = orc.step_faces39(n3, [-1000, -11, -11, -6, -11, -11], f0, g0, lm0, c, 0, steps)
    e_f = np.linalg.norm(f1 - f_o) / np.linalg.norm(f_o)
    e_g = np.linalg.norm(g1 - g_o) / np.linalg.norm(g_o)
    print(f"Fortran loop D3Q39 faces: relL2 f {e_f:.2e} g {e_g:.2e}")
    assert f"node_updates {steps * n}" in r.stdout
    assert np.linalg.norm(f_o - f0) / np.linalg.norm(f0) > 1e-5
    assert e_f < 1e-11 and e_g < 1e-11


@pytest.mark.gpu
def test_fortran_loop_c1_matches_oracle(orc, tmp_path):
    L, steps = 64, 200
    out = tmp_path / "c1.bin"
    r = subprocess.run([DRIVER, str(out), str(L), str(steps), "0"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = L ** 3
    data = np.fromfile(out, dtype=np.float64)
    assert data.size == 4 * n
    rho_g, u_g, v_g, w_g = data.reshape(4, n)
    assert f"node_updates {steps * n}" in r.stdout
    u0 = 0.1 / math.sqrt(3.0)
    nu = u0 * (L / (2 * math.pi)) / 1600.0
    f0 = orc.init_tgv27((L, L, L), u0, 1e-4, 20260101)
    f_o, _, _ = orc.step_periodic27((L, L, L), f0, 1 / (6 * nu + 1), nsteps=steps, tol=1e-10,
                                     want_alpha=False)
    rho_o, ux_o, uy_o, uz_o = orc.macro27(f_o)
    e_u = np.linalg.norm(np.stack([u_g, v_g, w_g]) - np.stack([ux_o, uy_o, uz_o])) / \
        np.linalg.norm(np.stack([ux_o, uy_o, uz_o]))
    e_r = np.linalg.norm(rho_g - rho_o) / np.linalg.norm(rho_o)
    print(f"Fortran loop C1: relL2 u {e_u:.2e}, rho {e_r:.2e}")
    assert e_u < 1e-5 and e_r < 1e-7


@pytest.mark.gpu
def test_fortran_two_level_loop_matches_python_schedule(pkg, tmp_path):
    """The Fortran recursive level advance (amr_advance) and the Python
    Solver.advance issue the same library calls: identical coarse fields."""
    L, steps = 32, 12
    out = tmp_path / "amr.bin"
    r = subprocess.run([DRIVER, str(out), str(L), str(steps), "0", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    data = np.fromfile(out, dtype=np.float64).reshape(4, L ** 3)
    u0 = 0.1 / math.sqrt(3.0)
    nu = u0 * (L / (2 * math.pi)) / 1600.0
    s = pkg.Solver(precision="fp32s", nu=nu, alpha_tol=1e-10)
    s.define_level(0, (L, L, L))
    s.init_tgv(0, u0, 1e-4, 20260101)
    s.define_level(1, (L + 1, L + 1, L + 1), periodic=(0, 0, 0), lo=(L // 2, L // 2, L // 2))
    s.init_from_coarse(1)
    s.advance(steps)
    ref = np.stack(s.macro(0))
    assert np.abs(data - ref).max() < 1e-12
    assert np.all(np.isfinite(data))


@pytest.mark.gpu
def test_fortran_driver_writes_plot3d_and_microphone(orc, tmp_path):
    """The Fortran host writes the final D3Q39 state through
    uclbs_plot3d_write / uclbs_mic_*: the grid file equals the oracle's byte
    for byte (domain placed at [0, n dx], dx = 1e-3, no link table: IBLANK 1),
    the solution block's rho, u, v, w, T are within 1e-12 relative of the
    oracle's flow variables of the state the driver wrote, and the
    microphone file has its header and one sample line."""
    import o_output as oo
    n3, steps = (24, 16, 12), 6
    out = tmp_path / "d3q39.bin"
    r = subprocess.run([DRIVER39, str(out), *map(str, n3), str(steps), "1", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = n3[0] * n3[1] * n3[2]
    data = np.fromfile(out, dtype=np.float64)
    per = (39 + 39 + 13) * n
    f1, g1 = data[per:per + 39 * n].reshape(39, n), data[per + 39 * n:per + 78 * n].reshape(39, n)
    c = orc.consts39(mach=1.5, shifted=True)
    m = np.stack(list(orc.macro39(f1, g1, c)) + [np.full(n, 2.0)])
    plo, dx, dlo = oo.level_geom((0.0, 0.0, 0.0), [v * 1e-3 for v in n3], n3, (0, 0, 0), 0)
    blocks = [dict(dims=n3, start=oo.block_start(plo, dx, dlo, (0, 0, 0)), dx=dx,
                   blank=np.ones(n, np.int32), fields=m)]
    ts = f"{steps:09d}"
    assert (tmp_path / f"uclbs_grid_{ts}.x").read_bytes() == oo.grid_file(blocks)
    recs = oo.read_records((tmp_path / f"uclbs_data_{ts}.q").read_bytes())
    h = dict(mach=1.5, reynolds=0.0, time=float(steps), gamma=1.4, t_ref=0.0, stag_temperature=0.0,
             cv=0.0, cp=0.0, gas_r=0.0, timestep=steps)
    assert len(recs) == 4 and recs[2] == oo.header_record(h)
    got = np.frombuffer(recs[3], np.float64).reshape(6, n)
    for k in range(5):
        assert np.abs(got[k] - m[k]).max() <= 1e-12 * np.abs(m[k]).max(), k
    assert np.all(got[5] == 2.0)
    lines = (tmp_path / "microphone00001.txt").read_text().splitlines()
    assert lines[0] == oo.mic_header(63).rstrip("\n") and len(lines) == 2
    assert len(lines[1]) == 6 * 13 - 1
