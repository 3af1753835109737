"""Checkpoint / restart through the C ABI: a context restored from a
checkpoint continues bit for bit as the saved one (D3Q27 duct with a walled
sphere, FP32; D3Q39 shifted lattice with a sphere inside the start-up window,
FP64, where the restored step counter drives the ramps), and a checkpoint is
refused by a context set up differently (faces, link table, lattice) or when
the file is damaged."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 2147483647
DUCT_FACES = [-101, -10, -10, -201, -10, -10]


def duct(pkg, faces=DUCT_FACES, radius=5.3):
    shape = (40, 24, 24)
    s = pkg.Solver(precision="fp32s", nu=0.01)
    s.define_level(0, shape, periodic=(0, 0, 0))
    s.set_faces(0, faces)
    s.set_bc_states(inlets=[(1.0, 0.05, 0.0, 0.0)], outlets=[(1.0, 0.05, 0.0, 0.0)])
    tri, nrm = pkg.icosphere(3, radius, (14.5, 11.75, 12.25), 20260102)
    s.build_links(0, tri, nrm)
    s.init_uniform(0, 1.0, (0.05, 0.0, 0.0))
    return s


def wedge39(pkg, prec="fp64"):
    shape = (32, 24, 16)
    s = pkg.Solver(lattice="d3q39", precision=prec)
    s.set_gas(mach=1.5, shifted=1, turbulent=True)
    s.define_level(0, shape, periodic=(0, 0, 0))
    s.set_faces(0, [-1000, -1000, -11, -6, -1000, -11])
    tri, nrm = pkg.icosphere(4, 4.3, (12.5, 11.75, 7.5), 20260103)
    s.build_links(0, tri, nrm)
    s.init_fields(0, 1.0, 1.5, 0.0, 0.0, 1 / 1.4)
    s.set_timestep(0, 0)
    return s


def test_d3q27_restart_bit_exact(pkg, tmp_path):
    path = str(tmp_path / "duct.ckp")
    a = duct(pkg)
    a.step(0, 15)
    a.checkpoint_write(path)
    a.step(0, 10)
    b = duct(pkg)
    b.step(0, 3)                      # state to be overwritten
    b.checkpoint_read(path)
    b.step(0, 10)
    fa, fb = a.download_f(0), b.download_f(0)
    print(f"duct restart: max |diff| {np.abs(fa - fb).max():.1e}, file {os.path.getsize(path)} B")
    assert np.array_equal(fa, fb)


def test_d3q39_restart_bit_exact_in_startup(pkg, tmp_path):
    path = str(tmp_path / "w39.ckp")
    a = wedge39(pkg)
    a.step(0, 12)
    a.checkpoint_write(path)
    a.step(0, 10)
    b = wedge39(pkg)
    b.checkpoint_read(path)
    b.step(0, 10)
    fa, ga, la = a.download_state(0)
    fb, gb, lb = b.download_state(0)
    assert np.array_equal(fa, fb) and np.array_equal(ga, gb) and np.array_equal(la, lb)
    # without the restored counter the start-up ramps would differ
    c = wedge39(pkg)
    c.checkpoint_read(path)
    c.set_timestep(0, 0)
    c.step(0, 10)
    assert not np.array_equal(c.download_state(0)[0], fa)


def test_checkpoint_refusals(pkg, tmp_path):
    path = str(tmp_path / "duct.ckp")
    a = duct(pkg)
    a.step(0, 2)
    a.checkpoint_write(path)
    with pytest.raises(pkg.UclbsError, match="set-up"):
        duct(pkg, faces=[-101, -10, -1000, -201, -10, -10]).checkpoint_read(path)
    with pytest.raises(pkg.UclbsError, match="set-up"):
        duct(pkg, radius=5.0).checkpoint_read(path)      # another link table
    with pytest.raises(pkg.UclbsError, match="lattice"):
        wedge39(pkg).checkpoint_read(path)
    raw = bytearray(open(path, "rb").read())
    raw[len(raw) // 2] ^= 0x40
    bad = str(tmp_path / "bad.ckp")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(pkg.UclbsError, match="checksum"):
        duct(pkg).checkpoint_read(bad)
    with pytest.raises(pkg.UclbsError, match="not a checkpoint"):
        open(bad, "wb").write(b"x" * 64)
        duct(pkg).checkpoint_read(bad)
