"""D3Q39 link tables on the device against the oracle: bit-exact int32
equality of the whole reference-form table -- the plain lattice
(state_directions.f95 with the long-link derivation) on a walled box with a
sphere and on a box with a periodic axis, and the shifted lattice
(shifted_state_directions.f95, links keyed on the upstream node) with a
sphere, with a periodic axis, and with a body cut by a face (upstream nodes
beyond the face classified by the ray cast); a plain-lattice level that
holds a table with wall links steps (its wall rule is checked against the
oracle in test_gpu_walls39p.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 2147483647


def _solver(pkg, shape, faces, mach=0.3, shifted=0):
    s = pkg.Solver(lattice="d3q39", precision="fp32s")
    s.set_gas(mach=mach, shifted=shifted)
    s.define_level(0, shape, periodic=[int(faces[d] == P) for d in range(3)])
    s.set_faces(0, faces)
    return s


@pytest.mark.parametrize("faces", [[-1000, -10, -11, -6, -10, -11], [-1000, P, -10, -6, P, -10]])
def test_d3q39_links_bit_exact(pkg, orc, faces):
    shape = (40, 30, 28)
    tri, nrm = pkg.icosphere(5, 8.3, (17.5, 15.25, 13.75), 20260102)
    st_o, ns_o, nt_o = orc.build_links39(shape, faces, tri, nrm)
    s = _solver(pkg, shape, faces)
    ns_g, nt_g = s.build_links(0, tri, nrm)
    st_g = s.download_links(0)
    print(f"faces {faces}: solid {ns_g}, ties {nt_g}", flush=True)
    assert st_g.shape == (39, shape[0] * shape[1] * shape[2])
    assert ns_g == ns_o and nt_g == nt_o
    assert np.array_equal(st_g, st_o)
    # the long-link derivation was exercised
    assert np.count_nonzero(st_o[15:21] < -1000) > 50 and np.count_nonzero(st_o[34] == 666) > 0


@pytest.mark.parametrize("faces,center", [
    ([-1000, -11, -11, -6, -11, -11], (17.5, 15.25, 13.75)),
    ([-1000, -11, P, -6, -11, P], (17.5, 15.25, 13.75)),
    ([-1000, -11, -11, -6, -11, -11], (17.5, 2.25, 13.75))])   # sphere cut by the y- face
def test_d3q39_shifted_links_bit_exact(pkg, orc, faces, center):
    shape = (40, 30, 28)
    tri, nrm = pkg.icosphere(5, 8.3, center, 20260102)
    st_o, ns_o, nt_o = orc.build_links39s(shape, faces, tri, nrm)
    s = _solver(pkg, shape, faces, mach=1.5, shifted=1)
    ns_g, nt_g = s.build_links(0, tri, nrm)
    st_g = s.download_links(0)
    print(f"shifted faces {faces} centre {center}: solid {ns_g}, wall links "
          f"{int(np.count_nonzero(st_o[1:] <= -1001))}, ties {nt_g}", flush=True)
    assert ns_g == ns_o and nt_g == nt_o
    assert np.array_equal(st_g, st_o)
    assert np.count_nonzero(st_o[1:] <= -1001) > 1000


def test_d3q39_plain_table_steps(pkg):
    shape = (20, 16, 16)
    tri, nrm = pkg.icosphere(3, 4.0, (9.5, 8.25, 7.75), 1)
    s = _solver(pkg, shape, [-1000, -10, -11, -6, -10, -11])
    s.build_links(0, tri, nrm)
    s.init_fields(0, 1.0, 0.3, 0.0, 0.0, 1 / 1.4)
    s.step(0, 5)
    st = s.stats()
    assert st["n_ibb_links"] > 0 and st["n_nonpos"] == 0
    assert np.isfinite(s.macro39(0)[0]).all()
