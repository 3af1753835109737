"""The captured multi-level advance (uclbs_advance with use_graph) after a
change of what its kernels read: new inlet/outlet tables
(uclbs_set_bc_states frees and reallocates them) and new face codes
(uclbs_set_faces; on D3Q39 the near-face node list is rebuilt).  The graph
must be captured again, so graph replays and eager launches give the same
fields bit for bit before and after each change."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _duct(pkg, L=32):
    u_in = 0.05
    s = pkg.Solver(precision="fp32s", nu=0.01)
    s.define_level(0, (L, L, L), periodic=(0, 0, 0))
    s.set_faces(0, [-101, -10, -10, -201, -10, -10])
    s.set_bc_states(inlets=[(1.0, u_in, 0.0, 0.0)], outlets=[(1.0, u_in, 0.0, 0.0)])
    tri, nrm = pkg.icosphere(3, 5.0, (10.5, 16.5, 16.5), 7)
    s.build_links(0, tri, nrm)
    s.init_uniform(0, 1.0, (u_in, 0.0, 0.0))
    return s


def _same(a, b, lvl=0):
    return np.array_equal(a.download_f(lvl), b.download_f(lvl))


def test_d3q27_graph_recaptured_after_bc_and_face_change(pkg):
    g, e = _duct(pkg), _duct(pkg)
    g.advance_native(4, graph=True)
    e.advance_native(4, graph=False)
    assert _same(g, e)
    # new inlet velocity and a second outlet state: the tables are reallocated
    for s in (g, e):
        s.set_bc_states(inlets=[(1.0, 0.08, 0.0, 0.0)],
                        outlets=[(1.0, 0.08, 0.0, 0.0), (1.0, 0.02, 0.0, 0.0)])
    g.advance_native(4, graph=True)
    e.advance_native(4, graph=False)
    assert _same(g, e)
    # the x+ face switches to outlet 2 (read by the face-coded nodes)
    for s in (g, e):
        s.set_faces(0, [-101, -10, -10, -202, -10, -10])
    g.advance_native(4, graph=True)
    e.advance_native(4, graph=False)
    assert _same(g, e)
    ref = _duct(pkg)
    ref.advance_native(12, graph=False)
    assert not _same(ref, e)  # the changes did change the flow


def test_d3q39_graph_recaptured_after_face_change(pkg):
    shape = (16, 12, 12)

    def make():
        s = pkg.Solver(lattice="d3q39", precision="fp32s")
        s.set_gas(mach=1.5, shifted=1)
        s.define_level(0, shape, periodic=(0, 0, 0))
        s.set_faces(0, [-1000, -11, -11, -6, -11, -11])
        nx, ny, nz = shape
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        s.init_fields(0, 1.0 + 0.01 * np.sin(2 * np.pi * x.ravel() / nx), 1.5, 0.0,
                      0.01 * np.sin(2 * np.pi * y.ravel() / ny), 1 / 1.4)
        s.set_timestep(0, 100)
        return s
    g, e = make(), make()
    g.advance_native(2, graph=True)
    e.advance_native(2, graph=False)
    assert all(np.array_equal(a, b) for a, b in zip(g.download_state(0), e.download_state(0)))
    for s in (g, e):
        s.set_faces(0, [-1000, -10, -10, -6, -10, -10])
    g.advance_native(2, graph=True)
    e.advance_native(2, graph=False)
    assert all(np.array_equal(a, b) for a, b in zip(g.download_state(0), e.download_state(0)))
