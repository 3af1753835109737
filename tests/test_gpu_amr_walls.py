"""A refined D3Q27 level around a solid body, on the GPU through the C ABI,
against the oracle's two-level schedule: the coarse duct (inlet, outlet,
bounce-back walls) and the refined box both carry device-built link tables
of the same sphere (the refined table at the level's spacing and origin,
bit-exact against the oracle's builder); per coarse step the coarse level
steps, the refined level takes two substeps with its ghost shell filled from
the coarse level at theta = 0 and 1/2 (oracle/o_amr.py) and is injected back.
FP64 storage, 6 coarse steps: coarse fluid nodes and refined valid nodes
relative L2 <= 1e-12 against the oracle (oracle/o_box27.c on each level);
wall links rebuilt on both levels (counts equal)."""
import numpy as np
import pytest

import o_amr

pytestmark = pytest.mark.gpu

COARSE = (24, 16, 16)
FACES = [-101, -10, -10, -201, -10, -10]
FLO, FHI = (14, 6, 6), (35, 25, 25)
U, NU = 0.04, 0.02


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_two_level_walls_match_oracle(pkg, orc):
    fvalid = tuple(FHI[d] - FLO[d] + 1 for d in range(3))
    fstored = tuple(v + 2 for v in fvalid)
    forigin = tuple(FLO[d] - 1 for d in range(3))
    fphys = tuple(0.5 * v for v in forigin)
    tri, nrm = pkg.icosphere(3, 3.1, (12.25, 7.75, 8.1), 20260102)
    st_c, ns_c, _ = orc.build_links27(COARSE, FACES, tri, nrm)
    st_f, ns_f, _ = orc.build_links27(fstored, [-10] * 6, tri, nrm, dx=0.5, origin=fphys)
    s = pkg.Solver(precision="fp64", nu=NU, alpha_tol=1e-12)
    s.define_level(0, COARSE, periodic=(0, 0, 0))
    s.set_faces(0, FACES)
    s.set_bc_states(inlets=[(1.0, U, 0.0, 0.0)], outlets=[(1.0, U, 0.0, 0.0)])
    assert s.build_links(0, tri, nrm)[0] == ns_c
    s.define_level(1, fvalid, periodic=(0, 0, 0), lo=FLO)
    assert s.build_links(1, tri, nrm, origin=fphys, dx=0.5)[0] == ns_f
    assert np.array_equal(s.download_links(0), st_c)
    assert np.array_equal(s.download_links(1), st_f)
    print(f"solid coarse {ns_c}, fine {ns_f}; wall links coarse "
          f"{int(np.count_nonzero(st_c[1:] <= -1001))}, fine {int(np.count_nonzero(st_f[1:] <= -1001))}",
          flush=True)
    fc = orc.uniform27(st_c, 1.0, (U, 0.0, 0.0))
    s.upload_f(0, fc)
    s.init_from_coarse(1)
    ff = o_amr.coarse_to_fine(fc, fc, COARSE, (0, 0, 0), fstored, forigin, 1.0)
    assert np.abs(s.download_f(1) - ff).max() < 1e-14
    ncoarse = 6
    s.advance(ncoarse)
    b0, b1 = 1 / (6 * NU + 1), 1 / (12 * NU + 1)
    gh = o_amr.ghost_nodes(fstored)
    ib = [0, 0]
    for _ in range(ncoarse):
        fp = fc
        fc, _, _, bc = orc.step_box27(COARSE, FACES, st_c, fc, b0, inlets=[(1.0, U, 0.0, 0.0)],
                                      outlets=[(1.0, U, 0.0, 0.0)], tol=1e-12)
        ib[0] += bc["n_ibb_links"]
        for theta in (0.0, 0.5):
            ff[:, gh] = o_amr.coarse_to_fine(fp, fc, COARSE, (0, 0, 0), fstored, forigin, theta,
                                             nodes=gh)
            ff, _, _, bf = orc.step_box27(fstored, [-10] * 6, st_f, ff, b1, tol=1e-12)
        fc = o_amr.fine_to_coarse(ff, fstored, forigin, FLO, FHI, fc, COARSE, (0, 0, 0))
    gc, gf = s.download_f(0), s.download_f(1)
    flc = st_c[0] >= 0
    valid = np.ones(fstored[::-1], bool)
    valid[[0, -1], :, :] = valid[:, [0, -1], :] = valid[:, :, [0, -1]] = False
    vf = valid.ravel() & (st_f[0] >= 0)
    e_c, e_f = rel_l2(gc[:, flc], fc[:, flc]), rel_l2(gf[:, vf], ff[:, vf])
    u0 = orc.uniform27(st_c, 1.0, (U, 0.0, 0.0))
    print(f"two levels with walls: relL2 coarse {e_c:.2e} fine {e_f:.2e}; moved "
          f"{rel_l2(fc[:, flc], u0[:, flc]):.2e}; wall links {s.stats()['n_ibb_links']}")
    assert rel_l2(fc[:, flc], u0[:, flc]) > 1e-4
    assert e_c < 1e-12 and e_f < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32s"])
def test_refined_boxes_with_walls_match_one_box(pkg, orc, prec):
    """The refined region split into uneven boxes (the sphere crossing box
    faces), its link table built box by box, equals the same region as one
    box bit for bit after 6 coarse steps: coarse level and refined valid
    nodes, wall-link counts.  The alpha tolerance is 1e-12 in both
    precisions: the one-box and many-box kernels run the same arithmetic,
    but the compiler may round its intermediate products differently in
    each, which the early exits of a looser tolerance (1e-6 in FP32) pass
    on at the 1e-8 level of the FP32 storage."""
    from test_gpu_boxes import _tiling
    fvalid = tuple(FHI[d] - FLO[d] + 1 for d in range(3))
    fphys = tuple(0.5 * (FLO[d] - 1) for d in range(3))
    tri, nrm = pkg.icosphere(3, 3.1, (12.25, 7.75, 8.1), 20260102)
    boxes = _tiling(FLO, FHI, (5, 7, 9), 11)
    runs = []
    for many in (False, True):
        s = pkg.Solver(precision=prec, nu=NU, alpha_tol=1e-12)
        s.define_level(0, COARSE, periodic=(0, 0, 0))
        s.set_faces(0, FACES)
        s.set_bc_states(inlets=[(1.0, U, 0.0, 0.0)], outlets=[(1.0, U, 0.0, 0.0)])
        s.build_links(0, tri, nrm)
        if many:
            s.define_level_boxes(1, boxes)
            s.build_links(1, tri, nrm, origin=(0.0, 0.0, 0.0), dx=0.5)
        else:
            s.define_level(1, fvalid, periodic=(0, 0, 0), lo=FLO)
            s.build_links(1, tri, nrm, origin=fphys, dx=0.5)
        s.init_uniform(0, 1.0, (U, 0.0, 0.0))
        s.init_from_coarse(1)
        s.advance(6)
        fc, ff = s.download_f(0), s.download_f(1)
        if many:
            fine = s.stored_to_dense(1, ff, fvalid[::-1], origin=FLO)
        else:
            fine = ff.reshape(27, fvalid[2] + 2, fvalid[1] + 2, fvalid[0] + 2)[:, 1:-1, 1:-1, 1:-1]
        runs.append((fc, fine, s.stats()["n_ibb_links"]))
        s.close()
    print(f"{prec}: {len(boxes)} boxes; wall links {runs[0][2]} vs {runs[1][2]}; max|d| coarse "
          f"{np.abs(runs[0][0] - runs[1][0]).max():.1e} fine {np.abs(runs[0][1] - runs[1][1]).max():.1e}")
    assert len(boxes) > 8 and runs[0][2] == runs[1][2] > 0
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
