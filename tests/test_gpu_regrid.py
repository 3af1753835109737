"""Regrid on the GPU through the C ABI: the GGM indicator, remaking a refined
level on new boxes, and the tag -> cluster -> remake -> re-identify loop.
  GGM (uclbs_tag_ggm) against the oracle (oracle/o_ggm.py) on a duct with a
    sphere after 12 steps, link table from the device builder: every scheme
    and quantity 1, 2, 7; FP64 ggm max abs <= 1e-12 x max ggm, FP32 storage
    <= 1e-9 x max ggm; tags equal wherever ggm is not within 1e-9 of the
    limit
  remake (uclbs_regrid_level), FP64: a refined level of uneven boxes moved
    and re-tiled keeps the old populations bit for bit at every new node
    that was a valid node before, and every other new node equals a fresh
    uclbs_init_from_coarse of the same boxes; stepping before its link table
    is built again fails; a finer level is stale until regridded (its valid
    nodes then kept bit for bit) and clear_level drops it
  the loop (Solver.refine_by_ggm), FP32 storage: level 1 created around the
    sphere from level 0's tags, links built, stepped; regridded twice as the
    flow develops, each time covering every tag, taking over the old nodes,
    and stepping on with finite populations
  a refined level of many boxes: its GGM (dense over the boxes' stored
    nodes, ghost shells filled from level 0 first) equals the same region as
    one box bit for bit and the oracle's restatement to 1e-12 x max ggm;
    level 2 made and regridded from level 1's tags (three levels)."""
import numpy as np
import pytest

import o_ggm

pytestmark = pytest.mark.gpu

COARSE = (24, 16, 16)
FACES = [-101, -10, -10, -201, -10, -10]
FLO, FHI = (14, 6, 6), (35, 25, 25)
U, NU = 0.04, 0.02


def sphere(pkg):
    return pkg.icosphere(3, 3.1, (12.25, 7.75, 8.1), 20260102)


def duct(pkg, prec, shape=COARSE):
    s = pkg.Solver(precision=prec, nu=NU, alpha_tol=1e-12)
    s.define_level(0, shape, periodic=(0, 0, 0))
    s.set_faces(0, FACES)
    s.set_bc_states(inlets=[(1.0, U, 0.0, 0.0)], outlets=[(1.0, U, 0.0, 0.0)])
    return s


@pytest.mark.parametrize("prec", ["fp64", "fp32s"])
def test_ggm_matches_oracle(pkg, orc, prec):
    tri, nrm = sphere(pkg)
    s = duct(pkg, prec)
    s.build_links(0, tri, nrm)
    s.init_uniform(0, 1.0, (U, 0.0, 0.0))
    s.step(0, 12)
    f = s.download_f(0)
    st = s.download_links(0).astype(np.int64)
    cx, cy, cz, _, _ = orc.lattice27()
    worst = 0.0
    for scheme in (0, 1, 2):
        for which in (1, 2, 7):
            _, ggm_o = o_ggm.ggm27(f, COARSE, cx, cy, cz, which, scheme, state=st)
            limit = float(np.quantile(ggm_o[ggm_o > 0], 0.9))
            ggm_g, tags_g, cnt = s.tag_ggm(0, which, scheme, limit)
            scale = float(np.abs(ggm_o).max())
            err = float(np.abs(ggm_g - ggm_o).max()) / scale
            worst = max(worst, err)
            assert err < (1e-12 if prec == "fp64" else 1e-9), (scheme, which, err)
            tags_o = o_ggm.tags(ggm_o, limit)
            clear = np.abs(ggm_o - limit) > 1e-9 * scale
            assert np.array_equal(tags_g.astype(bool)[clear], tags_o[clear])
            assert cnt == int(tags_g.sum()) and cnt > 0
    print(f"{prec}: GGM vs oracle, worst max|d| / max ggm {worst:.2e}")


def stored_positions(s, lvl):
    """Level position (x, y, z) of every stored node of a level of many boxes."""
    out = []
    for off, e, lo in s.box_layout(lvl):
        z, y, x = np.mgrid[0:e[2], 0:e[1], 0:e[0]]
        out.append(np.stack([x.ravel() + lo[0] - 1, y.ravel() + lo[1] - 1, z.ravel() + lo[2] - 1], 1))
    return np.concatenate(out)


def valid_lookup(boxes, pos, layout):
    """Stored index of the valid node at each position (-1 where none)."""
    idx = np.full(len(pos), -1, np.int64)
    for (lo, hi), (off, e, _) in zip(boxes, layout):
        inside = np.all((pos >= lo) & (pos <= hi), axis=1)
        loc = pos[inside] - (np.array(lo) - 1)
        idx[inside] = off + loc[:, 0] + e[0] * (loc[:, 1] + e[1] * loc[:, 2])
    return idx


def test_remake_keeps_old_nodes_and_interpolates_new(pkg):
    from test_gpu_boxes import _tiling
    tri, nrm = sphere(pkg)
    s = duct(pkg, "fp64")
    s.build_links(0, tri, nrm)
    old_boxes = _tiling(FLO, FHI, (5, 7, 9), 11)
    s.define_level_boxes(1, old_boxes)
    s.build_links(1, tri, nrm, origin=(0.0, 0.0, 0.0), dx=0.5)
    s.init_uniform(0, 1.0, (U, 0.0, 0.0))
    s.init_from_coarse(1)
    s.advance(4)
    f_old = s.download_f(1)
    old_layout = s.box_layout(1)
    f0 = s.download_f(0)
    new_boxes = _tiling((18, 4, 6), (39, 23, 27), (6, 8, 10), 5)
    copied = s.regrid_level(1, new_boxes)
    f_new = s.download_f(1)
    pos = stored_positions(s, 1)
    src = valid_lookup(old_boxes, pos, old_layout)
    kept = src >= 0
    assert copied == int(kept.sum()) > 1000 and (~kept).sum() > 1000
    assert np.array_equal(f_new[:, kept], f_old[:, src[kept]])
    ref = pkg.Solver(precision="fp64", nu=NU, alpha_tol=1e-12)
    ref.define_level(0, COARSE, periodic=(0, 0, 0))
    ref.upload_f(0, f0)
    ref.define_level_boxes(1, new_boxes)
    ref.init_from_coarse(1)
    f_ref = ref.download_f(1)
    ref.close()
    assert np.array_equal(f_new[:, ~kept], f_ref[:, ~kept])
    with pytest.raises(pkg.UclbsError, match="build or set its links"):
        s.step(1, 1)
    s.build_links(1, tri, nrm, origin=(0.0, 0.0, 0.0), dx=0.5)
    s.advance(3)
    f1 = s.download_f(1)
    print(f"remake: {len(old_boxes)} -> {len(new_boxes)} boxes, {copied} nodes kept, "
          f"{int((~kept).sum())} interpolated; wall links {s.stats()['n_ibb_links']}")
    assert np.isfinite(f1).all() and s.stats()["n_ibb_links"] > 0
    s.close()


def test_finer_level_is_stale_until_regridded(pkg):
    s = pkg.Solver(precision="fp64", nu=NU, alpha_tol=1e-12)
    s.define_level(0, (24, 20, 20), periodic=(1, 1, 1))
    s.init_tgv(0, 0.05, 1e-3, 3)
    l1 = [((8, 8, 8), (23, 31, 31)), ((24, 8, 8), (39, 31, 31))]
    l2 = [((20, 20, 20), (43, 43, 43))]
    s.define_level_boxes(1, l1)
    s.init_from_coarse(1)
    s.define_level_boxes(2, l2)
    s.init_from_coarse(2)
    s.advance(2)
    f2 = s.download_f(2)
    s.regrid_level(1, [((6, 8, 8), (41, 33, 31))])
    with pytest.raises(pkg.UclbsError, match="regridded"):
        s.advance(1)
    copied = s.regrid_level(2, l2)
    lay = s.box_layout(2)
    pos = stored_positions(s, 2)
    src = valid_lookup(l2, pos, lay)
    kept = src >= 0
    assert copied == int(kept.sum()) == 24 ** 3
    assert np.array_equal(s.download_f(2)[:, kept], f2[:, src[kept]])
    s.advance(2)
    assert np.isfinite(s.download_f(2)).all()
    s.clear_level(2)
    assert 2 not in s.shapes
    s.advance(1)
    assert np.isfinite(s.download_f(1)).all()
    s.close()


def test_tag_cluster_remake_loop(pkg):
    tri, nrm = sphere(pkg)
    s = duct(pkg, "fp32s")
    s.build_links(0, tri, nrm)
    s.init_uniform(0, 1.0, (U, 0.0, 0.0))
    s.step(0, 20)
    history = []
    for rnd in range(3):
        ggm, tags, cnt = s.tag_ggm(0, 7, 0, 0.0)
        limit = float(np.quantile(ggm[ggm > 0], 0.95))
        had = 1 in s.shapes
        boxes = s.refine_by_ggm(1, quantity=7, scheme=0, limit=limit, efficiency=0.7, blocking=2,
                                buffer=1, max_size=16)
        _, tags, cnt = s.tag_ggm(0, 7, 0, limit)
        assert boxes and cnt > 0
        # every tagged coarse node's fine image lies in a new box
        t = tags.reshape(COARSE[::-1])
        zz, yy, xx = np.nonzero(t[1:-1, 1:-1, 1:-1])
        pts = np.stack([xx + 1, yy + 1, zz + 1], 1) * 2
        inside = np.zeros(len(pts), bool)
        for lo, hi in boxes:
            inside |= np.all((pts >= lo) & (pts <= hi), axis=1)
        assert inside.all()
        copied = None
        if had:
            n = ctypes_copied(s)
            copied = n
            assert n > 0
        s.build_links(1, tri, nrm, origin=(0.0, 0.0, 0.0), dx=0.5)
        s.advance(5)
        f1 = s.download_f(1)
        assert np.isfinite(f1).all()
        history.append((len(boxes), int(cnt), copied))
    print("tag -> cluster -> remake rounds (boxes, tags, nodes kept):", history)
    s.close()


def ctypes_copied(s):
    import ctypes
    n = ctypes.c_int64()
    s._check(s.lib.uclbs_regrid_copied(s.ctx, ctypes.byref(n)))
    return n.value


def test_ggm_on_many_boxes_equals_one_box_and_oracle(pkg, orc):
    """The refined region around the sphere as one box and as 26 uneven boxes
    (FP64, alpha tol 1e-12: the two runs hold the same bits): after 4 coarse
    steps the GGM of level 1 (ghost shells filled from level 0 first) is the
    same on both, bit for bit, and equals the oracle's dense restatement of
    the many-box level (oracle/o_ggm.py boxes_dense) and its one-box form to
    1e-12 x max ggm."""
    from test_gpu_boxes import _tiling
    fvalid = tuple(FHI[d] - FLO[d] + 1 for d in range(3))
    fphys = tuple(0.5 * (FLO[d] - 1) for d in range(3))
    tri, nrm = sphere(pkg)
    boxes = _tiling(FLO, FHI, (5, 7, 9), 11)
    cx, cy, cz, _, _ = orc.lattice27()
    runs = []
    for many in (False, True):
        s = duct(pkg, "fp64")
        s.build_links(0, tri, nrm)
        if many:
            s.define_level_boxes(1, boxes)
            s.build_links(1, tri, nrm, origin=(0.0, 0.0, 0.0), dx=0.5)
        else:
            s.define_level(1, fvalid, periodic=(0, 0, 0), lo=FLO)
            s.build_links(1, tri, nrm, origin=fphys, dx=0.5)
        s.init_uniform(0, 1.0, (U, 0.0, 0.0))
        s.init_from_coarse(1)
        s.advance(4)
        ggm, tags, cnt = s.tag_ggm(1, 7, 0, 1e-5)
        f1, st1 = s.download_f(1), s.download_links(1).astype(np.int64)
        if many:
            ref, valid = o_ggm.boxes_dense(f1, st1, boxes, cx, cy, cz, 7, 0)
            dense = s.stored_to_dense(1, ggm, fvalid[::-1], origin=FLO)
        else:
            fs = tuple(v + 2 for v in fvalid)
            _, ref = o_ggm.ggm27(f1, fs, cx, cy, cz, 7, 0, state=st1)
            ref = ref.reshape(fs[::-1])[1:-1, 1:-1, 1:-1].ravel()
            ggm = ggm.reshape(fs[::-1])[1:-1, 1:-1, 1:-1].ravel()
            dense = ggm.reshape(fvalid[::-1])
            valid = np.ones(ggm.size, bool)
        scale = float(np.abs(ref).max())
        err = float(np.abs(ggm[valid] - ref[valid]).max()) / scale if many else \
            float(np.abs(ggm - ref).max()) / scale
        print(f"{'many boxes' if many else 'one box'}: GGM vs oracle max|d| / max {err:.2e}, "
              f"{cnt} tags")
        assert err < 1e-12 and cnt > 0
        runs.append(dense)
        s.close()
    assert np.array_equal(runs[0], runs[1])


def test_three_level_tag_cluster_remake(pkg):
    """Level 1 from level 0's tags, then level 2 from the tags of level 1 (a
    level of many boxes), link tables built on both, stepped; level 2 is
    regridded again after more steps.  The new level-2 boxes nest in level
    1 (their ghost stencils fall on level-1 valid nodes) and cover every
    level-1 tag kept on nodes whose neighbours are all valid."""
    tri, nrm = sphere(pkg)
    s = duct(pkg, "fp32s")
    s.build_links(0, tri, nrm)
    s.init_uniform(0, 1.0, (U, 0.0, 0.0))
    s.step(0, 20)
    ggm0, _, _ = s.tag_ggm(0, 7, 0, 0.0)
    b1 = s.refine_by_ggm(1, quantity=7, limit=float(np.quantile(ggm0[ggm0 > 0], 0.9)),
                         blocking=2, buffer=1, max_size=16)
    s.build_links(1, tri, nrm, origin=(0.0, 0.0, 0.0), dx=0.5)
    s.advance(4)
    history = []
    for rnd in range(2):
        ggm1, _, _ = s.tag_ggm(1, 7, 0, 0.0)
        lim = float(np.quantile(ggm1[ggm1 > 0], 0.97))
        had = 2 in s.shapes
        b2 = s.refine_by_ggm(2, quantity=7, limit=lim, blocking=2, buffer=1, max_size=16)
        assert b2
        s.build_links(2, tri, nrm, origin=(0.0, 0.0, 0.0), dx=0.25)
        s.advance(3)
        f2 = s.download_f(2)
        assert np.isfinite(f2).all() and np.isfinite(s.download_f(1)).all()
        history.append((len(b1), len(b2), had))
    print("three levels (level-1 boxes, level-2 boxes, level 2 regridded):", history)
    s.close()
