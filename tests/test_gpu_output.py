"""Plot3D files and microphones written by the library (through the C ABI)
against the oracle's restatement of the reference's output path:
  - D3Q27 sphere in a walled box, FP64 storage, placed domain: the grid file
    equals the oracle's byte for byte (REAL coordinates of get_real_coords,
    IBLANK from the device-built link table); the solution file's records
    match: header record bytes equal; rho, u, v, w within 1e-13 relative of
    the oracle's moments of the downloaded populations; T = 1/3; alpha equal
    to the downloaded alpha
  - asynchronous write: the files hold the state at the call while the
    stream goes on stepping; a small subrecord limit gives the oracle's bytes
  - D3Q39 with a refined level: one block per level over its valid region
    (ghost shell of four excluded), rho, u, v, w, T within 1e-12 relative of
    the oracle's flow variables, alpha 2 when not kept; max_level clips
  - a level of two boxes: one block per box
  - microphones (D3Q27 with walls and a sphere; D3Q39 on a refined level):
    corners, fluid flags and weights equal the oracle's exactly; header
    lines equal; each sample line's F12.5 values within 1e-5 of the oracle's
    weighted sums of the oracle's flow variables of the state at the sample
    (F12.5 resolution), samples taken between steps and flushed at once."""
import os

import numpy as np
import pytest

import o_output as oo

pytestmark = pytest.mark.gpu

WALLS = [-10] * 6
SEED_GEOM = 20260102
HDR = dict(reynolds=100.0, t_ref=300.0, stag_temperature=301.5, cv=2.5, cp=3.5, gas_r=1.0)


def _sphere27(pkg, shape, prec="fp64"):
    s = pkg.Solver(precision=prec, nu=0.02, alpha_tol=1e-12)
    s.define_level(0, shape, periodic=(0, 0, 0))
    s.set_faces(0, WALLS)
    tri, nrm = pkg.icosphere(3, 4.3, (shape[0] / 2 + 0.5, shape[1] / 2 - 0.5, shape[2] / 2 + 0.25),
                             SEED_GEOM)
    s.build_links(0, tri, nrm)
    s.init_uniform(0, 1.0, (0.04, 0.01, 0.0))
    s.keep_alpha(True)
    return s


def _header(h, mach, gamma, time, ts):
    return dict(h, mach=mach, gamma=gamma, time=time, timestep=ts)


def _fields27(orc, f, alpha):
    rho, u, v, w = orc.macro27(f)
    return np.stack([rho, u, v, w, np.full_like(rho, 1.0 / 3.0), alpha])


def _check_data(path, blocks, h, tol, exact_rows=()):
    recs = oo.read_records(open(path, "rb").read())
    want = oo.read_records(oo.data_file(blocks, h))
    assert len(recs) == len(want) == 2 + 2 * len(blocks)
    assert recs[0] == want[0] and recs[1] == want[1]
    for k, b in enumerate(blocks):
        assert recs[2 + 2 * k] == oo.header_record(h)
        got = np.frombuffer(recs[3 + 2 * k], np.float64).reshape(6, -1)
        ref = np.asarray(b["fields"])
        for r in range(6):
            if r in exact_rows:
                assert np.array_equal(got[r], ref[r], equal_nan=True), r
            else:
                scale = np.nanmax(np.abs(ref[r])) or 1.0
                d = np.abs(got[r] - ref[r])
                assert np.all((d <= tol * scale) | (np.isnan(got[r]) & np.isnan(ref[r]))), r


def test_plot3d_d3q27_sphere(pkg, orc, tmp_path):
    shape = (24, 20, 18)
    s = _sphere27(pkg, shape)
    s.step(0, 6)
    problo, probhi = (-1.5, 2.0, 0.25), (-1.5 + 2.4, 4.0, 0.25 + 1.8)
    s.set_placement(problo, probhi)
    plo, dx, dlo = oo.level_geom(problo, probhi, shape, (0, 0, 0), 0)
    code = s.download_links(0)[0]
    f = s.download_f(0)
    alpha = s.alpha(0)
    blocks = [dict(dims=shape, start=oo.block_start(plo, dx, dlo, (0, 0, 0)), dx=dx,
                   blank=oo.iblank(code), fields=_fields27(orc, f, alpha))]
    assert (oo.iblank(code) == 0).sum() > 100
    h = _header(HDR, 0.1, 1.4, 0.35, 7)
    s.plot3d(str(tmp_path), 7, time=0.35, mach=0.1, gamma=1.4, **HDR)
    grid = (tmp_path / "uclbs_grid_000000007.x").read_bytes()
    assert grid == oo.grid_file(blocks)
    _check_data(tmp_path / "uclbs_data_000000007.q", blocks, h, 1e-13, exact_rows=(4, 5))
    # asynchronous: the call's state, while the stream goes on stepping
    os.mkdir(tmp_path / "a")
    s.plot3d(str(tmp_path / "a"), 8, time=0.35, mach=0.1, gamma=1.4, wait=False, **HDR)
    s.step(0, 3)
    s.output_wait()
    assert np.abs(s.download_f(0) - f).max() > 1e-6
    assert (tmp_path / "a" / "uclbs_grid_000000008.x").read_bytes() == grid
    _check_data(tmp_path / "a" / "uclbs_data_000000008.q", blocks, dict(h, timestep=8), 1e-13,
                exact_rows=(4, 5))
    # subrecords
    os.mkdir(tmp_path / "b")
    f = s.download_f(0)
    blocks[0]["fields"] = _fields27(orc, f, s.alpha(0))
    s.plot3d(str(tmp_path / "b"), 9, time=0.5, mach=0.1, gamma=1.4, max_subrecord=1000, **HDR)
    gb = (tmp_path / "b" / "uclbs_grid_000000009.x").read_bytes()
    assert gb == oo.grid_file(blocks, maxsub=1000) and gb != grid
    _check_data(tmp_path / "b" / "uclbs_data_000000009.q", blocks, _header(HDR, 0.1, 1.4, 0.5, 9),
                1e-13, exact_rows=(4, 5))
    s.close()


def test_plot3d_d3q39_two_levels(pkg, orc, tmp_path):
    from test_gpu_amr39 import CSHAPE, FLO, FVALID, FSTORED, G, T0, fields
    c = orc.consts39(mach=0.3, shifted=False)
    s = pkg.Solver(lattice="d3q39", precision="fp32s")
    s.set_gas(mach=0.3, shifted=0)
    s.define_level(0, CSHAPE)
    s.init_fields(0, *fields(CSHAPE, 0.3))
    s.define_level(1, FVALID, periodic=(0, 0, 0), lo=FLO)
    s.init_from_coarse(1)
    s.advance(2)
    plo0, dx0, dlo0 = oo.level_geom((0, 0, 0), CSHAPE, CSHAPE, (0, 0, 0), 0)
    plo1, dx1, dlo1 = oo.level_geom((0, 0, 0), CSHAPE, CSHAPE, (0, 0, 0), 1)
    f0, g0, _ = s.download_state(0)
    f1, g1, _ = s.download_state(1)
    m0 = np.stack(list(orc.macro39(f0, g0, c)) + [np.full(f0.shape[1], 2.0)])
    m1 = np.stack(list(orc.macro39(f1, g1, c)) + [np.full(f1.shape[1], 2.0)])
    m1 = m1.reshape(6, FSTORED[2], FSTORED[1], FSTORED[0])[:, G:-G, G:-G, G:-G].reshape(6, -1)
    nv1 = int(np.prod(FVALID))
    blocks = [dict(dims=CSHAPE, start=oo.block_start(plo0, dx0, dlo0, (0, 0, 0)), dx=dx0,
                   blank=np.ones(m0.shape[1], np.int32), fields=m0),
              dict(dims=FVALID, start=oo.block_start(plo1, dx1, dlo1, FLO), dx=dx1,
                   blank=np.ones(nv1, np.int32), fields=m1)]
    s.plot3d(str(tmp_path), 12, time=1.0, cv=c.cv, gas_r=1.0)
    assert (tmp_path / "uclbs_grid_000000012.x").read_bytes() == oo.grid_file(blocks)
    h = dict(HDR, mach=0.3, gamma=1.4, time=1.0, timestep=12, cv=c.cv, reynolds=0.0, t_ref=0.0,
             stag_temperature=0.0, cp=0.0)
    _check_data(tmp_path / "uclbs_data_000000012.q", blocks, h, 1e-12, exact_rows=(5,))
    s.plot3d(str(tmp_path), 13, max_level=0)
    recs = oo.read_records((tmp_path / "uclbs_grid_000000013.x").read_bytes())
    assert len(recs) == 3 and recs[2] == oo.read_records(oo.grid_file(blocks[:1]))[2]
    s.close()


def test_plot3d_level_of_boxes(pkg, orc, tmp_path):
    shape = (16, 6, 6)
    boxes = [((0, 0, 0), (6, 5, 5)), ((7, 0, 0), (15, 5, 5))]
    f0 = orc.init_tgv27(shape, 0.05, 1e-3, 3)
    s = pkg.Solver(precision="fp64", nu=0.01, alpha_tol=1e-12)
    s.define_level_boxes(0, boxes)
    s.upload_f(0, s.dense_to_stored(0, f0.reshape(27, shape[2], shape[1], shape[0])))
    s.step(0, 2)
    f = s.download_f(0)
    rho, u, v, w = orc.macro27(f)
    plo, dx, dlo = oo.level_geom((0, 0, 0), shape, shape, (0, 0, 0), 0)
    blocks = []
    for (off, e, lo), (blo, bhi) in zip(s.box_layout(0), boxes):
        dims = tuple(bhi[d] - blo[d] + 1 for d in range(3))
        sl = lambda a: a[off:off + e[0] * e[1] * e[2]].reshape(e[2], e[1], e[0])[1:-1, 1:-1, 1:-1].ravel()  # noqa: E731
        fl = np.stack([sl(a) for a in (rho, u, v, w)] + [np.full(int(np.prod(dims)), 1 / 3),
                                                           np.full(int(np.prod(dims)), 2.0)])
        blocks.append(dict(dims=dims, start=oo.block_start(plo, dx, dlo, blo), dx=dx,
                           blank=np.ones(fl.shape[1], np.int32), fields=fl))
    s.plot3d(str(tmp_path), 3)
    assert (tmp_path / "uclbs_grid_000000003.x").read_bytes() == oo.grid_file(blocks)
    _check_data(tmp_path / "uclbs_data_000000003.q", blocks,
                dict({k: 0.0 for k in HDR}, mach=0.0, gamma=0.0, time=0.0, timestep=3), 1e-13,
                exact_rows=(4, 5))
    s.close()


def _mic_lines(path):
    lines = open(path).read().splitlines(keepends=True)
    return lines[0], lines[1:]


def _check_mic_lines(lines, want_lines):
    assert len(lines) == len(want_lines)
    for got, want in zip(lines, want_lines):
        g, w = got.rstrip("\n"), want.rstrip("\n")
        assert len(g) == len(w) and len(g) % 13 == 12
        gv = [float(g[i:i + 12]) for i in range(0, len(g), 13)]
        wv = [float(w[i:i + 12]) for i in range(0, len(w), 13)]
        assert all(g[i + 12] == " " for i in range(0, len(g) - 12, 13))
        assert np.allclose(gv, wv, rtol=0, atol=1.01e-5), (got, want)


def test_microphones_d3q27(pkg, orc, tmp_path):
    shape = (24, 20, 18)
    s = _sphere27(pkg, shape)
    problo, probhi = (-1.0, -1.0, 0.0), (-1.0 + 24 * 0.25, -1.0 + 20 * 0.25, 18 * 0.25)
    s.set_placement(problo, probhi)
    plo, dx, dlo = oo.level_geom(problo, probhi, shape, (0, 0, 0), 0)
    code = s.download_links(0)[0].reshape(shape[2], shape[1], shape[0])
    inside = lambda G: all(0 <= G[d] < shape[d] for d in range(3))  # noqa: E731
    fluid = lambda G: inside(G) and code[G[2], G[1], G[0]] >= 0  # noqa: E731
    cen = [plo[d] + (shape[d] / 2 + (0.5, -0.5, 0.25)[d]) * dx[d] for d in range(3)]
    locs = [(0.6, 0.7, 1.3),                       # open flow, every field
            (cen[0] - 4.1 * dx[0], cen[1] + 0.3, cen[2] + 0.1),  # at the sphere
            (plo[0] + 0.3 * dx[0], 0.9, 2.2),       # cell at x index 0
            (probhi[0] + 0.1, 1.0, 1.0),            # outside the domain
            (probhi[0], 1.0, 1.0),                  # on probhi: outside the boxes
            (1.1, 1.2, 1.3)]                        # records nothing
    types = [63, 32 + 8, 4 + 1, 63, 63, 0]
    valid = s.mic_setup(str(tmp_path), locs, types, [0] * len(locs))
    want_valid, corners = [], []
    for i, p in enumerate(locs):
        v, pos, fl, w = oo.mic_setup(p, plo, dx, dlo, plo, probhi, inside, fluid)
        want_valid.append(v)
        corners.append((pos, fl, w))
        if v == 1:
            gpos, gw, gnode = s.mic_weights(i)
            assert np.array_equal(gpos, pos) and np.array_equal(gw, w)
            assert np.array_equal(gnode >= 0, fl)
    assert valid.tolist() == want_valid == [1, 1, 1, 0, 2, 1]
    assert not corners[1][1].all() and corners[1][1].any()   # some corners solid
    assert corners[2][2][0] == 0.0 and corners[2][1][0]       # fluid, x index 0: weight 0
    want = {i: [] for i in range(len(locs))}
    for k in range(3):
        s.step(0, 2)
        rho, u, v, w = orc.macro27(s.download_f(0))
        p = rho * (1 / 3 + (u * u + v * v + w * w) / 2)
        flow = np.stack([p, np.full_like(rho, 1 / 3), rho, u, v, w])
        for i, (pos, fl, wt) in enumerate(corners):
            if want_valid[i] == 1:
                idx = [int(g[0] + shape[0] * (g[1] + shape[1] * g[2])) if f else 0
                       for g, f in zip(pos, fl)]
                want[i].append(oo.mic_line(types[i], fl, wt, flow[:, idx]))
        s.mic_sample()
    s.mic_flush()
    for i in range(len(locs)):
        path = tmp_path / f"microphone{i + 1:05d}.txt"
        if want_valid[i] == 0:
            assert not path.exists()
            continue
        head, lines = _mic_lines(path)
        assert head == oo.mic_header(types[i])
        _check_mic_lines(lines, [x for x in want[i] if x])
    assert len(_mic_lines(tmp_path / "microphone00001.txt")[1]) == 3
    s.close()


def test_microphones_d3q39_refined_level(pkg, orc, tmp_path):
    from test_gpu_amr39 import CSHAPE, FLO, FHI, FSTORED, FORIGIN, FVALID, fields
    c = orc.consts39(mach=0.3, shifted=False)
    s = pkg.Solver(lattice="d3q39", precision="fp64")
    s.set_gas(mach=0.3, shifted=0)
    s.define_level(0, CSHAPE)
    s.init_fields(0, *fields(CSHAPE, 0.3))
    s.define_level(1, FVALID, periodic=(0, 0, 0), lo=FLO)
    s.init_from_coarse(1)
    plo, dx, dlo = oo.level_geom((0, 0, 0), CSHAPE, CSHAPE, (0, 0, 0), 1)
    owned = lambda G: all(FLO[d] <= G[d] <= FHI[d] for d in range(3))  # noqa: E731
    locs = [(7.3, 5.6, 6.2), (10.9, 8.4, 7.7), (2.0, 2.0, 2.0)]
    levels = [1, 5, 1]     # 5 clips to the finest level; the last is outside level 1's box
    valid = s.mic_setup(str(tmp_path), locs, [63, 16 + 2, 63], levels)
    corners = [oo.mic_setup(p, plo, dx, dlo, (0, 0, 0), CSHAPE, owned, owned) for p in locs]
    assert valid.tolist() == [x[0] for x in corners] == [1, 1, 2]
    for i in range(2):
        gpos, gw, _ = s.mic_weights(i)
        assert np.array_equal(gpos, corners[i][1]) and np.array_equal(gw, corners[i][3])
    want = [[], []]
    for k in range(2):
        s.advance(1)
        f, g, _ = s.download_state(1)
        rho, u, v, w, T = orc.macro39(f, g, c)
        flow = np.stack([rho * (T + (u * u + v * v + w * w) / 2), T, rho, u, v, w])
        for i in range(2):
            _, pos, fl, wt = corners[i]
            idx = [int((gp[0] - FORIGIN[0]) + FSTORED[0] * ((gp[1] - FORIGIN[1]) + FSTORED[1] *
                                                             (gp[2] - FORIGIN[2]))) for gp in pos]
            want[i].append(oo.mic_line([63, 18][i], fl, wt, flow[:, idx]))
        s.mic_sample()
    s.mic_flush()
    for i in range(2):
        head, lines = _mic_lines(tmp_path / f"microphone{i + 1:05d}.txt")
        assert head == oo.mic_header([63, 18][i])
        _check_mic_lines(lines, want[i])
    head, lines = _mic_lines(tmp_path / "microphone00003.txt")
    assert head == oo.mic_header(63) and lines == []
    s.close()
