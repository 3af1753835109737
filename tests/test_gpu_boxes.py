"""Levels of many boxes, the batched streaming launch, the native multi-level
advance and its hipGraph replay, on the GPU through the C ABI.

The reference for a level of many boxes is the same region held as one box
(itself checked against the oracle in test_gpu_d3q27 / test_gpu_amr): every
node update is the same arithmetic on the same inputs, and ghost copies and
injections are exact, so the fields must agree bit for bit, in FP64 and in
FP32-shifted storage:
  level 0 periodic domain split into uneven boxes == one periodic box;
  refined level split into uneven boxes == the same refined region as one box;
  hipGraph replays == eager launches, replay after replay with eager work and
    host reads in between;
  autotuned tile shape: same fields, state untouched by the tuning.
A three-level hierarchy of uneven boxes (edges from {8, 16, 24, 32}) stays
finite and conserves level-0 mass to 1e-4 over 8 coarse steps (nodal injection,
no flux correction, as the reference)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _split(lo, hi, edges, rng):
    """Segments [a, b] covering lo..hi with lengths drawn from edges (the last
    one takes the rest)."""
    out, a = [], lo
    while a <= hi:
        e = int(rng.choice(edges))
        b = min(hi, a + e - 1)
        if hi - b < min(edges):
            b = hi
        out.append((a, b))
        a = b + 1
    return out


def _tiling(lo, hi, edges, seed):
    """Uneven boxes tiling lo..hi: x segments per (z, y) slab drawn anew."""
    rng = np.random.default_rng(seed)
    boxes = []
    for z0, z1 in _split(lo[2], hi[2], edges, rng):
        for y0, y1 in _split(lo[1], hi[1], edges, rng):
            for x0, x1 in _split(lo[0], hi[0], edges, rng):
                boxes.append(((x0, y0, z0), (x1, y1, z1)))
    return boxes


def _tgv_state(orc, shape, seed=5):
    u0 = 0.05
    return orc.init_tgv27(shape, u0, 1e-3, seed)  # (27, n) x fastest


@pytest.mark.parametrize("prec", ["fp64", "fp32s"])
def test_level0_boxes_match_single_box(pkg, orc, prec):
    shape = (30, 22, 18)
    f0 = _tgv_state(orc, shape)
    ref = pkg.Solver(precision=prec, nu=0.005, alpha_tol=1e-12)
    ref.define_level(0, shape)
    ref.upload_f(0, f0)
    s = pkg.Solver(precision=prec, nu=0.005, alpha_tol=1e-12)
    boxes = _tiling((0, 0, 0), tuple(n - 1 for n in shape), (4, 6, 8, 11), 11)
    assert len(boxes) > 20
    s.define_level_boxes(0, boxes)
    dense = f0.reshape(27, shape[2], shape[1], shape[0])
    s.upload_f(0, s.dense_to_stored(0, dense))
    ref.step(0, 6)
    s.step(0, 6)
    fr = ref.download_f(0).reshape(27, shape[2], shape[1], shape[0])
    fs = s.stored_to_dense(0, s.download_f(0), (shape[2], shape[1], shape[0]))
    assert np.array_equal(fr, fs)
    sr, ss = ref.stats(), s.stats()
    assert sr["node_updates"] == ss["node_updates"] and sr["newton_iters"] == ss["newton_iters"]


def _two_level(pkg, orc, prec, fine_boxes):
    coarse = (20, 18, 16)
    s = pkg.Solver(precision=prec, nu=0.01, alpha_tol=1e-12)
    s.define_level(0, coarse)
    s.upload_f(0, _tgv_state(orc, coarse, 7))
    if fine_boxes is None:
        s.define_level(1, (18, 16, 14), periodic=(0, 0, 0), lo=(10, 9, 8))
    else:
        s.define_level_boxes(1, fine_boxes)
    s.init_from_coarse(1)
    return s


@pytest.mark.parametrize("prec", ["fp64", "fp32s"])
def test_refined_boxes_match_single_refined_box(pkg, orc, prec):
    lo, hi = (10, 9, 8), (27, 24, 21)
    boxes = _tiling(lo, hi, (3, 5, 7), 13)
    assert len(boxes) > 8
    ref = _two_level(pkg, orc, prec, None)
    s = _two_level(pkg, orc, prec, boxes)
    ref.advance(3)
    s.advance(3)
    assert np.array_equal(ref.download_f(0), s.download_f(0))
    fr = ref.download_f(1)                       # stored box with its ghost shell
    ext = tuple(hi[d] - lo[d] + 3 for d in range(3))
    fr = fr.reshape(27, ext[2], ext[1], ext[0])[:, 1:-1, 1:-1, 1:-1]
    fs = s.stored_to_dense(1, s.download_f(1), fr.shape[1:], origin=lo)
    assert np.array_equal(fr, fs)


def _three_levels(pkg, orc, prec, seed=17):
    s = pkg.Solver(precision=prec, nu=0.01, alpha_tol=1e-10)
    shape = (48, 40, 32)
    s.define_level_boxes(0, _tiling((0, 0, 0), tuple(n - 1 for n in shape), (8, 16, 24), seed))
    f0 = _tgv_state(orc, shape, 3).reshape(27, shape[2], shape[1], shape[0])
    s.upload_f(0, s.dense_to_stored(0, f0))
    s.define_level_boxes(1, _tiling((24, 20, 16), (71, 59, 47), (8, 16, 24), seed + 1))
    s.init_from_coarse(1)
    s.define_level_boxes(2, _tiling((64, 56, 48), (111, 95, 79), (8, 16, 24, 32), seed + 2))
    s.init_from_coarse(2)
    return s, shape


def test_graph_replays_match_eager(pkg, orc):
    eager, _ = _three_levels(pkg, orc, "fp32s")
    graph, _ = _three_levels(pkg, orc, "fp32s")
    for rep in range(3):
        eager.advance_native(2, graph=False)
        graph.advance_native(2, graph=True)      # one replay of the captured pair
        graph.macro(0)                            # host work between replays
        for lvl in range(3):
            assert np.array_equal(eager.download_f(lvl), graph.download_f(lvl)), (rep, lvl)
    # the Python schedule gives the same fields as the native one
    py, _ = _three_levels(pkg, orc, "fp32s")
    py.advance(6)
    for lvl in range(3):
        assert np.array_equal(py.download_f(lvl), eager.download_f(lvl)), lvl
    assert eager.stats()["node_updates"] == graph.stats()["node_updates"]


def test_autotune_keeps_state_and_fields(pkg, orc):
    a, _ = _three_levels(pkg, orc, "fp32s", seed=29)
    b, _ = _three_levels(pkg, orc, "fp32s", seed=29)
    f_before = a.download_f(1)
    t = a.autotune(1)
    assert t in (64, 128, 256)
    assert np.array_equal(f_before, a.download_f(1))
    a.advance_native(2)
    b.advance_native(2)
    for lvl in range(3):
        assert np.array_equal(a.download_f(lvl), b.download_f(lvl))


def test_three_level_hierarchy_conserves_and_stays_finite(pkg, orc):
    s, shape = _three_levels(pkg, orc, "fp32s", seed=41)
    r0 = s.macro(0)[0]
    m0 = float(s.stored_to_dense(0, r0, (shape[2], shape[1], shape[0])).sum())
    s.advance_native(8, graph=True)
    rho, ux, uy, uz = s.macro(0)
    dense = s.stored_to_dense(0, rho, (shape[2], shape[1], shape[0]))
    assert np.isfinite(dense).all()
    m1 = float(dense.sum())
    assert abs(m1 - m0) / m0 < 1e-4
    for lvl in (1, 2):
        assert np.isfinite(s.macro(lvl)[0]).all()
    st = s.stats()
    assert st["n_noconv"] == 0 and st["n_nonpos"] == 0


def test_level0_boxes_tgv_init_matches_single_box(pkg):
    shape = (32, 24, 20)
    ref = pkg.Solver(precision="fp32s", nu=0.005, alpha_tol=1e-10)
    ref.define_level(0, shape)
    ref.init_tgv(0, 0.05, 1e-3, 99)
    s = pkg.Solver(precision="fp32s", nu=0.005, alpha_tol=1e-10)
    s.define_level_boxes(0, _tiling((0, 0, 0), tuple(n - 1 for n in shape), (5, 8, 12), 23))
    s.init_tgv(0, 0.05, 1e-3, 99)
    dshape = (shape[2], shape[1], shape[0])
    fr = ref.download_f(0).reshape((27,) + dshape)
    assert np.array_equal(fr, s.stored_to_dense(0, s.download_f(0), dshape))
    # the ghost shell holds the wrapped domain values as well
    assert np.array_equal(s.download_f(0), s.dense_to_stored(0, fr))
    ref.step(0, 4)
    s.step(0, 4)
    assert np.array_equal(ref.download_f(0).reshape((27,) + dshape),
                          s.stored_to_dense(0, s.download_f(0), dshape))
