"""Lattice tables and the seeded RNG of the oracle, pinned by closed forms and
published known-answer vectors."""
import numpy as np


def test_d3q27_tables(orc):
    cx, cy, cz, opp, w = orc.lattice27()
    c = np.stack([cx, cy, cz], 1)
    # 27 distinct velocities of {-1,0,1}^3, rest first
    assert len({tuple(v) for v in c}) == 27 and tuple(c[0]) == (0, 0, 0)
    # opposite table: c_opp = -c (reference opp_27, linkwise.f95:36-37)
    assert np.array_equal(c[opp], -c)
    assert np.array_equal(opp[opp], np.arange(27))
    # weights: tensor products of (1/6, 2/3, 1/6) -> isotropy up to 4th order
    w1 = {0: 2 / 3, 1: 1 / 6, -1: 1 / 6}
    assert np.allclose(w, [w1[a] * w1[b] * w1[d] for a, b, d in c], rtol=0, atol=1e-16)
    assert abs(w.sum() - 1) < 1e-15
    for a in range(3):
        for b in range(3):
            assert abs((w * c[:, a] * c[:, b]).sum() - (a == b) / 3) < 1e-15
    assert abs((w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() - 1 / 9) < 1e-15
    assert abs((w * c[:, 0] ** 4).sum() - 1 / 3) < 1e-15


def test_d3q39_tables(orc):
    cx, cy, cz, opp = orc.lattice39()
    c = np.stack([cx, cy, cz], 1)
    assert len({tuple(v) for v in c}) == 39
    assert np.array_equal(c[opp], -c)
    # speeds: 1 rest, 6 x |1|, 8 x sqrt3, 6 x |2|, 12 x sqrt8, 6 x |3|
    mag = (c ** 2).sum(1)
    assert sorted(np.bincount(mag).tolist()) == sorted([1, 6, 8, 6, 12, 6] + [0] * (10 - 6))


def test_philox_known_answers(orc):
    # Random123 philox4x32-10 known-answer vectors
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        assert orc.philox(ctr, key) == list(want)


def test_uniform_range_and_mean(orc):
    u = np.array([orc.uniform(n, c, 20260101) for n in range(4000) for c in range(3)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.var() - 1 / 12) < 0.01
