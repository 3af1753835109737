"""The plain-lattice D3Q39 wall rule of the oracle (oracle/o_walls39.c,
o39p_*), pinned on the CPU:
  - F_i, G_i of every direction equal the reference's own case expressions
    (interpolative_bounceback_bc.f95:834-2838, evaluated in 40-digit
    arithmetic by tests/golden/make_ibb39_plain.py) to 1e-12 relative at 8
    random states
  - the Hermite projection alone equals them for the cases that carry no
    slip (22 of 38 for f, 27 for g), and misses them by more than 1e-6
    for every case that does
  - the derivative kinds the rule uses (sign of c_i along each axis: wall
    form, temperature 0; local difference along the others) equal the
    kinds of the reference's cases
  - the library's compiled slip table equals the fixture's
  - a fluid at rest around a sphere stays at rest (5 steps, 1e-13), and a
    Mach 0.3 stream past it stays finite with no repairs (20 steps)."""
import json
import os
import re

import numpy as np
import pytest

from conftest import PKG_DIR, ROOT

FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "ibb39_plain.json")))
T0 = 1 / 1.4


def _inputs(vec):
    x = dict(zip(FIX["input_names"], vec["inputs"]))
    D = [x[n] for n in FIX["derivatives"]]
    du = np.array([[D[0], D[3], D[6]], [D[1], D[4], D[7]], [D[2], D[5], D[8]]])
    return (x["temp_tgt"], [x["u_tgt"], x["v_tgt"], x["w_tgt"]], du, D[9:], x["beta_1"],
            x["beta_2"], x["const_vol"])


def test_neq_parts_equal_reference_cases(orc):
    worst = 0.0
    for vec in FIX["vectors"]:
        args = _inputs(vec)
        for i in range(1, 39):
            F, G = orc.wall_neq39p(i, *args)
            Fr, Gr = vec["FG"][i - 1]
            worst = max(worst, abs(F - Fr) / abs(Fr), abs(G - Gr) / abs(Gr))
    print(f"plain wall rule vs the reference's case expressions: max rel {worst:.2e}")
    assert worst < 1e-12


def test_hermite_projection_is_the_slip_free_part(orc):
    nf = [len(t) for t in FIX["slips_f"]]
    ng = [len(t) for t in FIX["slips_g"]]
    assert sum(n == 0 for n in nf) == 22 and sum(n == 0 for n in ng) == 27
    for vec in FIX["vectors"]:
        args = _inputs(vec)
        for i in range(1, 39):
            F, G = orc.wall_neq39p(i, *args, slips=False)
            Fr, Gr = vec["FG"][i - 1]
            ef, eg = abs(F - Fr) / abs(Fr), abs(G - Gr) / abs(Gr)
            assert (ef < 1e-12) == (nf[i - 1] == 0) and (eg < 1e-12) == (ng[i - 1] == 0), i
            if nf[i - 1]:
                assert ef > 1e-6, i
            if ng[i - 1]:
                assert eg > 1e-6, i


def test_derivative_kinds_follow_the_link_signs(orc):
    cx, cy, cz = orc.velocities39(False)
    for i in range(1, 39):
        c = (cx[i], cy[i], cz[i])
        want = []
        for comp in "uvwT":
            for ax in range(3):
                k = "+" if c[ax] > 0 else ("-" if c[ax] < 0 else "0")
                want.append(k + ("T" if comp == "T" and k != "0" else ""))
        # FIX order: du/dx du/dy du/dz dv/dx ... dT/dz
        assert FIX["kinds"][i - 1] == want, i


def test_compiled_slip_table_matches_fixture():
    text = open(os.path.join(PKG_DIR, "csrc", "ibb39_slips.h")).read()
    rows = re.findall(r"\{([-0-9.e]+), (-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+), (-?\d+)\}",
                      text)
    start = [int(v) for v in re.search(r"kIbbSlipStart\[77\] = \{([^}]*)\}", text).group(1).split(",")]
    want = [m for part in ("slips_f", "slips_g") for t in FIX[part] for m in t]
    assert len(rows) == len(want) == start[-1]
    for r, w in zip(rows, want):
        assert float(r[0]) == w[0] and [int(v) for v in r[1:]] == w[1:]
    lens = [len(t) for part in ("slips_f", "slips_g") for t in FIX[part]]
    assert np.array_equal(np.diff(start), lens)


def _sphere(pkg, orc, shape, faces):
    tri, nrm = pkg.icosphere(4, 4.3, (12.5, 11.75, 7.5), 20260103)
    st, ns, _ = orc.build_links39(shape, faces, tri, nrm)
    return st, ns


def test_rest_state_and_stream_around_sphere(pkg, orc):
    shape = (32, 24, 16)
    walls = [-10] * 6
    c = orc.consts39(mach=0.3, shifted=False)
    st, ns = _sphere(pkg, orc, shape, walls)
    f0, g0, lm0 = orc.init39(shape, 1.0, 0.0, 0.0, 0.0, T0, c)
    f1, g1, _, _, so, nwl, _ = orc.step_walls39(shape, walls, st, f0, g0, lm0, c, 0, 5)
    assert nwl > 1000 and so["n_repair"] == 0
    assert np.abs(f1 - f0).max() < 1e-13 and np.abs(g1 - g0).max() < 1e-13
    faces = [-1000, -11, -11, -6, -11, -11]
    st, _ = _sphere(pkg, orc, shape, faces)
    f0, g0, lm0 = orc.init39(shape, 1.0, 0.3, 0.0, 0.0, T0, c)
    f1, g1, _, _, so, nwl, _ = orc.step_walls39(shape, faces, st, f0, g0, lm0, c, 0, 20)
    rho, u, _, _, T = orc.macro39(f1, g1, c)
    fl = st[0] >= 0
    assert np.isfinite(f1).all() and so["n_repair"] == 0 and so["lm_noconv"] == 0
    assert 0.9 < rho[fl].min() and rho[fl].max() < 1.15 and np.abs(u[fl]).max() < 0.45
