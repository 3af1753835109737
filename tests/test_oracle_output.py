"""Output formats of the oracle (oracle/o_output.py) pinned on the CPU:
  - Plot3D grid and solution files equal, byte for byte, the files a Fortran
    program writes with the runtime's own unformatted sequential records
    from the same inputs (tests/golden/plot3d_layout.f90, built with the
    image's Fortran compiler), REAL coordinates from start + dx (i - lo)
    included; F12.5 microphone values and the list-directed header line
    equal that runtime's text (NaN included; the Inf spelling differs between
    Fortran runtimes and is not pinned)
  - subrecords: a record longer than the limit splits as gfortran documents
    (leading marker negative when more follow, trailing marker negative when
    one precedes), hand-written bytes
  - microphone weights: golden values worked out by hand from
    mic_setup.f95:86-213 (distances with + rdx on the far corners, solid
    corners out of the sum, corners at x index 0 weighing 0), and the
    invalid / outside-the-boxes cases
  - level geometry: spacing and first node of a refined box."""
import math
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT
import o_output as oo

FLANG = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang"
                                     if os.path.exists("/opt/rocm/bin/amdflang") else None)


def _inputs(rng):
    blocks = []
    for dims, start, dx in (((3, 2, 4), (-1.5, 0.1, 2.0 / 3.0), (0.1, 0.1, 0.1)),
                            ((5, 1, 2), (1.0 / 7.0, -3.3, 1e-3), (0.05, 0.3, 1.0 / 3.0))):
        nv = int(np.prod(dims))
        code = rng.integers(-3, 4, nv).astype(np.int32)
        code[::5] = -666
        code[1] = 1001
        blocks.append(dict(dims=dims, start=start, dx=dx, code=code,
                           blank=oo.iblank(code), fields=rng.standard_normal((6, nv))))
    h = dict(mach=1.5, reynolds=12345.6, time=0.125, gamma=1.4, t_ref=300.0, stag_temperature=435.0,
             cv=2.5, cp=3.5, gas_r=1.0 / 1.4, timestep=4200)
    return blocks, h


def test_records_and_text_match_fortran_runtime(tmp_path):
    if FLANG is None:
        pytest.skip("no Fortran compiler in this image")
    exe = tmp_path / "plot3d_layout"
    subprocess.run([FLANG, os.path.join(ROOT, "tests", "golden", "plot3d_layout.f90"), "-o",
                    str(exe)], check=True, capture_output=True, cwd=tmp_path)
    rng = np.random.default_rng(7)
    blocks, h = _inputs(rng)
    vals = np.array([0.5, -0.123456, 1234567.891234, -99999.5, 1e7, -1e-6, 2.5e-6, 999999.999996,
                     12.345675, -0.000005, 3.0, float("nan")])
    inp = [struct.pack("<i", len(blocks))]
    inp.append(np.asarray([b["dims"] for b in blocks], np.int32).tobytes())
    inp.append(np.asarray([b["start"] for b in blocks], np.float64).tobytes())
    inp.append(np.asarray([b["dx"] for b in blocks], np.float64).tobytes())
    inp.append(struct.pack("<9d", h["mach"], h["reynolds"], h["time"], h["gamma"], h["t_ref"],
                           h["stag_temperature"], h["cv"], h["cp"], h["gas_r"]))
    inp.append(struct.pack("<i", h["timestep"]))
    for b in blocks:
        inp.append(b["code"].tobytes() + np.ascontiguousarray(b["fields"]).tobytes())
    inp.append(struct.pack("<i", len(vals)) + vals.tobytes())
    (tmp_path / "in.bin").write_bytes(b"".join(inp))
    subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path)], check=True, timeout=60)
    assert (tmp_path / "grid.x").read_bytes() == oo.grid_file(blocks)
    assert (tmp_path / "data.q").read_bytes() == oo.data_file(blocks, h)
    lines = (tmp_path / "mic.txt").read_text().splitlines(keepends=True)
    assert lines[0] == oo.mic_header(63) and lines[1] == oo.mic_header(32 + 8)
    assert lines[2:2 + len(vals)] == [oo.f12_5(v) + "\n" for v in vals]
    six = [" ".join(oo.f12_5(v) for v in vals[i:i + 6]) + "\n" for i in range(0, len(vals) - 5, 6)]
    assert lines[2 + len(vals):] == six
    # the blocks of the grid file: x fastest, REAL, the IBLANK codes
    recs = oo.read_records(oo.grid_file(blocks))
    assert struct.unpack("<i", recs[0])[0] == 2 and len(recs) == 4
    x = np.frombuffer(recs[2], np.float32, 24)
    assert x[1] == np.float32(-1.5 + 0.1) and x[3] == np.float32(-1.5)


def test_subrecords_hand_written():
    payload = bytes(range(20))
    got = oo.record(payload, maxsub=7)
    want = (struct.pack("<i", -7) + payload[:7] + struct.pack("<i", 7)
            + struct.pack("<i", -7) + payload[7:14] + struct.pack("<i", -7)
            + struct.pack("<i", 6) + payload[14:] + struct.pack("<i", -6))
    assert got == want
    assert oo.record(payload, maxsub=20) == struct.pack("<i", 20) + payload + struct.pack("<i", 20)
    assert oo.read_records(got + oo.record(b"abc")) == [payload, b"abc"]


def _unit_level(solid=()):
    plo, dx, dlo = oo.level_geom((0.0, 0.0, 0.0), (16.0, 16.0, 16.0), (16, 16, 16), (0, 0, 0), 0)
    return plo, dx, dlo, (lambda G: all(0 <= g <= 15 for g in G)), (lambda G: tuple(G) not in solid)


def test_mic_weights_golden():
    plo, dx, dlo, owned, fluid = _unit_level()
    v, pos, fl, w = oo.mic_setup((2.25, 3.5, 4.75), plo, dx, dlo, plo, (16.0,) * 3, owned, fluid)
    # distances squared from (0.25, 0.5, 0.75) with + 1 on each far axis
    d2 = [0.875, 2.375, 2.875, 3.375, 4.375, 4.875, 5.375, 6.875]
    d = [math.sqrt(x) for x in d2]
    assert v == 1 and fl.all()
    assert pos.tolist() == [[2, 3, 4], [3, 3, 4], [2, 4, 4], [2, 3, 5], [3, 4, 4], [3, 3, 5],
                            [2, 4, 5], [3, 4, 5]]
    assert np.allclose(w, np.array(d) / sum(d), rtol=0, atol=1e-15)
    # a solid corner (the sixth) drops out of the sum
    plo, dx, dlo, owned, fluid = _unit_level(solid={(3, 3, 5)})
    v, pos, fl, w = oo.mic_setup((2.25, 3.5, 4.75), plo, dx, dlo, plo, (16.0,) * 3, owned, fluid)
    tot = sum(d) - d[5]
    assert v == 1 and not fl[5] and w[5] == 0.0
    assert np.allclose(np.delete(w, 5), np.delete(np.array(d), 5) / tot, rtol=0, atol=1e-15)
    # corners at x index 0 weigh 0, their distances stay in the sum
    plo, dx, dlo, owned, fluid = _unit_level()
    v, pos, fl, w = oo.mic_setup((0.25, 3.5, 4.75), plo, dx, dlo, plo, (16.0,) * 3, owned, fluid)
    assert v == 1 and [j for j in range(8) if w[j] == 0.0] == [0, 2, 3, 6]
    assert np.allclose(w[[1, 4, 5, 7]], np.array(d)[[1, 4, 5, 7]] / sum(d), rtol=0, atol=1e-15)


def test_mic_invalid_cases():
    plo, dx, dlo, owned, fluid = _unit_level()
    assert oo.mic_setup((16.5, 3.0, 3.0), plo, dx, dlo, plo, (16.0,) * 3, owned, fluid)[0] == 0
    assert oo.mic_setup((3.0, -0.1, 3.0), plo, dx, dlo, plo, (16.0,) * 3, owned, fluid)[0] == 0
    # x = probhi exactly: inside the bounds, its cell outside the boxes
    assert oo.mic_setup((16.0, 3.0, 3.0), plo, dx, dlo, plo, (16.0,) * 3, owned, fluid)[0] == 2
    allsolid = lambda G: False  # noqa: E731
    assert oo.mic_setup((3.5, 3.5, 3.5), plo, dx, dlo, plo, (16.0,) * 3, owned, allsolid)[0] == 0


def test_mic_line_and_header():
    fl = np.array([True] * 7 + [False])
    w = np.array([0.5, 0.25, 0.25, 0, 0, 0, 0, 0.0])
    cv = np.arange(48, dtype=float).reshape(6, 8)
    line = oo.mic_line(32 + 8 + 1, fl, w, cv)
    # p = 0.5*0 + 0.25*1 + 0.25*2; rho = 0.5*16 + 0.25*17 + 0.25*18; w-vel: row 5
    assert line == "     0.75000     16.75000     40.75000\n"
    assert oo.mic_line(0, fl, w, cv) == ""
    assert oo.mic_header(16 + 4) == " Temp,       ;U-vel,      ;\n"


def test_level_geometry_refined_box():
    plo, dx, dlo = oo.level_geom((-1.0, 0.0, 2.0), (1.0, 3.0, 4.0), (8, 12, 4), (0, 0, 0), 2)
    assert np.array_equal(dx, [2.0 / 32, 3.0 / 48, 2.0 / 16])
    s = oo.block_start(plo, dx, dlo, (5, 6, 7))
    assert s == [-1.0 + 5 * 2.0 / 32, 6 * 3.0 / 48, 2.0 + 7 * 2.0 / 16]
