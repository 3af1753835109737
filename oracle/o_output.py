"""Plot3D files and microphones, Python restatement.  TEST INFRASTRUCTURE ONLY.

Restates the reference's output path:
  plot3d_output.f95:76-305   grid file uclbs_grid_<ts>.x (block count, block
                             extents, per block x, y, z as REAL and IBLANK)
                             and solution file uclbs_data_<ts>.q (extents with
                             6 variables and a 0, per block a header record
                             and rho, u, v, w, T, alpha as REAL(8)); one
                             block per box, its valid region
  amr_processes.f95:1155-1168 get_real_coords: problo + (i - domain lo) dx
  mic_setup.f95:45-225       the eight corners of a microphone's cell and
                             their weights
  mic_initialize.f95:17-70   the header line of a microphone file
  mic_write.f95:45-190       the weighted sums and their F12.5 line
Records follow gfortran's unformatted sequential layout: a 4-byte length
before and after each subrecord; records longer than the subrecord limit
(2147483639 bytes by default) split into subrecords, the leading marker
negative when more follow, the trailing marker negative when one precedes.
"""
import math
import struct

import numpy as np

MAX_SUBRECORD = 2147483639
CORNERS = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1))
HEAD = ("P,          ;", "Temp,       ;", "Density,    ;", "U-vel,      ;", "V-vel,      ;",
        "W-vel,      ;")


def record(payload: bytes, maxsub: int = MAX_SUBRECORD) -> bytes:
    """One logical record, split into subrecords of at most maxsub bytes."""
    n = len(payload)
    out, pos, first = [], 0, True
    while True:
        k = min(n - pos, maxsub)
        last = pos + k == n
        out += [struct.pack("<i", k if last else -k), payload[pos:pos + k],
                struct.pack("<i", k if first else -k)]
        pos += k
        first = False
        if last:
            return b"".join(out)


def read_records(data: bytes):
    """Payloads of the logical records of a file (subrecords joined)."""
    recs, pos = [], 0
    while pos < len(data):
        parts = []
        while True:
            (lead,) = struct.unpack_from("<i", data, pos)
            k = abs(lead)
            parts.append(data[pos + 4:pos + 4 + k])
            (trail,) = struct.unpack_from("<i", data, pos + 4 + k)
            assert abs(trail) == k, "record markers disagree"
            pos += 8 + k
            if lead >= 0:
                break
        recs.append(b"".join(parts))
    return recs


def level_geom(problo, probhi, n0, lo0, lvl):
    """(first domain node, spacing, domain corner index) of level lvl: AMReX
    geometry at refinement ratio 2 over the level-0 domain of n0 cells."""
    plo = np.asarray(problo, np.float64)
    dx = (np.asarray(probhi, np.float64) - plo) / (np.asarray(n0, np.float64) * float(1 << lvl))
    dlo = np.asarray(lo0, np.int64) * (1 << lvl)
    return plo, dx, dlo


def block_start(plo, dx, dlo, lo):
    """get_real_coords of a box's lower corner."""
    return [plo[d] + (float(lo[d]) - float(dlo[d])) * dx[d] for d in range(3)]


def block_coords(start, dx, dims):
    """x, y, z of a block, x fastest, as REAL (start + dx i in REAL(8))."""
    nx, ny, nz = dims
    a = np.arange(nx, dtype=np.float64)
    b = np.arange(ny, dtype=np.float64)
    c = np.arange(nz, dtype=np.float64)
    x = np.broadcast_to((start[0] + dx[0] * a)[None, None, :], (nz, ny, nx))
    y = np.broadcast_to((start[1] + dx[1] * b)[None, :, None], (nz, ny, nx))
    z = np.broadcast_to((start[2] + dx[2] * c)[:, None, None], (nz, ny, nx))
    return [v.astype(np.float32).ravel() for v in (x, y, z)]


def iblank(code):
    """1 on fluid nodes (code >= 0), -1 on code -666, 0 otherwise."""
    code = np.asarray(code)
    return np.where(code >= 0, 1, np.where(code == -666, -1, 0)).astype(np.int32)


def grid_file(blocks, maxsub=MAX_SUBRECORD) -> bytes:
    """blocks: dicts with dims, start, dx and blank (IBLANK, x fastest)."""
    out = [record(struct.pack("<i", len(blocks)), maxsub),
           record(np.asarray([b["dims"] for b in blocks], np.int32).tobytes(), maxsub)]
    for b in blocks:
        x, y, z = block_coords(b["start"], b["dx"], b["dims"])
        out.append(record(x.tobytes() + y.tobytes() + z.tobytes() +
                          np.asarray(b["blank"], np.int32).tobytes(), maxsub))
    return b"".join(out)


def header_record(h) -> bytes:
    """mach, 0, reynolds, time, gamma, 0, t_ref, 0 (integer), stag_temperature,
    cv, cp, R, R, mach, time, timestep (integer)."""
    return (struct.pack("<7d", h["mach"], 0.0, h["reynolds"], h["time"], h["gamma"], 0.0, h["t_ref"])
            + struct.pack("<i", 0)
            + struct.pack("<7d", h["stag_temperature"], h["cv"], h["cp"], h["gas_r"], h["gas_r"],
                          h["mach"], h["time"])
            + struct.pack("<i", h["timestep"]))


def data_file(blocks, h, maxsub=MAX_SUBRECORD) -> bytes:
    """blocks: dicts with dims and fields (6 x nodes: rho, u, v, w, T, alpha)."""
    dims = np.asarray([b["dims"] for b in blocks], np.int32).ravel()
    out = [record(struct.pack("<i", len(blocks)), maxsub),
           record(np.concatenate([dims, np.asarray([6, 0], np.int32)]).tobytes(), maxsub)]
    for b in blocks:
        out.append(record(header_record(h), maxsub))
        out.append(record(np.asarray(b["fields"], np.float64).tobytes(), maxsub))
    return b"".join(out)


def f12_5(v: float) -> str:
    """Fortran F12.5 editing (asterisks when the value does not fit)."""
    if math.isnan(v):
        return "%12s" % "NaN"
    if math.isinf(v):
        return "%12s" % ("Infinity" if v > 0 else "-Infinity")
    s = "%12.5f" % v
    return "*" * 12 if len(s) > 12 else s


def mic_header(mask: int) -> str:
    """List-directed write of the header: a leading blank, then the labels."""
    return " " + "".join(HEAD[k] for k in range(6) if mask & (32 >> k)) + "\n"


def mic_setup(p, plo, dx, dlo, glo, ghi, owned, fluid):
    """Corners and weights of a microphone at p (mic_setup.f95:45-225).

    plo, dx, dlo: the level's geometry (level_geom); glo, ghi: the domain's
    physical bounds; owned(G): whether level position G lies in one of the
    level's boxes; fluid(G): whether the corner at G is a fluid node.  Returns
    (valid, corners 8 x 3, fluid flags 8, weights 8) with valid 0 (invalid),
    1 (sampled) or 2 (valid, outside the level's boxes)."""
    pos = np.zeros((8, 3), np.int64)
    fl = np.zeros(8, bool)
    w = np.zeros(8)
    if any(p[d] > ghi[d] or p[d] < glo[d] for d in range(3)):
        return 0, pos, fl, w
    rdx = dx[0]
    st = [int(dlo[d]) + math.floor((p[d] - plo[d]) / rdx) for d in range(3)]
    lo = [plo[d] + (float(st[d]) - float(dlo[d])) * dx[d] for d in range(3)]
    if not owned(st):
        return 2, pos, fl, w
    dist = np.zeros(8)
    total = 0.0
    for j, cj in enumerate(CORNERS):
        pos[j] = [st[d] + cj[d] for d in range(3)]
        fl[j] = fluid(tuple(int(v) for v in pos[j]))
        if not fl[j]:
            continue
        s = 0.0
        for d in range(3):
            t = (p[d] - lo[d] + rdx) if cj[d] else (p[d] - lo[d])
            s += t * t
        dist[j] = math.sqrt(s)
        total += dist[j]
    if not fl.any():
        return 0, pos, fl, w
    for j in range(8):
        w[j] = dist[j] / total if (fl[j] and pos[j][0] - dlo[0] > 0) else 0.0
    return 1, pos, fl, w


def mic_line(mask, fl, w, corner_values):
    """One line of a microphone file: corner_values is 6 x 8 in the order
    p, T, rho, u, v, w; the selected fields' weighted sums, F12.5, one blank
    apart (no line when nothing is selected)."""
    vals = []
    for k in range(6):
        if not mask & (32 >> k):
            continue
        v = 0.0
        for j in range(8):
            if fl[j]:
                v = v + w[j] * corner_values[k][j]
        vals.append(f12_5(v))
    return " ".join(vals) + "\n" if vals else ""
