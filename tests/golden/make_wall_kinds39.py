"""Writes wall_kinds39.json: the stencil kind of every (direction, axis) pair
of the reference's shifted-lattice wall derivatives, read from the text of
derivatives/modules/derivatives.f95 (first_order_shifted_vel_cond_unknown at
:816 and first_order_shifted_temp_cond_unknown at :1236): for Fortran
direction i = 1..39 and axis card = 1..3, +1 where the case calls dudx_ibb
with winds = .false., -1 with winds = .true., 0 where it calls dudx_local;
for the temperature function also whether its dudx_ibb call passes the target
as both operands (a zero gradient).  Run from the repository root with the
reference tree at /root/reference."""
import json
import os
import re
import sys

SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference",
                   "derivatives", "modules", "derivatives.f95")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "wall_kinds39.json")


def parse(lines, name):
    start = next(k for k, s in enumerate(lines) if re.match(rf"\s*function\s+{name}\b", s, re.I))
    end = next(k for k in range(start + 1, len(lines)) if re.match(r"\s*end\s+function", lines[k], re.I))
    kinds = {}
    zero_arg = {}
    case = card = None
    winds = None
    for s in lines[start:end]:
        s = s.split("!")[0]
        m = re.match(r"\s*case\s*\(\s*(\d+)\s*\)", s, re.I)
        if m:
            case, card, winds = int(m.group(1)), None, None
            continue
        m = re.search(r"card\s*==\s*(\d)", s)
        if m:
            card, winds = int(m.group(1)), None
            continue
        m = re.search(r"winds\s*=\s*\.(true|false)\.", s, re.I)
        if m:
            winds = m.group(1).lower() == "true"
            continue
        m = re.search(r"dqda\s*=\s*(dudx_ibb|dudx_local)\s*\(([^)]*)\)", s, re.I)
        if m:
            assert case is not None and card is not None and (case, card) not in kinds, (case, card)
            if m.group(1).lower() == "dudx_local":
                kinds[(case, card)] = 0
            else:
                assert winds is not None, (case, card)
                kinds[(case, card)] = -1 if winds else 1
                args = [a.strip() for a in m.group(2).split(",")]
                zero_arg[(case, card)] = args[0] == args[1]
    assert len(kinds) == 39 * 3, len(kinds)
    table = [[kinds[(i, a)] for a in (1, 2, 3)] for i in range(1, 40)]
    zeros = [[zero_arg.get((i, a)) for a in (1, 2, 3)] for i in range(1, 40)]
    return table, zeros


if __name__ == "__main__":
    with open(SRC) as fh:
        lines = fh.read().splitlines()
    vel, _ = parse(lines, "first_order_shifted_vel_cond_unknown")
    temp, tzero = parse(lines, "first_order_shifted_temp_cond_unknown")
    with open(OUT, "w") as fh:
        json.dump({"source": "derivatives/modules/derivatives.f95:816-1235,1236-",
                   "velocity_kinds": vel, "temperature_kinds": temp,
                   "temperature_ibb_zero": tzero}, fh, indent=None)
        fh.write("\n")
    print("wrote", OUT)
