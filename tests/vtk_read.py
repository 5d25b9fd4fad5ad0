"""Reader for the reference's full-precision ('unformatted') legacy VTK frames
(VTK_OUTPUT_UNFORMATTED.F90:17-281): POINTS then POINT_DATA scalars, one value
per line, list-directed."""
from __future__ import annotations

import numpy as np


def read_unformatted_vtk(path: str) -> dict:
    with open(path) as f:
        lines = f.read().splitlines()
    out = {}
    i = 0
    n = None
    while i < len(lines):
        s = lines[i].strip()
        if s.startswith("POINTS"):
            n = int(s.split()[1])
            pts = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)])
            out["points"] = pts
            i += 1 + n
            continue
        if s.startswith("SCALARS"):
            name = s.split()[1]
            # next line: LOOKUP_TABLE
            vals = lines[i + 2:i + 2 + n]
            out[name] = np.array([float(v.split()[0]) for v in vals])
            i += 2 + n
            continue
        i += 1
    return out
