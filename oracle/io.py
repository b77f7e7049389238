"""Restatement of the reference's text outputs and normals file (test infrastructure only).

ImageSaver::savetxt (image_save.cpp:5-47) and load_noraml (scene_normals.cpp:3-13).  An
iostream prints a float as printf prints the value promoted to double: "fixed <<
setprecision(p)" is "%.pf", default formatting "%g".  Epsilon is the double 1e-6
(global_constant.h:6), compared with the float promoted to double.
"""
from __future__ import annotations

import numpy as np

EPSILON = 1e-6


def savetxt_map(w: int, h: int, data, precision: int = 2) -> str:
    """Text that savetxt(filename, w, h, h_data, precision) appends (image_save.cpp:5-24)."""
    a = np.asarray(data, np.float32).reshape(-1)
    lines = []
    for r in range(h):
        cells = []
        for c in range(w):
            v = a[(h - 1 - r) * w + c]
            cells.append("0," if float(v) < EPSILON else "%.*f," % (precision, float(v)))
        lines.append("".join(cells) + "\n")
    return "".join(lines)


def intercept_sum(w: int, h: int, data) -> np.float32:
    """The float sum of savetxt(..., totalEnergy) (image_save.cpp:26-44): rows from the last."""
    a = np.asarray(data, np.float32).reshape(-1)
    s = np.float32(0.0)
    for r in range(h):
        for c in range(w):
            v = a[(h - 1 - r) * w + c]
            if float(v) >= EPSILON:
                s = np.float32(s + v)
    return s


def savetxt_ratio(w: int, h: int, data, total_energy: float) -> str:
    """Text that savetxt(filename, w, h, h_data, totalEnergy) appends (image_save.cpp:26-47)."""
    ratio = np.float32(intercept_sum(w, h, data) / np.float32(total_energy))
    return "%g\n" % float(ratio)


def load_normals(text: str) -> np.ndarray:
    """load_noraml (scene_normals.cpp:3-13): a count, then that many "x y z" normals."""
    tok = text.split()
    n = int(tok[0])
    vals = [float(t) for t in tok[1: 1 + 3 * n]]
    if len(vals) < 3 * n:
        raise ValueError("fewer normals than the count")
    return np.array(vals, np.float32).reshape(n, 3)
