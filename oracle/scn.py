"""Restatement of the reference's .scn parser (test infrastructure only).

SceneFileProc::SceneFileRead(SolarScene*, std::string), scene_file_proc.cpp:41-286, with
readHelioParamter (:3-11) and the key map of the constructor (:16-28):
  * a line is skipped when its first character is '#' or it is empty;
  * the first whitespace-separated token selects the block (ground / Recv / Grid) or a key;
  * a key unknown to the map reads as "pos" (std::map::operator[] inserts the default
    value 0 = StringValue::pos, :155, :187, :232);
  * a number that cannot be read is stored as 0 (std::istream extraction failure);
  * after a "helio x y z" line the next line holds the size (sx, sy, sz);
  * heliostat lines before both "gap" and "matrix" are set are ignored (:246-249), and so are
    heliostat lines of a grid whose heliostat type is not 0 (:253); neither consumes the
    size line, which then reads as a key line;
  * "gap" and "matrix" are forgotten at every Grid block (:136-137);
  * a grid's heliostats start at the sum of the "n" of the grids before it (:133, :199-201).
The first receiver and every grid are kept.  Where the reference prints a warning and goes
on into undefined behaviour (a grid listing a number of heliostats other than its "n",
:112-115 and :250-252; receiver type outside 0..2, :96-101; grid type other than 0,
:123-126; heliostat type other than 0 with heliostats) this restatement raises ValueError.
"""
from __future__ import annotations

import re

import numpy as np

KEYS = {"pos": 0, "size": 1, "norm": 2, "face": 3, "end": 4, "gap": 5, "matrix": 6,
        "helio": 7, "inter": 8, "n": 9, "type": 10}


_INT = re.compile(r"[+-]?\d+")


def _nums(tokens, n, conv=float):
    """istream >> a >> b >> ...: a failed extraction stores 0 and fails every later one.
    An int extraction stops at the first non-digit, which stays in the stream."""
    out = [conv(0)] * n
    toks = list(tokens)
    for i in range(n):
        if not toks:
            break
        t = toks.pop(0)
        if conv is int:
            m = _INT.match(t)
            if not m:
                break
            out[i] = int(m.group())
            if m.end() < len(t):
                toks.insert(0, t[m.end():])
        else:
            try:
                out[i] = float(t)
            except ValueError:
                break
    return out


def parse(text: str) -> dict:
    lines = text.split("\n")
    mode = "none"
    gap = [0.0, 0.0]
    mat = [1, 1]
    gap_set = matrix_set = False
    P = {"pos": [], "size": [], "gap": [], "matrix": [], "recv_type": 0,
         "recv_pos": [0.0] * 3, "recv_size": [0.0] * 3, "recv_normal": [0.0, 0.0, 1.0],
         "recv_face": 0, "ground": [0.0, 0.0]}
    grids = []
    n_recv = 0
    in_recv = in_grid = False
    next_first = 0
    i = 0
    while i < len(lines):
        line = lines[i]
        i += 1
        if line == "" or line[0] == "#":
            continue
        tok = line.split()
        head = tok[0] if tok else ""
        rest = tok[1:]
        if head == "ground":
            mode = "ground"
            P["ground"] = _nums(rest, 2)
            continue
        if head == "Recv":
            mode = "recv"
            t = _nums(rest, 1, int)[0]
            if not 0 <= t <= 2:
                raise ValueError(f"receiver type {t} is not defined")
            in_recv = n_recv == 0
            if in_recv:
                P.update(recv_type=t, recv_pos=[0.0] * 3, recv_size=[0.0] * 3,
                         recv_normal=[0.0, 0.0, 1.0], recv_face=0)
            continue
        if head == "Grid":
            mode = "grid"
            gap_set = matrix_set = False
            grids.append({"type": _nums(rest, 1, int)[0], "pos": [0.0] * 3, "size": [0.0] * 3,
                          "interval": [1.0] * 3, "n": 0, "helio_type": 0, "first": next_first,
                          "listed": 0, "skipped": 0})
            in_grid = True
            continue
        key = KEYS.get(head, 0)
        if mode == "recv":
            if not in_recv:
                continue
            if key == 0:
                P["recv_pos"] = _nums(rest, 3)
            elif key == 1:
                P["recv_size"] = _nums(rest, 3)
            elif key == 2:
                P["recv_normal"] = _nums(rest, 3)
            elif key == 3:
                P["recv_face"] = _nums(rest, 1, int)[0]
            elif key == 4:
                n_recv += 1
                in_recv = False
        elif mode == "grid":
            g = grids[-1]
            if g["type"] != 0:
                raise ValueError(f"grid type {g['type']} is not supported")
            if not in_grid:
                continue
            if key == 0:
                g["pos"] = _nums(rest, 3)
            elif key == 1:
                g["size"] = _nums(rest, 3)
            elif key == 8:
                g["interval"] = _nums(rest, 3)
            elif key == 9:
                g["n"] = _nums(rest, 1, int)[0]
                next_first += g["n"]
            elif key == 10:
                g["helio_type"] = _nums(rest, 1, int)[0]
            elif key == 4:
                in_grid = False
                mode = "helio"
        elif mode == "helio":
            g = grids[-1]
            if key == 5:
                gap = _nums(rest, 2)
                gap_set = True
            elif key == 6:
                mat = _nums(rest, 2, int)
                matrix_set = True
            elif key == 7:
                if not (gap_set and matrix_set) or g["helio_type"] != 0:
                    g["skipped"] += 1
                    continue
                P["pos"].append(_nums(rest, 3))
                nxt = lines[i] if i < len(lines) else ""
                i += 1
                P["size"].append(_nums(nxt.split(), 3))
                P["gap"].append(list(gap))
                P["matrix"].append(list(mat))
                g["listed"] += 1
    if not grids:
        raise ValueError("scene file has no grid")
    if n_recv == 0:
        raise ValueError("scene file has no receiver")
    for k, g in enumerate(grids):
        if g["helio_type"] != 0 and (g["n"] > 0 or g["skipped"] > 0):
            raise ValueError(f"grid {k}: heliostat type {g['helio_type']} is not supported")
        if g["n"] < 0 or g["listed"] != g["n"]:
            raise ValueError(f"grid {k} declares {g['n']} heliostats and lists {g['listed']}")
    H = len(P["pos"])
    f32 = lambda v: np.array(v, np.float32)
    return {
        "helio_pos": np.array(P["pos"], np.float32).reshape(H, 3),
        "helio_size": np.array(P["size"], np.float32).reshape(H, 3),
        "helio_gap": np.array(P["gap"], np.float32).reshape(H, 2),
        "helio_matrix": np.array(P["matrix"], np.int32).reshape(H, 2),
        "recv_type": P["recv_type"], "recv_pos": f32(P["recv_pos"]),
        "recv_size": f32(P["recv_size"]),
        "recv_normal": f32(P["recv_normal"]), "recv_face": P["recv_face"],
        "grid_pos": f32(grids[0]["pos"]),
        "grid_size": f32(grids[0]["size"]),
        "grid_interval": f32(grids[0]["interval"]),
        "grids": [{"pos": f32(g["pos"]), "size": f32(g["size"]), "interval": f32(g["interval"]),
                   "first_helio": g["first"], "n_helio": g["n"]} for g in grids],
        "ground_length": np.float32(P["ground"][0]), "ground_width": np.float32(P["ground"][1]),
    }


def parse_file(path: str) -> dict:
    with open(path, "r") as f:
        return parse(f.read())
