"""Restatement of the reference's .scn parser (test infrastructure only).

SceneFileProc::SceneFileRead(SolarScene*, std::string), scene_file_proc.cpp:41-286, with
readHelioParamter (:3-11) and the key map of the constructor (:16-28):
  * a line is skipped when its first character is '#' or it is empty;
  * the first whitespace-separated token selects the block (ground / Recv / Grid) or a key;
  * a key unknown to the map reads as "pos" (std::map::operator[] inserts the default
    value 0 = StringValue::pos, :155, :187, :232);
  * a number that cannot be read is stored as 0 (std::istream extraction failure);
  * after a "helio x y z" line the next line holds the size (sx, sy, sz);
  * heliostat lines before both "gap" and "matrix" are set are ignored (:221-224).
Only the first receiver and the first grid are kept.
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
         "recv_face": 0, "grid_pos": [0.0] * 3, "grid_size": [0.0] * 3,
         "grid_interval": [1.0] * 3}
    n_recv = n_grid = 0
    in_recv = False
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
            continue
        if head == "Recv":
            mode = "recv"
            in_recv = n_recv == 0
            if in_recv:
                P["recv_type"] = _nums(rest, 1, int)[0]
            continue
        if head == "Grid":
            mode = "grid"
            gap_set = matrix_set = False
            if n_grid >= 1:
                raise ValueError("more than one grid")
            continue
        key = KEYS.get(head, 0)
        if mode == "recv":
            if not in_recv:
                if key == 4:
                    n_recv += 1
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
            if key == 0:
                P["grid_pos"] = _nums(rest, 3)
            elif key == 1:
                P["grid_size"] = _nums(rest, 3)
            elif key == 8:
                P["grid_interval"] = _nums(rest, 3)
            elif key == 4:
                n_grid += 1
                mode = "helio"
        elif mode == "helio":
            if key == 5:
                gap = _nums(rest, 2)
                gap_set = True
            elif key == 6:
                mat = _nums(rest, 2, int)
                matrix_set = True
            elif key == 7:
                if not (gap_set and matrix_set):
                    continue
                P["pos"].append(_nums(rest, 3))
                nxt = lines[i] if i < len(lines) else ""
                i += 1
                P["size"].append(_nums(nxt.split(), 3))
                P["gap"].append(list(gap))
                P["matrix"].append(list(mat))
    H = len(P["pos"])
    return {
        "helio_pos": np.array(P["pos"], np.float32).reshape(H, 3),
        "helio_size": np.array(P["size"], np.float32).reshape(H, 3),
        "helio_gap": np.array(P["gap"], np.float32).reshape(H, 2),
        "helio_matrix": np.array(P["matrix"], np.int32).reshape(H, 2),
        "recv_type": P["recv_type"], "recv_pos": np.array(P["recv_pos"], np.float32),
        "recv_size": np.array(P["recv_size"], np.float32),
        "recv_normal": np.array(P["recv_normal"], np.float32), "recv_face": P["recv_face"],
        "grid_pos": np.array(P["grid_pos"], np.float32),
        "grid_size": np.array(P["grid_size"], np.float32),
        "grid_interval": np.array(P["grid_interval"], np.float32),
    }


def parse_file(path: str) -> dict:
    with open(path, "r") as f:
        return parse(f.read())
