"""Text outputs and the normals file (CPU): ImageSaver::savetxt (image_save.cpp:5-47) and
load_noraml (scene_normals.cpp:3-13).

The library's writers must append exactly the text the restatement in oracle/io.py gives;
the restatement is pinned by reading its text back: the first line is the map's last row,
values below 1e-6 are written as "0", every other value is within half a unit of the last
printed digit, and the ratio line is the float sum of the printed pixels over the total.
The normals loader reads the reference's own userData/normal.txt (tests/golden).
"""
import os

import numpy as np
import pytest

import oracle.io as oio
import onemirror_amd as om

HERE = os.path.dirname(os.path.abspath(__file__))


def _map(seed=0, h=37, w=53):
    rng = np.random.default_rng(seed)
    m = rng.gamma(1.5, 300.0, (h, w)).astype(np.float32)
    m[rng.random((h, w)) < 0.3] = 0.0
    m[0, :3] = np.float32([1e-6, 9.99e-7, 1.0000001e-6])  # around the threshold
    m[1, 0] = np.float32(-5.0)
    m[2, 0] = np.float32(123456.789)
    return m


@pytest.mark.parametrize("precision", [0, 2, 6])
def test_map_text_matches_the_restatement(tmp_path, precision):
    m = _map(precision)
    path = str(tmp_path / "map.txt")
    om.save_map_txt(path, m, precision)
    om.save_map_txt(path, m[::-1].copy(), precision)  # appends
    want = oio.savetxt_map(m.shape[1], m.shape[0], m, precision) + \
        oio.savetxt_map(m.shape[1], m.shape[0], m[::-1].copy(), precision)
    assert open(path).read() == want


def test_map_text_reads_back(tmp_path):
    m = _map(3)
    h, w = m.shape
    text = oio.savetxt_map(w, h, m, 2)
    rows = text.split("\n")[:-1]
    assert len(rows) == h
    back = np.array([[float(x) for x in r.split(",")[:-1]] for r in rows])[::-1]
    assert back.shape == (h, w) and all(r.endswith(",") for r in rows)
    small = m.astype(np.float64) < 1e-6
    assert np.all(back[small] == 0.0)
    assert np.all(np.abs(back[~small] - m[~small]) <= 0.005 + 1e-9 * np.abs(m[~small]))
    assert rows[-1].startswith("0,0,%.2f," % float(m[0, 2]))


def test_ratio_line(tmp_path):
    m = _map(5)
    h, w = m.shape
    total = float(m.astype(np.float64).sum() * 1.37)
    path = str(tmp_path / "ratio.txt")
    om.save_ratio_txt(path, m, total)
    om.ImageSaver.savetxt(path, w, h, m, np.float32(total))
    om.ImageSaver.savetxt(str(tmp_path / "csv.txt"), w, h, m)  # the int overload: the map
    line = oio.savetxt_ratio(w, h, m, total)
    assert open(path).read() == line + line
    assert open(str(tmp_path / "csv.txt")).read() == oio.savetxt_map(w, h, m, 2)
    kept = m.astype(np.float64)
    kept = kept[kept >= 1e-6].sum()
    assert abs(float(line) - kept / total) < 1e-5
    assert abs(float(line) - om.lib().om_intercept_ratio(m.ctypes.data_as(om._f32p), w, h, total)) < 1e-6


def test_reference_normals_file():
    path = os.path.join(HERE, "golden", "normal.txt")
    n = om.load_normals(path)
    assert n.shape == (9, 3)
    assert np.array_equal(n, oio.load_normals(open(path).read()))
    assert np.array_equal(n, np.tile(np.float32([0.64424, 0.436177, -0.62825]), (9, 1)))


def test_normals_file_errors(tmp_path):
    p = tmp_path / "n.txt"
    p.write_text("3\n0 1 0\n1 0 0\n")
    with pytest.raises(om.OMError, match="fewer normals"):
        om.load_normals(str(p))
    with pytest.raises(ValueError):
        oio.load_normals(p.read_text())
    p.write_text("x\n")
    with pytest.raises(om.OMError):
        om.load_normals(str(p))
    with pytest.raises(om.OMError):
        om.load_normals(str(tmp_path / "missing.txt"))
    p.write_text("0\n")
    assert om.load_normals(str(p)).shape == (0, 3)
