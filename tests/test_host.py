"""Host side of the product library (no GPU): the C ABI exports, the .scn parser and the
scene construction (aiming, grid, receiver, sun tables, flux scale) against the oracle,
and the error paths of the ABI."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import oracle.scn
import onemirror_amd as om
from onemirror_amd import fields as F

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SCN = os.path.join(HERE, "golden", "face2face_shadow.scn")


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "onemirror.h")).read()
    names = sorted(set(re.findall(r"\b(om_[a-z_]+)\s*\(", hdr)))
    assert len(names) >= 20
    L = ctypes.CDLL(om.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


class HostScene:
    """Host-only product scene (om_scene_create with ctx = NULL)."""

    def __init__(self, desc, prm):
        keep = []
        D = om._desc_struct(desc, keep)
        P = om._params_struct(prm)
        h = ctypes.c_void_p()
        rc = om.lib().om_scene_create(None, ctypes.byref(D), ctypes.byref(P), ctypes.byref(h))
        self.rc = rc
        self.h = h
        if rc == 0:
            info = om.SceneInfo()
            om.lib().om_scene_get_info(h, ctypes.byref(info))
            self.info = info.as_dict()

    def export(self, what, n, dt):
        b = np.zeros(n, np.uint8)
        rc = om.lib().om_scene_export(self.h, what.encode(), ctypes.c_void_p(b.ctypes.data), n)
        assert rc == 0, om.lib().om_last_error()
        return b.view(dt)

    def __del__(self):
        if self.rc == 0:
            om.lib().om_scene_destroy(self.h)


def _bits(a):
    return np.ascontiguousarray(a).ravel().view(np.uint32)


def _compare_scene(desc, prm):
    g = HostScene(desc, prm)
    assert g.rc == 0, om.lib().om_last_error()
    o = oracle.OracleScene(desc, prm)
    H = o.n_helio
    oh = o.helio()
    for k in ("normal", "v0", "v1", "v3"):
        assert np.array_equal(_bits(g.export("helio_" + k, 12 * H, np.float32)), _bits(oh[k])), k
    og = o.grid()
    assert g.info["grid_cells"] == o.cells and g.info["grid_nnz"] == o.nnz
    assert np.array_equal(g.export("grid_start", 4 * (o.cells + 1), np.int32), og["cell_start"])
    assert np.array_equal(g.export("grid_items", 4 * o.nnz, np.int32), og["cell_items"])
    rv = g.export("receiver", 60, np.float32)
    orv = o.receiver()
    assert np.array_equal(_bits(rv[:12]), _bits(orv["vertices"].ravel()))
    assert np.array_equal(_bits(rv[12:]), _bits(orv["focus"]))
    assert (g.info["res_x"], g.info["res_y"]) == orv["res"]
    nb = int(prm.get("buie_bins", 1024))
    b = g.export("buie", 4 * (3 * nb + 2), np.float32)
    ob = o.buie_table()
    assert np.array_equal(_bits(b), _bits(np.concatenate([ob["cdf_x"], ob["cdf_y"], ob["slope"]])))
    s = g.export("sun", 32, np.float32)
    os_ = o.sun()
    assert np.array_equal(_bits(s), _bits(np.concatenate(
        [os_["sun_dir"], np.float32([os_["int465"], os_["A"], os_["B"], os_["C"], os_["inv_gp1"]])])))
    scale, total = o.flux_scale()
    assert np.float32(g.info["flux_scale"]) == np.float32(scale)
    assert np.float32(g.info["total_energy"]) == np.float32(total)
    if prm["micro_mode"] == F.LATTICE:
        assert g.info["total_rays"] == o.total_rays
        assert np.array_equal(g.export("micro_count", 4 * H, np.int32), o.micro()["count"])
        assert np.array_equal(g.export("ray_offset", 8 * (H + 1), np.int64), o.micro()["ray_offset"])


def test_scene_small_matches_oracle():
    d, p = F.config_small()
    _compare_scene(d, F.with_lattice(p))
    _compare_scene(d, p)


def test_scene_mid_matches_oracle():
    d, p = F.config_mid()
    _compare_scene(d, F.with_lattice(p))


def test_scene_cylinder_matches_oracle():
    d, p = F.config_small_cylinder()
    _compare_scene(d, F.with_lattice(p))
    _compare_scene(d, p)
    g = HostScene(d, p)
    assert (g.info["res_x"], g.info["res_y"]) == (128, 64)
    # unrolled pixel area (2 pi r / 128) x (h / 64)
    assert abs(g.info["flux_scale"] * (2 * np.pi * 2.0 / 128) * (4.0 / 64) /
               (1000.0 * 0.125 ** 2 * 0.88 / 64) - 1.0) < 1e-5


def test_scene_large_matches_oracle():
    d, p = F.config_large()
    _compare_scene(d, F.with_lattice(p))


def test_scene_faceted_heliostats_match_oracle():
    d, p = F.config_small()
    d = dict(d)
    d["helio_matrix"] = np.tile(np.int32([2, 3]), (100, 1))
    d["helio_gap"] = np.tile(np.float32([0.05, 0.04]), (100, 1))
    _compare_scene(d, F.with_lattice(p))


def test_scene_face2face_matches_oracle():
    d = om.scn_parse(SCN)
    d["sun_dir"] = F.sun_direction(40.0, 160.0)
    for face, nrm in ((0, [0, 0, 1]), (1, [1, 0, 0.2]), (2, [0.3, 0, -1]), (3, [-1, 0, 0])):
        dd = dict(d)
        dd["recv_face"], dd["recv_normal"] = face, np.float32(nrm)
        p = dict(om.default_params())
        p.update(helio_pixel_length=0.05, rays_per_sample=16)
        _compare_scene(dd, p)


def _same_desc(a: dict, b: dict):
    """Descriptions equal key by key, the grid lists entry by entry."""
    for k in a:
        if k == "grids":
            assert len(a[k]) == len(b[k]), k
            for ga, gb in zip(a[k], b[k]):
                for f in ga:
                    assert np.array_equal(np.asarray(ga[f]), np.asarray(gb[f])), (k, f, ga[f], gb[f])
        else:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (k, a[k], b[k])


def test_scn_parser_matches_restatement(tmp_path):
    a = oracle.scn.parse_file(SCN)
    b = om.scn_parse(SCN)
    _same_desc(a, b)
    assert len(b["grids"]) == 1 and b["grids"][0]["n_helio"] == 1
    # quirks: unknown key reads as "pos"; a blank-only line is the key "" (= pos);
    # heliostats before gap/matrix are skipped; failed numbers read as 0; size on next line
    txt = "\n".join([
        "# comment", "ground 10 10", "ngrid 1", "",
        "Recv 0", "pos 1 2 3", "size 4 4 0.5", "norm 0 0 1", "face 2", "end",
        "Grid 0", "pos -50 0 -50", "size 100 20 100", "inter 5 5 5", "n 2", "type 0", "end",
        "helio 1 1 1", "2 0.1 2",
        "gap 0.1 0.2", "matrix 2 x",
        "helio 3 1.5 -7", "\t2.5 0.1 2.5",
        "bogus 9 9 9",
        "   ",
        "helio 4 1.5 abc", "3 0.1 3",
    ])
    p = tmp_path / "quirks.scn"
    p.write_text(txt)
    a = oracle.scn.parse_file(str(p))
    b = om.scn_parse(str(p))
    _same_desc(a, b)
    assert b["helio_pos"].shape == (2, 3)
    assert np.array_equal(b["helio_matrix"], np.int32([[2, 0], [2, 0]]))
    assert np.array_equal(b["helio_pos"][1], np.float32([4, 1.5, 0]))
    assert b["recv_face"] == 2


def test_abi_error_paths(tmp_path):
    L = om.lib()
    d, p = F.config_small()
    # bad parameters
    for bad in (dict(rays_per_sample=0), dict(helio_pixel_length=-1.0), dict(sunshape=7),
                dict(micro_mode=5), dict(buie_bins=0, sunshape=1)):
        pp = dict(p)
        pp.update(bad)
        assert HostScene(d, pp).rc == 1, bad
    # receiver face out of range, zero grid interval
    dd = dict(d)
    dd["recv_face"] = 9
    assert HostScene(dd, p).rc == 1
    dd = dict(d)
    dd["grid_interval"] = np.float32([0, 1, 1])
    assert HostScene(dd, p).rc == 1
    # missing file; files the reference reads into undefined behaviour
    desc = om.SceneDesc()
    assert L.om_scn_parse(b"/nonexistent/x.scn", ctypes.byref(desc)) == 2
    bad = tmp_path / "bad.scn"
    for text, rc, msg in (
            ("Recv 0\nend\nGrid 0\nn 2\nend\ngap 0 0\nmatrix 1 1\nhelio 0 1 0\n1 1 1\n", 3, b"declares 2"),
            ("Recv 0\nend\nGrid 0\nn 1\ntype 1\nend\ngap 0 0\nmatrix 1 1\nhelio 0 1 0\n1 1 1\n", 6,
             b"heliostat type 1"),
            ("Recv 5\nend\nGrid 0\nend\n", 3, b"receiver type 5"),
            ("Recv 0\nend\nGrid 1\npos 0 0 0\nend\n", 6, b"grid type 1"),
            ("Grid 0\nend\n", 3, b"no receiver"),
            ("Recv 0\nend\n", 3, b"no grid")):
        bad.write_text(text)
        assert L.om_scn_parse(str(bad).encode(), ctypes.byref(desc)) == rc, text
        assert msg in L.om_last_error(), (text, L.om_last_error())
        with pytest.raises(ValueError):
            oracle.scn.parse(text)
    # a host-only scene cannot be traced
    g = HostScene(d, F.with_lattice(p))
    f = om.Flux()
    buf = np.zeros(16, np.uint64)
    f.counts = f.energy = buf.ctypes.data
    f.capacity = 1 << 20
    assert L.om_trace_plain(g.h, ctypes.byref(f), None) == 1
    assert b"host-only" in L.om_last_error()
    # status strings
    assert L.om_status_string(7) == b"buffer too small"


def test_scene_extent_checks():
    """Scenes the walk cannot index safely are refused: a mirror reaching outside the grid
    (its rays would start outside the cell arrays), coordinates beyond 2^20 m, cell edges
    outside [2^-20, 2^20] m, a zero sun vector."""
    d, p = F.config_small()
    assert HostScene(d, p).rc == 0
    dd = dict(d)
    pos = np.array(d["helio_pos"], np.float32)
    pos[3, 1] = 0.2  # a 2 m mirror centred 0.2 m above the grid floor (y = 0)
    dd["helio_pos"] = pos
    g = HostScene(dd, p)
    assert g.rc == 6 and b"heliostat 3 reaches outside the grid on axis y" in om.lib().om_last_error()
    dd = dict(d)
    dd["grid_pos"] = np.float32([-128.0, -2.0e6, -106.0])
    assert HostScene(dd, p).rc == 6
    dd = dict(d)
    dd["grid_interval"] = np.float32([4.0, 2.0, 1e-7])
    assert HostScene(dd, p).rc == 6
    dd = dict(d)
    dd["sun_dir"] = np.float32([0, 0, 0])
    assert HostScene(dd, p).rc == 1


def test_intercept_ratio_matches_oracle():
    rng = np.random.default_rng(0)
    flux = rng.gamma(1.0, 5.0, (64, 48)).astype(np.float32)
    flux[rng.random(flux.shape) < 0.4] = 0.0
    flux[0, :5] = np.float32(9e-7)
    assert om.intercept_ratio(flux, 12345.0) == oracle.intercept_ratio(flux, 12345.0)
