"""Multi-grid fields and the .scn format (CPU): the oracle's grids, the library's scene
build over several grids, the .scn parser on multi-grid files, and the .scn writer.

The reference keeps every Grid block of a .scn file (scene_file_proc.cpp:110-138,185-219),
each with the heliostats listed after it (start_helio_pos_, num_helios_), and traces the
heliostats of grid 0 over grid 0 (Gtest.cu:12,21-23).  A heliostat's rays walk the grid
that holds it (recthelio_tracing.cpp:46-79).  Grid k > 0 does not work in the reference:
its CSR holds global heliostat ids while the vertex array it indexes is re-based to the
grid's first heliostat (grid.cpp:53, steps_for_raytracing.cu:161-177).  The oracle and the
library index one global vertex array.

How the oracle's grids are pinned (its single-grid path is pinned by the golden vectors
and the closed-form checks of the other oracle tests):
  * grid 0 of a two-grid field traces exactly as the field of grid 0's heliostats alone;
  * grid 1 traces exactly as a one-grid field holding every heliostat in grid 1's box: the
    other sub-field lies more than a cell outside that box, so the box bins none of it,
    and the heliostat indices (hence the Philox streams) are the same.
"""
import ctypes

import numpy as np
import pytest

import oracle
import oracle.scn
import onemirror_amd as om
from onemirror_amd import fields as F

N_EACH = 50


def _one_grid(desc: dict, k: int) -> dict:
    """The description with grid k's box as its only grid, holding every heliostat."""
    d = dict(desc)
    g = desc["grids"][k]
    d.update(grid_pos=g["pos"], grid_size=g["size"], grid_interval=g["interval"])
    d.pop("grids")
    return d


def _same_trace(a: dict, b: dict, records: bool = True):
    assert np.array_equal(a["counts"], b["counts"])
    assert np.array_equal(a["energy"], b["energy"])
    if records:
        for k in ("shade", "block", "pixel"):
            assert np.array_equal(a[k], b[k]), k


@pytest.fixture(scope="module")
def two():
    d, p = F.config_two_grids(n_each=N_EACH)
    return d, F.with_lattice(p)


def test_two_grid_field_is_separated(two):
    """Each grid holds its own sub-field and bins none of the other."""
    d, p = two
    o = oracle.OracleScene(d, p)
    for k in range(2):
        items = o.grid(k)["cell_items"]
        assert items.size > 0
        assert items.min() >= k * N_EACH and items.max() < (k + 1) * N_EACH
    w = oracle.OracleScene(_one_grid(d, 1), p)
    items = w.grid(0)["cell_items"]
    assert items.min() >= N_EACH  # grid 1's box bins no heliostat of grid 0


def test_oracle_grid0_traces_as_its_sub_field(two):
    d, p = two
    y = oracle.OracleScene(d, p)
    x = oracle.OracleScene(F.subset(_one_grid(d, 0), slice(0, N_EACH)), p)
    ty, tx = y.trace_f32(0, N_EACH, records=True), x.trace_f32(0, N_EACH, records=True)
    _same_trace(ty, tx)
    st = ty["stats"]
    assert st["shaded"] > 0 and st["blocked"] > 0 and st["hit"] > 0


def test_oracle_grid1_traces_as_one_grid_over_its_box(two):
    d, p = two
    y = oracle.OracleScene(d, p)
    w = oracle.OracleScene(_one_grid(d, 1), p)
    ty = y.trace_f32(N_EACH, 2 * N_EACH, records=True)
    tw = w.trace_f32(N_EACH, 2 * N_EACH, records=True)
    _same_trace(ty, tw)
    # every blocker of a grid-1 ray is a grid-1 heliostat
    for k in ("shade", "block"):
        b = ty[k][ty[k] >= 0]
        assert b.size > 0 and b.min() >= N_EACH
    # the FP64 model walks the same grids (its float64 sums run in thread order)
    fy, fw = y.trace_f64(N_EACH, 2 * N_EACH, records=True), w.trace_f64(N_EACH, 2 * N_EACH, records=True)
    for k in ("counts", "shade", "block", "pixel"):
        assert np.array_equal(fy[k], fw[k]), k
    assert np.allclose(fy["energy"], fw["energy"], rtol=1e-12, atol=0.0)


def test_oracle_whole_field_is_the_sum_of_its_grids(two):
    d, p = two
    y = oracle.OracleScene(d, p)
    t = y.trace_f32(records=True)
    a, b = y.trace_f32(0, N_EACH), y.trace_f32(N_EACH, 2 * N_EACH)
    assert np.array_equal(t["counts"], a["counts"] + b["counts"])
    assert np.array_equal(t["energy"], a["energy"] + b["energy"])


class HostScene:
    """A host-only library scene (no context): aiming, grids and receiver, no tracing."""

    def __init__(self, desc, prm):
        self.keep = []
        d = om._desc_struct(desc, self.keep)
        p = om._params_struct(prm)
        self.h = ctypes.c_void_p()
        self.rc = om.lib().om_scene_create(None, ctypes.byref(d), ctypes.byref(p), ctypes.byref(self.h))

    def __del__(self):
        if self.h:
            om.lib().om_scene_destroy(self.h)

    def info(self):
        i = om.SceneInfo()
        assert om.lib().om_scene_get_info(self.h, ctypes.byref(i)) == 0
        return i.as_dict()

    def export(self, what, dtype, n):
        buf = np.zeros(n, dtype)
        rc = om.lib().om_scene_export(self.h, what.encode(), buf.ctypes.data, buf.nbytes)
        assert rc == 0, (what, om.lib().om_last_error())
        return buf


def test_library_grids_match_the_oracle(two):
    d, p = two
    o = oracle.OracleScene(d, p)
    g = HostScene(d, dict(p, grid_select=om.GRIDS_ALL))
    assert g.rc == 0, om.lib().om_last_error()
    info = g.info()
    assert info["n_grids"] == 2 and info["n_helio"] == 2 * N_EACH
    assert info["grid_cells"] == o.cells and info["grid_nnz"] == o.nnz
    for k in range(2):
        og = o.grid(k)
        gn = g.export(f"grid_num.{k}", np.int32, 3)
        assert np.array_equal(gn, og["grid_num"])
        cells = int(np.prod(gn.astype(np.int64)))
        assert np.array_equal(g.export(f"grid_start.{k}", np.int32, cells + 1), og["cell_start"])
        assert np.array_equal(g.export(f"grid_items.{k}", np.int32, og["cell_items"].size), og["cell_items"])
        assert list(g.export(f"grid_helio.{k}", np.int32, 2)) == [k * N_EACH, N_EACH]
    bad = np.zeros(4, np.int32)
    assert om.lib().om_scene_export(g.h, b"grid_num.2", bad.ctypes.data, bad.nbytes) == 1
    assert om.lib().om_scene_export(g.h, b"grid_num.x", bad.ctypes.data, bad.nbytes) == 1
    # the same aims and flux normalisation as the oracle
    om_n = g.export("helio_normal", np.float32, 6 * N_EACH).reshape(-1, 3)
    oracle_n = o.helio()["normal"]
    assert np.array_equal(om_n, oracle_n)
    scale, total = o.flux_scale()
    assert np.float32(info["flux_scale"]) == np.float32(scale)
    assert np.float32(info["total_energy"]) == np.float32(total)


@pytest.mark.parametrize("k", [0, 1])
def test_grid_select_is_the_sub_field(two, k):
    """grid_select = k: grid k and its heliostats alone (k = 0 is the reference's trace)."""
    d, p = two
    sub = F.subset(_one_grid(d, k), slice(k * N_EACH, (k + 1) * N_EACH))
    o = oracle.OracleScene(sub, p)
    g = HostScene(d, dict(p, grid_select=k))
    assert g.rc == 0, om.lib().om_last_error()
    info = g.info()
    assert info["n_grids"] == 1 and info["n_helio"] == N_EACH
    og = o.grid(0)
    assert np.array_equal(g.export("grid_items", np.int32, og["cell_items"].size), og["cell_items"])
    assert np.array_equal(g.export("helio_normal", np.float32, 3 * N_EACH).reshape(-1, 3), o.helio()["normal"])
    assert np.float32(info["total_energy"]) == np.float32(o.flux_scale()[1])


def test_grid_descriptions_are_checked(two):
    d, p = two
    for bad in ({"first_helio": 1}, {"n_helio": N_EACH + 1}, {"n_helio": -1}, {"interval": np.float32([0, 1, 1])}):
        dd = dict(d)
        dd["grids"] = [dict(d["grids"][0], **bad), d["grids"][1]]
        assert HostScene(dd, dict(p, grid_select=om.GRIDS_ALL)).rc == 1, bad
    assert HostScene(d, dict(p, grid_select=2)).rc == 1
    assert HostScene(d, dict(p, grid_select=-2)).rc == 1
    # a heliostat outside the grid that holds it
    dd = dict(d)
    dd["grids"] = [d["grids"][1], dict(d["grids"][0], first_helio=N_EACH)]
    dd["grids"][0] = dict(dd["grids"][0], first_helio=0)
    g = HostScene(dd, dict(p, grid_select=om.GRIDS_ALL))
    assert g.rc == 6 and b"outside the grid" in om.lib().om_last_error()


def _scn_text(desc: dict) -> str:
    """A multi-grid .scn file in the reference's layout (written here by hand)."""
    out = ["# two grids", "ground 200 200", "ngrid 2", "",
           "Recv 0", "pos %r %r %r" % tuple(float(x) for x in desc["recv_pos"]),
           "size %r %r %r" % tuple(float(x) for x in desc["recv_size"]),
           "norm %r %r %r" % tuple(float(x) for x in desc["recv_normal"]), "face 0", "end"]
    for g in desc["grids"]:
        out += ["", "Grid 0", "pos %r %r %r" % tuple(float(x) for x in g["pos"]),
                "size %r %r %r" % tuple(float(x) for x in g["size"]),
                "inter %r %r %r" % tuple(float(x) for x in g["interval"]),
                "n %d" % g["n_helio"], "type 0", "end", "gap 0 0", "matrix 1 1"]
        for i in range(g["first_helio"], g["first_helio"] + g["n_helio"]):
            out += ["helio %r %r %r" % tuple(float(x) for x in desc["helio_pos"][i]),
                    "\t%r %r %r" % tuple(float(x) for x in desc["helio_size"][i])]
    return "\n".join(out) + "\n"


def test_multigrid_scn_parses_as_the_restatement(tmp_path, two):
    d, _ = two
    path = tmp_path / "two.scn"
    path.write_text(_scn_text(d))
    a, b = oracle.scn.parse_file(str(path)), om.scn_parse(str(path))
    assert len(b["grids"]) == 2 and b["ground_length"] == 200
    for k in ("helio_pos", "helio_size", "helio_gap", "helio_matrix", "recv_pos", "recv_size", "grid_pos"):
        assert np.array_equal(a[k], b[k]), k
        if k.startswith("helio") or k.startswith("recv"):
            assert np.array_equal(b[k], np.asarray(d[k], b[k].dtype)), k
    for ga, gb, gd in zip(a["grids"], b["grids"], d["grids"]):
        for f in ("pos", "size", "interval", "first_helio", "n_helio"):
            assert np.array_equal(np.asarray(ga[f]), np.asarray(gb[f])), f
            assert np.array_equal(np.asarray(gb[f]), np.asarray(gd[f])), f
    # gap and matrix are forgotten at each Grid block: without them the second grid's
    # heliostat lines are skipped and its count no longer matches
    text = _scn_text(d).split("\n")
    second = [i for i, ln in enumerate(text) if ln == "Grid 0"][1]
    j = text.index("gap 0 0", second)
    del text[j:j + 2]
    path.write_text("\n".join(text))
    with pytest.raises(om.OMError, match="grid 1 declares 50 heliostats and lists 0"):
        om.scn_parse(str(path))
    with pytest.raises(ValueError, match="grid 1 declares"):
        oracle.scn.parse_file(str(path))


def _written_equals(desc: dict, back: dict):
    for k in ("helio_pos", "helio_size", "helio_gap", "helio_matrix", "recv_pos", "recv_size", "recv_normal"):
        a = np.asarray(desc[k])
        assert np.array_equal(back[k].view(np.uint32 if back[k].dtype == np.float32 else back[k].dtype),
                              a.astype(back[k].dtype).view(np.uint32 if back[k].dtype == np.float32
                                                           else back[k].dtype)), k
    assert back["recv_type"] == int(desc.get("recv_type", 0))
    assert back["recv_face"] == int(desc.get("recv_face", 0))
    grids = om.grids_of(desc)
    assert len(back["grids"]) == len(grids)
    for ga, gb in zip(grids, back["grids"]):
        for f in ("pos", "size", "interval"):
            assert np.array_equal(np.asarray(ga[f], np.float32).view(np.uint32), gb[f].view(np.uint32)), f
        assert (ga["first_helio"], ga["n_helio"]) == (gb["first_helio"], gb["n_helio"])


@pytest.mark.parametrize("name", ["small", "mid", "large", "small_cylinder", "two_grids", "batched0"])
def test_scn_writer_round_trip(tmp_path, name):
    """A generated field written as .scn reads back bit for bit, through the library's
    parser and the restatement of SceneFileRead."""
    if name == "batched0":
        descs, _ = F.config_batched(n_scenes=2, h_range=(500, 600))
        desc = descs[0]
    else:
        desc, _ = getattr(F, "config_" + name)()
    path = str(tmp_path / f"{name}.scn")
    om.scn_write(path, desc)
    back = om.scn_parse(path)
    _written_equals(desc, back)
    ref = oracle.scn.parse_file(path)
    for k in ("helio_pos", "helio_size", "helio_gap", "helio_matrix"):
        assert np.array_equal(ref[k], back[k]), k
    # and once more: writing what was read gives the same file
    path2 = str(tmp_path / f"{name}_2.scn")
    om.scn_write(path2, back)
    assert open(path).read() == open(path2).read()


def test_scn_writer_round_trips_the_reference_scene(tmp_path):
    import os
    src = os.path.join(os.path.dirname(__file__), "golden", "face2face_shadow.scn")
    a = om.scn_parse(src)
    path = str(tmp_path / "f2f.scn")
    om.scn_write(path, a)
    b = om.scn_parse(path)
    _written_equals(a, b)
    assert b["ground_length"] == a["ground_length"] == 1024


def test_scn_writer_keeps_gap_and_matrix_per_heliostat(tmp_path):
    d, _ = F.config_small()
    n = d["helio_pos"].shape[0]
    d = dict(d)
    d["helio_gap"] = np.float32(np.c_[np.arange(n) % 3 * 0.05, -0.0 * np.ones(n)])
    d["helio_matrix"] = np.int32(np.c_[1 + np.arange(n) % 2, 1 + np.arange(n) // 50])
    path = str(tmp_path / "facets.scn")
    om.scn_write(path, d)
    _written_equals(d, om.scn_parse(path))
    assert np.array_equal(oracle.scn.parse_file(path)["helio_gap"], d["helio_gap"])


def test_scn_writer_refuses_what_cannot_be_read_back(tmp_path):
    d, _ = F.config_small()
    bad = dict(d)
    bad["helio_pos"] = d["helio_pos"].copy()
    bad["helio_pos"][3, 0] = np.inf
    with pytest.raises(om.OMError):
        om.scn_write(str(tmp_path / "x.scn"), bad)
    two, _ = F.config_two_grids()
    two = dict(two)
    two["grids"] = [dict(two["grids"][0], n_helio=10), two["grids"][1]]
    with pytest.raises(om.OMError):
        om.scn_write(str(tmp_path / "y.scn"), two)
    with pytest.raises(om.OMError):
        om.scn_write("/nonexistent/dir/z.scn", d)
