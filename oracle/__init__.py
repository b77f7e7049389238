"""CPU oracle for the heliostat-field ray tracer -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import this package, and only as the checker or as the host baseline.  The product
(``onemirror_amd``) never imports it and has no CPU fallback.

The compiled part (``liboracle.so``, built by ``make -C oracle``) restates the reference
routines of ``Hu_smooth/ClassedSunRay`` in plain C; see ``om_oracle.h`` for the function
list and citations.  ``scn.py`` restates the ``.scn`` parser.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


class GridDescC(ctypes.Structure):
    _fields_ = [
        ("pos", ctypes.c_float * 3),
        ("size", ctypes.c_float * 3),
        ("interval", ctypes.c_float * 3),
        ("first_helio", ctypes.c_int),
        ("n_helio", ctypes.c_int),
    ]


class SceneDescC(ctypes.Structure):
    _fields_ = [
        ("n_helio", ctypes.c_int),
        ("helio_pos", _f32p),
        ("helio_size", _f32p),
        ("helio_gap", _f32p),
        ("helio_matrix", _i32p),
        ("recv_type", ctypes.c_int),
        ("recv_pos", ctypes.c_float * 3),
        ("recv_size", ctypes.c_float * 3),
        ("recv_normal", ctypes.c_float * 3),
        ("recv_face", ctypes.c_int),
        ("recv_pixel_length", ctypes.c_float),
        ("grid_pos", ctypes.c_float * 3),
        ("grid_size", ctypes.c_float * 3),
        ("grid_interval", ctypes.c_float * 3),
        ("sun_dir", ctypes.c_float * 3),
        ("dni", ctypes.c_float),
        ("recv_res_x", ctypes.c_int),
        ("recv_res_y", ctypes.c_int),
        ("n_grids", ctypes.c_int),
        ("grids", ctypes.POINTER(GridDescC)),
        ("helio_normal", _f32p),
    ]


class ParamsC(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("scene_index", ctypes.c_uint32),
        ("sunshape", ctypes.c_int),
        ("pillbox_half_angle", ctypes.c_float),
        ("csr", ctypes.c_float),
        ("buie_bins", ctypes.c_int),
        ("slope_sigma", ctypes.c_float),
        ("micro_mode", ctypes.c_int),
        ("helio_pixel_length", ctypes.c_float),
        ("poisson_k", ctypes.c_int),
        ("rays_per_sample", ctypes.c_int),
        ("reflectivity", ctypes.c_float),
    ]


class StatsC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                ("rays", "shaded", "blocked", "hit", "dda_steps", "dda_steps_eff",
                 "cand_tests", "self_skips", "tag_rays", "tag_shade_steps_eff",
                 "tag_shade_tests", "tag_shade_self_skips", "self_hits")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


def lib():
    """Load liboracle.so (built by ``make -C oracle``); raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
    L = ctypes.CDLL(path)
    L.oc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    L.oc_philox_batch.argtypes = [_u32p, _u32p, _u32p, ctypes.c_int]
    L.oc_math_batch.argtypes = [ctypes.c_int, _f32p, _f32p, ctypes.c_int]
    L.oc_erfinv_d_batch.argtypes = [_f64p, _f64p, ctypes.c_int]
    L.oc_buie_table_d.argtypes = [ctypes.c_double, ctypes.c_int, _f64p, _f64p, _f64p, _f64p]
    L.oc_buie_theta_d_batch.argtypes = [ctypes.c_double, ctypes.c_int, _f64p, _f64p, ctypes.c_int]
    L.oc_scene_build.argtypes = [ctypes.POINTER(SceneDescC), ctypes.POINTER(ParamsC),
                                 ctypes.POINTER(ctypes.c_void_p)]
    L.oc_scene_build.restype = ctypes.c_int
    L.oc_scene_free.argtypes = [ctypes.c_void_p]
    L.oc_scene_counts.argtypes = [ctypes.c_void_p, _i64p]
    L.oc_scene_helio.argtypes = [ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p]
    L.oc_scene_grid.argtypes = [ctypes.c_void_p, _i32p, _i32p, _i32p]
    L.oc_scene_grid_k.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32p, _i32p, _i32p, _i64p]
    L.oc_scene_grid_k.restype = ctypes.c_int
    L.oc_master_normal.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _f32p, _f32p]
    L.oc_master_normal.restype = ctypes.c_int
    L.oc_facet_heliostat.argtypes = [ctypes.c_float, ctypes.c_float, _f32p, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, _f32p, _f32p, _f32p]
    L.oc_facet_heliostat.restype = ctypes.c_int
    L.oc_scene_receiver.argtypes = [ctypes.c_void_p, _f32p, _f32p, _f32p]
    L.oc_scene_sun.argtypes = [ctypes.c_void_p, _f32p, _f32p]
    L.oc_scene_buie_table.argtypes = [ctypes.c_void_p, _f32p, _f32p, _f32p]
    L.oc_scene_buie_table.restype = ctypes.c_int
    L.oc_scene_micro.argtypes = [ctypes.c_void_p, _i32p, _i64p]
    L.oc_scene_poisson.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, ctypes.c_int]
    L.oc_scene_poisson.restype = ctypes.c_int
    L.oc_scene_poisson_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, _f32p, ctypes.c_int, _i32p, _f32p]
    L.oc_scene_poisson_grid.restype = ctypes.c_int
    L.oc_scene_set_literal_self.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.oc_scene_set_ray_stream.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.oc_scene_flux_scale.argtypes = [ctypes.c_void_p, _f32p, _f32p]
    L.oc_scene_set_tags.argtypes = [ctypes.c_void_p, _i32p]
    L.oc_trace_f32.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _u32p, _u64p,
                               _i32p, _i32p, _i32p, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(StatsC)]
    L.oc_trace_f64.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _u32p, _f64p,
                               _i32p, _i32p, _i32p, ctypes.c_int, ctypes.POINTER(StatsC)]
    L.oc_trace_pooled.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_uint64, _f64p, ctypes.c_int]
    L.oc_energy_to_float.argtypes = [_u64p, ctypes.c_int, _f32p]
    L.oc_trimmed_mean.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_float]
    L.oc_intercept_ratio.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float]
    L.oc_intercept_ratio.restype = ctypes.c_float
    L.oc_dda_probe.argtypes = [ctypes.c_void_p, _f32p, _f32p, ctypes.c_int, _i32p,
                               ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.oc_dda_probe.restype = ctypes.c_int
    L.oc_local2world.argtypes = [_f32p, _f32p, _f32p]
    L.oc_rotate_y.argtypes = [_f32p, _f32p, _f32p, _f32p]
    L.oc_ray_parallelogram.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.oc_ray_parallelogram.restype = ctypes.c_int
    L.oc_cylinder_pixel_batch.argtypes = [_f32p, _f32p, ctypes.c_int, _f32p, ctypes.c_float,
                                          ctypes.c_float, ctypes.c_int, ctypes.c_int, _i32p]
    L.oc_cylinder_pixel_batch.restype = ctypes.c_int
    L.oc_cone_pixel_batch.argtypes = [_f32p, _f32p, ctypes.c_int, _f32p, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int, _i32p]
    L.oc_cone_pixel_batch.restype = ctypes.c_int
    L.oc_diff_splat64.argtypes = [ctypes.c_void_p, _f64p, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_uint8), _f64p, _f64p, ctypes.c_int]
    L.oc_diff_splat64.restype = ctypes.c_int
    L.oc_sample_ray.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                _f32p, _f32p]
    _LIB = L
    return L


def _p(a: Optional[np.ndarray], t):
    if a is None:
        return ctypes.cast(None, t)
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in np.asarray(v, np.float32)])


# ---------------------------------------------------------------------------------------
# math
# ---------------------------------------------------------------------------------------
_MATH = {"u01": 0, "sin2pi": 1, "cos2pi": 2, "sin_small": 3, "cos_small": 4, "log": 5,
         "exp": 6, "erfinv": 7, "gauss": 8, "sqrt": 9, "div": 10, "atan2": 11}


def erfinv64(x) -> np.ndarray:
    """The FP64 model's erfinv (om_oracle.c erfinv_d) element-wise."""
    xd = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(xd)
    lib().oc_erfinv_d_batch(_p(xd, _f64p), _p(out, _f64p), int(xd.size))
    return out


def buie_table64(csr: float, ng: int) -> dict:
    """The FP64 model's Buie table (om_oracle.c oc_buie_table_d): cx, cy [ng + 1], slope [ng]
    and the tail's int465, A, B, C, inv_gp1."""
    cx, cy, sl, c = np.zeros(ng + 1), np.zeros(ng + 1), np.zeros(ng), np.zeros(5)
    lib().oc_buie_table_d(float(csr), int(ng), _p(cx, _f64p), _p(cy, _f64p), _p(sl, _f64p), _p(c, _f64p))
    return dict(cx=cx, cy=cy, slope=sl, **dict(zip(("int465", "A", "B", "C", "inv_gp1"), c.tolist())))


def buie_theta64(csr: float, ng: int, u) -> np.ndarray:
    """The FP64 model's sunshape angle in mrad of the draws u (om_oracle.c buie_theta_d)."""
    ud = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty_like(ud)
    lib().oc_buie_theta_d_batch(float(csr), int(ng), _p(ud, _f64p), _p(out, _f64p), int(ud.size))
    return out


def math(fn: str, x) -> np.ndarray:
    """Evaluate one of the project-written float functions element-wise."""
    if fn == "u01":
        xi = np.ascontiguousarray(x, dtype=np.uint32)
        xf = xi.view(np.float32)
    else:
        xf = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(xf)
    lib().oc_math_batch(_MATH[fn], _p(xf, _f32p), _p(out, _f32p), int(xf.size))
    if fn in ("div", "atan2"):  # pairs (in[2i], in[2i+1])
        return out[: xf.size // 2].copy()
    return out


def philox(ctr, key) -> np.ndarray:
    """Philox4x32-10 of counters ``ctr`` ([n,4] uint32) under ``key`` (2 uint32)."""
    c = np.ascontiguousarray(np.atleast_2d(ctr), dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.empty_like(c)
    lib().oc_philox_batch(_p(c, _u32p), _p(k, _u32p), _p(out, _u32p), int(c.shape[0]))
    return out


# ---------------------------------------------------------------------------------------
# scene
# ---------------------------------------------------------------------------------------
def _desc_struct(desc: dict, keep: list):
    H = int(desc["helio_pos"].shape[0])
    arrs = {
        "helio_pos": np.ascontiguousarray(desc["helio_pos"], np.float32),
        "helio_size": np.ascontiguousarray(desc["helio_size"], np.float32),
        "helio_gap": np.ascontiguousarray(desc["helio_gap"], np.float32),
        "helio_matrix": np.ascontiguousarray(desc["helio_matrix"], np.int32),
    }
    keep.extend(arrs.values())
    d = SceneDescC()
    d.n_helio = H
    d.helio_pos = _p(arrs["helio_pos"], _f32p)
    d.helio_size = _p(arrs["helio_size"], _f32p)
    d.helio_gap = _p(arrs["helio_gap"], _f32p)
    d.helio_matrix = _p(arrs["helio_matrix"], _i32p)
    d.recv_type = int(desc.get("recv_type", 0))
    d.recv_pos = _f3(desc["recv_pos"])
    d.recv_size = _f3(desc["recv_size"])
    d.recv_normal = _f3(desc["recv_normal"])
    d.recv_face = int(desc.get("recv_face", 0))
    d.recv_pixel_length = float(desc["recv_pixel_length"])
    d.grid_pos = _f3(desc["grid_pos"])
    d.grid_size = _f3(desc["grid_size"])
    d.grid_interval = _f3(desc["grid_interval"])
    d.sun_dir = _f3(desc["sun_dir"])
    d.dni = float(desc["dni"])
    d.recv_res_x = int(desc.get("recv_res_x", 0))
    d.recv_res_y = int(desc.get("recv_res_y", 0))
    grids = desc.get("grids") or []
    if grids:
        arr = (GridDescC * len(grids))()
        for i, g in enumerate(grids):
            arr[i].pos = _f3(g["pos"])
            arr[i].size = _f3(g["size"])
            arr[i].interval = _f3(g["interval"])
            arr[i].first_helio = int(g["first_helio"])
            arr[i].n_helio = int(g["n_helio"])
        keep.append(arr)
        d.n_grids = len(grids)
        d.grids = ctypes.cast(arr, ctypes.POINTER(GridDescC))
    if desc.get("helio_normal") is not None:
        n = np.ascontiguousarray(desc["helio_normal"], np.float32).reshape(H, 3)
        keep.append(n)
        d.helio_normal = _p(n, _f32p)
    return d


def _params_struct(prm: dict):
    p = ParamsC()
    p.seed = int(prm["seed"])
    p.scene_index = int(prm.get("scene_index", 0))
    p.sunshape = int(prm["sunshape"])
    p.pillbox_half_angle = float(prm.get("pillbox_half_angle", 4.65e-3))
    p.csr = float(prm.get("csr", 0.1))
    p.buie_bins = int(prm.get("buie_bins", 1024))
    p.slope_sigma = float(prm["slope_sigma"])
    p.micro_mode = int(prm["micro_mode"])
    p.helio_pixel_length = float(prm["helio_pixel_length"])
    p.poisson_k = int(prm.get("poisson_k", 8))
    p.rays_per_sample = int(prm["rays_per_sample"])
    p.reflectivity = float(prm.get("reflectivity", 0.88))
    return p


class OracleScene:
    """An oracle scene built from a scene description and tracing parameters (the same
    dictionaries the product takes)."""

    def __init__(self, desc: dict, prm: dict):
        L = lib()
        keep: list = []
        d = _desc_struct(desc, keep)
        p = _params_struct(prm)
        h = ctypes.c_void_p()
        rc = L.oc_scene_build(ctypes.byref(d), ctypes.byref(p), ctypes.byref(h))
        if rc != 0:
            raise ValueError(f"oc_scene_build failed: {rc}")
        self._h = h
        self.desc, self.prm = desc, prm
        if prm.get("literal_self"):
            L.oc_scene_set_literal_self(self._h, 1)
        c = np.zeros(8, np.int64)
        L.oc_scene_counts(self._h, _p(c, _i64p))
        (self.n_helio, self.cells, self.nnz, self.res_x, self.res_y, self.total_micro,
         self.total_rays, self.max_micro) = [int(v) for v in c]

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.oc_scene_free(self._h)
            self._h = None

    def set_tags(self, tags=None):
        """Heliostats whose rays and shading-walk work oc_stats' tag_* fields count."""
        if tags is None:
            lib().oc_scene_set_tags(self._h, None)
            return
        t = np.ascontiguousarray(tags, np.int32)
        assert t.shape == (self.n_helio,)
        lib().oc_scene_set_tags(self._h, _p(t, _i32p))

    # --- accessors ---------------------------------------------------------------------
    def helio(self):
        H = self.n_helio
        n, v0, v1, v3 = (np.zeros((H, 3), np.float32) for _ in range(4))
        lib().oc_scene_helio(self._h, _p(n, _f32p), _p(v0, _f32p), _p(v1, _f32p), _p(v3, _f32p))
        return {"normal": n, "v0": v0, "v1": v1, "v3": v3}

    def grid(self, k: int = 0):
        """CSR of grid k: cell_start [cells+1], cell_items [nnz] (global heliostat ids)."""
        gn = np.zeros(3, np.int32)
        nnz = np.zeros(1, np.int64)
        if lib().oc_scene_grid_k(self._h, int(k), None, None, _p(gn, _i32p), _p(nnz, _i64p)) != 0:
            raise IndexError(f"no grid {k}")
        cells = int(np.prod(gn.astype(np.int64)))
        cs = np.zeros(cells + 1, np.int32)
        ci = np.zeros(max(int(nnz[0]), 1), np.int32)
        lib().oc_scene_grid_k(self._h, int(k), _p(cs, _i32p), _p(ci, _i32p), _p(gn, _i32p), None)
        return {"cell_start": cs, "cell_items": ci[: int(nnz[0])], "grid_num": gn}

    def receiver(self):
        v = np.zeros((4, 3), np.float32)
        f = np.zeros(3, np.float32)
        pl = np.zeros(1, np.float32)
        lib().oc_scene_receiver(self._h, _p(v, _f32p), _p(f, _f32p), _p(pl, _f32p))
        return {"vertices": v, "focus": f, "pixel_length": float(pl[0]),
                "res": (self.res_x, self.res_y)}

    def sun(self):
        s = np.zeros(3, np.float32)
        c = np.zeros(6, np.float32)
        lib().oc_scene_sun(self._h, _p(s, _f32p), _p(c, _f32p))
        return {"sun_dir": s, "int465": c[0], "A": c[1], "B": c[2], "C": c[3],
                "inv_gp1": c[4], "gamma": c[5]}

    def buie_table(self):
        n = int(self.prm.get("buie_bins", 1024))
        x = np.zeros(n + 1, np.float32)
        y = np.zeros(n + 1, np.float32)
        k = np.zeros(n, np.float32)
        lib().oc_scene_buie_table(self._h, _p(x, _f32p), _p(y, _f32p), _p(k, _f32p))
        return {"cdf_x": x, "cdf_y": y, "slope": k}

    def micro(self):
        c = np.zeros(self.n_helio, np.int32)
        o = np.zeros(self.n_helio + 1, np.int64)
        lib().oc_scene_micro(self._h, _p(c, _i32p), _p(o, _i64p))
        return {"count": c, "ray_offset": o}

    def poisson(self, h: int) -> np.ndarray:
        cap = self.max_micro
        pts = np.zeros((max(cap, 1), 2), np.float32)
        n = lib().oc_scene_poisson(self._h, int(h), _p(pts, _f32p), cap)
        if n < 0:
            raise ValueError("scene has no Poisson microsamples")
        return pts[:n].copy()

    def poisson_grid(self, h: int) -> dict:
        """Heliostat h's Poisson cell grid before compaction: 'grid' [rows][cols][2] (a point
        per cell, or -2 cell for none; the padding ring included), 'r', 'cell', and 'n' the
        accepted point count."""
        dims = np.zeros(2, np.int32)
        rc = np.zeros(2, np.float32)
        ncell = lib().oc_scene_poisson_grid(self._h, int(h), None, 0, _p(dims, _i32p), _p(rc, _f32p))
        if ncell < 0:
            raise ValueError("no such heliostat")
        g = np.zeros((int(dims[1]), int(dims[0]), 2), np.float32)
        n = lib().oc_scene_poisson_grid(self._h, int(h), _p(g, _f32p), int(ncell), _p(dims, _i32p),
                                        _p(rc, _f32p))
        return {"grid": g, "r": float(rc[0]), "cell": float(rc[1]), "n": int(n)}

    def set_ray_stream(self, scene_index: int):
        """Rays from Philox stream scene_index; the microsamples keep the build's stream (the
        draws of a scene fork)."""
        lib().oc_scene_set_ray_stream(self._h, int(scene_index))

    def set_literal_self(self, on: bool = True):
        """Walks test the ray's own heliostat like any candidate (the reference's literal
        Intersect loop) instead of skipping it."""
        lib().oc_scene_set_literal_self(self._h, int(bool(on)))

    def flux_scale(self):
        s = np.zeros(1, np.float32)
        t = np.zeros(1, np.float32)
        lib().oc_scene_flux_scale(self._h, _p(s, _f32p), _p(t, _f32p))
        return float(s[0]), float(t[0])

    def sample_ray(self, h, m, l):
        a = np.zeros(3, np.float32)
        b = np.zeros(3, np.float32)
        lib().oc_sample_ray(self._h, int(h), int(m), int(l), _p(a, _f32p), _p(b, _f32p))
        return a, b

    def dda_probe(self, orig, direction, self_index=-1, cap=4096):
        o = np.ascontiguousarray(orig, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        cells = np.zeros(cap, np.int32)
        hit = ctypes.c_int(-1)
        n = lib().oc_dda_probe(self._h, _p(o, _f32p), _p(d, _f32p), int(self_index),
                               _p(cells, _i32p), cap, ctypes.byref(hit))
        return cells[: min(n, cap)].copy(), int(hit.value)

    # --- tracing -----------------------------------------------------------------------
    def _range(self, h0, h1):
        h1 = self.n_helio if h1 is None else h1
        off = self.micro()["ray_offset"]
        return h0, h1, off

    def trace_f32(self, h0: int = 0, h1: Optional[int] = None, records: bool = False,
                  early_exit: bool = True, nthreads: int = 0) -> dict:
        h0, h1, off = self._range(h0, h1)
        npx = self.res_x * self.res_y
        counts = np.zeros(npx, np.uint32)
        energy = np.zeros(npx, np.uint64)
        recs = [None, None, None]
        if records:
            nr = int(off[h1] - off[h0])
            recs = [np.full(nr, -1, np.int32) for _ in range(3)]
        st = StatsC()
        lib().oc_trace_f32(self._h, int(h0), int(h1), _p(counts, _u32p), _p(energy, _u64p),
                           _p(recs[0], _i32p), _p(recs[1], _i32p), _p(recs[2], _i32p),
                           int(bool(early_exit)), int(nthreads), ctypes.byref(st))
        out = {"counts": counts.reshape(self.res_y, self.res_x),
               "energy": energy.reshape(self.res_y, self.res_x), "stats": st.as_dict()}
        if records:
            out.update(shade=recs[0], block=recs[1], pixel=recs[2])
        return out

    def trace_f64(self, h0: int = 0, h1: Optional[int] = None, records: bool = False,
                  nthreads: int = 0) -> dict:
        h0, h1, off = self._range(h0, h1)
        npx = self.res_x * self.res_y
        counts = np.zeros(npx, np.uint32)
        energy = np.zeros(npx, np.float64)
        recs = [None, None, None]
        if records:
            nr = int(off[h1] - off[h0])
            recs = [np.full(nr, -1, np.int32) for _ in range(3)]
        st = StatsC()
        lib().oc_trace_f64(self._h, int(h0), int(h1), _p(counts, _u32p), _p(energy, _f64p),
                           _p(recs[0], _i32p), _p(recs[1], _i32p), _p(recs[2], _i32p),
                           int(nthreads), ctypes.byref(st))
        out = {"counts": counts.reshape(self.res_y, self.res_x),
               "energy": energy.reshape(self.res_y, self.res_x), "stats": st.as_dict()}
        if records:
            out.update(shade=recs[0], block=recs[1], pixel=recs[2])
        return out

    def diff_splat64(self, theta, vis, h0: int = 0, h1: Optional[int] = None, nthreads: int = 0):
        """FP64 splat map and energy of the differentiable model for heliostats [h0, h1) at
        parameters theta [H, 4], visibility vis (one byte per ray of the range)."""
        h1 = self.n_helio if h1 is None else h1
        th = np.ascontiguousarray(theta, np.float64).reshape(-1, 4)
        v = np.ascontiguousarray(vis, np.uint8)
        off = self.micro()["ray_offset"]
        assert v.size == int(off[h1] - off[h0])
        S = np.zeros(self.res_x * self.res_y, np.float64)
        E = np.zeros(1, np.float64)
        rc = lib().oc_diff_splat64(self._h, _p(th, _f64p), int(h0), int(h1),
                                   v.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                   _p(S, _f64p), _p(E, _f64p), int(nthreads))
        if rc != 0:
            raise ValueError("differentiable model needs a rectangle receiver")
        return S.reshape(self.res_y, self.res_x), float(E[0])

    def trace_pooled(self, h0=0, h1=None, groups=128, lights=None, pool_seed=12345,
                     nthreads=0) -> np.ndarray:
        h1 = self.n_helio if h1 is None else h1
        lights = int(self.prm["rays_per_sample"]) if lights is None else int(lights)
        e = np.zeros(h1 - h0, np.float64)
        rc = lib().oc_trace_pooled(self._h, int(h0), int(h1), int(groups), lights,
                                   int(pool_seed), _p(e, _f64p), int(nthreads))
        if rc != 0:
            raise ValueError("pool smaller than rays per sample")
        return e


def master_normal(X, Y, Z, sun_dir) -> np.ndarray:
    out = np.zeros(3, np.float32)
    s = np.ascontiguousarray(sun_dir, np.float32)
    if lib().oc_master_normal(float(X), float(Y), float(Z), _p(s, _f32p), _p(out, _f32p)) != 0:
        raise ValueError("degenerate master normal")
    return out


def facet_heliostat(a, b, mirror_normal, X, Y, Z) -> dict:
    n = np.ascontiguousarray(mirror_normal, np.float32)
    pos, nrm, size = (np.zeros((9, 3), np.float32) for _ in range(3))
    if lib().oc_facet_heliostat(float(a), float(b), _p(n, _f32p), float(X), float(Y), float(Z),
                                _p(pos, _f32p), _p(nrm, _f32p), _p(size, _f32p)) != 0:
        raise ValueError("degenerate facet heliostat")
    return {"pos": pos, "normal": nrm, "size": size}


def energy_to_float(energy: np.ndarray) -> np.ndarray:
    e = np.ascontiguousarray(energy, np.uint64)
    out = np.empty(e.shape, np.float32)
    lib().oc_energy_to_float(_p(e, _u64p), int(e.size), _p(out, _f32p))
    return out


def trimmed_mean(img: np.ndarray, kernel_radius: int, trimmed_ratio: float) -> np.ndarray:
    a = np.ascontiguousarray(img, np.float32)
    h, w = a.shape
    out = np.empty_like(a)
    lib().oc_trimmed_mean(_p(a, _f32p), _p(out, _f32p), w, h, int(kernel_radius),
                          float(trimmed_ratio))
    return out


def intercept_ratio(flux: np.ndarray, total_energy: float) -> float:
    a = np.ascontiguousarray(flux, np.float32)
    h, w = a.shape
    return float(lib().oc_intercept_ratio(_p(a, _f32p), w, h, float(np.float32(total_energy))))


def local2world(d, n):
    out = np.zeros(3, np.float32)
    lib().oc_local2world(_p(np.ascontiguousarray(d, np.float32), _f32p),
                         _p(np.ascontiguousarray(n, np.float32), _f32p), _p(out, _f32p))
    return out


def rotate_y(o, old, new):
    out = np.zeros(3, np.float32)
    lib().oc_rotate_y(*[_p(np.ascontiguousarray(a, np.float32), _f32p) for a in (o, old, new)],
                      _p(out, _f32p))
    return out


def ray_parallelogram(orig, d, A, B, C):
    tuv = np.zeros(3, np.float32)
    r = lib().oc_ray_parallelogram(*[_p(np.ascontiguousarray(a, np.float32), _f32p)
                                     for a in (orig, d, A, B, C)], _p(tuv, _f32p))
    return bool(r), tuv


def cone_pixel(orig, d, base, r0, k, h, inv_area, res_x, res_y) -> np.ndarray:
    """Truncated-cone receiver pixel (or -1) of each ray: orig, d are [n, 3]."""
    o = np.ascontiguousarray(orig, np.float32).reshape(-1, 3)
    dd = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(base, np.float32)
    out = np.empty(o.shape[0], np.int32)
    lib().oc_cone_pixel_batch(_p(o, _f32p), _p(dd, _f32p), int(o.shape[0]), _p(b, _f32p), float(r0), float(k),
                              float(h), float(inv_area), int(res_x), int(res_y), _p(out, _i32p))
    return out


def cylinder_pixel(orig, d, base, r2, h, res_x, res_y) -> np.ndarray:
    """Cylinder receiver pixel (or -1) of each ray: orig, d are [n, 3]."""
    o = np.ascontiguousarray(orig, np.float32).reshape(-1, 3)
    dd = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
    b = np.ascontiguousarray(base, np.float32)
    out = np.empty(o.shape[0], np.int32)
    lib().oc_cylinder_pixel_batch(_p(o, _f32p), _p(dd, _f32p), int(o.shape[0]), _p(b, _f32p),
                                  float(r2), float(h), int(res_x), int(res_y), _p(out, _i32p))
    return out
