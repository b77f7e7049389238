"""onemirror_amd -- heliostat-field Monte-Carlo ray tracer on AMD MI355X (gfx950).

Python binding of the C ABI in include/onemirror.h (libonemirror_amd.so, built in-tree by
``python -m onemirror_amd.build``).  Device memory comes from PyTorch-ROCm tensors; every
computation runs in the HIP kernels of the library.  There is no CPU fallback: if the
library is missing or no GPU is present, the calls raise.

Names mirror the reference's interface (Hu_smooth/ClassedSunRay):
  SolarScene                    SolarScene + SceneFileProc::SceneFileRead + InitContent
  results_with_given_normal     Gtest.cu:10-63
  recthelio_ray_tracing         recthelio_tracing.cpp:46-79 (per-heliostat accumulate)
  ImageSmoother.image_smooth    image_smooth.cu:60-115
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libonemirror_amd.so")
_LIB = None

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class OMError(RuntimeError):
    pass


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("pos", ctypes.c_float * 3),
        ("size", ctypes.c_float * 3),
        ("interval", ctypes.c_float * 3),
        ("first_helio", ctypes.c_int),
        ("n_helio", ctypes.c_int),
    ]


class SceneDesc(ctypes.Structure):
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
        ("grids", ctypes.POINTER(GridDesc)),
        ("ground_length", ctypes.c_float),
        ("ground_width", ctypes.c_float),
        ("helio_normal", _f32p),
    ]


class Params(ctypes.Structure):
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
        ("smooth_radius", ctypes.c_int),
        ("smooth_ratio", ctypes.c_float),
        ("accumulate", ctypes.c_int),
        ("pack_shift", ctypes.c_int),
        ("grid_select", ctypes.c_int),
        ("literal_self", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("ray_compact", ctypes.c_int),
        ("redo_cap", ctypes.c_int),
    ]


ACC_PACKED, ACC_DEVICE, ACC_WINDOW = 0, 1, 2
PREC_FP32, PREC_FP64, PREC_MIXED = 0, 1, 2
GRIDS_ALL = -1


class Flux(ctypes.Structure):
    _fields_ = [("counts", _vp), ("energy", _vp), ("flux", _vp), ("flux_smooth", _vp),
                ("capacity", ctypes.c_int64)]


class SceneInfo(ctypes.Structure):
    _fields_ = [
        ("n_helio", ctypes.c_int),
        ("res_x", ctypes.c_int),
        ("res_y", ctypes.c_int),
        ("grid_num", ctypes.c_int * 3),
        ("grid_cells", ctypes.c_int64),
        ("grid_nnz", ctypes.c_int64),
        ("total_micro", ctypes.c_int64),
        ("total_rays", ctypes.c_int64),
        ("max_micro", ctypes.c_int),
        ("pixel_length", ctypes.c_float),
        ("flux_scale", ctypes.c_float),
        ("total_energy", ctypes.c_float),
        ("n_grids", ctypes.c_int),
        ("pixel_area", ctypes.c_float),
    ]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["grid_num"] = tuple(self.grid_num)
        return d


class Result(ctypes.Structure):
    _fields_ = [("flux", _f32p), ("flux_smooth", _f32p), ("ratio", ctypes.c_float),
                ("ratio_smooth", ctypes.c_float), ("total_energy", ctypes.c_float),
                ("rays", ctypes.c_int64)]


class Launch(ctypes.Structure):
    _fields_ = [("workgroup", ctypes.c_int), ("rays_per_thread", ctypes.c_int)]


class TuneTrial(ctypes.Structure):
    _fields_ = [("workgroup", ctypes.c_int), ("rays_per_thread", ctypes.c_int), ("ms", ctypes.c_float)]


class Tune(ctypes.Structure):
    _fields_ = [("best", Launch), ("best_ms", ctypes.c_float), ("rays_per_s", ctypes.c_double),
                ("n_trials", ctypes.c_int), ("trials", TuneTrial * 12)]


_EXPORTS = [
    "om_status_string", "om_last_error", "om_default_params", "om_ctx_create", "om_ctx_destroy",
    "om_scene_from_scn", "om_scn_parse", "om_scene_desc_free", "om_scene_create",
    "om_scene_destroy", "om_scene_get_info", "om_scene_set_sun", "om_scene_export",
    "om_trace_plain", "om_trace_heliostats", "om_trace_records", "om_flux_finalize",
    "om_smooth_trimmed", "om_results_with_given_normal", "om_intercept_ratio", "om_math_eval",
    "om_batch_create", "om_batch_run", "om_batch_info", "om_batch_destroy", "om_trace_batched",
    "om_autotune", "om_autotune_accumulate", "om_diff_mask_words", "om_diff_forward", "om_diff_backward",
    "om_diff_visibility", "om_scene_fork", "om_sun_direction", "om_dni_hottel", "om_scn_write",
    "om_master_normal", "om_facet_heliostat", "om_facet_scene_desc", "om_savetxt_map", "om_savetxt_ratio",
    "om_load_normals", "om_trace_plain_noalloc", "om_ctx_abort", "om_ctx_progress", "om_scene_redo_count",
    "om_batch_redo_count",
]


def lib():
    """Load libonemirror_amd.so; raises OMError if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise OMError(f"{LIB_PATH} is missing: build it with `python -m onemirror_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.om_status_string.restype = ctypes.c_char_p
    L.om_status_string.argtypes = [ctypes.c_int]
    L.om_last_error.restype = ctypes.c_char_p
    L.om_default_params.argtypes = [ctypes.POINTER(Params)]
    L.om_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.om_ctx_destroy.argtypes = [_vp]
    L.om_ctx_abort.argtypes = [_vp, ctypes.c_int]
    L.om_ctx_progress.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.om_scene_from_scn.argtypes = [_vp, ctypes.c_char_p, _f32p, ctypes.c_float,
                                    ctypes.POINTER(Params), ctypes.POINTER(_vp)]
    L.om_scn_parse.argtypes = [ctypes.c_char_p, ctypes.POINTER(SceneDesc)]
    L.om_scn_write.argtypes = [ctypes.c_char_p, ctypes.POINTER(SceneDesc)]
    L.om_savetxt_map.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int]
    L.om_savetxt_ratio.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_float]
    L.om_load_normals.argtypes = [ctypes.c_char_p, _f32p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.om_master_normal.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _f32p, _f32p]
    L.om_facet_heliostat.argtypes = [ctypes.c_float, ctypes.c_float, _f32p, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, _f32p, _f32p, _f32p]
    L.om_facet_scene_desc.argtypes = [ctypes.c_float, ctypes.c_float, _f32p, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, _f32p, _f32p, _f32p, _f32p, _i32p, ctypes.POINTER(SceneDesc)]
    L.om_scene_desc_free.argtypes = [ctypes.POINTER(SceneDesc)]
    L.om_scene_desc_free.restype = None
    L.om_scene_create.argtypes = [_vp, ctypes.POINTER(SceneDesc), ctypes.POINTER(Params),
                                  ctypes.POINTER(_vp)]
    L.om_scene_destroy.argtypes = [_vp]
    L.om_scene_get_info.argtypes = [_vp, ctypes.POINTER(SceneInfo)]
    L.om_scene_set_sun.argtypes = [_vp, _f32p, ctypes.c_float]
    L.om_scene_fork.argtypes = [_vp, _f32p, ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(_vp)]
    L.om_sun_direction.argtypes = [ctypes.c_float, ctypes.c_float, _f32p]
    L.om_dni_hottel.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_float]
    L.om_dni_hottel.restype = ctypes.c_float
    L.om_scene_export.argtypes = [_vp, ctypes.c_char_p, _vp, ctypes.c_size_t]
    L.om_trace_plain.argtypes = [_vp, ctypes.POINTER(Flux), _vp]
    L.om_trace_plain_noalloc.argtypes = [_vp, ctypes.POINTER(Flux), _vp]
    L.om_autotune_accumulate.argtypes = [_vp, ctypes.POINTER(Flux), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)]
    L.om_trace_heliostats.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Flux), _vp]
    L.om_scene_redo_count.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64)]
    L.om_trace_records.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                                   ctypes.c_int64, _vp]
    L.om_flux_finalize.argtypes = [_vp, ctypes.POINTER(Flux), _vp]
    L.om_smooth_trimmed.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, _vp]
    L.om_results_with_given_normal.argtypes = [_vp, ctypes.POINTER(Result)]
    L.om_intercept_ratio.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float]
    L.om_intercept_ratio.restype = ctypes.c_float
    L.om_math_eval.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]
    L.om_batch_create.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(Flux),
                                  ctypes.POINTER(Launch), ctypes.POINTER(_vp)]
    L.om_batch_run.argtypes = [_vp, _vp]
    L.om_batch_info.argtypes = [_vp, ctypes.POINTER(Launch), ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_int64)]
    L.om_batch_redo_count.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.om_batch_destroy.argtypes = [_vp]
    L.om_trace_batched.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(Flux),
                                   ctypes.POINTER(Launch), _vp]
    L.om_autotune.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(Flux),
                              ctypes.c_int, ctypes.POINTER(Tune)]
    L.om_diff_mask_words.argtypes = [_vp, ctypes.POINTER(ctypes.c_int64)]
    L.om_diff_forward.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.om_diff_backward.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.om_diff_visibility.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int64, _vp]
    for name in _EXPORTS:
        if name not in ("om_status_string", "om_last_error", "om_default_params",
                        "om_scene_desc_free", "om_intercept_ratio", "om_dni_hottel"):
            getattr(L, name).restype = ctypes.c_int
    _LIB = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        L = lib()
        raise OMError(f"{what}: {L.om_status_string(rc).decode()} ({L.om_last_error().decode()})")


def _stream_ptr(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def default_params() -> dict:
    p = Params()
    lib().om_default_params(ctypes.byref(p))
    return {n: getattr(p, n) for n, _ in Params._fields_}


def _params_struct(prm: dict) -> Params:
    base = default_params()
    base.update(prm)
    p = Params()
    for n, t in Params._fields_:
        setattr(p, n, float(base[n]) if t is ctypes.c_float else int(base[n]))
    return p


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in np.asarray(v, np.float32)])


def _desc_struct(desc: dict, keep: list) -> SceneDesc:
    arrs = [np.ascontiguousarray(desc[k], t) for k, t in
            (("helio_pos", np.float32), ("helio_size", np.float32), ("helio_gap", np.float32),
             ("helio_matrix", np.int32))]
    keep.extend(arrs)
    d = SceneDesc()
    d.n_helio = int(arrs[0].shape[0])
    d.helio_pos = arrs[0].ctypes.data_as(_f32p)
    d.helio_size = arrs[1].ctypes.data_as(_f32p)
    d.helio_gap = arrs[2].ctypes.data_as(_f32p)
    d.helio_matrix = arrs[3].ctypes.data_as(_i32p)
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
        arr = (GridDesc * len(grids))()
        for i, g in enumerate(grids):
            arr[i].pos = _f3(g["pos"])
            arr[i].size = _f3(g["size"])
            arr[i].interval = _f3(g["interval"])
            arr[i].first_helio = int(g["first_helio"])
            arr[i].n_helio = int(g["n_helio"])
        keep.append(arr)
        d.n_grids = len(grids)
        d.grids = ctypes.cast(arr, ctypes.POINTER(GridDesc))
    d.ground_length = float(desc.get("ground_length", 0.0))
    d.ground_width = float(desc.get("ground_width", 0.0))
    if desc.get("helio_normal") is not None:
        n = np.ascontiguousarray(desc["helio_normal"], np.float32).reshape(-1, 3)
        if n.shape[0] != d.n_helio:
            raise ValueError("helio_normal needs one normal per heliostat")
        keep.append(n)
        d.helio_normal = n.ctypes.data_as(_f32p)
    return d


def grids_of(desc: dict) -> list:
    """The grids of a description: its "grids" list, or the one grid of grid_pos/size/interval
    holding every heliostat."""
    if desc.get("grids"):
        return list(desc["grids"])
    return [{"pos": np.asarray(desc["grid_pos"], np.float32), "size": np.asarray(desc["grid_size"], np.float32),
             "interval": np.asarray(desc["grid_interval"], np.float32), "first_helio": 0,
             "n_helio": int(np.asarray(desc["helio_pos"]).shape[0])}]


def scn_write(path: str, desc: dict):
    """Write a description as a .scn file that SceneFileProc::SceneFileRead (and scn_parse)
    reads back to the same description (om_scn_write)."""
    keep: list = []
    full = dict(desc)
    for k, v in (("recv_pixel_length", 0.01), ("sun_dir", (0.0, -1.0, 0.0)), ("dni", 1000.0)):
        full.setdefault(k, v)
    d = _desc_struct(full, keep)
    _check(lib().om_scn_write(path.encode(), ctypes.byref(d)), "om_scn_write")


def clear_heliostats(desc: dict, prm: dict) -> np.ndarray:
    """int32 [n]: 1 for the heliostats whose shading walks the trace kernels skip (no ray
    toward the sun from the mirror can meet another heliostat), from a host-only scene
    (om_scene_create without a context; om_scene_export "helio_clear")."""
    keep: list = []
    d = _desc_struct(desc, keep)
    p = _params_struct(prm)
    h = _vp()
    _check(lib().om_scene_create(None, ctypes.byref(d), ctypes.byref(p), ctypes.byref(h)), "om_scene_create")
    try:
        out = np.zeros(int(np.asarray(desc["helio_pos"]).shape[0]), np.int32)
        _check(lib().om_scene_export(h, b"helio_clear", out.ctypes.data_as(_vp), out.nbytes), "om_scene_export")
        return out
    finally:
        lib().om_scene_destroy(h)


def master_normal(X: float, Y: float, Z: float, sun_dir) -> np.ndarray:
    """Master normal of the faceted heliostat at (X, Y, Z) for a sun direction: the bisector
    of the directions to the point (0, 13.01, 0) and to the sun (main.cpp:84-91)."""
    out = np.zeros(3, np.float32)
    _check(lib().om_master_normal(float(X), float(Y), float(Z), _f3(sun_dir), out.ctypes.data_as(_f32p)),
           "om_master_normal")
    return out


def facet_heliostat(a: float, b: float, mirror_normal, X: float, Y: float, Z: float) -> dict:
    """The nine facets of the paraboloid heliostat y = x^2/a^2 + z^2/b^2 turned to
    mirror_normal at (X, Y, Z) (scene_file_proc.cpp:288-339): pos, normal (not normalised),
    size, [9, 3] each."""
    pos, nrm, size = (np.zeros((9, 3), np.float32) for _ in range(3))
    _check(lib().om_facet_heliostat(float(a), float(b), _f3(mirror_normal), float(X), float(Y), float(Z),
                                    pos.ctypes.data_as(_f32p), nrm.ctypes.data_as(_f32p),
                                    size.ctypes.data_as(_f32p)), "om_facet_heliostat")
    return {"pos": pos, "normal": nrm, "size": size}


def facet_scene(a: float, b: float, X: float, Y: float, Z: float, sun_dir, dni: float,
                mirror_normal=None) -> dict:
    """Description of the reference's faceted scene (scene_file_proc.cpp:288-534): the nine
    facets with their normals given, its receiver and grid; the master normal defaults to
    the driver's for this sun (main.cpp:84-91)."""
    if mirror_normal is None:
        mirror_normal = master_normal(X, Y, Z, sun_dir)
    pos, nrm, size, gap = (np.zeros(n, np.float32) for n in (27, 27, 27, 18))
    mat = np.zeros(18, np.int32)
    d = SceneDesc()
    _check(lib().om_facet_scene_desc(float(a), float(b), _f3(mirror_normal), float(X), float(Y), float(Z),
                                     pos.ctypes.data_as(_f32p), nrm.ctypes.data_as(_f32p),
                                     size.ctypes.data_as(_f32p), gap.ctypes.data_as(_f32p),
                                     mat.ctypes.data_as(_i32p), ctypes.byref(d)), "om_facet_scene_desc")
    return {
        "helio_pos": pos.reshape(9, 3), "helio_size": size.reshape(9, 3), "helio_gap": gap.reshape(9, 2),
        "helio_matrix": mat.reshape(9, 2), "helio_normal": nrm.reshape(9, 3),
        "recv_type": d.recv_type, "recv_pos": np.array(d.recv_pos[:], np.float32),
        "recv_size": np.array(d.recv_size[:], np.float32), "recv_normal": np.array(d.recv_normal[:], np.float32),
        "recv_face": d.recv_face, "recv_pixel_length": np.float32(d.recv_pixel_length),
        "grid_pos": np.array(d.grid_pos[:], np.float32), "grid_size": np.array(d.grid_size[:], np.float32),
        "grid_interval": np.array(d.grid_interval[:], np.float32),
        "ground_length": np.float32(d.ground_length), "ground_width": np.float32(d.ground_width),
        "sun_dir": np.asarray(sun_dir, np.float32), "dni": np.float32(dni),
        "mirror_normal": np.asarray(mirror_normal, np.float32),
    }


def scn_parse(path: str) -> dict:
    """Parse a .scn file (SceneFileProc::SceneFileRead) into a description dict; "grids" lists
    every Grid block with its heliostat range."""
    d = SceneDesc()
    _check(lib().om_scn_parse(path.encode(), ctypes.byref(d)), "om_scn_parse")
    H = d.n_helio
    try:
        grids = [{"pos": np.array(d.grids[k].pos[:], np.float32), "size": np.array(d.grids[k].size[:], np.float32),
                  "interval": np.array(d.grids[k].interval[:], np.float32),
                  "first_helio": d.grids[k].first_helio, "n_helio": d.grids[k].n_helio}
                 for k in range(d.n_grids)]
        out = {
            "helio_pos": np.ctypeslib.as_array(d.helio_pos, (max(H, 1), 3))[:H].copy(),
            "helio_size": np.ctypeslib.as_array(d.helio_size, (max(H, 1), 3))[:H].copy(),
            "helio_gap": np.ctypeslib.as_array(d.helio_gap, (max(H, 1), 2))[:H].copy(),
            "helio_matrix": np.ctypeslib.as_array(d.helio_matrix, (max(H, 1), 2))[:H].copy(),
            "recv_type": d.recv_type, "recv_pos": np.array(d.recv_pos[:], np.float32),
            "recv_size": np.array(d.recv_size[:], np.float32),
            "recv_normal": np.array(d.recv_normal[:], np.float32), "recv_face": d.recv_face,
            "recv_pixel_length": np.float32(d.recv_pixel_length),
            "grid_pos": np.array(d.grid_pos[:], np.float32),
            "grid_size": np.array(d.grid_size[:], np.float32),
            "grid_interval": np.array(d.grid_interval[:], np.float32),
            "sun_dir": np.array(d.sun_dir[:], np.float32), "dni": np.float32(d.dni),
            "recv_res_x": d.recv_res_x, "recv_res_y": d.recv_res_y,
            "grids": grids, "ground_length": np.float32(d.ground_length),
            "ground_width": np.float32(d.ground_width),
        }
    finally:
        lib().om_scene_desc_free(ctypes.byref(d))
    return out


class Context:
    """A context on one GPU (om_ctx)."""

    def __init__(self, device: int = 0):
        import torch
        if not torch.cuda.is_available():
            raise OMError("no GPU: the tracer runs only on the HIP device")
        torch.cuda.set_device(device)
        h = _vp()
        _check(lib().om_ctx_create(int(device), ctypes.byref(h)), "om_ctx_create")
        self._h = h
        self.device = device

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.om_ctx_destroy(self._h)
            self._h = None

    def abort(self, on: bool = True):
        """Stop the running persistent launches of this context, and the later ones until
        abort(False) (om_ctx_abort); their maps hold only the work traced before."""
        _check(lib().om_ctx_abort(self._h, int(bool(on))), "om_ctx_abort")

    def progress(self) -> tuple:
        """(chunks taken, chunks in all) of the persistent launch that runs now or ran last
        (om_ctx_progress); callable from another thread while it runs."""
        t, n = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().om_ctx_progress(self._h, ctypes.byref(t), ctypes.byref(n)), "om_ctx_progress")
        return t.value, n.value

    def math_eval(self, fn: int, x):
        """Evaluate a device float function (parity self-test of the kernel arithmetic)."""
        import torch
        xin = torch.as_tensor(np.ascontiguousarray(x, np.float32)).cuda(self.device)
        out = torch.empty_like(xin)
        _check(lib().om_math_eval(self._h, int(fn), _vp(xin.data_ptr()), _vp(out.data_ptr()),
                                  int(xin.numel()), _stream_ptr()), "om_math_eval")
        torch.cuda.synchronize()
        return out.cpu().numpy()


_DEFAULT_CTX: Optional[Context] = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


class FluxMaps:
    """Device maps of one receiver (torch tensors, row-major [res_y, res_x])."""

    def __init__(self, res_x: int, res_y: int, device=None, smooth: bool = True):
        import torch
        dev = device if device is not None else torch.device("cuda")
        self.counts = torch.zeros((res_y, res_x), dtype=torch.int32, device=dev)
        self.energy = torch.zeros((res_y, res_x), dtype=torch.int64, device=dev)
        self.flux = torch.zeros((res_y, res_x), dtype=torch.float32, device=dev)
        self.flux_smooth = (torch.zeros((res_y, res_x), dtype=torch.float32, device=dev)
                            if smooth else None)

    def struct(self, with_flux: bool = True) -> Flux:
        f = Flux()
        f.counts = self.counts.data_ptr()
        f.energy = self.energy.data_ptr()
        f.flux = self.flux.data_ptr() if with_flux else None
        f.flux_smooth = (self.flux_smooth.data_ptr()
                         if (with_flux and self.flux_smooth is not None) else None)
        f.capacity = self.counts.numel()
        return f


class SolarScene:
    """A traced scene: description + parameters, built on the GPU (om_scene)."""

    def __init__(self, desc: dict, params: dict, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.desc, self.params = desc, dict(params)
        keep: list = []
        d = _desc_struct(desc, keep)
        p = _params_struct(params)
        h = _vp()
        _check(lib().om_scene_create(self.ctx._h, ctypes.byref(d), ctypes.byref(p),
                                     ctypes.byref(h)), "om_scene_create")
        self._h = h
        self._refresh()

    @classmethod
    def from_scn(cls, path: str, sun_dir, dni: float, params: dict,
                 ctx: Optional[Context] = None) -> "SolarScene":
        desc = scn_parse(path)
        desc["sun_dir"] = np.asarray(sun_dir, np.float32)
        desc["dni"] = np.float32(dni)
        return cls(desc, params, ctx)

    def _refresh(self):
        info = SceneInfo()
        _check(lib().om_scene_get_info(self._h, ctypes.byref(info)), "om_scene_get_info")
        self.info = info.as_dict()
        self.res_x, self.res_y = info.res_x, info.res_y

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.om_scene_destroy(self._h)
            self._h = None

    def set_sun(self, sun_dir, dni: float):
        _check(lib().om_scene_set_sun(self._h, _f3(sun_dir), float(dni)), "om_scene_set_sun")
        self._refresh()

    def fork(self, sun_dir, dni: float, scene_index: Optional[int] = None) -> "SolarScene":
        """This field under another sun (om_scene_fork): shares the grid, microsamples and
        tiles on the device, re-aims the heliostats.  The fork keeps its root alive."""
        h = _vp()
        idx = self.params.get("scene_index", 0) if scene_index is None else int(scene_index)
        _check(lib().om_scene_fork(self._h, _f3(sun_dir), float(dni), idx, ctypes.byref(h)), "om_scene_fork")
        f = SolarScene.__new__(SolarScene)
        f.ctx = self.ctx
        f.desc = dict(self.desc, sun_dir=np.asarray(sun_dir, np.float32), dni=np.float32(dni))
        f.params = dict(self.params, scene_index=idx)
        f.root = getattr(self, "root", None) or self
        f._h = h
        f._refresh()
        return f

    def export(self, what: str, nbytes: int, dtype) -> np.ndarray:
        buf = np.zeros(max(nbytes, 1), np.uint8)
        _check(lib().om_scene_export(self._h, what.encode(), _vp(buf.ctypes.data), buf.nbytes),
               f"om_scene_export({what})")
        return buf[:nbytes].view(dtype)

    def new_maps(self, smooth: bool = True) -> FluxMaps:
        return FluxMaps(self.res_x, self.res_y, smooth=smooth)

    def trace_plain(self, maps: Optional[FluxMaps] = None, stream=None,
                    finalize: bool = True) -> FluxMaps:
        maps = maps or self.new_maps()
        f = maps.struct(with_flux=finalize)
        _check(lib().om_trace_plain(self._h, ctypes.byref(f), _stream_ptr(stream)),
               "om_trace_plain")
        return maps

    def trace_plain_noalloc(self, maps: FluxMaps, stream=None, finalize: bool = True) -> FluxMaps:
        """trace_plain into caller maps without allocating device memory: the workspace is
        the library's static per-GPU array (om_trace_plain_noalloc).  Calls that overlap on
        one GPU race."""
        f = maps.struct(with_flux=finalize)
        _check(lib().om_trace_plain_noalloc(self._h, ctypes.byref(f), _stream_ptr(stream)),
               "om_trace_plain_noalloc")
        return maps

    def trace_heliostats(self, h0: int, h1: int, maps: FluxMaps, stream=None) -> FluxMaps:
        f = maps.struct(with_flux=False)
        _check(lib().om_trace_heliostats(self._h, int(h0), int(h1), ctypes.byref(f),
                                         _stream_ptr(stream)), "om_trace_heliostats")
        return maps

    def finalize(self, maps: FluxMaps, stream=None) -> FluxMaps:
        f = maps.struct(with_flux=True)
        _check(lib().om_flux_finalize(self._h, ctypes.byref(f), _stream_ptr(stream)),
               "om_flux_finalize")
        return maps

    def autotune_accumulate(self, maps: Optional[FluxMaps] = None, reps: int = 3) -> dict:
        """Time OM_ACC_PACKED and OM_ACC_WINDOW on this scene (identical maps) and keep the
        faster for later traces (om_autotune_accumulate).  Times in ms, -1 for a mode the
        scene cannot use."""
        maps = maps or self.new_maps(smooth=False)
        f = maps.struct(with_flux=False)
        ms = (ctypes.c_float * 2)()
        best = ctypes.c_int()
        _check(lib().om_autotune_accumulate(self._h, ctypes.byref(f), int(reps), ms, ctypes.byref(best)),
               "om_autotune_accumulate")
        self.params["accumulate"] = best.value
        return {"accumulate": {ACC_PACKED: "packed", ACC_WINDOW: "window"}[best.value],
                "packed_ms": float(ms[0]), "window_ms": float(ms[1])}

    def ray_offset(self) -> np.ndarray:
        n = self.info["n_helio"]
        return self.export("ray_offset", 8 * (n + 1), np.int64)

    def redo_count(self) -> int:
        """OM_PREC_MIXED: rays of the last trace (or records call) whose FP32 decisions fell
        within their error bounds and were traced by the FP64 trace (om_scene_redo_count)."""
        n = ctypes.c_int64(0)
        _check(lib().om_scene_redo_count(self._h, ctypes.byref(n)), "om_scene_redo_count")
        return int(n.value)

    def trace_records(self, h0: int = 0, h1: Optional[int] = None, stream=None) -> dict:
        import torch
        h1 = self.info["n_helio"] if h1 is None else h1
        off = self.ray_offset()
        n = int(off[h1] - off[h0])
        t = [torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda") for _ in range(3)]
        _check(lib().om_trace_records(self._h, int(h0), int(h1), _vp(t[0].data_ptr()),
                                      _vp(t[1].data_ptr()), _vp(t[2].data_ptr()), n,
                                      _stream_ptr(stream)), "om_trace_records")
        torch.cuda.synchronize()
        return {"shade": t[0][:n].cpu().numpy(), "block": t[1][:n].cpu().numpy(),
                "pixel": t[2][:n].cpu().numpy()}


def results_with_given_normal(scene: SolarScene, nonsmooth_path: Optional[str] = None,
                              smooth_path: Optional[str] = None) -> dict:
    """Gtest.cu:10-63: trace, scale, intercept ratio, smooth, ratio (synchronous).  With
    paths, the ratio lines are appended to them as the reference appends to
    solarenergy::result_save_nonsmooth_path / result_save_smooth_path (Gtest.cu:47,59)."""
    n = scene.res_x * scene.res_y
    fl = np.zeros(n, np.float32)
    fs = np.zeros(n, np.float32)
    r = Result()
    r.flux = fl.ctypes.data_as(_f32p)
    r.flux_smooth = fs.ctypes.data_as(_f32p)
    _check(lib().om_results_with_given_normal(scene._h, ctypes.byref(r)),
           "om_results_with_given_normal")
    out = {"flux": fl.reshape(scene.res_y, scene.res_x),
           "flux_smooth": fs.reshape(scene.res_y, scene.res_x), "ratio": r.ratio,
           "ratio_smooth": r.ratio_smooth, "total_energy": r.total_energy, "rays": r.rays}
    for path, key in ((nonsmooth_path, "flux"), (smooth_path, "flux_smooth")):
        if path:
            save_ratio_txt(path, out[key], r.total_energy)
    return out


class ImageSaver:
    """ImageSaver (image_save.h:8-13): appending text writers of host maps."""

    @staticmethod
    def savetxt(filename: str, w: int, h: int, h_data, precision_or_total=2):
        """The reference's two overloads: an int selects the map writer with that precision,
        a float the ratio line against that total energy."""
        m = np.ascontiguousarray(h_data, np.float32).reshape(h, w)
        if isinstance(precision_or_total, (int, np.integer)) and not isinstance(precision_or_total, bool):
            save_map_txt(filename, m, int(precision_or_total))
        else:
            save_ratio_txt(filename, m, float(precision_or_total))


def save_map_txt(path: str, flux, precision: int = 2):
    """om_savetxt_map: append the map as the reference's CSV (last row first)."""
    m = np.ascontiguousarray(flux, np.float32)
    h, w = m.shape
    _check(lib().om_savetxt_map(path.encode(), w, h, m.ctypes.data_as(_f32p), int(precision)), "om_savetxt_map")


def save_ratio_txt(path: str, flux, total_energy: float):
    """om_savetxt_ratio: append the intercept ratio line."""
    m = np.ascontiguousarray(flux, np.float32)
    h, w = m.shape
    _check(lib().om_savetxt_ratio(path.encode(), w, h, m.ctypes.data_as(_f32p), float(total_energy)),
           "om_savetxt_ratio")


def load_normals(path: str) -> np.ndarray:
    """load_noraml (scene_normals.cpp:3-13): the [N, 3] normals of a normals file."""
    n = ctypes.c_int()
    _check(lib().om_load_normals(path.encode(), None, 0, ctypes.byref(n)), "om_load_normals")
    out = np.zeros((max(n.value, 1), 3), np.float32)
    _check(lib().om_load_normals(path.encode(), out.ctypes.data_as(_f32p), n.value, ctypes.byref(n)),
           "om_load_normals")
    return out[: n.value]


def recthelio_ray_tracing(scene: SolarScene, helio: int, maps: FluxMaps, stream=None) -> int:
    """recthelio_ray_tracing (recthelio_tracing.cpp:46-79): accumulate one heliostat into
    the maps; returns its microsample count."""
    scene.trace_heliostats(helio, helio + 1, maps, stream)
    cnt = scene.export("micro_count", 4 * scene.info["n_helio"], np.int32)
    return int(cnt[helio])


class ImageSmoother:
    """ImageSmoother (image_smooth.cuh:17-20)."""

    @staticmethod
    def image_smooth(d_array, kernel_radius: int, trimmed_ratio: float, width: int, height: int,
                     out=None, ctx: Optional[Context] = None, stream=None):
        """Trimmed mean of a device float map; in place when ``out`` is None."""
        ctx = ctx or default_context()
        dst = d_array if out is None else out
        _check(lib().om_smooth_trimmed(ctx._h, _vp(d_array.data_ptr()), _vp(dst.data_ptr()),
                                       int(width), int(height), int(kernel_radius),
                                       float(trimmed_ratio), _stream_ptr(stream)),
               "om_smooth_trimmed")
        return dst


def sun_direction(zenith_deg: float, azimuth_deg: float) -> np.ndarray:
    """Light travel direction of the sun (main.cpp:71-73), om_sun_direction."""
    out = np.zeros(3, np.float32)
    _check(lib().om_sun_direction(float(zenith_deg), float(azimuth_deg), out.ctypes.data_as(_f32p)),
           "om_sun_direction")
    return out


def dni_hottel(zenith_deg: float, month: int, day: int, elevation_m: float) -> float:
    """Clear-sky DNI of getDNI (SPA.cpp:1174-1189), om_dni_hottel, W/m^2."""
    return float(lib().om_dni_hottel(float(zenith_deg), int(month), int(day), float(elevation_m)))


def intercept_ratio(flux: np.ndarray, total_energy: float) -> float:
    a = np.ascontiguousarray(flux, np.float32)
    h, w = a.shape
    return float(lib().om_intercept_ratio(a.ctypes.data_as(_f32p), w, h, float(total_energy)))


# ---------------------------------------------------------------------------------------
# batched mode (no reference counterpart: one persistent launch over many scenes)
# ---------------------------------------------------------------------------------------
def _batch_arrays(scenes, maps):
    n = len(scenes)
    if maps is None:
        maps = [sc.new_maps(smooth=True) for sc in scenes]
    hs = (_vp * n)(*[sc._h for sc in scenes])
    fl = (Flux * n)(*[m.struct(with_flux=m.flux is not None) for m in maps])
    return maps, hs, fl


BATCH_PER_SCENE = 1024  # Batch workgroup: the window trace of each scene in turn (OM_BATCH_PER_SCENE)
BATCH_WINDOW = 2048     # Batch workgroup: one persistent launch with LDS windows (OM_BATCH_WINDOW)


class Batch:
    """Many scenes traced by one persistent launch (workgroup 128/256/512, or BATCH_WINDOW
    for the form that keeps each scene's map window in LDS), or with
    workgroup=BATCH_PER_SCENE by the window trace of each scene in turn, replayed from a
    hipGraph (om_batch_*).

    ``maps[i]`` receives the maps of ``scenes[i]``, equal to what ``scenes[i].trace_plain``
    gives; they are overwritten by every ``run``."""

    def __init__(self, scenes, maps=None, workgroup: int = 0, rays_per_thread: int = 0):
        self.scenes = list(scenes)
        self.maps, self._hs, self._fl = _batch_arrays(self.scenes, maps)
        lp = Launch(int(workgroup), int(rays_per_thread))
        h = _vp()
        ctx = self.scenes[0].ctx
        _check(lib().om_batch_create(ctx._h, self._hs, len(self.scenes), self._fl, ctypes.byref(lp),
                                     ctypes.byref(h)), "om_batch_create")
        self._h = h
        L, it, ry = Launch(), ctypes.c_int64(), ctypes.c_int64()
        lib().om_batch_info(h, ctypes.byref(L), ctypes.byref(it), ctypes.byref(ry))
        self.workgroup, self.rays_per_thread = L.workgroup, L.rays_per_thread
        self.items, self.rays = it.value, ry.value

    def run(self, stream=None):
        _check(lib().om_batch_run(self._h, _stream_ptr(stream)), "om_batch_run")
        return self.maps

    def redo_count(self):
        """Mixed-precision persistent batches (om_batch_redo_count): the rays of the last run
        traced again in FP64, and the scenes whose redo list filled up.  Waits for the device."""
        r, f = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().om_batch_redo_count(self._h, ctypes.byref(r), ctypes.byref(f)), "om_batch_redo_count")
        return r.value, f.value

    def __del__(self):
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.om_batch_destroy(self._h)
            self._h = None


def trace_batched(scenes, maps=None, workgroup: int = 0, rays_per_thread: int = 0, stream=None):
    """One-shot batch (no graph); returns the maps.  Synchronous."""
    scenes = list(scenes)
    maps, hs, fl = _batch_arrays(scenes, maps)
    lp = Launch(int(workgroup), int(rays_per_thread))
    _check(lib().om_trace_batched(scenes[0].ctx._h, hs, len(scenes), fl, ctypes.byref(lp),
                                  _stream_ptr(stream)), "om_trace_batched")
    return maps


def autotune_launch(scenes, maps=None, reps: int = 2) -> dict:
    """Time every (workgroup, rays per thread) candidate on the batch (om_autotune)."""
    scenes = list(scenes)
    maps, hs, fl = _batch_arrays(scenes, maps)
    t = Tune()
    _check(lib().om_autotune(scenes[0].ctx._h, hs, len(scenes), fl, int(reps), ctypes.byref(t)),
           "om_autotune")
    return {"workgroup": t.best.workgroup, "rays_per_thread": t.best.rays_per_thread,
            "ms": t.best_ms, "rays_per_s": t.rays_per_s,
            "trials": [(t.trials[i].workgroup, t.trials[i].rays_per_thread, t.trials[i].ms)
                       for i in range(t.n_trials)]}


def autotune(configs, grid_cells=(64, 128, 256), grid_top: Optional[float] = None,
             sample: int = 4, reps: int = 2) -> dict:
    """Autotuner of batched mode: grid resolution (cells per axis), workgroup size and rays
    per work item.  ``configs`` is a list of (desc, params).  The grid is part of a scene, so
    it is chosen first, on up to ``sample`` of the scenes rebuilt at each resolution; the
    launch parameters are timed with ``om_autotune`` for each.  Returns the fastest and every
    trial in rays/s."""
    from . import fields as F
    configs = list(configs)
    idx = np.linspace(0, len(configs) - 1, min(sample, len(configs))).round().astype(int)
    rows = []
    best = None
    for n in grid_cells:
        scenes = []
        for i in sorted(set(idx.tolist())):
            d, prm = dict(configs[i][0]), configs[i][1]
            top = grid_top if grid_top is not None else float(d["grid_pos"][1] + d["grid_size"][1])
            d.update(F.grid_for(d["helio_pos"], d["helio_size"], top, int(n)))
            scenes.append(SolarScene(d, prm))
        r = autotune_launch(scenes, reps=reps)
        rows.append({"grid_cells": int(n), **r})
        if best is None or r["rays_per_s"] > best["rays_per_s"]:
            best = rows[-1]
        del scenes
    return {"grid_cells": best["grid_cells"], "workgroup": best["workgroup"],
            "rays_per_thread": best["rays_per_thread"], "rays_per_s": best["rays_per_s"],
            "trials": rows}
