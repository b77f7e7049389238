"""The C ABI seen from compiled C++ (CPU; no GPU needed).

tests/abi_layout.cpp includes include/onemirror.h, is built with g++ (not the HIP compiler)
and linked against libonemirror_amd.so.  It reports the size, alignment and field offsets of
every public struct, which must equal the Python ctypes mirrors field by field (a drift in
either fails here); and the results of the entry points that need no GPU, which must equal
the same calls made through ctypes.  The C++ driver of INTEGRATION.md (the reference's entry
points written on top of the ABI) must compile and link against the library as printed there.
"""
import ctypes
import json
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import onemirror_amd as om

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "onemirror_amd")
SCN = os.path.join(ROOT, "tests", "golden", "face2face_shadow.scn")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(om.LIB_PATH),
                                reason="needs g++ and the built library")

MIRRORS = {
    "om_grid_desc": om.GridDesc, "om_scene_desc": om.SceneDesc, "om_params": om.Params,
    "om_flux": om.Flux, "om_scene_info": om.SceneInfo, "om_result": om.Result,
    "om_launch": om.Launch, "om_tune": om.Tune,
}


def _link_flags():
    return ["-L", LIBDIR, "-lonemirror_amd", f"-Wl,-rpath,{LIBDIR}"]


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    exe = str(d / "abi_layout")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", INCLUDE,
                    os.path.join(ROOT, "tests", "abi_layout.cpp"), "-o", exe] + _link_flags(), check=True)
    out = subprocess.run([exe, SCN], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


@pytest.mark.parametrize("name", list(MIRRORS))
def test_struct_layout_matches_ctypes(report, name):
    c = report["layout"][name]
    py = MIRRORS[name]
    assert ctypes.sizeof(py) == c["size"], (name, ctypes.sizeof(py), c["size"])
    assert ctypes.alignment(py) == c["align"], name
    py_fields = [f[0] for f in py._fields_]
    assert py_fields == list(c["fields"]), (name, py_fields, list(c["fields"]))
    for f, (off, size) in c["fields"].items():
        desc = getattr(py, f)
        assert (desc.offset, desc.size) == (off, size), (name, f, (desc.offset, desc.size), (off, size))


def test_host_calls_match_ctypes(report):
    calls = report["calls"]
    p = om.default_params()
    for k, v in calls["default_params"].items():
        assert p[k] == pytest.approx(v, rel=1e-7, abs=0), k
    assert calls["status_size"] == om.lib().om_status_string(7).decode()
    assert calls["sun_direction"][0] == 0
    np.testing.assert_array_equal(np.float32(calls["sun_direction"][1:]), om.sun_direction(30.0, 135.0))
    assert np.float32(calls["dni_hottel"]) == np.float32(om.dni_hottel(40.0, 6, 21, 1500.0))
    h = calls["scn_host_scene"]
    assert h["status"] == 0 and h["n_helio"] >= 1 and h["res_x"] > 0 and h["res_y"] > 0
    desc = om.scn_parse(SCN)
    assert h["n_helio"] == int(np.asarray(desc["helio_pos"]).shape[0])
    assert calls["null_info_status"] == 1  # OM_ERR_INVALID, no exit


def test_integration_driver_compiles_and_links(tmp_path):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = re.search(r"```cpp\n(.*?)```", text, re.S)
    assert block, "INTEGRATION.md has no C++ driver"
    src = tmp_path / "driver.cpp"
    # every function of the driver is referenced, so that the link resolves its ABI calls
    names = re.findall(r"^\w[\w\s\*]*?\b(\w+)\(", block.group(1), re.M)
    refs = "".join(f"    (void)&{n};\n" for n in names if n not in ("check",))
    src.write_text(block.group(1) + "\nint main() {\n" + refs + "    return 0;\n}\n")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE,
                    "-I", os.path.join(rocm, "include"), str(src), "-o", str(tmp_path / "driver"),
                    "-L", os.path.join(rocm, "lib"), "-lamdhip64"] + _link_flags(), check=True)
