"""The faceted heliostat of the annual driver and given heliostat normals (CPU).

The reference's annual driver (main.cpp:56-116) traces one paraboloid heliostat made of
nine flat facets (SceneFileProc::SceneFileRead(SolarScene*, a, b, mirrorNormal, X, Y, Z),
scene_file_proc.cpp:288-534) whose normals are given to InitContent(normals)
(solar_scene.cpp:101-115), turned to a master normal that bisects the directions to the
receiver point and to the sun (main.cpp:84-91).

Pins (the library's builder is compared bit for bit with the oracle's restatement; both
are pinned by what the geometry fixes):
  * the law of reflection: the sun reflected about the master normal points at
    (0, 13.01, 0);
  * the centre facet sits at (X, Y, Z) with the master normal, exactly;
  * every facet centre, taken back through the matrix, lies on y = x^2/a^2 + z^2/b^2 at
    the lattice points +-(sqrt(2)/3 + 0.1);
  * opposite facets add up to twice their height along the master normal;
  * for a master normal with |n1| = |n2| the matrix is a rotation: distances from the
    centre and the tilt of each facet normal from the master normal are those of the
    paraboloid (elsewhere its horizontal column has the length
    sqrt(n0^2 + n2^2) / sqrt(n0^2 + n1^2) that the reference's normalisation gives it).
"""
import ctypes

import numpy as np
import pytest

import oracle
import onemirror_amd as om
from onemirror_amd import fields as F

E = np.sqrt(2.0) / 3 + 0.1
LX = np.array([-E, 0, E, -E, 0, E, -E, 0, E])
LZ = np.array([-E, -E, -E, 0, 0, 0, E, E, E])


def _cases(n=40, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        zen, az = rng.uniform(5, 75), rng.uniform(60, 300)
        sun = F.sun_direction(90.0 - zen, az)
        X, Y, Z = rng.uniform(-60, 60), rng.uniform(0.5, 3), rng.uniform(60, 200)
        a, b = rng.uniform(8, 40, 2)
        out.append((float(a), float(b), sun, float(X), float(Y), float(Z)))
    return out


@pytest.mark.parametrize("case", range(40))
def test_builder_matches_the_restatement(case):
    a, b, sun, X, Y, Z = _cases()[case]
    m = om.master_normal(X, Y, Z, sun)
    assert np.array_equal(m.view(np.uint32), oracle.master_normal(X, Y, Z, sun).view(np.uint32))
    g, o = om.facet_heliostat(a, b, m, X, Y, Z), oracle.facet_heliostat(a, b, m, X, Y, Z)
    for k in ("pos", "normal", "size"):
        assert np.array_equal(g[k].view(np.uint32), o[k].view(np.uint32)), k


def test_master_normal_reflects_the_sun_onto_the_receiver_point():
    for a, b, sun, X, Y, Z in _cases(seed=1):
        n = om.master_normal(X, Y, Z, sun).astype(np.float64)
        s = np.asarray(sun, np.float64)
        assert abs(np.linalg.norm(n) - 1) < 1e-6
        r = s - 2 * np.dot(s, n) * n
        want = np.array([-X, 13.01 - Y, -Z])
        want /= np.linalg.norm(want)
        assert np.abs(r - want).max() < 2e-6, (r, want)


def _matrix(n):
    """The matrix of scene_file_proc.cpp:302-325 in float64."""
    t = np.sqrt(n[0] ** 2 + n[1] ** 2)
    if n[2] > 0:
        c0 = np.array([n[2] / t, 0.0, -n[0] / t])
    elif n[2] < 0:
        c0 = np.array([-n[2] / t, 0.0, n[0] / t])
    else:
        c0 = np.array([0.0, 0.0, 1.0])
    return np.column_stack([c0, n, np.cross(c0, n)])


def test_facets_lie_on_the_paraboloid():
    for a, b, sun, X, Y, Z in _cases(seed=2):
        n = om.master_normal(X, Y, Z, sun)
        g = om.facet_heliostat(a, b, n, X, Y, Z)
        p = g["pos"].astype(np.float64)
        # centre facet: exactly at (X, Y, Z), exactly the master normal
        assert np.array_equal(g["pos"][4], np.float32([X, Y, Z]))
        assert np.array_equal(g["normal"][4], n)
        M = _matrix(n.astype(np.float64))
        # float32 positions round to a few 1e-7 of the coordinates; the matrix is not a
        # rotation, so its condition number scales what that does to the local coordinates
        tol = 2e-6 * (1 + max(abs(X), abs(Y), abs(Z))) * np.linalg.cond(M)
        loc = np.linalg.solve(M, (p - np.array([X, Y, Z])).T).T
        assert np.abs(loc[:, 0] - LX).max() < tol
        assert np.abs(loc[:, 2] - LZ).max() < tol
        assert np.abs(loc[:, 1] - (loc[:, 0] ** 2 / a ** 2 + loc[:, 2] ** 2 / b ** 2)).max() < tol
        # opposite facets
        for i in range(4):
            hsum = p[i] + p[8 - i] - 2 * np.array([X, Y, Z])
            h = 2 * (LX[i] ** 2 / a ** 2 + LZ[i] ** 2 / b ** 2)
            assert np.abs(hsum - h * n).max() < tol
        # facet normals: the paraboloid's, through the matrix
        nl = np.column_stack([-2 * LX / a ** 2, np.ones(9), -2 * LZ / b ** 2])
        assert np.abs(g["normal"] - (M @ nl.T).T).max() < 1e-5
        assert np.array_equal(g["size"], np.tile(np.float32([1.4142136 / 3, 0.1, 1.4142136 / 3]), (9, 1)))


def test_balanced_master_normal_turns_rigidly():
    rng = np.random.default_rng(3)
    for _ in range(20):
        th = rng.uniform(0.2, np.pi - 0.2)
        s = np.float32(np.sin(th) / np.sqrt(2.0))
        n = np.float32([np.cos(th), s, s * rng.choice([-1, 1])])
        M = _matrix(n.astype(np.float64))
        assert np.abs(M.T @ M - np.eye(3)).max() < 1e-6
        a, b = rng.uniform(8, 40, 2)
        g = om.facet_heliostat(a, b, n, 1.0, 2.0, 3.0)
        d = np.linalg.norm(g["pos"].astype(np.float64) - [1.0, 2.0, 3.0], axis=1)
        want = np.linalg.norm(np.column_stack([LX, LX ** 2 / a ** 2 + LZ ** 2 / b ** 2, LZ]), axis=1)
        assert np.abs(d - want).max() < 1e-5
        nn = g["normal"].astype(np.float64)
        cos = nn @ n.astype(np.float64) / np.linalg.norm(nn, axis=1)
        want = 1.0 / np.sqrt(1 + 4 * LX ** 2 / a ** 4 + 4 * LZ ** 2 / b ** 4)
        assert np.abs(cos - want).max() < 1e-6


def test_flat_limit_gives_the_master_normal_everywhere():
    n = om.master_normal(12.0, 1.2, 107.0, F.sun_direction(40.0, 170.0))
    g = om.facet_heliostat(1e6, 1e6, n, 12.0, 1.2, 107.0)
    nn = g["normal"] / np.linalg.norm(g["normal"], axis=1)[:, None]
    assert np.abs(nn - n).max() < 1e-6


def test_builder_refuses_degenerate_input():
    with pytest.raises(om.OMError):
        om.facet_heliostat(0.0, 10.0, np.float32([0, 1, 0]), 0, 0, 0)
    with pytest.raises(om.OMError):
        om.facet_heliostat(10.0, 10.0, np.float32([0, 0, 1]), 0, 0, 0)  # n0 = n1 = 0
    with pytest.raises(om.OMError):  # light travelling straight back to the receiver point
        om.master_normal(0.0, 0.0, 0.0, np.float32([0.0, 1.0, 0.0]))


def test_facet_scene_is_the_reference_scene():
    sun = F.sun_direction(50.0, 160.0)
    d = om.facet_scene(21.0, 21.0, 12.0, 1.2, 107.0, sun, 900.0)
    assert d["helio_pos"].shape == (9, 3)
    assert np.array_equal(d["recv_pos"], np.float32([0, 13, -1.5]))
    assert np.array_equal(d["recv_size"], np.float32([1.4, 1.4, 3.0]))
    assert np.array_equal(d["grid_pos"], np.float32([-5.5, -5.5, 94.5]))
    assert np.array_equal(d["grid_interval"], np.float32([110, 110, 110]))
    assert np.array_equal(d["helio_gap"], np.zeros((9, 2), np.float32))
    assert np.array_equal(d["helio_matrix"], np.ones((9, 2), np.int32))
    assert d["recv_pixel_length"] == np.float32(0.01) and d["ground_length"] == 1024


class HostScene:
    def __init__(self, desc, prm):
        self.keep = []
        dd = om._desc_struct(desc, self.keep)
        p = om._params_struct(prm)
        self.h = ctypes.c_void_p()
        self.rc = om.lib().om_scene_create(None, ctypes.byref(dd), ctypes.byref(p), ctypes.byref(self.h))

    def __del__(self):
        if self.h:
            om.lib().om_scene_destroy(self.h)

    def export(self, what, n):
        buf = np.zeros(n, np.float32)
        assert om.lib().om_scene_export(self.h, what.encode(), buf.ctypes.data, buf.nbytes) == 0
        return buf.reshape(-1, 3)


def test_given_normals_are_kept_and_normalised():
    """InitContent(normals): each heliostat takes its given normal, normalised, whatever
    the sun; its vertices follow as for an aimed heliostat."""
    sun = F.sun_direction(50.0, 160.0)
    d = om.facet_scene(21.0, 21.0, 12.0, 1.2, 107.0, sun, 900.0)
    p = dict(F.config_small()[1], micro_mode=F.LATTICE, helio_pixel_length=0.01, rays_per_sample=8)
    g, o = HostScene(d, p), oracle.OracleScene(d, p)
    assert g.rc == 0, om.lib().om_last_error()
    oh = o.helio()
    for k, name in (("normal", "helio_normal"), ("v0", "helio_v0"), ("v1", "helio_v1"), ("v3", "helio_v3")):
        assert np.array_equal(g.export(name, 27).view(np.uint32), oh[k].view(np.uint32)), k
    n = d["helio_normal"].astype(np.float64)
    assert np.abs(oh["normal"] - n / np.linalg.norm(n, axis=1)[:, None]).max() < 1e-7
    d2 = dict(d, sun_dir=F.sun_direction(20.0, 250.0))
    assert np.array_equal(HostScene(d2, p).export("helio_normal", 27), g.export("helio_normal", 27))
    # zero normals are refused
    bad = dict(d, helio_normal=np.zeros((9, 3), np.float32))
    assert HostScene(bad, p).rc == 1
