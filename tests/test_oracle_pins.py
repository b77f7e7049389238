"""Pins of oracle functions that the GPU comparison alone cannot fix (the GPU and the oracle
could share a mistake): the Poisson-disk sampler, the early exit of the grid walk, and the
reference's literal self-intersection rule.

Poisson-disk sampler (poisson_sample + copy_if, poisson_disk.cu:93-201,
compact_sample.cu:6-19, compact_sample.cuh:13-22), on heliostats of the three plain fields
(2 x 2 m at pitch 0.125, 10 x 10 m at pitch 0.3125, 12 x 10 m), checked against what the
construction guarantees:
  * every pair of points, over the whole cell grid before compaction, is at least r apart
    (brute force in FP64; a neighbour window too small or a wrong cell offset breaks it);
  * every point lies in its own cell [(x-1) c, x c] x [(y-1) c, y c] (the padding offset);
  * every phase-0 cell (x, y both multiples of 3) holds exactly one point: phase 0 places
    one unconditionally;
  * the compacted list is exactly the grid's points strictly inside (0, w) x (0, h), in
    cell order (the stable copy_if);
  * the accepted count lies between the phase-0 cells wholly inside the area and Groemer's
    bound for points r apart in a convex region of area A and perimeter P,
    N <= 2 A / (sqrt(3) r^2) + P / (2 r) + 1;
  * saturation: no empty cell of the interior has half or more of its area farther than r
    from every point.  k = 8 darts, each uniform in the cell, all miss an open half with
    probability 2^-8, and the open area only shrinks in later phases.  With k = 4 such
    cells appear (measured: 2 in 21,902 empty cells), and with k = 1 there are hundreds;
  * the mean count over heliostats lies in [0.85, 1.0447] of the target w h / pitch^2.
    The upper end is the random-sequential-adsorption jamming density: disks of radius r/2
    cover 0.547 of the plane, i.e. 0.6965 / r^2 points, which is 1.0447 / pitch^2 at
    r = pitch / sqrt(1.5).

Early exit of the walk (NotCollision, rectgrid_raytracing.cu:38-117): the walk that stops
once an axis has passed the occupied cell range gives the same records, maps and counts as
the reference's walk to the grid exit, on S and on slices of the Mid and Lg fields.

Literal self-intersection (Intersect, rectgrid_raytracing.cu:21-36; t >= 1e-6 at
global_function.cuh:102): the reference tests a ray's own heliostat like any candidate.  The
oracle's literal mode must differ from its default (own heliostat skipped) exactly on the
rays whose walk meets their own heliostat first, and the count of those rays must equal the
self_hits counter.
"""
import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F

FIELDS = {"S": F.config_small, "Mid": F.config_mid, "Lg": F.config_large}


def _poisson_scene(name, n=12):
    d, p = FIELDS[name]()
    return oracle.OracleScene(F.subset(d, slice(0, n)), p), F.subset(d, slice(0, n)), p


def _open_fraction(G, r, c, rng, ns=128):
    """Fraction of each empty interior cell's area farther than r from every point
    (Monte Carlo, ns samples per cell; points within two cells are the only ones in reach)."""
    hi, wi = G.shape[:2]
    filled = G[..., 0] > -c
    pad = np.full((hi + 4, wi + 4, 2), -1e9)
    pad[2:-2, 2:-2][filled] = G[filled]
    ys, xs = np.nonzero(~filled)
    keep = (ys >= 1) & (ys <= hi - 2) & (xs >= 1) & (xs <= wi - 2)
    ys, xs = ys[keep], xs[keep]
    nb = np.stack([pad[ys + 2 + dy, xs + 2 + dx] for dy in range(-2, 3) for dx in range(-2, 3)], 1)
    s = (rng.random((len(ys), ns, 2)) + np.stack([xs - 1, ys - 1], 1)[:, None, :]) * c
    d2 = ((s[:, :, None, :] - nb[:, None, :, :]) ** 2).sum(-1).min(-1)
    return (d2 >= r * r).mean(1)


@pytest.mark.parametrize("name", ["S", "Mid", "Lg"])
def test_poisson_disk_invariants(name):
    o, d, p = _poisson_scene(name)
    pitch = float(p["helio_pixel_length"])
    rng = np.random.default_rng(5)
    ratios = []
    for h in range(o.n_helio):
        g = o.poisson_grid(h)
        G, r, c = g["grid"], g["r"], g["cell"]
        assert abs(r - pitch / np.sqrt(1.5)) < 1e-6 * r and abs(c - r / np.sqrt(2)) < 1e-6 * c
        hi, wi = G.shape[:2]
        filled = G[..., 0] > -c
        ys, xs = np.nonzero(filled)
        pts = G[filled].astype(np.float64)
        # minimum distance over all points of the grid
        dd = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
        np.fill_diagonal(dd, np.inf)
        assert dd.min() >= r * (1 - 1e-6), (name, h, dd.min() / r)
        # each point in its own cell (padding ring of one cell)
        tol = 1e-6 * c * max(hi, wi)  # FP32 rounding of (u + x - 1) * c
        assert np.all(pts[:, 0] >= (xs - 1) * c - tol) and np.all(pts[:, 0] <= xs * c + tol)
        assert np.all(pts[:, 1] >= (ys - 1) * c - tol) and np.all(pts[:, 1] <= ys * c + tol)
        # phase 0: one point in every cell whose indices are both multiples of 3
        assert filled[0::3, 0::3].all(), (name, h)
        # compaction: the points strictly inside the mirror, in cell order
        size = d["helio_size"][h]
        gap, mat = d["helio_gap"][h], d["helio_matrix"][h]
        sub_w = np.float32((np.float32(size[0]) - np.float32(gap[0]) * np.float32(mat[1] - 1)) / np.float32(mat[1]))
        sub_h = np.float32((np.float32(size[2]) - np.float32(gap[1]) * np.float32(mat[0] - 1)) / np.float32(mat[0]))
        w = np.float32(sub_w * np.float32(mat[1]))
        hh = np.float32(sub_h * np.float32(mat[0]))
        flat = G.reshape(-1, 2)
        inside = (flat[:, 0] > 0) & (flat[:, 1] > 0) & (flat[:, 0] < w) & (flat[:, 1] < hh)
        got = o.poisson(h)
        assert np.array_equal(got, flat[inside]), (name, h)
        assert g["n"] == len(got)
        # count band: phase-0 cells wholly inside, Groemer's packing bound
        cx = np.arange(0, wi, 3)
        cy = np.arange(0, hi, 3)
        in_x = ((cx - 1) * c > 0) & (cx * c < w)
        in_y = ((cy - 1) * c > 0) & (cy * c < hh)
        lower = int(in_x.sum()) * int(in_y.sum())
        upper = 2 * float(w) * float(hh) / (np.sqrt(3) * r * r) + 2 * (float(w) + float(hh)) / (2 * r) + 1
        assert lower <= len(got) <= upper, (name, h, lower, len(got), upper)
        # saturation: no empty cell left half open
        f = _open_fraction(G, r, c, rng)
        assert int((f >= 0.5).sum()) == 0, (name, h, float(f.max()))
        ratios.append(len(got) / (float(w) * float(hh) / pitch ** 2))
    assert 0.85 <= float(np.mean(ratios)) <= 1.0447, (name, np.mean(ratios))


@pytest.fixture(scope="module")
def scenes():
    out = {}
    for name, fn in FIELDS.items():
        d, p = fn()
        out[name] = oracle.OracleScene(d, p)
    return out


SLICES = {"S": (0, 100), "Mid": (2000, 2010), "Lg": (12000, 12008)}


@pytest.mark.parametrize("name", ["S", "Mid", "Lg"])
def test_early_exit_walk_equals_full_walk(scenes, name):
    o = scenes[name]
    h0, h1 = SLICES[name]
    a = o.trace_f32(h0, h1, records=True, early_exit=True)
    b = o.trace_f32(h0, h1, records=True, early_exit=False)
    for k in ("shade", "block", "pixel", "counts", "energy"):
        assert np.array_equal(a[k], b[k]), (name, k)
    for k in ("rays", "shaded", "blocked", "hit"):
        assert a["stats"][k] == b["stats"][k], (name, k)
    # the full walk goes to the grid exit: many more cells
    assert b["stats"]["dda_steps"] > 5 * a["stats"]["dda_steps"]


def _own(o, h0, h1):
    off = o.micro()["ray_offset"]
    return np.repeat(np.arange(h0, h1, dtype=np.int32), np.diff(off[h0:h1 + 1]))


@pytest.mark.parametrize("name", ["S", "Mid", "Lg"])
def test_literal_self_intersection_differs_only_on_self_hits(scenes, name):
    o = scenes[name]
    h0, h1 = SLICES[name]
    a = o.trace_f32(h0, h1, records=True)
    o.set_literal_self(True)
    try:
        b = o.trace_f32(h0, h1, records=True)
    finally:
        o.set_literal_self(False)
    own = _own(o, h0, h1)
    changed = (a["shade"] != b["shade"]) | (a["block"] != b["block"]) | (a["pixel"] != b["pixel"])
    stopped_by_self = (b["shade"] == own) | ((b["shade"] == a["shade"]) & (b["block"] == own))
    assert np.array_equal(changed, stopped_by_self), name
    assert int(changed.sum()) == b["stats"]["self_hits"]
    assert a["stats"]["self_hits"] == 0
    # every changed ray lost its receiver hit or its blocker
    assert np.all(b["pixel"][changed] == -1)
    # the literal rule's rays that survive reach the same pixels
    assert np.array_equal(a["pixel"][~changed], b["pixel"][~changed])
