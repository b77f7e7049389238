"""Abort and progress of the persistent launches (om_ctx_abort, om_ctx_progress), forced from
the host: the kernels read a flag the host sets, and every launch ends normally.

Pins:
  * with the abort word set before a window trace (FP32, FP64 and mixed geometry) or a persistent
    batch (LDS windows, and work items), no chunk is traced: the maps are all zero, and
    progress reports the launch's total with at most one chunk per wave (or per workgroup
    for work items) taken;
  * after om_ctx_abort(ctx, 0) the same calls give the maps of a context that never aborted
    (for the window traces, the oracle's: the FP64 model's counts for FP64 and mixed), and
    progress reports every chunk taken (for the mixed trace, its own launch's: the FP64
    fallback of a full redo list does not run);
  * a watchdog thread that polls progress while a stream of 40 traces of the 5,000-heliostat
    field runs, and sets the abort word once it has seen progress, ends the stream early (its
    wall time is under half of the stream's without abort); the progress it read never
    exceeded the total, and a trace after clearing the word gives the unaborted maps.
"""
import threading
import time

import numpy as np
import pytest

import oracle
from onemirror_amd import fields as F

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def om():
    import onemirror_amd
    return onemirror_amd


def _maps(m):
    import torch
    torch.cuda.synchronize()
    return m.counts.cpu().numpy().view(np.uint32).copy(), m.energy.cpu().numpy().view(np.uint64).copy()


@pytest.mark.parametrize("precision", ["fp32", "fp64", "mixed"])
def test_window_trace_aborted_then_cleared(om, precision):
    ctx = om.Context(0)
    d, p = F.config_mid()
    d = F.subset(d, slice(2000, 2300))
    p = dict(p, accumulate=om.ACC_WINDOW)
    if precision != "fp32":
        p = dict(p, precision=om.PREC_FP64 if precision == "fp64" else om.PREC_MIXED)
    g = om.SolarScene(d, p, ctx=ctx)
    m = g.new_maps(smooth=False)
    print(f"abort: window trace {precision}", flush=True)
    ctx.abort(True)
    try:
        g.trace_plain(m)
        counts, energy = _maps(m)
        taken, total = ctx.progress()
    finally:
        ctx.abort(False)
    assert total > 0 and counts.sum() == 0 and energy.sum() == 0
    assert 0 <= taken <= min(total, 256 * 16)
    g.trace_plain(m)
    counts, energy = _maps(m)
    assert ctx.progress() == (total, total)
    o = oracle.OracleScene(d, p)
    want = o.trace_f32() if precision == "fp32" else o.trace_f64()
    assert np.array_equal(counts.ravel(), want["counts"].ravel())
    if precision == "fp32":
        assert np.array_equal(energy.ravel(), want["energy"].ravel())


@pytest.mark.parametrize("wg", [2048, 256])
def test_persistent_batch_aborted_then_cleared(om, wg):
    ctx = om.Context(0)
    d, p = F.config_small()
    scenes = [om.SolarScene(dict(d, sun_dir=F.sun_direction(el, 180.0)), dict(p, scene_index=i), ctx=ctx)
              for i, el in enumerate((45.0, 30.0, 60.0))]
    ref = om.Context(0)
    want = []
    for i, s in enumerate(scenes):
        r = om.SolarScene(dict(d, sun_dir=F.sun_direction((45.0, 30.0, 60.0)[i], 180.0)),
                          dict(p, scene_index=i), ctx=ref)
        want.append(_maps(r.trace_plain()))
    b = om.Batch(scenes, workgroup=wg)
    print(f"abort: batch workgroup {wg}", flush=True)
    ctx.abort(True)
    try:
        maps = b.run()
        got = [_maps(mm) for mm in maps]
        taken, total = ctx.progress()
    finally:
        ctx.abort(False)
    assert total == b.items
    assert all(c.sum() == 0 and e.sum() == 0 for c, e in got)
    assert 0 <= taken <= min(total, 256 * 16)
    maps = b.run()
    got = [_maps(mm) for mm in maps]
    assert ctx.progress() == (total, total)
    for i, (c, e) in enumerate(got):
        assert np.array_equal(c, want[i][0]) and np.array_equal(e, want[i][1]), i


def test_watchdog_aborts_a_running_stream(om):
    import torch
    ctx = om.Context(0)
    d, p = F.config_mid()
    g = om.SolarScene(d, dict(p, accumulate=om.ACC_WINDOW), ctx=ctx)
    m = g.new_maps(smooth=False)
    g.trace_plain(m)
    want = _maps(m)
    n = 40
    t0 = time.perf_counter()
    for _ in range(n):
        g.trace_plain(m)
    torch.cuda.synchronize()
    full = time.perf_counter() - t0
    print(f"watchdog: {n} traces take {full * 1e3:.1f} ms", flush=True)
    seen = []

    def watchdog():
        end = time.perf_counter() + 5.0
        while time.perf_counter() < end:
            t, tot = ctx.progress()
            seen.append((t, tot))
            if 0 < t < tot or len(seen) >= 200:
                ctx.abort(True)
                return
            time.sleep(2e-4)
        ctx.abort(True)

    t0 = time.perf_counter()
    for _ in range(n):
        g.trace_plain(m)
    th = threading.Thread(target=watchdog)
    th.start()
    torch.cuda.synchronize()
    aborted = time.perf_counter() - t0
    th.join()
    ctx.abort(False)
    print(f"watchdog: aborted stream {aborted * 1e3:.1f} ms after {len(seen)} polls, last {seen[-1]}", flush=True)
    assert all(0 <= t <= tot for t, tot in seen)
    assert aborted < 0.5 * full
    g.trace_plain(m)
    got = _maps(m)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
