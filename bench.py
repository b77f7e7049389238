#!/usr/bin/env python3
"""Throughput of the heliostat-field ray tracer on one MI355X.

One step = one full trace of the field (sunshape + slope-error sampling, shading DDA,
reflection, blocking DDA, receiver hit, fixed-point accumulation) followed by the flux
scaling and the trimmed-mean smoothing, all on the GPU with the scene resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mid|small]

Prints one JSON line: rays/s, the roofline of the trace kernel (VALU lane-ops per ray from
the cost model below times rays/s, against 256 CU x 128 lanes x 2.4 GHz: a CU issues two
64-lane FP32 VALU instructions per clock, which the VALU issue probe (--valu-probe) sustains
and AMD's 157.3 TFLOP/s FP32 vector peak states), the counter-based VALU utilisation of the
same kernel from its committed rocprof run, and the CPU oracle's rays/s on a stated sample
of the same field on the host cores.

VALU cost model per ray (lane operations; add/mul/fma/cmp/cvt/int = 1, div/sqrt/rcp = 1,
sincos = 30, log = 25, Philox4x32-10 = 100):
    F0 = 569 fixed: Philox 100, four uniforms 12, sunshape lookup 15, sun-ray sincos pair 60
         and direction 21, two walk set-ups 114, perturbed normal 128 (Gaussian 40, sincos
         pair 60, frame 18, normalize 10), reflect + normalize 21, microsample centre 24,
         receiver hit + pixel 64, energy + accumulation 10
    + 12 per DDA cell step + 58 per candidate parallelogram test + 3 per own-heliostat skip,
with the step/test counts measured by the oracle on the same sample of the field; the
shading walks the kernels skip (heliostats no sun ray can leave toward another heliostat)
are taken out: their steps and tests, and one walk set-up (57) per ray.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/s (w/ DDA shading+blocking) at ≤1e-5 flux err vs FP64; batched & fwd+bwd/s"
P_VALU = 256 * 128 * 2.4e9         # lane-op/s: two 64-lane VALU instructions per CU clock
P_VALU64 = 256 * 64 * 2.4e9        # FP64 lane-op/s: one 64-lane FP64 instruction per CU clock
F0, C_STEP, C_TEST, C_SKIP = 569.0, 12.0, 58.0, 3.0
C_SETUP = 57.0  # one walk set-up, the half of F0's two that a skipped shading walk saves
GEOMETRY = {
    "fp32": "FP32: the reference's ray geometry, bit-exact against its host restatement",
    "fp64": ("FP64: sampling, frames, walks, intersections and receiver in double (om_params.precision = "
             "OM_PREC_FP64), energies in exact fixed point, FP32 flux map"),
    "mixed": ("mixed (om_params.precision = OM_PREC_MIXED): FP32 sampling, walks and receiver hit, the "
              "direction as an FP32 deviation from the FP64 model's nominal reflection, every decision "
              "certified by an error bound; rays with a decision inside its bound traced again in FP64: the "
              "FP64 model's shading, blocking and pixel for every ray; energies in exact fixed point, FP32 "
              "flux map"),
}


def model_ops(st: dict) -> tuple:
    """Lane-ops per ray of the cost model from the oracle's counts, with the shading walks of
    the heliostats the kernels skip (tagged in the oracle) taken out; (ops, S, T)."""
    n = max(1, st["rays"])
    S = (st["dda_steps_eff"] - st.get("tag_shade_steps_eff", 0)) / n
    T = (st["cand_tests"] - st.get("tag_shade_tests", 0)) / n
    K = (st["self_skips"] - st.get("tag_shade_self_skips", 0)) / n
    return F0 - C_SETUP * st.get("tag_rays", 0) / n + C_STEP * S + C_TEST * T + C_SKIP * K, S, T


# VALU issue probes of om_math_eval (fn 12..14): VALU instructions per lane of each kernel,
# counted in the assembly of this build (the loop's instructions times its trips, plus the
# set-up and the final sum)
VALU_PROBES = {12: ("fma_f32", 2048 + 27), 13: ("fma_f32+add_u32", 4096 + 42), 14: ("fma_f32+exp_f32", 2560 + 27),
               15: ("fma_f64", 2048 + 27)}


def valu_probe(mi: int, reps: int = 5) -> dict:
    """Runs each VALU issue probe over mi Mi threads and reports its time (HIP events on the
    launch stream) and the VALU wave-instructions it issues per CU cycle at 2.4 GHz."""
    import torch

    import onemirror_amd as om

    n = mi << 20
    x = torch.rand(n, device="cuda")
    y = torch.empty_like(x)
    ctx = om.default_context()
    stream = torch.cuda.current_stream()
    out = {}
    for fn, (name, per_lane) in VALU_PROBES.items():
        def run():
            om._check(om.lib().om_math_eval(ctx._h, fn, om._vp(x.data_ptr()), om._vp(y.data_ptr()), n,
                                            om._stream_ptr()), "om_math_eval")
        run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        instr = n / 64 * per_lane
        out[name] = {"ms": ms, "valu_wave_instructions": instr, "valu_per_lane": per_lane,
                     "per_cu_cycle_at_2.4GHz": instr / (ms * 1e-3) / (256 * 2.4e9)}
        print(f"# VALU probe {name}: {ms:.3f} ms, {instr:.4g} wave-instructions, "
              f"{out[name]['per_cu_cycle_at_2.4GHz']:.3f} per CU cycle at 2.4 GHz", file=sys.stderr, flush=True)
    return out


ATOMIC_PROBES = {17: "global_u64_no_return", 18: "global_u64_returning", 19: "lds_u64_returning"}
ATOM_ITERS = 64


def atomic_probe(mi: int, reps: int = 5) -> dict:
    """Runs each atomic throughput probe (om_math_eval 17..19: 64 atomic adds of 64 bits per
    thread at hashed addresses, in device memory over mi / 2 Mi words or in a window-sized LDS
    array per workgroup) over mi Mi threads, and reports its time (HIP events on the launch
    stream) and the atomic operations per second it sustains."""
    import torch

    import onemirror_amd as om

    n = mi << 20
    x = torch.rand(n, device="cuda")
    y = torch.zeros_like(x)
    ctx = om.default_context()
    stream = torch.cuda.current_stream()
    out = {}
    for fn, name in ATOMIC_PROBES.items():
        def run():
            om._check(om.lib().om_math_eval(ctx._h, fn, om._vp(x.data_ptr()), om._vp(y.data_ptr()), n,
                                            om._stream_ptr()), "om_math_eval")
        run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        ops = n * ATOM_ITERS
        out[name] = {"ms": ms, "atomics": ops, "atomics_per_s": ops / (ms * 1e-3),
                     "footprint": "LDS, 18,400 words per workgroup" if fn == 19 else f"{n // 2} words of device memory"}
        print(f"# atomic probe {name}: {ms:.3f} ms, {ops:.4g} atomics, {ops / (ms * 1e-3):.4g} per s",
              file=sys.stderr, flush=True)
    return out


def host_cores() -> dict:
    """Host cores this process may run on, the machine's count, the CPU model, and the
    threads the CPU baseline uses: every core it may run on, or fewer where OMP_NUM_THREADS
    sets this process's share of a machine shared by several GPUs (16 per GPU on the GPU
    boxes here)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    share = os.environ.get("OMP_NUM_THREADS")
    threads = avail
    if share and share.isdigit() and int(share) > 0:
        threads = min(avail, int(share))
    return {"threads": max(1, threads), "available": avail, "machine": os.cpu_count(), "model": model,
            "omp_num_threads": share}


def host_threads() -> int:
    return host_cores()["threads"]


def flux_error_vs_fp64(scene, maps, osc, f64: dict, nthreads: int, block: int = 2000) -> dict:
    """FP32 flux map of a GPU trace (maps, the whole field) against the FP64 model of the same
    rays (f64: the oracle's trace of the same Philox draws with double geometry and double
    sums): per-pixel relative error over the pixels the FP64 map lights (median, p99, max,
    share above 1e-5) and over the pixels above 1 % of the peak, the total-power error, the
    error of the accumulation alone (the FP32 map against its own exact integer sums), and
    the share of rays whose receiver pixel, shading and blocking heliostat differ from the
    FP64 model's on one block of heliostats (GPU records against FP64 records)."""
    import numpy as np
    import torch

    torch.cuda.synchronize()
    scale = float(np.float32(scene.info["flux_scale"]))
    flux32 = maps.flux.cpu().numpy().astype(np.float64).ravel()
    energy = maps.energy.cpu().numpy().view(np.uint64).ravel()
    flux64 = f64["energy"].ravel() * scale
    lit = flux64 > 0
    rel = np.abs(flux32[lit] - flux64[lit]) / flux64[lit]
    big = flux64 > 0.01 * flux64.max()
    relb = np.abs(flux32[big] - flux64[big]) / flux64[big]
    exact = energy.astype(np.float64) * 2.0 ** -31 * scale
    nz = exact > 0
    acc = float(np.max(np.abs(flux32[nz] - exact[nz]) / exact[nz])) if nz.any() else 0.0
    H = scene.info["n_helio"]
    h0 = min(block, max(0, H - 40))
    h1 = min(H, h0 + 40)
    rg = scene.trace_records(h0, h1)
    r64 = osc.trace_f64(h0, h1, records=True, nthreads=nthreads)
    nr = rg["pixel"].size
    return {
        "pixels_lit": int(lit.sum()),
        "median": float(np.median(rel)), "p99": float(np.quantile(rel, 0.99)), "max": float(rel.max()),
        "share_above_1e-5": float((rel > 1e-5).mean()),
        "bright_pixels": int(big.sum()), "bright_median": float(np.median(relb)),
        "bright_p99": float(np.quantile(relb, 0.99)), "bright_share_above_1e-5": float((relb > 1e-5).mean()),
        "total_power": float(abs(flux32.sum() - flux64.sum()) / flux64.sum()),
        "accumulation_only_max": acc,
        "pixel_flip_rate": float((rg["pixel"] != r64["pixel"]).sum() / nr),
        "shade_flip_rate": float((rg["shade"] != r64["shade"]).sum() / nr),
        "block_flip_rate": float((rg["block"] != r64["block"]).sum() / nr),
        "flip_sample": f"heliostats [{h0}, {h1}), {nr} rays",
        "per_pixel_1e-5_met": bool((rel > 1e-5).sum() == 0),
    }


def time_trace(scene, maps, steps: int, warmup: int) -> dict:
    """Whole steps (clear maps, trace every heliostat, flux scaling and smoothing) timed
    between synchronisations, and the trace launch (with its decode) by HIP events on the
    launch stream."""
    import numpy as np
    import torch

    H, R = scene.info["n_helio"], scene.info["total_rays"]
    stream = torch.cuda.current_stream()

    def step(ev0=None, ev1=None):
        maps.counts.zero_()
        maps.energy.zero_()
        if ev0 is not None:
            ev0.record(stream)
        scene.trace_heliostats(0, H, maps)
        if ev1 is not None:
            ev1.record(stream)
        scene.finalize(maps)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        step(e0, e1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"ms_per_step": 1e3 * dt / steps, "trace_ms": float(np.mean([e0.elapsed_time(e1) for e0, e1 in evs])),
            "rays_s": R / (dt / steps), "hits": int(maps.counts.sum().item())}


def copy_probe(mib: int):
    """An elementwise copy reading and writing mib MiB, three times: the reference kernel that
    calibrates the HBM byte counters."""
    import torch
    n = mib * (1 << 20) // 4
    xs = torch.rand(n, device="cuda")
    ys = torch.empty_like(xs)
    for _ in range(3):
        torch.mul(xs, 1.0, out=ys)
    torch.cuda.synchronize()
    del xs, ys
    print(f"# copy probe: 3 x elementwise copy of {mib} MiB", file=sys.stderr, flush=True)


def roofline(run: dict, workload: str, R: int, ops_ray: float, S: float, T: float) -> dict:
    """VALU roofline of a run's trace kernel: the cost model's lane-ops per ray times rays
    over the kernel time, against the VALU peak of its arithmetic (FP32: two 64-lane
    instructions per CU clock; FP64: one), with the kernel's counters from its committed
    rocprof run (VALU instructions per CU busy cycle against the issue probe of the same
    arithmetic) and its HBM bytes per launch."""
    fp64 = run["kernel"].startswith("trace_window64") or run["kernel"].startswith("trace64")
    # the mixed kernel issues FP32 and FP64 instructions: priced against the FP32 peak
    peak = P_VALU64 if fp64 else P_VALU
    achieved = R * ops_ray / (run["trace_ms"] * 1e-3)
    counters = None
    vp = os.path.join(ROOT, "profiles", "valu_probe_kernel.json")
    if os.path.exists(vp):
        with open(vp) as f:
            vj = json.load(f)
        tw = vj.get(f"{run['kernel']}_{workload}")
        probe = vj["probes"].get("probe_fma_f64" if fp64 else "probe_fma")
        if tw and probe:
            counters = {"valu_per_cu_busy_cycle": tw["valu_per_cu_busy_cycle"],
                        "probe_valu_per_cu_busy_cycle": probe["valu_per_cu_busy_cycle"],
                        "source": "profiles/valu_probe_kernel.json"}
            if fp64:
                f64i = sum(tw.get(k, 0.0) for k in ("SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
                                                    "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64"))
                counters["fp64_share"] = f64i / tw["SQ_INSTS_VALU"]
                counters["fp64_util"] = f64i * 64 / (tw["avg_ns"] * 1e-9) / P_VALU64
            else:
                counters["valu_util"] = tw["SQ_INSTS_VALU"] * 64 / (tw["avg_ns"] * 1e-9) / P_VALU
                counters["of_probe"] = tw["valu_per_cu_busy_cycle"] / probe["valu_per_cu_busy_cycle"]
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"{run['kernel']}_{workload}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            pj = json.load(f)
        if pj.get("rays") == R:
            traffic = pj["hbm_bytes_per_launch"]
            ap = os.path.join(ROOT, "profiles", "atomic_probe.json")
            if counters is not None and os.path.exists(ap) and pj.get("tcc_atomic"):
                # the kernel's atomics per second against the probes': memory-side atomics
                # (the hits outside the LDS window) and LDS atomics (the hits inside it)
                with open(ap) as f:
                    aj = json.load(f)["probes"]
                mem = pj["tcc_atomic"] / (pj["avg_ns"] * 1e-9)
                lds = max(run["hits"] - pj["tcc_atomic"], 0.0) / (pj["avg_ns"] * 1e-9) \
                    if run["kernel"].startswith("trace_window") else None
                counters["atomics"] = {
                    "memory_side_per_s": mem,
                    "probe_memory_side_per_s": aj["global_u64_returning"]["atomics_per_s"],
                    "memory_side_of_probe": mem / aj["global_u64_returning"]["atomics_per_s"],
                    "lds_per_s": lds,
                    "probe_lds_per_s": aj["lds_u64_returning"]["atomics_per_s"],
                    "lds_of_probe": None if lds is None else lds / aj["lds_u64_returning"]["atomics_per_s"],
                    "source": "profiles/atomic_probe.json, TCC_ATOMIC_sum of the kernel's profile"}
    return {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "Tlane-op/s",
            "frac": achieved / peak, "traffic": traffic, "ops_per_ray": ops_ray, "dda_steps_per_ray": S,
            "tests_per_ray": T, "arithmetic": "fp64" if fp64 else ("fp32+fp64" if "mixed" in run["kernel"] else "fp32"),
            "counters": counters}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mid",
                    choices=["mid", "small", "large", "batched", "diff", "annual", "facet_annual", "mid_grids"])
    ap.add_argument("--scenes", type=int, default=64, help="batched: number of scenes")
    ap.add_argument("--grid-tune", action="store_true",
                    help="batched: tune the grid resolution (64, 128, 256) instead of 128^3")
    ap.add_argument("--no-modes", action="store_true",
                    help="mid: skip the batched and fwd+bwd summaries added to the line")
    ap.add_argument("--cpu-every", type=int, default=0,
                    help="CPU baseline traces one block of 40 heliostats in every n "
                         "(default: the whole field, one block in 5 for large)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="all", choices=["all", "both", "fp32", "fp64", "mixed"],
                    help="ray geometry: fp32 (the reference's, bit-exact against its host restatement), "
                         "fp64 (the FP64 model: per-pixel flux within 1e-5 of FP64), mixed (FP32 walks "
                         "with certified decisions, double directions and receiver, uncertain rays "
                         "traced in FP64: the FP64 model's outcome per ray), both (fp32 and fp64) or "
                         "all (default: the value is the mixed run, the metric's accuracy at the "
                         "least cost; fp64 and fp32 under modes)")
    ap.add_argument("--ray-compact", type=int, default=0, choices=[0, 1],
                    help="fp32 window kernel: 1 packs the rays that pass the blocking test into full "
                         "waves before the receiver stage (om_params.ray_compact)")
    ap.add_argument("--valu-probe", type=int, default=0, metavar="MI",
                    help="first run the VALU issue probes (om_math_eval 12..14) over MI Mi threads: "
                         "kernels with a known VALU instruction count that calibrate the VALU peak "
                         "and counters")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run an elementwise copy of MIB MiB (reads and writes MIB MiB "
                         "each): the reference kernel that calibrates HBM byte counters")
    ap.add_argument("--atomic-probe", type=int, default=0, metavar="MI",
                    help="first run the atomic throughput probes (om_math_eval 17..19) over MI Mi "
                         "threads: 64-bit atomic adds in device memory (without and with the old "
                         "value returned) and in a window-sized LDS array, the ceilings the trace "
                         "kernels' atomic rates are set against")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(every GPU box here has one GPU)", file=sys.stderr)
        return 2

    if args.workload in ("batched", "diff", "annual", "facet_annual"):
        out = {"batched": bench_batched, "diff": bench_diff, "annual": bench_annual,
               "facet_annual": bench_facet_annual}[args.workload](args)
        print(json.dumps(out), flush=True)
        return 0

    import numpy as np
    import torch

    import onemirror_amd as om
    from onemirror_amd import fields as F

    if args.workload == "mid":
        desc, prm = F.config_mid()
        wl = ("mid: 5000 heliostats 10x10 m, radii 100-800 m, flat 25.6 m receiver, "
              "Buie CSR 0.1 (2048-bin CDF), slope 2.5 mrad, 32x32 Poisson x 40 rays, "
              "128^3 grid, 256x256 map, trimmed mean 5x5/20%")
    elif args.workload == "mid_grids":
        desc, prm = F.config_mid_grids()
        prm = dict(prm, grid_select=om.GRIDS_ALL)
        wl = ("mid_grids: the mid field's heliostats west and east of the tower in two 128^3 "
              "grids of their own (.scn multi-grid scene), every grid traced")
    elif args.workload == "large":
        desc, prm = F.config_large()
        wl = ("large: 25000 heliostats 12x10 m on a surround field to 1.5 km, cylinder "
              "receiver r 8 m h 20 m at 200 m (512x256 unrolled map), Buie CSR 0.05, "
              "slope 2 mrad, ~1000 Poisson microsamples x 40 rays, 256^3 grid")
    else:
        desc, prm = F.config_small()
        wl = ("small: 100 heliostats 2x2 m, pillbox 4.65 mrad, slope 2 mrad, 16x16 Poisson "
              "x 64 rays, 64^3 grid, 128x128 map")
    precisions = {"all": ["fp32", "fp64", "mixed"], "both": ["fp32", "fp64"]}.get(args.precision, [args.precision])
    head = "mixed" if "mixed" in precisions else ("fp64" if "fp64" in precisions else "fp32")
    runs = {}
    probe = atomics = None
    for pr in precisions:
        p = {"fp64": dict(prm, precision=om.PREC_FP64), "mixed": dict(prm, precision=om.PREC_MIXED)}.get(
            pr, dict(prm, ray_compact=args.ray_compact))
        t0 = time.time()
        scene = om.SolarScene(desc, p)
        info = scene.info
        H, R = info["n_helio"], info["total_rays"]
        print(f"# {pr} scene built in {time.time() - t0:.1f}s: {H} heliostats, {R} rays, "
              f"grid {info['grid_num']} nnz {info['grid_nnz']}", file=sys.stderr, flush=True)
        maps = scene.new_maps(smooth=True)
        # accumulation autotuner: per-tile packed atomics or the LDS window (identical maps)
        acc = scene.autotune_accumulate(maps, reps=3)
        kernel = ("trace_window" if acc["accumulate"] == "window" else "trace") + \
            {"fp64": "64", "mixed": "_mixed" if acc["accumulate"] == "window" else "64"}.get(pr, "") + "_kernel"
        print(f"# {pr} accumulation autotune: packed {acc['packed_ms']:.3f} ms, window {acc['window_ms']:.3f} ms "
              f"-> {acc['accumulate']}", file=sys.stderr, flush=True)
        if probe is None and args.valu_probe > 0:
            probe = valu_probe(args.valu_probe)
        if pr == precisions[0] and args.copy_probe > 0:
            copy_probe(args.copy_probe)
        if pr == precisions[0] and args.atomic_probe > 0:
            atomics = atomic_probe(args.atomic_probe)
        t = time_trace(scene, maps, args.steps, args.warmup)
        if pr == "mixed":
            t["redo_rays"] = scene.redo_count()
        print(f"# {pr}: {args.steps} steps, {t['ms_per_step']:.3f} ms/step, trace kernel {t['trace_ms']:.3f} ms, "
              f"{t['rays_s']:.4g} rays/s, receiver hits {t['hits']}"
              + (f", {t['redo_rays']} rays traced in FP64" if pr == "mixed" else ""), file=sys.stderr, flush=True)
        runs[pr] = dict(scene=scene, maps=maps, acc=acc, kernel=kernel, **t)
    scene = runs[head]["scene"]
    info = scene.info
    H, R = info["n_helio"], info["total_rays"]

    # oracle sample: step/test counts for the cost model and the CPU baseline (the FP32
    # restatement of the reference's tracer)
    import oracle
    osc = oracle.OracleScene(desc, prm)
    every = args.cpu_every if args.cpu_every > 0 else (5 if args.workload == "large" else 1)
    cores = host_cores()
    nthreads = cores["threads"]
    blk = 40  # heliostats per block: blocks of neighbours keep the threads busy
    blocks = [(h, min(H, h + blk)) for h in range(0, H, blk * every)]
    if args.no_cpu:
        blocks = blocks[:: max(1, len(blocks) // 8)]
    st_tot = {"rays": 0, "dda_steps_eff": 0, "cand_tests": 0, "self_skips": 0, "tag_rays": 0,
              "tag_shade_steps_eff": 0, "tag_shade_tests": 0, "tag_shade_self_skips": 0}
    osc.set_tags(scene.export("helio_clear", 4 * H, np.int32))  # shading walks the kernels skip
    osc.trace_f32(0, min(H, 4), nthreads=nthreads)  # warm-up
    t1 = time.perf_counter()
    for h0, h1 in blocks:
        r = osc.trace_f32(h0, h1, nthreads=nthreads)
        for k in st_tot:
            st_tot[k] += r["stats"][k]
    cpu_t = time.perf_counter() - t1
    cpu_rays = st_tot["rays"]
    nsample = sum(h1 - h0 for h0, h1 in blocks)
    ops_ray, S, T = model_ops(st_tot)
    # every run's flux against the FP64 model of the same rays (one FP64 trace of the field)
    errs, fp64_model = {}, None
    if args.workload in ("mid", "small", "mid_grids") and not args.no_cpu:
        t2 = time.perf_counter()
        f64 = osc.trace_f64(nthreads=nthreads)
        fp64_model = {"value": R / (time.perf_counter() - t2), "unit": "rays/s", "cores": nthreads,
                      "seconds": time.perf_counter() - t2,
                      "sample": f"the whole field ({R} rays); oracle FP64 model, OpenMP"}
        for pr, run in runs.items():
            errs[pr] = flux_error_vs_fp64(run["scene"], run["maps"], osc, f64, nthreads)
            e = errs[pr]
            print(f"# {pr} flux vs FP64 (whole field): per-pixel median {e['median']:.2g}, p99 {e['p99']:.2g}, "
                  f"{100 * e['share_above_1e-5']:.1f} % of lit pixels above 1e-5, total power {e['total_power']:.2g}, "
                  f"pixel flips {e['pixel_flip_rate']:.2g} of rays", file=sys.stderr, flush=True)
    rooflines = {pr: roofline(run, args.workload, R, ops_ray, S, T) for pr, run in runs.items()}
    hr = runs[head]
    out = {
        "metric": METRIC,
        "value": hr["rays_s"],
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hr["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32+f64"}.get(head, "f32"),
        "data": "synthetic",
        "config": {"workload": wl, "precision": head, "heliostats": H, "rays": R,
                   "grid": list(info["grid_num"]),
                   "map": [info["res_y"], info["res_x"]], "parallelism": "single GPU",
                   "trace_kernel_ms": hr["trace_ms"], "receiver_hits": hr["hits"], "kernel": hr["kernel"],
                   "accumulate": hr["acc"],
                   "geometry": GEOMETRY[head]},
        "roofline": rooflines[head],
        "flux_err_vs_fp64": errs.get(head),
        "valu_probe": probe,
        "atomic_probe": atomics,
        "cpu_baseline": None if args.no_cpu else {
            "value": cpu_rays / cpu_t, "unit": "rays/s", "cores": nthreads, "kind": "port",
            "host_cores": cores,
            "sample": (f"the whole field ({nsample} heliostats, {cpu_rays} rays)" if every == 1 else
                       f"{len(blocks)} blocks of {blk} neighbouring heliostats, one in every "
                       f"{every} of the same field ({nsample} heliostats, {cpu_rays} rays)")
                      + "; oracle FP32 tracer (the restatement of the reference's) with the same "
                        "early-exit walk, OpenMP",
            "seconds": cpu_t, "fp64_model": fp64_model},
    }
    if head == "mixed":
        out["config"]["redo_rays"] = hr["redo_rays"]
    for pr, name in (("fp64", "fp64"), ("fp32", "fp32_reference")):
        if pr == head or pr not in runs:
            continue
        r = runs[pr]
        out.setdefault("modes", {})[name] = {
            "value": r["rays_s"], "unit": "rays/s", "ms_per_step": r["ms_per_step"],
            "trace_kernel_ms": r["trace_ms"], "kernel": r["kernel"], "accumulate": r["acc"],
            "geometry": GEOMETRY[pr], "flux_err_vs_fp64": errs.get(pr), "roofline": rooflines[pr]}
    if args.workload == "mid" and not args.no_modes:
        # the batched and differentiable forms of the metric and the annual runs, on their own
        # workloads (a few timed steps each)
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup = 3, 1
        b = bench_batched(sub, with_cpu=False)
        d = bench_diff(sub)
        sub_a = argparse.Namespace(**vars(args))
        sub_a.steps, sub_a.warmup = 1, 1
        an = bench_annual(sub_a)
        fa = bench_facet_annual(sub_a)
        out.setdefault("modes", {}).update({
            "batched": {"value": b["value"], "unit": "rays/s", "workload": b["config"]["workload"],
                        "roofline": b["roofline"],
                        **{k: b["config"][k] for k in ("precision", "scenes", "rays", "batch_ms", "plain_ms",
                                                       "plain_rays_per_s", "batched_vs_plain",
                                                       "grid_cells", "workgroup", "rays_per_thread",
                                                       "redo_rays", "full_lists", "fp32_reference")}},
            "annual": {"value": an["value"], "unit": "rays/s", "workload": an["config"]["workload"],
                       **{k: an["config"][k] for k in ("precision", "time_steps", "rays", "mode", "batch_ms",
                                                        "plain_ms", "time_steps_per_s", "annual_ratio",
                                                        "fp32_reference")}},
            "facet_annual": {"value": fa["value"], "unit": "rays/s", "workload": fa["config"]["workload"],
                             **{k: fa["config"][k] for k in ("precision", "time_steps", "rays", "mode", "batch_ms",
                                                              "plain_ms", "time_steps_per_s", "annual_ratio",
                                                              "fp32_reference")}},
            "fwd_bwd": {"value": d["config"]["fwd_bwd_per_s"], "unit": "fwd+bwd/s",
                        "rays_per_s": d["value"], "workload": d["config"]["workload"],
                        **{k: d["config"][k] for k in ("fwd_bwd_ms", "forward_ms", "backward_ms",
                                                       "plain_trace_ms", "fwd_bwd_vs_plain",
                                                       "plain_trace_mixed_ms", "fwd_bwd_vs_plain_mixed")}},
        })
    print(json.dumps(out), flush=True)
    return 0


def with_fp32_reference(fn, args) -> dict:
    """fn's line in the mixed precision (the metric's accuracy), with fn's FP32 run of the
    same workload under config.fp32_reference."""
    out = fn(args, "mixed")
    ref = fn(args, "fp32")
    out["config"]["fp32_reference"] = {
        "value": ref["value"], "unit": ref["unit"], "ms_per_step": ref["ms_per_step"],
        **{k: ref["config"][k] for k in ("mode", "batch_ms", "plain_ms", "time_steps_per_s")}}
    return out


def bench_annual(args) -> dict:
    return with_fp32_reference(annual_once, args)


def bench_facet_annual(args) -> dict:
    return with_fp32_reference(facet_annual_once, args)


def annual_once(args, precision: str) -> dict:
    """Annual driver (main.cpp:56-116): the mid field at every time step of the reference's
    schedule with the sun above 15 degrees, each step a fork of the field, traced by the
    batched launch and by one plain (accumulation-autotuned) trace per step; the faster is
    the value."""
    import torch

    import onemirror_amd as om
    from onemirror_amd import annual as A
    from onemirror_amd import fields as F

    d, p = F.config_mid()
    p = dict(p, precision={"fp32": om.PREC_FP32, "mixed": om.PREC_MIXED}[precision])
    t0 = time.time()
    root = om.SolarScene(d, p)
    acc = root.autotune_accumulate(reps=2)
    steps = A.annual_steps()
    runs = {m: A.AnnualRun(root, steps, batch_size=64, mode=m) for m in ("batch", "plain")}
    R = runs["batch"].rays
    print(f"# annual {precision}: {len(steps)} time steps (zenith <= {A.ZENITH_MAX}), {R} rays, forks built in "
          f"{time.time() - t0:.1f}s; accumulation {acc['accumulate']}", file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    ms = {}
    for m, run in runs.items():
        for _ in range(max(1, args.warmup)):
            run.trace()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            run.trace()
        e1.record(stream)
        torch.cuda.synchronize()
        ms[m] = e0.elapsed_time(e1) / args.steps
        print(f"# annual {precision} {m}: {ms[m]:.1f} ms per year ({R / (ms[m] * 1e-3):.4g} rays/s)", file=sys.stderr, flush=True)
    best = min(ms, key=ms.get)
    res = runs[best].results()
    print(f"# annual intercept ratio {res['annual_ratio']:.4f}, intercepted {res['intercepted_w_sum'] / 1e6:.1f} "
          f"MW summed over the steps", file=sys.stderr, flush=True)
    return {
        "metric": METRIC, "value": R / (ms[best] * 1e-3), "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms[best], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": ("annual: the mid field at the reference schedule (months 1-12, days 1-28 "
                                "every 3, hours 6-18 every 2) at 30.3N 120.2E UTC+8, 40 m, zenith <= 75 deg "
                                "(approximate solar positions), clear-sky DNI of getDNI; one step = the "
                                "whole year"),
                   "precision": precision, "geometry": GEOMETRY[precision],
                   "time_steps": len(steps), "rays": R, "mode": best,
                   "batch_ms": ms["batch"], "plain_ms": ms["plain"],
                   "time_steps_per_s": len(steps) / (ms[best] * 1e-3),
                   "accumulate": acc["accumulate"], "annual_ratio": res["annual_ratio"],
                   "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
    }


def facet_annual_once(args, precision: str) -> dict:
    """The reference driver's own run (main.cpp:56-116): the faceted paraboloid heliostat
    rebuilt and traced for every time step of the year with the reference's default
    parameters, batched and one plain trace per step; the faster is the value."""
    import torch

    import onemirror_amd as om
    from onemirror_amd import annual as A
    from onemirror_amd import fields as F

    site = F.facet_site()
    t0 = time.time()
    steps = A.annual_steps()
    p = dict(om.default_params(), precision={"fp32": om.PREC_FP32, "mixed": om.PREC_MIXED}[precision])
    # accumulation autotune on the middle step's scene, kept for every step's plain trace
    mid = steps[len(steps) // 2]
    probe = om.SolarScene(om.facet_scene(site["a"], site["b"], site["X"], site["Y"], site["Z"], mid["sun_dir"],
                                         mid["dni"]), p)
    acc = probe.autotune_accumulate(reps=3)
    del probe
    print(f"# facet accumulation autotune: packed {acc['packed_ms']:.3f} ms, window {acc['window_ms']:.3f} ms "
          f"-> {acc['accumulate']}", file=sys.stderr, flush=True)
    p_plain = dict(p, accumulate=om.ACC_WINDOW if acc["accumulate"] == "window" else om.ACC_PACKED)
    runs = {m: A.AnnualRun.facets(site["a"], site["b"], site["X"], site["Y"], site["Z"], steps,
                                  p if m == "batch" else p_plain, batch_size=len(steps), mode=m)
            for m in ("batch", "plain")}
    R = runs["batch"].rays
    print(f"# facet annual {precision}: {len(steps)} time steps, {R} rays ({R // len(steps)} a step), scenes built in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    stream = torch.cuda.current_stream()
    ms = {}
    for m, run in runs.items():
        for _ in range(max(1, args.warmup)):
            run.trace()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            run.trace()
        e1.record(stream)
        torch.cuda.synchronize()
        ms[m] = e0.elapsed_time(e1) / args.steps
        print(f"# facet annual {precision} {m}: {ms[m]:.1f} ms per year ({R / (ms[m] * 1e-3):.4g} rays/s)", file=sys.stderr,
              flush=True)
    best = min(ms, key=ms.get)
    res = runs[best].results()
    print(f"# facet annual intercept ratio {res['annual_ratio']:.4f}", file=sys.stderr, flush=True)
    return {
        "metric": METRIC, "value": R / (ms[best] * 1e-3), "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms[best], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": ("facet_annual: the annual driver's faceted paraboloid heliostat (9 facets "
                                "0.47 m, a = b = %.2f m) at (12, 1.2, 107) m turned to each step's master "
                                "normal, 1.4 m receiver, reference defaults (Buie CSR 0.1, 1024 bins, slope "
                                "1 mrad, lattice 0.01 m x 2048 rays); the reference schedule at 30.3N 120.2E, "
                                "zenith <= 75 deg; one step = the whole year" % site["a"]),
                   "precision": precision, "geometry": GEOMETRY[precision],
                   "time_steps": len(steps), "rays": R, "mode": best, "batch_ms": ms["batch"],
                   "plain_ms": ms["plain"], "time_steps_per_s": len(steps) / (ms[best] * 1e-3),
                   "accumulate": acc, "annual_ratio": res["annual_ratio"], "parallelism": "single GPU"},
        "roofline": None, "cpu_baseline": None,
    }


def oracle_sample(configs, blocks_per_scene: int, blk: int = 40):
    """Oracle FP32 trace of the first ``blk`` heliostats of each scene (one block per scene):
    step/test counts for the cost model and the CPU baseline."""
    import oracle
    nthreads = host_threads()
    import onemirror_amd as om
    tot = {"rays": 0, "dda_steps_eff": 0, "cand_tests": 0, "self_skips": 0, "tag_rays": 0,
           "tag_shade_steps_eff": 0, "tag_shade_tests": 0, "tag_shade_self_skips": 0}
    secs = 0.0
    nhel = 0
    for d, p in configs:
        o = oracle.OracleScene(d, p)
        o.set_tags(om.clear_heliostats(d, p))  # shading walks the kernels skip
        H = o.n_helio
        for b in range(blocks_per_scene):
            h0 = (b * H) // max(1, blocks_per_scene)
            h1 = min(H, h0 + blk)
            t = time.perf_counter()
            r = o.trace_f32(h0, h1, nthreads=nthreads)
            secs += time.perf_counter() - t
            nhel += h1 - h0
            for k in tot:
                tot[k] += r["stats"][k]
    return tot, secs, nhel, nthreads


def bench_batched(args, with_cpu: bool = True) -> dict:
    """Batched mode: the scenes of config_batched (128^3 grids; --grid-tune picks 64, 128 or
    256 cells per axis by timing the persistent kernel on a sample) in one om_batch replayed
    from a hipGraph, with the launch the autotuner picks: one persistent launch over every
    scene, or the window trace of each scene in turn; the plain path (one om_trace_plain per
    scene, back to back on the stream) timed on the same scenes.  The value is the mixed
    trace's (OM_PREC_MIXED: the FP64 model's outcome of every ray, the metric's accuracy);
    the same batch in FP32 reference geometry is reported beside it."""
    import torch

    import onemirror_amd as om
    from onemirror_amd import fields as F

    cfg = F.config_batched(n_scenes=args.scenes)
    t0 = time.time()
    grid_cells = 128
    tune_grid = None
    if args.grid_tune:
        tune_grid = om.autotune(cfg, grid_cells=(64, 128, 256), sample=4, reps=1)
        grid_cells = tune_grid["grid_cells"]
        print(f"# grid autotune ({time.time() - t0:.1f}s): "
              + ", ".join(f"{r['grid_cells']}^3 {r['rays_per_s']:.3g} rays/s" for r in tune_grid["trials"]),
              file=sys.stderr, flush=True)
    t1 = time.time()
    configs = []
    for d, p in cfg:
        d = dict(d)
        d.update(F.grid_for(d["helio_pos"], d["helio_size"], float(d["grid_pos"][1] + d["grid_size"][1]),
                            grid_cells))
        configs.append((d, p))
    def run(precision):
        """The batch of the scenes in this precision: autotuned launch, graph replays, and
        the plain path (one accumulation-autotuned om_trace_plain per scene) on the same
        scenes."""
        pc = {"fp32": om.PREC_FP32, "mixed": om.PREC_MIXED}[precision]
        t1 = time.time()
        scenes = [om.SolarScene(d, dict(p, precision=pc)) for d, p in configs]
        torch.cuda.synchronize()
        R = sum(s.info["total_rays"] for s in scenes)
        Hs = [s.info["n_helio"] for s in scenes]
        print(f"# {precision}: {len(scenes)} scenes built in {time.time() - t1:.1f}s: {sum(Hs)} heliostats "
              f"({min(Hs)}..{max(Hs)}), {R} rays, grids {grid_cells}^3", file=sys.stderr, flush=True)
        maps = [s.new_maps(smooth=True) for s in scenes]
        tl = om.autotune_launch(scenes, maps, reps=2)
        print(f"# {precision} launch autotune: workgroup {tl['workgroup']} x {tl['rays_per_thread']} rays/thread "
              f"({', '.join(f'{w}x{r} {ms:.1f} ms' for w, r, ms in tl['trials'])})", file=sys.stderr, flush=True)
        b = om.Batch(scenes, maps, workgroup=tl["workgroup"], rays_per_thread=tl["rays_per_thread"])
        stream = torch.cuda.current_stream()
        for _ in range(max(1, args.warmup)):
            b.run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            b.run()
        e1.record(stream)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        batch_ms = e0.elapsed_time(e1) / args.steps
        redo, full = b.redo_count()
        for s, m in zip(scenes, maps):
            s.autotune_accumulate(m, reps=2)
            s.trace_plain(m)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(args.steps):
            for s, m in zip(scenes, maps):
                s.trace_plain(m)
        e1.record(stream)
        torch.cuda.synchronize()
        plain_ms = e0.elapsed_time(e1) / args.steps
        print(f"# {precision} batched {batch_ms:.2f} ms ({R / (batch_ms * 1e-3):.4g} rays/s), plain per scene "
              f"{plain_ms:.2f} ms ({R / (plain_ms * 1e-3):.4g} rays/s)"
              + (f", {redo} rays traced again in FP64, {full} full lists" if precision == "mixed" else ""),
              file=sys.stderr, flush=True)
        del b, maps, scenes
        torch.cuda.synchronize()
        return {"rays": R, "heliostats": sum(Hs), "dt": dt, "batch_ms": batch_ms, "plain_ms": plain_ms,
                "tl": tl, "redo_rays": redo, "full_lists": full}

    runs = {pr: run(pr) for pr in ("mixed", "fp32")}
    head = runs["mixed"]
    R, batch_ms, plain_ms, tl, dt = head["rays"], head["batch_ms"], head["plain_ms"], head["tl"], head["dt"]
    Hs_total = head["heliostats"]
    sample_cfg = configs[:: max(1, len(configs) // 8)]
    if not with_cpu:
        sample_cfg = []
    st, cpu_t, nhel, nthreads = oracle_sample(sample_cfg, 1) if sample_cfg else (
        {"rays": 0, "dda_steps_eff": 0, "cand_tests": 0, "self_skips": 0}, 0.0, 0, 0)
    ops_ray, S, T = model_ops(st)
    achieved = R * ops_ray / (batch_ms * 1e-3) / 1e12
    peak = P_VALU / 1e12
    # HBM bytes and VALU issue of the batch kernel from its committed rocprof run, when that
    # run traced this batch (the launch the autotuner picked)
    traffic, counters = None, None
    if tl["workgroup"] == 2048:
        prof = os.path.join(ROOT, "profiles", "trace_batch_window_mixed_kernel_batched.json")
        vp = os.path.join(ROOT, "profiles", "valu_probe_kernel.json")
        if os.path.exists(prof) and os.path.exists(vp):
            with open(prof) as f:
                pj = json.load(f)
            with open(vp) as f:
                probe = json.load(f)["probes"]["probe_fma"]
            if pj.get("rays") == R:
                traffic = pj["hbm_bytes_per_launch"]
                counters = {"valu_per_cu_busy_cycle": pj["valu_per_cu_busy_cycle"],
                            "probe_valu_per_cu_busy_cycle": probe["valu_per_cu_busy_cycle"],
                            "of_probe": pj["valu_per_cu_busy_cycle"] / probe["valu_per_cu_busy_cycle"],
                            "valu_util": pj["SQ_INSTS_VALU"] * 64 / (pj["avg_ns"] * 1e-9) / P_VALU,
                            "memory_side_atomics": pj["tcc_atomic"],
                            "source": "profiles/trace_batch_window_mixed_kernel_batched.json"}
    return {
        "metric": METRIC, "value": R / dt, "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"batched: {len(configs)} scenes, heliostats log-uniform 500-20000, "
                               "sun elevation 10-80 deg azimuth 90-270 deg, rays per scene "
                               "proportional to heliostats in [1e7, 2e8], Buie CSR 0.1, 256x256 maps",
                   "precision": "mixed", "geometry": GEOMETRY["mixed"],
                   "scenes": len(configs), "heliostats": Hs_total, "rays": R,
                   "grid_cells": grid_cells, "workgroup": tl["workgroup"],
                   "rays_per_thread": tl["rays_per_thread"], "batch_ms": batch_ms,
                   "plain_ms": plain_ms, "plain_rays_per_s": R / (plain_ms * 1e-3),
                   "batched_vs_plain": plain_ms / batch_ms,
                   "redo_rays": head["redo_rays"], "full_lists": head["full_lists"],
                   "fp32_reference": {
                       "value": runs["fp32"]["rays"] / runs["fp32"]["dt"], "unit": "rays/s",
                       "batch_ms": runs["fp32"]["batch_ms"], "plain_ms": runs["fp32"]["plain_ms"],
                       "batched_vs_plain": runs["fp32"]["plain_ms"] / runs["fp32"]["batch_ms"],
                       "workgroup": runs["fp32"]["tl"]["workgroup"],
                       "rays_per_thread": runs["fp32"]["tl"]["rays_per_thread"],
                       "geometry": GEOMETRY["fp32"]},
                   "grid_tuning": tune_grid["trials"] if tune_grid else None,
                   "launch_tuning": tl["trials"], "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "Tlane-op/s",
                     "frac": achieved / peak, "traffic": traffic, "ops_per_ray": ops_ray,
                     "dda_steps_per_ray": S, "tests_per_ray": T, "counters": counters},
        "cpu_baseline": None if not sample_cfg else {"value": st["rays"] / cpu_t, "unit": "rays/s", "cores": nthreads,
                         "kind": "port",
                         "sample": f"the first 40 heliostats of {len(sample_cfg)} of the scenes "
                                   f"({nhel} heliostats, {st['rays']} rays); oracle FP32 tracer, OpenMP",
                         "seconds": cpu_t},
    }


def bench_diff(args) -> dict:
    """Differentiable mode on the Mid field: forward splat (trace + Gaussian splat of the
    parametrised aims), the spillage + softmax-peak loss, and the adjoint pass, timed together
    per step; the plain trace of the same field timed beside it."""
    import numpy as np
    import torch

    import onemirror_amd as om
    import onemirror_amd.diff as D
    from onemirror_amd import fields as F

    desc, prm = F.config_mid()
    scene = om.SolarScene(desc, prm)
    H, R = scene.info["n_helio"], scene.info["total_rays"]
    rng = np.random.default_rng(4)
    th = np.zeros((H, 4), np.float32)
    th[:, :2] = rng.normal(0.0, 0.3, (H, 2))
    th[:, 2:] = rng.normal(0.0, 0.3, (H, 2))
    p = torch.tensor(th, device="cuda", requires_grad=True)
    with torch.no_grad():
        S0, _, _ = D.forward(scene, p)
    tau = float(0.05 * S0.max())

    def step():
        p.grad = None
        S, E = D.splat(scene, p)
        L = D.spill_peak_loss(S, E, tau, 1.0)
        L.backward()
        return L

    for _ in range(max(1, args.warmup)):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    t = time.perf_counter()
    e[0].record(stream)
    for _ in range(args.steps):
        L = step()
    e[1].record(stream)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    fb_ms = e[0].elapsed_time(e[1]) / args.steps
    # forward alone and the plain trace of the same field
    e[0].record(stream)
    for _ in range(args.steps):
        with torch.no_grad():
            D.forward(scene, p)
    e[1].record(stream)
    maps = scene.new_maps(smooth=False)
    scene.autotune_accumulate(maps, reps=2)  # the plain path as the plain bench runs it
    scene.trace_heliostats(0, H, maps)
    e[2].record(stream)
    for _ in range(args.steps):
        scene.trace_heliostats(0, H, maps)
    e[3].record(stream)
    torch.cuda.synchronize()
    fwd_ms = e[0].elapsed_time(e[1]) / args.steps
    plain_ms = e[2].elapsed_time(e[3]) / args.steps
    # the plain trace at the metric's accuracy (OM_PREC_MIXED) of the same field
    mixed = om.SolarScene(desc, dict(prm, precision=om.PREC_MIXED))
    mmaps = mixed.new_maps(smooth=False)
    mixed.autotune_accumulate(mmaps, reps=2)
    mixed.trace_heliostats(0, H, mmaps)
    e[2].record(stream)
    for _ in range(args.steps):
        mixed.trace_heliostats(0, H, mmaps)
    e[3].record(stream)
    torch.cuda.synchronize()
    plain_mixed_ms = e[2].elapsed_time(e[3]) / args.steps
    del mixed, mmaps
    print(f"# fwd+bwd {fb_ms:.2f} ms ({1e3 / fb_ms:.2f}/s), forward {fwd_ms:.2f} ms, plain trace "
          f"{plain_ms:.2f} ms (FP32), {plain_mixed_ms:.2f} ms (mixed), loss {float(L.detach()):.6f}",
          file=sys.stderr, flush=True)
    return {
        "metric": METRIC, "value": R / dt, "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "diff: the mid field (5000 heliostats, 2.01e8 rays), 4 FP32 parameters "
                               "per heliostat (aim offsets, tracking deltas), Gaussian splat sigma 1 px "
                               "on 256x256, loss spillage + softmax peak; one step = forward + loss + "
                               "adjoint",
                   "heliostats": H, "rays": R, "fwd_bwd_per_s": 1e3 / fb_ms, "fwd_bwd_ms": fb_ms,
                   "forward_ms": fwd_ms, "backward_ms": fb_ms - fwd_ms, "plain_trace_ms": plain_ms,
                   "fwd_bwd_vs_plain": plain_ms / fb_ms, "plain_trace_mixed_ms": plain_mixed_ms,
                   "fwd_bwd_vs_plain_mixed": plain_mixed_ms / fb_ms, "parallelism": "single GPU"},
        "roofline": None,
        "cpu_baseline": None,
    }


if __name__ == "__main__":
    sys.exit(main())
