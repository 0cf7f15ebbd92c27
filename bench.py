#!/usr/bin/env python3
"""Benchmark of the FLUXOS-OVERLAND time step on one MI355X.

Headline workload (the single large domain the metric is quoted on): a 4096 x
4096 basin, dx = 2 m, diamond-square fractal DEM (roughness 0.6, relief 8 m)
over a 0.3 % slope, melt ramping 0 -> 60 mm/day over 6 h, two solutes, on the
FP32-mixed path with the fast ADE.  A step is one loop iteration of the
reference's main.cpp:271-406; a cell-update is one interior cell advanced by
one step that ran the solver (hpall != 0, main.cpp:385).

The run is spun up on the device, untimed, from t = 0 to the wet state at
t = --start-t (3600 s: the melt has wetted the basin, every iteration runs the
solver), then W warm-up iterations run from there, then EXACTLY K timed
iterations (--steps K), all of which must run the solver.  The line's value
and the FP64 water ledger of the window come from those K iterations, timed
without instrumentation; the kernel times of the roofline come from the next
K iterations, with HIP events around every launch on the library's own
stream.  --steps 0 times the whole 6-hour run from t = 0 instead (one launch
in four timed, in the run).

Beside it, from the same snapshot at t = --start-t: the FP64 parity path over
a window of its own, FP32 depth error against FP64 next to FP64's own rounding
sensitivity, and the reference's CPU path (the oracle, OpenMP) on a bounded
sample with a per-phase split; then the 64-scenario batched ensemble, the
served request stream (p50 / p99 time-to-result) and the 16384^2 domain.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# the whole-run mode brackets one launch in four of each kernel with HIP events
# (every launch costs about 2% of the step time in event records)
PROFILE_EVERY_FULL_RUN = 4
METRIC = "cell-updates/s FP32-mixed vs FP64 at <=1e-3 m h error; p50/p99 scenario latency"
PHASES = ("cfl", "forcing", "hydro", "ade")


# algorithmic bytes per cell-update (read d|z, qx, qy, zb; write d|z, qx, qy;
# plus c read+write per solute) -- flags, masks, persistent fluxes and
# wet-time tracer are implementation overhead and not counted
def b_alg(precision: str, nchem: int) -> int:
    return (28 + 8 * nchem) if precision == "fp32" else (56 + 16 * nchem)


def measured_traffic(n: int, kernel: str):
    """HBM bytes per launch of a kernel at this workload size and the revision
    they were measured at, from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    committed as profiles/hbm_traffic_<n>x<n>.txt, or (None, None)."""
    f = ROOT / "profiles" / f"hbm_traffic_{n}x{n}.txt"
    if not f.exists():
        return None, None
    rev = None
    val = None
    for line in f.read_text().splitlines():
        if line.startswith("# revision:"):
            rev = line.split(":", 1)[1].strip()
        parts = line.split()
        if len(parts) == 2 and not line.startswith("#") and parts[0].split("<")[0] == kernel:
            val = int(parts[1])
    return val, rev


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spin_up(dom, t_spin: float, max_steps: int) -> tuple[float, int]:
    """Advance the device loop until simulated time t_spin (bounded)."""
    done = 0
    t = dom.clock()[0]
    print_next = 600.0 * (int(t // 600.0) + 1)
    while t < t_spin and done < max_steps:
        info = dom.advance(1e30, min(print_next, t_spin), max_steps - done)
        done += info.steps
        t = info.tim
        if t >= print_next:
            print_next += 600.0
        if info.steps == 0:
            break
    return t, done


def physical_cores() -> int | None:
    """Physical cores of the host (lscpu: distinct core/socket pairs)."""
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    pairs = {tuple(line.split(",")) for line in out.splitlines() if line and not line.startswith("#")}
    return len(pairs) or None


def cpu_threads() -> int:
    """Threads of the all-core CPU run: the process's CPU share (affinity,
    capped by OMP_NUM_THREADS when the environment sets it)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(snap: dict, sc, init, steps_all: int, steps_one: int, threads: int, host: dict) -> dict:
    """The oracle (the reference's FP64 OpenMP path restated in C: CFL, melt
    and ADE serial, the hydro loops in one parallel region, as main.cpp:300-397
    and hydrodynamics_calc.cpp:49) from the GPU run's snapshot: steps_all loop
    iterations on `threads` threads, then steps_one on one thread, each with
    its wall time split by phase.  host: the host's core counts, read before
    the OpenMP runtime bound this thread."""
    from oracle.oracle import OracleDomain
    from tests.helpers import STATE_D, STATE_F
    o = OracleDomain(sc.nrows, sc.ncols, sc.dxy, sc.hdry, sc.nchem, sc.d_coef, True)
    o.g["zb"][...] = init.zb
    o.g["ks"][...] = init.ks
    o.g["weir"][...] = init.weir
    for k in STATE_D + STATE_F:
        o.g[k][...] = snap[k]
    for k in range(sc.nchem):
        o.conc[k][...] = snap["conc"][k]
    t, dt, it = snap["clock"]
    o.tim, o.dtfl = t, dt
    o.it.value = it
    o.set_meteo(sc.meteo)
    cells = sc.nrows * sc.ncols
    bind = {0: "false", 1: "true", 2: "primary", 3: "close", 4: "spread"}.get(o.lib.orc_proc_bind(), "?")

    def run(nthr: int, nsteps: int) -> dict:
        o.lib.orc_set_threads(nthr)
        ph = np.zeros(4)
        solver = 0
        t0 = time.perf_counter()
        for _ in range(nsteps):
            flags, _ = o.step_timed(1e30, ph)
            solver += flags & 1
        el = time.perf_counter() - t0
        return {"value": round(cells * solver / el, 1), "threads": nthr, "iterations": nsteps,
                "solver_steps": solver, "wall_s": round(el, 3),
                "phase_ms_per_iteration": {k: round(v * 1e3 / nsteps, 2) for k, v in zip(PHASES, ph)}}

    allc = run(threads, steps_all)
    one = run(1, steps_one) if steps_one > 0 else None
    return {"value": allc["value"], "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": (f"{steps_all} loop iterations of the {sc.nrows}x{sc.ncols} two-solute basin on {threads} "
                       f"threads (OMP_PROC_BIND={bind}), then {steps_one} on one thread, from the GPU run's "
                       f"state at t={t:.0f}s; CFL, melt and ADE serial and the hydro loops parallel, as in "
                       f"the reference"),
            "all_threads": allc, "one_thread": one, "omp_proc_bind": bind, **host}


def snapshot(src) -> dict:
    """Every grid a domain needs to restart, in the reference layout."""
    out = {k: src.get(k) for k in ("z", "h", "qx", "qy", "fe1", "fe2", "fe3", "fn1", "fn2", "fn3",
                                   "twet", "ldry", "ldry_prev", "h0")}
    out["conc"] = [src.get_conc(k) for k in range(src.nchem)]
    out["clock"] = src.clock()
    return out


def restore(dst, snap: dict) -> None:
    for k, v in snap.items():
        if k in ("conc", "clock"):
            continue
        dst.set(k, v)
    for k, c in enumerate(snap["conc"]):
        dst.set_conc(k, c)
    t, dt, it = snap["clock"]
    dst.set_clock(t, dt, it)


def read_kernels(dom) -> dict:
    from fluxos_cpp_amd import capi
    kt = {}
    for kid, name in capi.KERNEL_NAMES.items():
        ms, n = dom.profile_read(kid)
        if n:
            kt[name] = (ms, n)
    return kt


def timed(dom, steps: int, every: int = 1):
    dom.synchronize()
    dom.profile(True, every)
    dom.synchronize()
    t0 = time.perf_counter()
    info = dom.advance(1e30, 1e30, steps)
    dom.synchronize()
    el = time.perf_counter() - t0
    kt = read_kernels(dom)
    dom.profile(False)
    return info, el, kt


def accuracy(sc, init, snap, spans: str) -> dict:
    """Depth error of the FP32-mixed path against the FP64 path, both advanced
    on the device from the same snapshot to equal simulated times, beside the
    FP64 path's own sensitivity to rounding: FP64 from the snapshot with the
    water surface of every wet cell moved by one ulp up or down (seeded).  The
    hydrodynamics do not depend on the solutes, so these runs carry none."""
    from fluxos_cpp_amd import capi
    ts = [snap["clock"][0] + float(x) for x in spans.split(",")]
    doms = {}
    for name, prec in (("fp64", capi.FP64_PARITY), ("fp32", capi.FP32_MIXED), ("fp64_ulp", capi.FP64_PARITY)):
        d = capi.HipDomain.from_scenario(sc, precision=prec, ade=False, init=init)
        restore(d, dict(snap, conc=[]))
        doms[name] = d
    z = doms["fp64_ulp"].get("z")
    wet = np.zeros_like(z, dtype=bool)
    wet[1:-1, 1:-1] = snap["h"][1:-1, 1:-1] > sc.hdry
    up = np.random.default_rng(11).random(z.shape) < 0.5
    z = np.where(wet, np.where(up, np.nextafter(z, np.inf), np.nextafter(z, -np.inf)), z)
    doms["fp64_ulp"].set("z", z)
    out = []
    for t in ts:
        steps = {}
        for name, d in doms.items():
            info = d.advance(1e30, t, 10 ** 7)
            steps[name] = info.steps
        h64 = doms["fp64"].get("h")[1:-1, 1:-1]
        row = {"after_s": round(t - snap["clock"][0], 3), "steps_fp64": steps["fp64"],
               "steps_fp32": steps["fp32"]}
        for name in ("fp32", "fp64_ulp"):
            dh = np.abs(doms[name].get("h")[1:-1, 1:-1] - h64)
            row[f"{name}_max_abs_h_m"] = float(dh.max())
            row[f"{name}_rms_h_m"] = float(np.sqrt((dh ** 2).mean()))
        w64 = int((doms["fp64"].get("ldry")[1:-1, 1:-1] == 0).sum())
        w32 = int((doms["fp32"].get("ldry")[1:-1, 1:-1] == 0).sum())
        row["wet_fp64"], row["wet_fp32"] = w64, w32
        row["within_1e-3_m"] = row["fp32_max_abs_h_m"] <= 1e-3
        out.append(row)
    for d in doms.values():
        d.close()
    return {"from_t": round(snap["clock"][0], 2), "spans_s": out}


def whole_run_accuracy(sc, init) -> list:
    """FP32-mixed against FP64, both from t = 0, at every print time of the
    whole run (no solutes: the hydrodynamics do not depend on them)."""
    from fluxos_cpp_amd import capi
    d64 = capi.HipDomain.from_scenario(sc, precision=capi.FP64_PARITY, ade=False, init=init)
    d32 = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=False, init=init)
    out = []
    pn = float(sc.print_step)
    while pn <= sc.t_end:
        d64.advance(float(sc.t_end), pn)
        d32.advance(float(sc.t_end), pn)
        h64 = d64.get("h")[1:-1, 1:-1]
        dh = np.abs(d32.get("h")[1:-1, 1:-1] - h64)
        out.append({"t_s": pn, "fp32_max_abs_h_m": float(dh.max()), "fp32_rms_h_m": float(np.sqrt((dh ** 2).mean())),
                    "max_h_m": float(h64.max()),
                    "wet_fp64": int((d64.get("ldry")[1:-1, 1:-1] == 0).sum()),
                    "wet_fp32": int((d32.get("ldry")[1:-1, 1:-1] == 0).sum())})
        pn += sc.print_step
    d64.close()
    d32.close()
    return out


def accuracy_configs(spans: str) -> dict:
    """FP32-mixed depth error against FP64 for each configuration's generator,
    from a wet snapshot (FP32-mixed spin-up on the device) over the given
    spans, beside FP64's one-ulp floor; C1 also over its whole run from t = 0."""
    from fluxos_cpp_amd import capi, gen
    cfgs = [("C1 256x256 (full size)", gen.scenario_c1(256), 3600.0),
            ("C3 member 4 (256x256, 14.8 mm/day)", gen.scenario_c3(4), 1800.0),
            ("C3 member 9 (288x288, 29.5 mm/day)", gen.scenario_c3(9), 1800.0),
            ("C3 member 8 (362x362, 66.0 mm/day)", gen.scenario_c3(8), 1800.0),
            ("C5 generator at 4096x4096 (dx 1 m, channels, melt multiplier)",
             gen.scenario_c5(4096, t_end=10800), 1800.0)]
    out = {}
    for name, sc, t0 in cfgs:
        t1 = time.perf_counter()
        init = capi.prepare(sc)
        d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=False, init=init)
        spin_up(d, t0, 10 ** 7)
        snap = snapshot(d)
        d.close()
        row = accuracy(sc, init, snap, spans)
        if name.startswith("C1"):
            row["whole_run_from_t0"] = whole_run_accuracy(sc, init)
        out[name] = row
        log(f"[bench] accuracy {name}: {time.perf_counter() - t1:.1f}s; "
            + ", ".join(f"{r['after_s']}s {r['fp32_max_abs_h_m']:.2e} (ulp {r['fp64_ulp_max_abs_h_m']:.2e})"
                        for r in row["spans_s"]))
    return out


def largest_domain(nsteps: int, t_spin: float, spin_wall: float = 60.0) -> dict:
    """The largest single domain: 16384 x 16384, dx = 1 m, fractal DEM with
    carved channels, a smooth melt-multiplier field (10-60 mm/day), four
    solutes, FP32-mixed.  Spun up on the device (untimed) to t_spin, then
    nsteps timed loop iterations (the spin-up stops after spin_wall seconds of
    wall time at the latest; the window's start time is reported)."""
    from fluxos_cpp_amd import capi, gen
    t0 = time.perf_counter()
    sc = gen.scenario_c5(16384)
    log(f"[bench] largest domain: generated in {time.perf_counter() - t0:.1f}s")
    init = capi.prepare(sc)
    log(f"[bench] largest domain: read_geo + initiation at {time.perf_counter() - t0:.1f}s")
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, init=init)
    del init
    v0 = d.mass_balance()[0]
    log(f"[bench] largest domain: on the device at {time.perf_counter() - t0:.1f}s")
    # spin-up in chunks of simulated time, at most spin_wall seconds of wall time
    t_sp, n_sp, pn = 0.0, 0, 600.0
    t1 = time.perf_counter()
    while t_sp < t_spin and time.perf_counter() - t1 < spin_wall:
        info = d.advance(1e30, min(pn, t_spin), 2000)
        n_sp += info.steps
        t_sp = info.tim
        if t_sp >= pn:
            pn += 600.0
        log(f"[bench] largest domain: spin-up t={t_sp:.1f}s, {n_sp} steps, {time.perf_counter() - t0:.1f}s")
        if info.steps == 0:
            break
    info, el, kt = timed(d, nsteps, 1)
    cells = sc.nrows * sc.ncols
    v, melt, weir = d.mass_balance()
    alg = {"k32_step": b_alg("fp32", 0), "k32_ade": 8 * sc.nchem}
    kern = {}
    for k, (ms, n) in kt.items():
        row = {"avg_us": round(ms / n * 1e3, 1), "launches_timed": n}
        if k in alg:
            gbs = cells * alg[k] / (ms / n * 1e-3) / 1e9
            row.update(bytes_per_cell=alg[k], achieved_gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
        kern[k] = row
    out = {"workload": "16384x16384, dx=1 m, fractal + carved channels, melt multiplier 10-60 mm/day, "
                       "4 solutes, FP32-mixed with fast ADE",
           "cells": cells, "window_from_t": round(t_sp, 1), "spinup_steps": n_sp,
           "steps": info.steps, "solver_steps": info.solver_steps,
           "cell_updates_per_s": round(cells * info.solver_steps / el, 1),
           "ms_per_step": round(el * 1e3 / max(1, info.steps), 3),
           "step_roofline_frac": round(cells * info.solver_steps * b_alg("fp32", sc.nchem) / el / 1e9
                                       / HBM_PEAK_GBS, 4),
           "kernels": kern, "tiling_ch_ach": list(d.tiling()), "device_bytes": d.device_bytes,
           "mass_rel_err": abs(v - v0 - melt - weir) / max(abs(v), 1e-300)}
    d.close()
    log(f"[bench] largest domain: {out['cell_updates_per_s']:.3e} cell-updates/s, {out['ms_per_step']} ms/step "
        f"from t={t_sp:.0f}s")
    return out


def ledger(v0, v1, melt0, melt1, weir0, weir1) -> dict:
    """FP64 water ledger between two readings: volume change against melt
    input plus net weir-face exchange."""
    dv, dm, dw = v1 - v0, melt1 - melt0, weir1 - weir0
    return {"volume_change_m3": dv, "melt_in_m3": dm, "weir_in_m3": dw,
            "rel_err": abs(dv - dm - dw) / max(abs(v1), 1e-300)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200,
                    help="timed loop iterations from the wet state at --start-t "
                         "(0 = the whole run, t = 0 .. t_end)")
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed iterations from the wet state before the timed ones")
    ap.add_argument("--start-t", type=float, default=3600.0,
                    help="simulated time of the wet state the timed window starts from [s]")
    ap.add_argument("--tiling", type=str, default="",
                    help="ch,ach column-block lengths (default: the autotuner's table)")
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--fp64-steps", type=int, default=20)
    ap.add_argument("--accuracy-spans", type=str, default="1,10,60,600",
                    help="simulated seconds after the snapshot at which depths are compared")
    ap.add_argument("--cpu-steps", type=int, default=12, help="oracle iterations on all threads")
    ap.add_argument("--cpu-steps-1t", type=int, default=2, help="oracle iterations on one thread")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = the process's CPU share")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--ensemble", type=int, default=64,
                    help="scenarios of the batched ensemble leg (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="arrival window of the served request stream (0 = skip); the stated "
                         "workload's window is 300 s")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate [1/s]")
    ap.add_argument("--serve-rate-sustained", type=float, default=2.0,
                    help="a second, lower arrival rate for the serving leg (0 = skip)")
    ap.add_argument("--c5-steps", type=int, default=100,
                    help="timed iterations of the 16384x16384 four-solute domain (0 = skip)")
    ap.add_argument("--c5-start-t", type=float, default=3600.0,
                    help="untimed spin-up of that domain [simulated s]")
    ap.add_argument("--accuracy-configs", action="store_true",
                    help="also measure FP32-mixed vs FP64 depth error for the C1, C3 and C5 generators")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (calibrates HBM counters)")
    args = ap.parse_args()

    if args.copy_calib_mib:
        import torch
        n = args.copy_calib_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_calib_mib} MiB x3 (read+write {2 * args.copy_calib_mib} MiB each)")
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the path is a "
              "single-device kernel (row-strip sharding across GPUs is not built)", file=sys.stderr)
        return 2

    # the CPU baseline's OpenMP threads stay on neighbouring cores; read by the
    # OpenMP runtime when the oracle library first loads
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    # the process's CPU share, read before the OpenMP runtime binds this thread
    # to one place (it is given back after the CPU baseline)
    affinity = os.sched_getaffinity(0)
    host = {"physical_cores_lscpu": physical_cores(), "affinity_cpus": len(affinity)}
    cpu_thr = args.cpu_threads or cpu_threads()

    from fluxos_cpp_amd import capi, gen
    capi.load()
    t0 = time.perf_counter()
    sc = gen.scenario_c2(args.n)
    init = capi.prepare(sc)
    log(f"[bench] generated {args.n}^2 basin in {time.perf_counter() - t0:.1f}s")
    cells = sc.nrows * sc.ncols

    d32 = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, init=init)
    if args.tiling:
        d32.set_tiling(*(int(x) for x in args.tiling.split(",")))
    tiling = d32.tiling()              # from the autotuner's table (fluxos_cpp_amd/tuning.json)
    mb_start = d32.mass_balance()
    if args.steps > 0:
        # untimed: t = 0 .. start_t on the device, the snapshot, then W warm-up
        # iterations from the wet state; timed: exactly K iterations
        t1 = time.perf_counter()
        t_sp, n_sp = spin_up(d32, args.start_t, 10 ** 7)
        log(f"[bench] untimed spin-up to t={t_sp:.1f}s in {n_sp} steps, {time.perf_counter() - t1:.1f}s")
        snap = snapshot(d32)
        if args.warmup > 0:
            d32.advance(1e30, 1e30, args.warmup)
            d32.synchronize()
        mb0 = d32.mass_balance()
        t_from = d32.clock()[0]
        # the timed window, uninstrumented: HIP events around every launch
        # cost about 3.5% of the step here
        d32.synchronize()
        t1 = time.perf_counter()
        info = d32.advance(1e30, 1e30, args.steps)
        d32.synchronize()
        el = time.perf_counter() - t1
        mb1 = d32.mass_balance()
        steps, solver_steps, t_reached = info.steps, info.solver_steps, info.tim
        if not (steps == args.steps and solver_steps == steps):
            print(f"bench.py: the timed window ran {steps} iterations, {solver_steps} of them with the "
                  f"solver (asked for {args.steps}); start it from a wet state (--start-t)", file=sys.stderr)
            return 3
        # the kernel times of the roofline: the next K iterations, with HIP
        # events around every launch
        info_p, el_p, kt = timed(d32, args.steps, 1)
        profiled = {"ms_per_step": round(el_p * 1e3 / max(1, info_p.steps), 4), "steps": info_p.steps,
                    "solver_steps": info_p.solver_steps, "from_t": round(t_reached, 3),
                    "to_t": round(info_p.tim, 3)}
        window = f"{args.steps} iterations from t={t_from:.1f}s (spun up untimed, then {args.warmup} warm-up)"
    else:
        # the whole run t = 0 .. t_end, device-resident between print times
        profiled = None                 # kernel times come from this run itself
        snap = None
        d32.synchronize()
        d32.profile(True, PROFILE_EVERY_FULL_RUN)
        d32.synchronize()
        steps = solver_steps = 0
        print_next = float(sc.print_step)
        mb0 = mb_start
        t_from = 0.0
        t1 = time.perf_counter()
        while True:
            info = d32.advance(float(sc.t_end), print_next)
            steps += info.steps
            solver_steps += info.solver_steps
            if info.status & 2 or info.steps == 0:
                break
            if info.status & 1:
                print_next += sc.print_step
        d32.synchronize()
        el = time.perf_counter() - t1
        t_reached = info.tim
        kt = read_kernels(d32)
        d32.profile(False)
        mb1 = d32.mass_balance()
        window = "the whole 6 simulated hours from t=0"
    value = cells * solver_steps / el
    ms_step = el * 1e3 / max(1, steps)
    log(f"[bench] fp32 window to t={t_reached:.1f}s: {steps} steps ({solver_steps} with the solver) in {el:.3f}s")
    log("[bench] fp32 kernels: " + ", ".join(f"{k}={v[0] / max(1, v[1]) * 1e3:.1f}us x{v[1]}" for k, v in kt.items()))

    # FP64 water ledger: the timed window, and the run since t = 0
    mass = {"window": ledger(mb0[0], mb1[0], mb0[1], mb1[1], mb0[2], mb1[2]),
            "since_t0": ledger(mb_start[0], mb1[0], mb_start[1], mb1[1], mb_start[2], mb1[2]),
            "volume_m3": mb1[0],
            # solute kept in the non-weir cells against the FP64 ledger of solute the
            # melt brought into water (the reference's scheme zeroes it in cells that
            # dry and in weir cells, so the two differ by those sinks)
            "solute": [dict(zip(("mass_g", "brought_in_g"), d32.solute_balance(k))) for k in range(sc.nchem)]}

    # roofline of the dominant kernel (algorithmic bytes: hydro 28 B per
    # cell-update, ADE 8 B per solute per cell-update), from the HIP events of
    # the timed launches
    alg = {"k32_step": b_alg("fp32", 0), "k32_ade": 8 * sc.nchem}
    prof_name = {"k32_step": "k32_sweep", "k32_ade": "k32_ade_sweep"}
    kernels = {}
    for name, (ms, n) in kt.items():
        avg_us = ms / n * 1e3
        row = {"avg_us": round(avg_us, 2), "launches_timed": n}
        if name in alg:
            gbs = cells * alg[name] / (avg_us * 1e-6) / 1e9
            row.update(bytes_per_cell=alg[name], achieved_gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
        kernels[name] = row
    dom_name = max(alg, key=lambda k: kt[k][0] if k in kt else 0.0)
    dk = kernels[dom_name]
    traffic, traffic_rev = measured_traffic(args.n, prof_name[dom_name])
    roofline = {"bound": "hbm", "achieved": dk["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dk["frac"], "traffic": traffic,
                "traffic_source": (f"rocprofv3 FETCH_SIZE/WRITE_SIZE per launch, profiles/hbm_traffic_"
                                   f"{args.n}x{args.n}.txt, kernels at revision {traffic_rev}"
                                   if traffic is not None else None),
                "kernel": prof_name[dom_name], "bytes_per_cell": dk["bytes_per_cell"], "avg_us": dk["avg_us"],
                "launches_timed": dk["launches_timed"]}
    # the whole step (every kernel of an iteration) on the algorithmic bytes
    step_frac = cells * b_alg("fp32", sc.nchem) * solver_steps / el / 1e9 / HBM_PEAK_GBS

    if snap is None:                    # whole-run mode: a snapshot for the legs below
        d32.close()
        d32 = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, init=init)
        spin_up(d32, args.start_t, 10 ** 7)
        snap = snapshot(d32)
    d32.close()

    fp64 = None
    err = None
    if not args.no_fp64:
        d64 = capi.HipDomain.from_scenario(sc, precision=capi.FP64_PARITY, ade=True, init=init)
        restore(d64, snap)
        d64.advance(1e30, 1e30, 2)
        info64, el64, kt64 = timed(d64, args.fp64_steps)
        log("[bench] fp64 kernels: " + ", ".join(f"{k}={v[0] / max(1, v[1]) * 1e3:.1f}us x{v[1]}" for k, v in kt64.items()))
        v64 = cells * info64.solver_steps / el64
        fp64 = {"value": round(v64, 1), "ms_per_step": round(el64 * 1e3 / max(1, info64.steps), 3),
                "steps": info64.steps, "solver_steps": info64.solver_steps,
                "window_from_t": round(snap["clock"][0], 1),
                "kernels_avg_us": {k: round(ms / n * 1e3, 1) for k, (ms, n) in kt64.items()},
                "fp32_over_fp64": round(value / v64, 1) if v64 > 0 else None}
        d64.close()
        t1 = time.perf_counter()
        err = accuracy(sc, init, snap, args.accuracy_spans)
        log(f"[bench] accuracy spans in {time.perf_counter() - t1:.1f}s")

    cpu = None
    if not args.no_cpu:
        t1 = time.perf_counter()
        try:
            cpu = cpu_baseline(snap, sc, init, args.cpu_steps, args.cpu_steps_1t, cpu_thr, host)
        finally:
            os.sched_setaffinity(0, affinity)
        log(f"[bench] cpu baseline in {time.perf_counter() - t1:.1f}s: {cpu['value']:.3e} cell-updates/s on "
            f"{cpu['cores']} threads")
    del snap

    ensemble = None
    if args.ensemble > 0:
        from fluxos_cpp_amd import serve
        t1 = time.perf_counter()
        scens = [gen.scenario_c3(i) for i in range(args.ensemble)]
        inits = [capi.prepare(s) for s in scens]
        log(f"[bench] ensemble: {len(scens)} scenarios generated in {time.perf_counter() - t1:.1f}s")
        ensemble = serve.run_ensemble(scens, inits)
        ensemble["workload"] = (f"{args.ensemble} scenarios, side log-uniform 256-2048, dx=2 m, fractal DEMs, "
                                "melt U(5,80) mm/day, 1 solute, 1 simulated hour each, one batch")
        ensemble["roofline_frac"] = round(ensemble["cell_updates_per_s"] * b_alg("fp32", 1) / 1e9 / HBM_PEAK_GBS, 4)
        log(f"[bench] ensemble: {ensemble['cell_updates_per_s']:.3e} cell-updates/s over {ensemble['wall_s']:.1f}s")
        del scens, inits

    serving = None
    if args.serve_seconds > 0:
        from fluxos_cpp_amd import serve
        serving = {}
        for label, rate in (("stated_rate", args.serve_rate), ("sustained_rate", args.serve_rate_sustained)):
            if rate <= 0:
                continue
            reqs = serve.c4_stream(rate=rate, seconds=args.serve_seconds)
            if not reqs:                                # a short window can draw no arrival
                serving[label] = {"requests": 0, "arrival_rate_per_s": rate,
                                  "arrival_window_s": args.serve_seconds}
                continue
            r = serve.serve(reqs)
            r["arrival_rate_per_s"] = rate
            r["arrival_window_s"] = args.serve_seconds
            mean_cells = float(np.mean([q.cells for q in reqs]))
            r["offered_cell_updates_per_s"] = rate * mean_cells * r.get("mean_steps", 0.0)
            serving[label] = r
            log(f"[bench] serving at {rate}/s: {r['requests']} requests, p50 {r.get('p50_s', float('nan')):.2f}s "
                f"p99 {r.get('p99_s', float('nan')):.2f}s, {r['cell_updates_per_s']:.3e} cell-updates/s")
        serving["workload"] = ("Poisson arrivals; side {512,1024,2048} p={0.6,0.3,0.1}; fractal DEMs; melt "
                               "U(5,80) mm/day; 1 solute; 30 simulated minutes each; HBM budget 200 GB; "
                               "continuous batching; time-to-result = arrival to final depth on the host")
        serving["reduced"] = args.serve_seconds < 300.0
        if serving["reduced"]:
            serving["note"] = (f"reduced: {args.serve_seconds:g}-s arrival window of the stated 300 s "
                               "(the full stream is ~23 GPU-minutes of serving at today's throughput)")

    acc_cfg = None
    if args.accuracy_configs:
        acc_cfg = accuracy_configs(args.accuracy_spans)

    c5 = None
    if args.c5_steps > 0:
        c5 = largest_domain(args.c5_steps, args.c5_start_t)

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 depth, dt, mass)", "data": "synthetic",
        "config": {"workload": f"single domain {args.n}x{args.n}, dx=2 m, fractal DEM + 0.3% slope, melt 0-60 "
                               f"mm/day over 6 h, 2 solutes, FP32-mixed with fast ADE; timed: {window}",
                   "cells": cells, "nchem": sc.nchem, "window_from_t": round(t_from, 1),
                   "t_end_s": round(t_reached, 1), "loop_iterations": steps, "solver_steps": solver_steps,
                   "tiling_ch_ach": list(tiling)},
        "roofline": roofline,
        "kernels": kernels,
        # the K iterations after the timed window, with HIP events around
        # every launch (the kernels' times above), and their own wall time
        "kernels_window": profiled,
        "step_roofline_frac": round(step_frac, 4),
        "mass_balance": mass,
        "fp64_parity": fp64, "fp32_vs_fp64": err, "fp32_vs_fp64_by_config": acc_cfg,
        "cpu_baseline": cpu,
        "ensemble": ensemble,
        "serving": serving,
        "largest_domain": c5,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
