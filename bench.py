#!/usr/bin/env python3
"""Benchmark of the FLUXOS-OVERLAND time step on one MI355X.

Headline workload (the single large domain the metric is quoted on): a 4096 x
4096 basin, dx = 2 m, diamond-square fractal DEM (roughness 0.6, relief 8 m)
over a 0.3 % slope, melt ramping 0 -> 60 mm/day over 6 h, two solutes, run on
the FP32-mixed path for the whole 6 simulated hours (about 10^5 loop
iterations), device-resident between print times, after W warm-up iterations
of a throwaway domain.  A step is one loop iteration of the reference's
main.cpp:271-406; a cell-update is one interior cell advanced by one step
that ran the solver (hpall != 0).

Beside it, from a snapshot of the same run at t = 3600 s: the FP64 parity path
and FP32-mixed over the same window, FP32 depth error against FP64 next to
FP64's own rounding sensitivity, and the reference's CPU path (the oracle,
OpenMP) on a bounded sample; then the 64-scenario batched ensemble and the
served request stream (p50 / p99 time-to-result).

Prints ONE JSON line: metric, value (cell-updates/s), roofline of the dominant
kernel (HIP events on the library's own stream over the timed run), and the
legs above.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# the timed FP32 runs bracket one launch in four of each kernel with HIP events
# (every launch costs about 2% of the step time in event records)
PROFILE_EVERY = 4
METRIC = "cell-updates/s FP32-mixed vs FP64 at <=1e-3 m h error; p50/p99 scenario latency"

# algorithmic bytes per cell-update (read d|z, qx, qy, zb; write d|z, qx, qy;
# plus c read+write per solute) -- flags, masks, persistent fluxes and
# wet-time tracer are implementation overhead and not counted
def b_alg(precision: str, nchem: int) -> int:
    return (28 + 8 * nchem) if precision == "fp32" else (56 + 16 * nchem)


def measured_traffic(n: int, kernel: str):
    """HBM bytes per launch of a kernel at this workload size, as measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters (profiles/hbm_traffic_<n>x<n>.txt),
    or None when no measurement for this size is committed."""
    f = ROOT / "profiles" / f"hbm_traffic_{n}x{n}.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 2 and not line.startswith("#") and parts[0].split("<")[0] == kernel:
            return int(parts[1])
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spin_up(dom, t_spin: float, max_steps: int) -> tuple[float, int]:
    """Advance the device loop until simulated time t_spin (bounded)."""
    done = 0
    t = 0.0
    print_next = 600.0
    while t < t_spin and done < max_steps:
        info = dom.advance(1e30, print_next, max_steps - done)
        done += info.steps
        t = info.tim
        if t >= print_next:
            print_next += 600.0
        if info.steps == 0:
            break
    return t, done


def cpu_baseline(dom, sc, nsteps: int, threads: int | None):
    """The oracle (the reference's FP64 OpenMP path restated in C) on a
    bounded sample: nsteps loop iterations of the same domain from the
    device's current state."""
    from oracle.oracle import OracleDomain
    from tests.helpers import STATE_D, STATE_F
    if threads:
        os.environ["OMP_NUM_THREADS"] = str(threads)
    o = OracleDomain(sc.nrows, sc.ncols, sc.dxy, sc.hdry, sc.nchem, sc.d_coef, True)
    from fluxos_cpp_amd import capi
    st = capi.prepare(sc)
    o.g["zb"][...] = st.zb
    o.g["ks"][...] = st.ks
    o.g["weir"][...] = st.weir
    for k in STATE_D:
        o.g[k][...] = dom.get(k)
    for k in STATE_F:
        o.g[k][...] = dom.get(k)
    for k in range(sc.nchem):
        o.conc[k][...] = dom.get_conc(k)
    t, dt, it = dom.clock()
    o.tim, o.dtfl = t, dt
    o.it.value = it
    o.set_meteo(sc.meteo)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        o.step(1e30)
    el = time.perf_counter() - t0
    return sc.nrows * sc.ncols * nsteps / el, el, o.lib.orc_num_threads()


def snapshot(src, fields_only=False) -> dict:
    """Every grid a domain needs to restart, in the reference layout."""
    from fluxos_cpp_amd import capi
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
    h = snap.get("h")
    for k, c in enumerate(snap["conc"]):
        dst.set_conc(k, c)
    t, dt, it = snap["clock"]
    dst.set_clock(t, dt, it)


def timed(dom, steps: int, every: int = 1):
    dom.synchronize()
    dom.profile(True, every)
    dom.synchronize()
    t0 = time.perf_counter()
    info = dom.advance(1e30, 1e30, steps)
    dom.synchronize()
    el = time.perf_counter() - t0
    from fluxos_cpp_amd import capi
    kt = {}
    for kid, name in capi.KERNEL_NAMES.items():
        ms, n = dom.profile_read(kid)
        if n:
            kt[name] = (ms, n)
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
        out.append(row)
    for d in doms.values():
        d.close()
    return {"from_t": round(snap["clock"][0], 2), "spans_s": out}


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
    info, el, kt = timed(d, nsteps, PROFILE_EVERY)
    cells = sc.nrows * sc.ncols
    v, melt, weir = d.mass_balance()
    out = {"workload": "16384x16384, dx=1 m, fractal + carved channels, melt multiplier 10-60 mm/day, "
                       "4 solutes, FP32-mixed with fast ADE",
           "cells": cells, "window_from_t": round(t_sp, 1), "spinup_steps": n_sp,
           "steps": info.steps, "solver_steps": info.solver_steps,
           "cell_updates_per_s": round(cells * info.solver_steps / el, 1),
           "ms_per_step": round(el * 1e3 / max(1, info.steps), 3),
           "kernels_avg_us": {k: round(ms / n * 1e3, 1) for k, (ms, n) in kt.items()},
           "tiling_ch_ach": list(d.tiling()), "device_bytes": d.device_bytes,
           "mass_rel_err": abs(v - v0 - melt - weir) / max(abs(v), 1e-300)}
    d.close()
    log(f"[bench] largest domain: {out['cell_updates_per_s']:.3e} cell-updates/s, {out['ms_per_step']} ms/step "
        f"from t={t_sp:.0f}s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed loop iterations (0 = the whole run, t = 0 .. t_end)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--tiling", type=str, default="",
                    help="ch,ach column-block lengths (default: the autotuner's table)")
    ap.add_argument("--start-t", type=float, default=0.0,
                    help="with --steps K: spin up (untimed) to this simulated time first")
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--fp64-steps", type=int, default=20)
    ap.add_argument("--spinup-t", type=float, default=3600.0)
    ap.add_argument("--spinup-max-steps", type=int, default=20000)
    ap.add_argument("--accuracy-spans", type=str, default="1,10,60",
                    help="simulated seconds after the snapshot at which depths are compared")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--ensemble", type=int, default=64,
                    help="scenarios of the batched ensemble leg (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="arrival window of the served request stream (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate [1/s]")
    ap.add_argument("--serve-rate-sustained", type=float, default=2.0,
                    help="a second, lower arrival rate for the serving leg (0 = skip)")
    ap.add_argument("--c5-steps", type=int, default=100,
                    help="timed iterations of the 16384x16384 four-solute domain (0 = skip)")
    ap.add_argument("--c5-start-t", type=float, default=3600.0,
                    help="untimed spin-up of that domain [simulated s]")
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

    from fluxos_cpp_amd import capi, gen
    capi.load()
    t0 = time.perf_counter()
    sc = gen.scenario_c2(args.n)
    init = capi.prepare(sc)
    log(f"[bench] generated {args.n}^2 basin in {time.perf_counter() - t0:.1f}s")
    cells = sc.nrows * sc.ncols

    # warm-up: W loop iterations of a throwaway domain (code objects, first touch)
    dw = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, init=init)
    dw.advance(float(sc.t_end), 1e30, args.warmup)
    dw.synchronize()
    dw.close()

    # timed: the whole run, t = 0 .. t_end, device-resident between print times
    # (PRINT_STEP as in the reference; the host only reads the clock there)
    d32 = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, init=init)
    if args.tiling:
        d32.set_tiling(*(int(x) for x in args.tiling.split(",")))
    tiling = d32.tiling()              # from the autotuner's table (fluxos_cpp_amd/tuning.json)
    v0 = d32.mass_balance()[0]
    if args.steps > 0:                  # K iterations, from t = 0 or from --start-t
        if args.start_t > 0:
            t_sp, n_sp = spin_up(d32, args.start_t, 10 ** 7)
            log(f"[bench] untimed spin-up to t={t_sp:.1f}s in {n_sp} steps")
        info, el, kt = timed(d32, args.steps, PROFILE_EVERY)
        steps, solver_steps, t_reached = info.steps, info.solver_steps, info.tim
    else:
        d32.synchronize()
        d32.profile(True, PROFILE_EVERY)
        d32.synchronize()
        steps = solver_steps = 0
        print_next = float(sc.print_step)
        t0 = time.perf_counter()
        while True:
            info = d32.advance(float(sc.t_end), print_next)
            steps += info.steps
            solver_steps += info.solver_steps
            if info.status & 2 or info.steps == 0:
                break
            if info.status & 1:
                print_next += sc.print_step
        d32.synchronize()
        el = time.perf_counter() - t0
        t_reached = info.tim
        kt = {}
        for kid, name in capi.KERNEL_NAMES.items():
            ms, n = d32.profile_read(kid)
            if n:
                kt[name] = (ms, n)
        d32.profile(False)
    value = cells * solver_steps / el
    ms_step = el * 1e3 / max(1, steps)
    log(f"[bench] fp32 run to t={t_reached:.1f}s: {steps} steps ({solver_steps} with the solver) in {el:.1f}s")
    log("[bench] fp32 kernels: " + ", ".join(f"{k}={v[0] / max(1, v[1]) * 1e3:.1f}us x{v[1]}" for k, v in kt.items()))

    # water ledger of the run so far (FP64): volume now against the start plus
    # melt input and weir-face exchange
    v, melt, weir = d32.mass_balance()
    mass = {"rel_err": abs(v - v0 - melt - weir) / max(abs(v), 1e-300), "volume_m3": v,
            "melt_m3": melt, "weir_m3": weir,
            # solute kept in the non-weir cells against the FP64 ledger of solute the
            # melt brought into water (the reference's scheme zeroes it in cells that
            # dry and in weir cells, so the two differ by those sinks)
            "solute": [dict(zip(("mass_g", "brought_in_g"), d32.solute_balance(k))) for k in range(sc.nchem)]}

    # roofline of the dominant kernel (algorithmic bytes: hydro 28 B per
    # cell-update, ADE 8 B per solute per cell-update)
    alg = {"k32_step": b_alg("fp32", 0), "k32_ade": 8 * sc.nchem}
    prof_name = {"k32_step": "k32_sweep", "k32_ade": "k32_ade_sweep"}
    kernels = {}
    for name, (ms, n) in kt.items():
        avg_us = ms / n * 1e3
        row = {"avg_us": round(avg_us, 2), "launches_timed": n, "timed_every": PROFILE_EVERY}
        if name in alg:
            gbs = cells * alg[name] / (avg_us * 1e-6) / 1e9
            row.update(bytes_per_cell=alg[name], achieved_gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
        kernels[name] = row
    dom_name = max(alg, key=lambda k: kt[k][0] if k in kt else 0.0)
    dk = kernels[dom_name]
    roofline = {"bound": "hbm", "achieved": dk["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": dk["frac"], "traffic": measured_traffic(args.n, prof_name[dom_name]),
                "kernel": prof_name[dom_name],
                "bytes_per_cell": dk["bytes_per_cell"], "avg_us": dk["avg_us"]}
    step_ms = sum(kernels[k]["avg_us"] for k in alg if k in kernels) * 1e-3   # the two kernels of every step
    step_frac = cells * b_alg("fp32", sc.nchem) / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS

    # a snapshot of the run at t = spinup_t for the comparisons below
    d32.close()
    d32 = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, init=init)
    t0 = time.perf_counter()
    t_sim, nsp = spin_up(d32, args.spinup_t, args.spinup_max_steps)
    log(f"[bench] snapshot: fp32 spin-up to t={t_sim:.1f}s in {nsp} steps, {time.perf_counter() - t0:.1f}s wall")
    snap = snapshot(d32)

    # FP64 parity path and FP32-mixed over the same window from the snapshot,
    # and the depth error at equal time
    fp64 = None
    err = None
    d64 = None
    if not args.no_fp64:
        restore(d32, snap)
        info_w, el_w, _ = timed(d32, args.fp64_steps)
        d64 = capi.HipDomain.from_scenario(sc, precision=capi.FP64_PARITY, ade=True, init=init)
        restore(d64, snap)
        info64, el64, kt64 = timed(d64, args.fp64_steps)
        log("[bench] fp64 kernels: " + ", ".join(f"{k}={v[0] / max(1, v[1]) * 1e3:.1f}us x{v[1]}" for k, v in kt64.items()))
        v64 = cells * info64.solver_steps / el64
        v32w = cells * info_w.solver_steps / el_w
        fp64 = {"value": round(v64, 1), "ms_per_step": round(el64 * 1e3 / max(1, info64.steps), 3),
                "steps": info64.steps, "window_from_t": round(snap["clock"][0], 1),
                "fp32_value_same_window": round(v32w, 1),
                "fp32_over_fp64": round(v32w / v64, 1) if v64 > 0 else None}   # None: window still dry
        d64.close()
        d64 = None
        err = accuracy(sc, init, snap, args.accuracy_spans)
    cpu = None
    if not args.no_cpu:
        restore(d32, snap)
        rate, cel, thr = cpu_baseline(d32, sc, args.cpu_steps, args.cpu_threads or None)
        cpu = {"value": round(rate, 1), "unit": "cell-updates/s", "cores": thr, "kind": "port",
               "sample": f"{args.cpu_steps} loop iterations of the {args.n}x{args.n} basin from the "
                         f"GPU run's state at t={snap['clock'][0]:.0f}s ({cel:.1f}s wall; hydro loops "
                         f"OpenMP, CFL/melt/ADE serial as in the reference)"}

    ensemble = None
    if args.ensemble > 0:
        from fluxos_cpp_amd import serve
        t0 = time.perf_counter()
        scens = [gen.scenario_c3(i) for i in range(args.ensemble)]
        inits = [capi.prepare(s) for s in scens]
        log(f"[bench] ensemble: {len(scens)} scenarios generated in {time.perf_counter() - t0:.1f}s")
        ensemble = serve.run_ensemble(scens, inits)
        ensemble["workload"] = (f"{args.ensemble} scenarios, side log-uniform 256-2048, dx=2 m, fractal DEMs, "
                                "melt U(5,80) mm/day, 1 solute, 1 simulated hour each, one batch")
        log(f"[bench] ensemble: {ensemble['cell_updates_per_s']:.3e} cell-updates/s over {ensemble['wall_s']:.1f}s")

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
            r["offered_cell_updates_per_s"] = rate * mean_cells * r["mean_steps"]
            serving[label] = r
            log(f"[bench] serving at {rate}/s: {r['requests']} requests, p50 {r['p50_s']:.2f}s "
                f"p99 {r['p99_s']:.2f}s, {r['cell_updates_per_s']:.3e} cell-updates/s")
        serving["workload"] = ("Poisson arrivals; side {512,1024,2048} p={0.6,0.3,0.1}; fractal DEMs; melt "
                               "U(5,80) mm/day; 1 solute; 30 simulated minutes each; HBM budget 200 GB; "
                               "continuous batching; time-to-result = arrival to final depth on the host")

    c5 = None
    if args.c5_steps > 0:
        c5 = largest_domain(args.c5_steps, args.c5_start_t)

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (f64 depth, dt, mass)", "data": "synthetic",
        "config": {"workload": f"single domain {args.n}x{args.n}, dx=2 m, fractal DEM + 0.3% slope, melt 0-60 "
                               f"mm/day over 6 h, 2 solutes, FP32-mixed with fast ADE, "
                               + ("the whole 6 simulated hours" if args.steps == 0 else
                                  f"{args.steps} iterations from t={args.start_t:.0f}s"),
                   "cells": cells, "nchem": sc.nchem, "t_end_s": round(t_reached, 1),
                   "loop_iterations": steps, "solver_steps": solver_steps, "tiling_ch_ach": list(tiling)},
        "roofline": roofline,
        "kernels": kernels,
        "step_roofline_frac": round(step_frac, 4),
        "mass_balance": mass,
        "fp64_parity": fp64, "fp32_vs_fp64": err,
        "cpu_baseline": cpu,
        "ensemble": ensemble,
        "serving": serving,
        "largest_domain": c5,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
