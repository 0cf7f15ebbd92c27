"""Ensembles and a request-serving loop over batches of FP32-mixed domains.

Both sit on the C ABI's batches (``fx_batch_*``): every batched step advances
each member by one iteration of the reference's loop (main.cpp:271-406) with
its own CFL step, in one launch per kernel over all members.

* ``run_ensemble`` -- a fixed set of scenarios admitted together and run to
  their end time (the 64-scenario ensemble).
* ``serve`` -- continuous batching of a request stream: requests are admitted
  as they arrive while their device memory fits the HBM budget, their state
  stays resident while they run, members that finish leave at once (their
  slot and memory are reused by the next arrivals), and every request's
  time-to-result (arrival -> final state copied back) is recorded.

A member halts at each print time (PRINT_STEP, as the reference's loop clips
dt there, main.cpp:321) and is resumed right away; the final print at the
scenario's end time is its result.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import capi, gen


@dataclass
class Request:
    rid: int
    arrival: float                       # seconds after the stream starts
    scenario: gen.Scenario
    init: capi.InitialState
    admitted: float = math.nan
    finished: float = math.nan
    steps: int = 0
    solver_steps: int = 0
    result: np.ndarray | None = None     # depth at the end time (padded grid)
    error: str = ""                      # why the request failed ("" = it completed)
    queued: bool = False                 # waited for memory or a slot after it arrived

    @property
    def cells(self) -> int:
        return self.scenario.nrows * self.scenario.ncols


def domain_bytes(sc: gen.Scenario, ade: bool = True) -> int:
    """Device bytes of an FP32-mixed domain for this scenario."""
    lib = capi.load()
    desc = capi.DomainDesc()
    desc.nrows, desc.ncols, desc.dxy, desc.nchem = sc.nrows, sc.ncols, sc.dxy, sc.nchem
    desc.hdry, desc.nodata = sc.hdry, gen.NODATA
    desc.gacc, desc.cfl, desc.cvdef, desc.nuem = 9.80665, 1.0, 0.07, 1.793e-6
    desc.ade_solver = 1 if ade else 0
    desc.precision = capi.FP32_MIXED
    desc.ade_mode = capi.ADE_FAST if ade else capi.ADE_OFF
    b = lib.fx_domain_bytes(ctypes.byref(desc))
    if b < 0:
        raise capi.FxError(int(b), lib.fx_last_error().decode())
    return int(b)


def _drive(batch: capi.Batch, t_end: dict, next_print: dict, print_step: float, on_done,
           steps_per_poll: int, on_error=None) -> None:
    """One run/poll round: resume members at intermediate prints, hand the
    finished ones to on_done(slot, tim, steps, solver_steps) and members whose
    state holds NaN/Inf (halted alone by the device) to on_error(slot, tim);
    without on_error such a member raises."""
    batch.run(steps_per_poll)
    for slot, kind, tim, it, steps, solver_steps in batch.poll():
        if kind == capi.EV_ERROR:
            if on_error is None:
                raise RuntimeError(f"member {slot}: NaN/Inf in the state")
            on_error(slot, tim)
            continue
        if kind == capi.EV_DONE or tim >= t_end[slot]:
            on_done(slot, tim, steps, solver_steps)
        else:
            next_print[slot] = next_print[slot] + print_step
            batch.resume(slot, next_print[slot])


def run_ensemble(scens: list[gen.Scenario], inits: list[capi.InitialState], ade: bool = True,
                 steps_per_poll: int = 16, tiling: tuple[int, int] | None = None) -> dict:
    """Admit every scenario into one batch and run each to its print at t_end
    (its result); a member leaves the batch as it finishes.  The timed region starts with
    every input resident on the device.  The members' tile shape is `tiling`,
    or the autotuner's batch entry."""
    from . import autotune
    nchem = max(sc.nchem for sc in scens)
    batch = capi.Batch(nchem=nchem, ade=ade, max_members=len(scens))
    doms, t_end, next_print = {}, {}, {}
    if tiling is None:
        tiling = autotune.choose_batch([sc.nrows * sc.ncols for sc in scens])
    for sc, st in zip(scens, inits):
        d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=ade, init=st, tiling=tiling)
        s = batch.admit(d, float(sc.t_end), float(sc.print_step))
        doms[s], t_end[s], next_print[s] = d, float(sc.t_end), float(sc.print_step)
    done: dict[int, int] = {}
    solver_done: dict[int, int] = {}
    failed: dict[int, float] = {}
    cells_of = {s: d.nrows * d.ncols for s, d in doms.items()}
    total_cells = sum(cells_of.values())
    timeline = []                        # (wall s, members still running, their cells)

    def finished(s, tim, steps, solver):
        # a finished member leaves the batch at once, so that the steps of the
        # others no longer launch its (idle) blocks
        done[s], solver_done[s] = steps, solver
        batch.evict(s)

    def errored(s, tim):
        failed[s] = tim
        batch.evict(s)

    try:
        batch.poll()
        t0 = time.perf_counter()
        while len(done) + len(failed) < len(doms):
            _drive(batch, t_end, next_print, float(scens[0].print_step), finished, steps_per_poll, errored)
            running = [s for s in doms if s not in done and s not in failed]
            timeline.append((time.perf_counter() - t0, len(running), sum(cells_of[s] for s in running)))
        el = time.perf_counter() - t0
        n, cells = len(doms), total_cells
        upd = sum(cells_of[s] * solver_done[s] for s in solver_done)
    finally:
        batch.close()
        for d in doms.values():
            d.close()
    out = {"scenarios": n, "cells": cells, "cell_updates": upd, "wall_s": el, "tiling": tiling,
           "cell_updates_per_s": upd / el, "max_steps": max(done.values(), default=0),
           "min_steps": min(done.values(), default=0), "failed_slots": sorted(failed),
           "occupancy": occupancy_summary(timeline, len(doms), total_cells)}
    return out


def occupancy_summary(timeline, members: int, total_cells: int) -> dict:
    """How full a batch stayed over its run: the time-weighted mean fraction
    of the members' cells still stepping, and the wall times at which the
    running members fell to 50 % and 10 % of the batch."""
    if not timeline:
        return {}
    t_prev, acc = 0.0, 0.0
    t50 = t10 = None
    for t, n, c in timeline:
        acc += (t - t_prev) * c
        t_prev = t
        if t50 is None and n <= members * 0.5:
            t50 = t
        if t10 is None and n <= members * 0.1:
            t10 = t
    wall = timeline[-1][0]
    return {"mean_active_cell_fraction": acc / (wall * total_cells) if wall > 0 else None,
            "wall_s_at_half_members_done": t50, "wall_s_at_90pct_members_done": t10,
            "samples": [[round(t, 3), n, c] for t, n, c in timeline[:: max(1, len(timeline) // 32)]]}


def poisson_arrivals(rate: float, seconds: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            return np.array(out)
        out.append(t)


def c4_stream(rate: float = 20.0, seconds: float = 300.0, seed: int = 4, pool: int = 4,
              t_end: int = 1800) -> list[Request]:
    """The served request stream: Poisson arrivals; per request a grid side
    from {512, 1024, 2048} with p = {0.6, 0.3, 0.1}, a fractal DEM (one of
    `pool` seeded DEMs per side, so inputs are prepared once) and its own
    melt rate drawn from U(5, 80) mm/day; t_end simulated seconds each."""
    rng = np.random.default_rng(seed + 1)
    base = {}
    for side in (512, 1024, 2048):
        for k in range(pool):
            s = 4000 + 10 * k + side
            sc = gen.Scenario(f"c4_{side}_{k}", gen.dem_fractal(side, side, 2.0, s), 2,
                              gen.meteo_table(np.arange(0, t_end + 1, 600), 20.0, [1.0]), 1, [0.01],
                              t_end, s)
            base[(side, k)] = (sc, capi.prepare(sc))
    reqs = []
    for i, t in enumerate(poisson_arrivals(rate, seconds, seed)):
        side = int(rng.choice([512, 1024, 2048], p=[0.6, 0.3, 0.1]))
        k = int(rng.integers(pool))
        rate_i = float(rng.uniform(5.0, 80.0))
        sc0, st = base[(side, k)]
        sc = dataclasses.replace(sc0, name=f"c4_req{i}", seed=i,
                                 meteo=gen.meteo_table(np.arange(0, t_end + 1, 600), rate_i, [1.0]))
        reqs.append(Request(i, float(t), sc, st))
    return reqs


def serve(requests: list[Request], budget_bytes: int = 200 * 10 ** 9, max_members: int = 1024,
          steps_per_poll: int = 8, keep_results: bool = False, prepare_hook=None) -> dict:
    """Continuous batching of a request stream (arrival times relative to the
    start).  Returns latency percentiles and throughput.

    Admission is in arrival order while the device bytes of the live members
    plus the next request fit ``budget_bytes``; a request that does not fit
    waits (queues) until finished members free their memory.  A request that
    alone needs more than the budget can never run: it is rejected at once.
    A member whose state turns non-finite is halted by the device on its own
    (FX_EV_ERROR), evicted and reported failed; the others keep running.
    ``prepare_hook(req, dom)``, if given, is called on each request's domain
    after its set-up and before admission."""
    from . import autotune
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    batch = capi.Batch(nchem=1, ade=True, max_members=max_members)
    live: dict[int, tuple[Request, capi.HipDomain, int]] = {}
    t_end: dict[int, float] = {}
    next_print: dict[int, float] = {}
    used = 0
    peak_used = 0
    peak_live = 0
    queued = 0                           # admissions that waited for memory or a slot
    finished: list[Request] = []
    failed: list[Request] = []
    t0 = time.perf_counter()

    def now() -> float:
        return time.perf_counter() - t0

    def release(slot):
        nonlocal used
        req, dom, nb = live.pop(slot)
        batch.evict(slot)
        used -= nb
        return req, dom

    def on_done(slot, tim, steps, solver_steps):
        req, dom = release(slot)
        h = dom.get("h")                              # the result, copied back to the host
        if keep_results:
            req.result = h
        dom.close()
        req.steps = steps
        req.solver_steps = solver_steps
        req.finished = now()
        finished.append(req)

    def on_error(slot, tim):
        req, dom = release(slot)
        dom.close()
        req.error = f"NaN/Inf in the state at t={tim:.3f}s"
        req.finished = now()
        failed.append(req)

    try:
        while pending or live:
            # admission: arrived requests, in order, while memory and slots allow
            while pending and pending[0].arrival <= now():
                req = pending[0]
                nb = domain_bytes(req.scenario)
                if nb > budget_bytes:                 # can never fit: reject, do not wait
                    pending.popleft()
                    req.error = f"needs {nb} device bytes, above the {budget_bytes}-byte budget"
                    req.finished = now()
                    failed.append(req)
                    continue
                if used + nb > budget_bytes or len(live) >= max_members:
                    if not req.queued:
                        req.queued = True
                        queued += 1
                    break
                pending.popleft()
                # tile shape for how full the device will be with this member
                sizes = [r.cells for r, _, _ in live.values()] + [req.cells]
                dom = capi.HipDomain.from_scenario(req.scenario, precision=capi.FP32_MIXED, ade=True,
                                                   init=req.init, tiling=autotune.choose_batch(sizes))
                try:
                    if prepare_hook is not None:
                        prepare_hook(req, dom)
                    slot = batch.admit(dom, float(req.scenario.t_end), float(req.scenario.print_step))
                except Exception:
                    dom.close()
                    raise
                live[slot] = (req, dom, nb)
                t_end[slot] = float(req.scenario.t_end)
                next_print[slot] = float(req.scenario.print_step)
                used += nb
                req.admitted = now()
            peak_used = max(peak_used, used)
            peak_live = max(peak_live, len(live))
            if not live:
                if pending:
                    time.sleep(max(0.0, pending[0].arrival - now()))
                continue
            _drive(batch, t_end, next_print, float(requests[0].scenario.print_step), on_done, steps_per_poll,
                   on_error)
    finally:
        wall = now()
        doms = [dom for _, dom, _ in live.values()]
        batch.close()                                 # members still in it leave with their state
        for dom in doms:
            dom.close()
    out = {"requests": len(finished), "failed": len(failed), "queued_for_memory": queued, "wall_s": wall,
           "mean_admission_wait_s": None, "cell_updates": 0, "cell_updates_per_s": 0.0,
           "peak_members": peak_live, "peak_device_bytes": int(peak_used),
           "failed_ids": [r.rid for r in failed], "errors": sorted({r.error for r in failed})}
    if not finished:
        return out
    lat = np.array([r.finished - r.arrival for r in finished])
    wait = np.array([r.admitted - r.arrival for r in finished])
    upd = sum(r.cells * r.solver_steps for r in finished)
    by_side = {}
    for side in sorted({r.scenario.nrows for r in finished}):
        ls = np.array([r.finished - r.arrival for r in finished if r.scenario.nrows == side])
        by_side[str(side)] = {"n": int(ls.size), "p50_s": float(np.percentile(ls, 50)),
                              "p99_s": float(np.percentile(ls, 99))}
    out.update({"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)),
                "mean_admission_wait_s": float(wait.mean()), "max_admission_wait_s": float(wait.max()),
                "cell_updates": int(upd), "cell_updates_per_s": upd / wall, "by_side": by_side,
                "mean_steps": float(np.mean([r.steps for r in finished]))})
    return out


def main(argv=None) -> int:
    """``python -m fluxos_cpp_amd.serve --ensemble N`` runs the first N C3
    scenarios in one batch; ``--stream RATE SECONDS`` serves a C4 Poisson
    stream.  The result goes to stdout as one JSON line."""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m fluxos_cpp_amd.serve")
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--ensemble", type=int, metavar="N", help="the C3 ensemble of N scenarios")
    g.add_argument("--stream", type=float, nargs=2, metavar=("RATE", "SECONDS"), help="a C4 arrival stream")
    args = ap.parse_args(argv)
    capi.load()
    if args.ensemble:
        scens = [gen.scenario_c3(i) for i in range(args.ensemble)]
        inits = [capi.prepare(sc) for sc in scens]
        out = run_ensemble(scens, inits)
    else:
        out = serve(c4_stream(rate=args.stream[0], seconds=args.stream[1]))
    print(json.dumps(out, default=float), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
