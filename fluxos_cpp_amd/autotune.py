"""Tile-shape autotuner of the FP32-mixed sweeps.

The step sweep and the ADE sweep walk their column blocks with one wave per
strip of rows; how many columns a block walks (``ch`` for the step sweep,
``ach`` for the ADE sweep) trades the halo columns every block reloads
against the number of waves in flight.  Results do not depend on it (every
cell is computed by one lane from the same inputs either way).

``tune(n)`` times every candidate on a synthetic basin of side n (the
fractal-DEM generator, two solutes) spun up on the device to t = 3600 s (the
melt has wetted the basin) and then run on, candidate after candidate, and
keeps the fastest for each kernel; ``main`` tunes a ladder of sizes and writes
``tuning.json`` beside this file, which ``HipDomain.from_scenario`` consults
(nearest size by cell count).  The time per block count is not smooth in
``ch``: blocks run in rounds of the device's wave slots, and a last round
that is nearly empty costs as much as a full one.

    python -m fluxos_cpp_amd.autotune [--sizes 512,1024,2048,4096,8192,16384] [--steps 40] [--merge]

``--merge`` replaces only the tuned sizes' entries of an existing table.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

PKG = Path(__file__).resolve().parent
TABLE = PKG / "tuning.json"
CANDIDATES_CH = (8, 12, 16, 20, 22, 23, 24, 26, 28, 32, 40, 48, 64)
CANDIDATES_ACH = (16, 32, 48, 64, 96, 128)
FX32_ACH_DEFAULT = 64


def _kernel_us(dom, steps: int) -> dict[str, float]:
    from . import capi
    dom.synchronize()
    dom.profile(True)
    dom.advance(1e30, 1e30, steps)
    out = {}
    for kid, name in capi.KERNEL_NAMES.items():
        ms, n = dom.profile_read(kid)
        if n:
            out[name] = ms / n * 1e3
    dom.profile(False)
    return out


def tune(n: int, steps: int = 40, t_wet: float = 3600.0) -> dict:
    from . import capi, gen
    sc = gen.scenario_c2(n)
    init = capi.prepare(sc)
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, init=init, tiling=None)
    pn = 600.0
    while True:                                   # print-clipped spin-up to the wet state
        info = d.advance(1e30, min(pn, t_wet))
        if info.tim >= t_wet or info.steps == 0:
            break
        if info.tim >= pn:
            pn += 600.0
    # candidates are timed back to back on the continuing run (the state
    # evolves by a few simulated seconds per candidate), twice, keeping the
    # faster of the two passes
    sweep = {ch: float("inf") for ch in CANDIDATES_CH}
    ade = {ach: float("inf") for ach in CANDIDATES_ACH}
    for _ in range(2):
        for ch in CANDIDATES_CH:
            d.set_tiling(ch, FX32_ACH_DEFAULT)
            d.advance(1e30, 1e30, 2)                  # warm-up of this shape
            sweep[ch] = min(sweep[ch], _kernel_us(d, steps)["k32_step"])
    best_ch = min(sweep, key=sweep.get)
    for _ in range(2):
        for ach in CANDIDATES_ACH:
            d.set_tiling(best_ch, ach)
            d.advance(1e30, 1e30, 2)
            ade[ach] = min(ade[ach], _kernel_us(d, steps)["k32_ade"])
    d.close()
    best_ach = min(ade, key=ade.get)
    return {"n": n, "cells": n * n, "ch": best_ch, "ach": best_ach,
            "k32_sweep_us": {str(k): round(v, 2) for k, v in sweep.items()},
            "k32_ade_us": {str(k): round(v, 2) for k, v in ade.items()},
            "workload": f"fractal {n}x{n}, 2 solutes, run on from t={t_wet:.0f}s, {steps} steps per shape, best of 2"}


BATCH_CANDIDATES = ((8, 32), (8, 64), (12, 64), (16, 64), (23, 64))


def tune_batch(count: int = 64, candidates=BATCH_CANDIDATES) -> dict:
    """Tile shape for batches: the ensemble of `count` C3 scenarios (sides
    256-2048) run whole under each candidate; the fastest by cell-updates/s.
    Small members weigh more here than in any single-domain entry."""
    from . import capi, gen, serve
    scens = [gen.scenario_c3(i) for i in range(count)]
    inits = [capi.prepare(s) for s in scens]
    rates = {}
    for t in candidates:
        rates[f"{t[0]},{t[1]}"] = serve.run_ensemble(scens, inits, tiling=t)["cell_updates_per_s"]
    best = max(rates, key=rates.get)
    ch, ach = (int(x) for x in best.split(","))
    return {"ch": ch, "ach": ach, "cell_updates_per_s": {k: round(v, 1) for k, v in rates.items()},
            "workload": f"{count} C3 scenarios in one batch, 1 simulated hour each"}


def load_table(path: Path = TABLE) -> list[dict]:
    if not path.exists():
        return []
    try:
        return json.loads(path.read_text()).get("entries", [])
    except (ValueError, OSError):
        return []


def choose_cells(cells: int, path: Path = TABLE) -> tuple[int, int] | None:
    """(ch, ach) of the tuned size nearest to `cells` (by log distance), or
    None without a table.  For a batch, pass the cells of the whole batch:
    what matters is how full the device is."""
    entries = load_table(path)
    if not entries:
        return None
    cells = max(1, int(cells))
    e = min(entries, key=lambda e: abs(math.log(e["cells"] / cells)))
    return int(e["ch"]), int(e["ach"])


def choose_batch(member_cells: list[int], path: Path = TABLE) -> tuple[int, int] | None:
    """(ch, ach) for the members of a batch: the table's "batch" entry when it
    has one (tuned on an ensemble), else the entry nearest the batch's total
    cell count (how full the device is)."""
    try:
        b = json.loads(Path(path).read_text()).get("batch") if Path(path).exists() else None
    except (ValueError, OSError):
        b = None
    if b:
        return int(b["ch"]), int(b["ach"])
    return choose_cells(sum(member_cells), path)


def choose(nrows: int, ncols: int, path: Path = TABLE) -> tuple[int, int] | None:
    """(ch, ach) for a single domain of nrows x ncols."""
    return choose_cells(nrows * ncols, path)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="512,1024,2048,4096,8192,16384")
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--batch", action="store_true", help="also tune the batch entry (about a minute per candidate)")
    ap.add_argument("--merge", action="store_true", help="keep the table's entries of the sizes not tuned now")
    args = ap.parse_args(argv)
    from . import capi
    capi.load()
    entries = []
    for n in (int(x) for x in args.sizes.split(",")):
        t0 = time.perf_counter()
        e = tune(n, args.steps)
        entries.append(e)
        print(f"[autotune] n={n}: ch={e['ch']} ach={e['ach']} sweep {e['k32_sweep_us']} ade {e['k32_ade_us']} "
              f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    if args.merge:
        tuned = {e["n"] for e in entries}
        entries = sorted([e for e in load_table(Path(args.out)) if e.get("n") not in tuned] + entries,
                         key=lambda e: e["cells"])
    table = {"entries": entries}
    if args.batch:
        table["batch"] = tune_batch()
        print(f"[autotune] batch: {table['batch']}", file=sys.stderr, flush=True)
    else:
        try:                                 # keep the batch entry already tuned
            old = json.loads(Path(args.out).read_text()).get("batch")
        except (ValueError, OSError):
            old = None
        if old:
            table["batch"] = old
    Path(args.out).write_text(json.dumps(table, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
