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

    python -m fluxos_cpp_amd.autotune [--sizes 512,1024,2048,4096,8192,16384] [--steps 40]
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
CANDIDATES_CH = (8, 12, 16, 20, 24, 28, 32, 40, 48, 64)
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


def choose(nrows: int, ncols: int, path: Path = TABLE) -> tuple[int, int] | None:
    """(ch, ach) for a single domain of nrows x ncols."""
    return choose_cells(nrows * ncols, path)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="512,1024,2048,4096,8192,16384")
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--out", default=str(TABLE))
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
    Path(args.out).write_text(json.dumps({"entries": entries}, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
