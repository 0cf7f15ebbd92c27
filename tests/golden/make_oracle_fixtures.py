"""Write the oracle's trajectories of the benchmark configurations as fixtures.

The oracle (oracle/fluxos_oracle.c, the reference's FP64 OpenMP path restated)
takes minutes on the CPU for the larger cases, so its results are recorded
here once, per print time, and the GPU tests compare the device's FP64 parity
path against them bit for bit.  For each print time a fixture holds the clock
(tim, solver-step counter, loop iterations) and a SHA-256 of each state grid
in the reference layout (padded, column-major; -0.0 written as +0.0), plus a
few sums for a readable failure message.

Cases (each a generator of fluxos_cpp_amd/gen.py at the size and duration
stated, FP64 with the exact ADE):
  c1_256        the plain-method configuration at its full size and length:
                256x256, 2 simulated hours, one solute
  c3_04         ensemble member 4 (256x256, 14.8 mm/day), 1 simulated hour
  c5_192        the largest-domain generator (fractal + carved channels,
                smooth melt multiplier, four solutes with distinct D) at
                192x192, dx = 1 m, 40 simulated minutes

Usage: python tests/golden/make_oracle_fixtures.py [case ...]
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

HERE = Path(__file__).resolve().parent
FIELDS = ("z", "h", "qx", "qy", "ldry")


def cases():
    from fluxos_cpp_amd import gen
    return {
        "c1_256": lambda: gen.scenario_c1(256),
        "c3_04": lambda: gen.scenario_c3(4),
        "c5_192": lambda: gen.scenario_c5(192, t_end=2400),
    }


def digest(a: np.ndarray) -> str:
    """SHA-256 of a padded grid's column-major bytes, zeros made positive."""
    a = np.asfortranarray(a)
    if a.dtype.kind == "f":
        a = a + a.dtype.type(0)          # -0.0 -> +0.0
    return hashlib.sha256(a.tobytes(order="F")).hexdigest()


def record(grids: dict, tim: float, it: int, steps: int) -> dict:
    out = {"tim": tim, "it": it, "steps": steps,
           "sha256": {k: digest(v) for k, v in grids.items()},
           "sum_h": float(np.sum(grids["h"][1:-1, 1:-1])),
           "wet": int((grids["ldry"][1:-1, 1:-1] == 0).sum())}
    return out


def oracle_grids(o) -> dict:
    g = {k: o.g[k] for k in FIELDS}
    for k in range(o.nchem):
        g[f"conc{k}"] = o.conc[k]
    return g


def run_case(name: str, make) -> dict:
    from oracle.oracle import OracleDomain
    sc = make()
    o = OracleDomain.from_scenario(sc)
    prints = []
    print_next = sc.print_step
    steps = 0
    t0 = time.perf_counter()
    while o.tim <= sc.t_end:                      # main.cpp:271
        o.step(print_next)
        steps += 1
        if o.tim >= print_next:                   # main.cpp:444
            prints.append(record(oracle_grids(o), o.tim, o.it.value, steps))
            print(f"[fixtures] {name}: print {print_next} at t={o.tim:.3f} it={o.it.value} steps={steps} "
                  f"({time.perf_counter() - t0:.0f}s)", flush=True)
            print_next += sc.print_step
    return {"case": name, "nrows": sc.nrows, "ncols": sc.ncols, "nchem": sc.nchem, "t_end": sc.t_end,
            "print_step": sc.print_step, "prints": prints}


def main(argv):
    todo = cases()
    names = argv or list(todo)
    for n in names:
        out = run_case(n, todo[n])
        (HERE / f"oracle_{n}.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main(sys.argv[1:])
