"""A basin driven by the reference's own melt record.

The forcing is the reference's Working_example file (tests/golden/
Qmelt_info_2009): real diurnal melt rates at irregular 1799/1800/1801 s
spacing, read through fx_read_forcing.  The file has no concentration
column, so the test supplies 1 mg/L.  add_meteo picks the first row whose
time is later than the clock (common.cpp:195-203); the other scenarios use
evenly spaced rows.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen
from oracle.oracle import run_reference_loop
from tests.helpers import matched, max_rel

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
T_END, PRINT = 86400, 3600


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scenario():
    rows = capi.read_forcing_file(GOLDEN / "Qmelt_info_2009")
    tab = np.zeros((rows.shape[0] + 1, 3), dtype=np.float32)       # row 0: the empty header slot
    tab[1:, :2] = rows
    tab[1:, 2] = 1.0
    return gen.Scenario("working_example_melt", gen.dem_c1(48, 5), 2, tab, 1, [0.01], T_END, 5,
                        print_step=PRINT)


def test_fp64_real_melt_record_bitwise():
    """One simulated day (about 40,700 loop iterations) on the FP64 path
    against the oracle: bit-identical state and equal wet counts at each of
    the 24 prints."""
    sc = _scenario()
    o, d = matched(sc)
    snaps = {}

    def grab(t, od):
        snaps[t] = {k: od.g[k].copy() for k in ("z", "h", "qx", "qy", "ldry")}
        snaps[t]["conc0"] = od.conc[0].copy()
        snaps[t]["wet"] = od.wet_count()

    dts, solver = run_reference_loop(o, sc.t_end, sc.print_step, on_print=grab)
    assert int(solver.sum()) > 10000 and snaps[T_END]["wet"] > 0
    seen = []

    def check(t, dd, info):
        ref = snaps[t]
        got = {k: dd.get(k) for k in ("z", "h", "qx", "qy", "ldry")}
        got["conc0"] = dd.get_conc(0)
        wet = int((got["ldry"][1:-1, 1:-1] == 0).sum())
        r = {k: max_rel(ref[k], got[k]) for k in got}
        _say("print", t, "it", info.it, "wet", wet, ref["wet"], r)
        assert wet == ref["wet"]
        assert all(v == 0.0 for v in r.values()), (t, r)
        seen.append(t)

    total, solver_steps = capi.run_loop(d, sc.t_end, sc.print_step, on_print=check)
    assert total == len(dts) and solver_steps == int(solver.sum())
    assert seen == sorted(snaps) and len(seen) == T_END // PRINT
    d.close()


def _day(sc, init, precision, nudge=None):
    """Depth after the day, from the shared initial state; ``nudge`` raises
    one cell's z by one ulp first."""
    d = capi.HipDomain.from_scenario(sc, precision=precision, init=init)
    if nudge is not None:
        z = d.get("z")
        z[nudge] = np.nextafter(z[nudge], np.inf)
        d.set("z", z)
    v0 = d.mass_balance()[0]
    capi.run_loop(d, sc.t_end, sc.print_step)
    h = d.get("h")[1:-1, 1:-1].copy()
    v, melt, weir = d.mass_balance()
    d.close()
    return h, (v0, v, melt, weir)


def test_fp32_real_melt_record_ledger_and_envelope():
    """FP32-mixed over the same day: its FP64 water ledger closes to 1e-9
    relative, and its depth differs from FP64 by no more than 3x what a
    one-ulp nudge to a single cell's z does to FP64 itself over the day
    (about 1.3 cm of a 3.2 cm maximum depth: once water moves, the flow
    amplifies any rounding difference)."""
    sc = _scenario()
    init = capi.prepare(sc)
    h64, _ = _day(sc, init, capi.FP64_PARITY)
    env = max(float(np.abs(_day(sc, init, capi.FP64_PARITY, cell)[0] - h64).max())
              for cell in ((24, 24), (10, 30)))
    h32, (v0, v, melt, weir) = _day(sc, init, capi.FP32_MIXED)
    err = float(np.abs(h64 - h32).max())
    rel = abs(v - v0 - melt - weir) / max(abs(v), 1e-300)
    _say(f"max|dh| {err:.3e} m of hmax {h64.max():.3e}; one-ulp FP64 envelope {env:.3e} m; "
         f"V0={v0:.9e} V={v:.9e} melt={melt:.9e} weir={weir:.9e} rel err={rel:.3e}")
    assert melt > 0 and rel <= 1e-9
    assert env > 0 and err <= 3 * env
