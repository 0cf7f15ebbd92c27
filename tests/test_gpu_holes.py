"""Interior NODATA holes and negative melt rates, on both paths.

read_geo turns a NODATA cell inside the domain into a weir cell whose bed is
the mean of its valid neighbours (read_functions.cpp:145-225), and add_meteo
takes the forcing rate as it is, so a negative rate (evaporation) lowers the
surface and can leave deficits (common.cpp:164-273).  The scenarios elsewhere
have NODATA only on the outer ring and non-negative rates.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen
from oracle.oracle import run_reference_loop
from tests.helpers import max_rel, matched

pytestmark = pytest.mark.gpu

HOLES = [(20, 20), (21, 20), (20, 21), (21, 21), (40, 12), (10, 45), (33, 33)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scenario(n=64, seed=3, t_end=1800):
    dem = gen.dem_c1(n, seed)
    for r, c in HOLES:
        dem[r, c] = -99999.0
    times = np.arange(0, t_end + 1, 300)
    rates = np.where(np.arange(len(times)) % 3 == 2, -25.0, 40.0)    # every third row evaporates
    return gen.Scenario("holes", dem, 2, gen.meteo_table(times, rates, [1.0]), 1, [0.01], t_end, seed,
                        print_step=300)


def test_fp64_holes_and_evaporation_bitwise():
    """Device-resident loop against the oracle's reference loop for 30
    simulated minutes: bit-identical state and identical wet counts at every
    print time."""
    sc = _scenario()
    assert (sc.meteo[1:, 1] < 0).any()
    o, d = matched(sc)
    assert int((o.weir[2:-2, 2:-2] == 1).sum()) == len(HOLES)
    snaps = {}

    def grab(t, od):
        snaps[t] = {k: od.g[k].copy() for k in ("z", "h", "qx", "qy", "ldry")}
        snaps[t]["conc0"] = od.conc[0].copy()
        snaps[t]["wet"] = od.wet_count()

    dts, solver = run_reference_loop(o, sc.t_end, sc.print_step, on_print=grab)
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
    assert seen == sorted(snaps) and len(seen) == 6
    d.close()


def test_fp32_holes_and_evaporation():
    """FP32-mixed from the same start: depth within 1e-3 m of FP64 shortly
    after the basin wets, the same wet flags, and its FP64 water ledger
    (negative melt included) closing to 1e-9 over the 30 minutes."""
    sc = _scenario()
    init = capi.prepare(sc)
    d64 = capi.HipDomain.from_scenario(sc, precision=capi.FP64_PARITY, init=init)
    d32 = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, init=init)
    v0 = d32.mass_balance()[0]
    for d in (d64, d32):
        t, pn = 0.0, 300.0
        while t < 700.0:
            info = d.advance(1e30, min(pn, 700.0))
            t = info.tim
            if t >= pn:
                pn += 300.0
    h64, h32 = d64.get("h")[1:-1, 1:-1], d32.get("h")[1:-1, 1:-1]
    err = float(np.abs(h64 - h32).max())
    w64 = int((d64.get("ldry")[1:-1, 1:-1] == 0).sum())
    w32 = int((d32.get("ldry")[1:-1, 1:-1] == 0).sum())
    _say(f"t={d64.clock()[0]:.3f}/{d32.clock()[0]:.3f} max|dh| {err:.3e} m, wet {w64}/{w32}")
    assert err <= 1e-3
    # every cell sits just below hdry and the solver gate is closed: both
    # paths report the wet flag the last solver step left (all wet)
    assert w64 == w32 == 64 * 64
    capi.run_loop(d32, sc.t_end, sc.print_step, timstart=700)
    v, melt, weir = d32.mass_balance()
    rel = abs(v - v0 - melt - weir) / max(abs(v), 1e-300)
    _say(f"V0={v0:.9e} V={v:.9e} melt={melt:.9e} weir={weir:.9e} rel err={rel:.3e}")
    assert rel <= 1e-9
    d64.close()
    d32.close()
