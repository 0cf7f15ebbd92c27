"""WINTRA soil release (WINTRAsolver_calc.cpp:26-57) on the device, through
the C ABI: the FP64 parity path bit for bit against the oracle (the
granular call and the device-resident loop), and the FP32-mixed path within
the FP32 bar against the FP64 one."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen
from oracle.oracle import run_reference_loop
from tests.helpers import compare, matched, max_rel, push_oracle_state, random_wet_state

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _soil(shape, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros(shape, order="F")
    s[1:-1, 1:-1] = rng.uniform(0.0, 50.0, (shape[0] - 2, shape[1] - 2))
    return s


def test_wintra_granular_random_states():
    sc = gen.scenario_c1(48, seed=3)
    for seed in range(3):
        _say("wintra random state seed", seed)
        o = random_wet_state(sc, seed)
        d = capi.HipDomain.from_scenario(sc)
        push_oracle_state(o, d)
        soil = _soil(o.zb.shape, seed + 10)
        o.soil[...] = soil
        d.set_wintra(True, 0.6)
        d.set("soil", soil)
        o.set_wintra(True, 0.6)
        o.dtfl = 0.37 + seed
        d.set_clock(o.tim, o.dtfl, o.it.value)
        o.wintra(0)
        d.wintrasolver_calc(0)
        r = {"soil": max_rel(o.soil, d.get("soil")), "conc0": max_rel(o.conc[0], d.get_conc(0))}
        assert r == {"soil": 0.0, "conc0": 0.0}, (seed, r)
        assert (d.get("soil") != soil).any()
        d.close()


def test_wintra_device_loop_matches_oracle():
    """fx_advance with WINTRA on against the oracle's loop: bit-identical
    state, solute and soil mass at every print time."""
    sc = gen.scenario_c1(64, seed=2, t_end=1800)
    o, d = matched(sc)
    soil = _soil(o.zb.shape, 5)
    o.soil[...] = soil
    o.set_wintra(True, 1.2)
    d.set_wintra(True, 1.2)
    d.set("soil", soil)
    snaps = {}

    def grab(t, od):
        snaps[t] = {k: od.g[k].copy() for k in ("z", "h", "qx", "qy", "soil")}
        snaps[t]["conc0"] = od.conc[0].copy()

    run_reference_loop(o, sc.t_end, sc.print_step, on_print=grab)
    seen = []

    def check(t, dd, info):
        got = {k: dd.get(k) for k in ("z", "h", "qx", "qy", "soil")}
        got["conc0"] = dd.get_conc(0)
        r = {k: max_rel(snaps[t][k], got[k]) for k in got}
        _say("print", t, "it", info.it, r)
        assert all(v == 0.0 for v in r.values()), (t, r)
        seen.append(t)

    capi.run_loop(d, sc.t_end, sc.print_step, on_print=check)
    assert seen == sorted(snaps) and len(seen) == 3
    assert (snaps[1800]["soil"] < soil).any()


def _still_basin(precision, n=96, H=0.05, soil0=20.0, rate=0.9):
    """A flat basin (NODATA ring) under H of water everywhere, uniform soil
    mass, no melt: every interior cell is wet at every step, so the release
    depends only on the dt sequence."""
    dem = np.full((n, n), 500.0, dtype=np.float32)
    dem[0, :] = dem[-1, :] = dem[:, 0] = dem[:, -1] = -99999.0
    hdry = 1e-3
    zb, ks, weir = capi.read_geo(dem, -99999.0, hdry)
    zb, z, h, ldry, c0 = capi.initiation(zb, True)
    z[1:-1, 1:-1] = zb[1:-1, 1:-1] + H
    h[1:-1, 1:-1] = H
    ldry[1:-1, 1:-1] = 0.0
    d = capi.HipDomain(n, n, 2, hdry, 1, (0.01,), True, precision)
    d.set_static(zb, ks, weir)
    d.set("z", z)
    d.set("h", h)
    d.set("ldry", ldry)
    d.set_conc(0, c0)
    d.set_clock(0.0, 0.0, 0)
    soil = np.zeros(zb.shape, order="F")
    soil[1:-1, 1:-1] = soil0
    d.set_wintra(True, rate)
    d.set("soil", soil)
    return d


def test_wintra_fp32_against_fp64():
    """FP32-mixed with the fast ADE on a basin that stays wet everywhere:
    the FP64 path's soil mass follows the closed form s0 prod(1 - r dt/3600)
    over its own dt sequence; the FP32 path's soil mass (held in FP64) follows
    the FP64 path to 1e-6, and the released solute, uniform in the interior,
    agrees to FP32 rounding away from the ring (where the walls' thrust and
    the ring's zero solute reach within 30 steps: one cell per step)."""
    n, steps, soil0, rate = 96, 30, 20.0, 0.9
    res = {}
    for prec in (capi.FP64_PARITY, capi.FP32_MIXED):
        d = _still_basin(prec, n=n, soil0=soil0, rate=rate)
        dts = []
        for _ in range(steps):
            info = d.advance(1e9, 1e9, 1)
            dts.append(info.dtfl)
        res[prec] = {"soil": d.get("soil"), "conc": d.get_conc(0), "dts": np.array(dts)}
        d.close()
    deep = (slice(40, n - 38), slice(40, n - 38))
    r64, r32 = res[capi.FP64_PARITY], res[capi.FP32_MIXED]
    closed = soil0 * np.prod(1.0 - rate * r64["dts"] / 3600.0)
    _say("dt fp64", r64["dts"][:3], "fp32", r32["dts"][:3], "closed-form soil", closed)
    assert np.allclose(r64["soil"][deep], closed, rtol=1e-13, atol=0)
    assert max_rel(r64["soil"][1:-1, 1:-1], r32["soil"][1:-1, 1:-1]) <= 1e-6
    c64, c32 = r64["conc"][deep], r32["conc"][deep]
    _say("deep conc fp64", c64.min(), c64.max(), "fp32", c32.min(), c32.max())
    assert c64.min() > 0 and max_rel(c64, c32) <= 1e-5


def test_wintra_needs_ade_and_stays_out_of_batches():
    sc = gen.scenario_c1(32, seed=1)
    d = capi.HipDomain.from_scenario(sc, ade=False)
    with pytest.raises(capi.FxError):
        d.set_wintra(True, 0.5)
    d.close()
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED)
    d.set_wintra(True, 0.5)
    b = capi.Batch(nchem=1, ade=True, max_members=4)
    with pytest.raises(capi.FxError):
        b.admit(d, 600.0, 600.0)
    b.close()
    d.close()
