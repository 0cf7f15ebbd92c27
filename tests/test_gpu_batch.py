"""Batched FP32-mixed domains (fx_batch_*): each member's result is the one it
gets alone.

Alone, a domain runs with fx_advance (one kernel sequence per step for that
domain); in a batch, every step is one launch per kernel over all members'
blocks, replayed from a captured graph, with each member on its own CFL step
and clock.  The arithmetic per cell is the same code, so the states must be
identical bit for bit -- the stated tolerance for "the same alone and in the
mix" is zero."""
from __future__ import annotations

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen

pytestmark = pytest.mark.gpu
PRINT = 600.0


def _scen(n, seed, rate):
    sc = gen.scenario_c1(n, seed=seed, t_end=1500)
    sc.meteo[1:, 1] = np.float32(rate)          # per-scenario melt rate [mm/day]
    return sc


def _alone(sc, t_end):
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True)
    pn = PRINT
    while True:
        info = d.advance(t_end, pn)
        if info.status & 2:
            break
        if info.status & 1:
            pn += PRINT
    return d, info


def _state(d):
    return {k: d.get(k) for k in ("h", "qx", "qy", "z")} | {"c0": d.get_conc(0), "clock": np.array(d.clock()[:2])}


def _drive(batch, next_print, done, limit=100000):
    """Run the batch until every member has finished; resume at prints."""
    for _ in range(limit):
        if len(done) == len(batch.members):
            return
        batch.run(16)
        for slot, kind, tim, it, steps, solver in batch.poll():
            assert kind != capi.EV_ERROR, f"member {slot} reported NaN/Inf"
            if kind == capi.EV_DONE:
                done.add(slot)
            else:
                next_print[slot] += PRINT
                batch.resume(slot, next_print[slot])
    raise AssertionError("batch did not finish")


def test_batch_members_match_alone():
    t_end = 1500.0
    specs = [(64, 1, 20.0), (100, 3, 45.0), (130, 4, 70.0), (77, 5, 33.0)]
    scens = [_scen(*s) for s in specs]
    alone = [_state(_alone(sc, t_end)[0]) for sc in scens]
    print("alone clocks:", [a["clock"].tolist() for a in alone], flush=True)

    batch = capi.Batch(nchem=1, ade=True, max_members=8)
    doms = [capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True) for sc in scens]
    slots = [batch.admit(d, t_end, PRINT) for d in doms]
    next_print = {s: PRINT for s in slots}
    done: set[int] = set()
    _drive(batch, next_print, done)
    n, cells, upd = batch.info()
    print(f"batch: {n} members, {cells} cells, {upd} cell-updates", flush=True)
    for spec, slot, d, a in zip(specs, slots, doms, alone):
        b = _state(d)
        for k in a:
            assert np.array_equal(a[k], b[k], equal_nan=True), f"{spec} field {k} differs"
    batch.close()


def test_batch_evict_and_late_admission():
    """A member leaves mid-run and another joins late: the remaining members'
    results are unchanged, and the late one matches its own alone run."""
    t_end = 1300.0
    scens = [_scen(90, 7, 50.0), _scen(70, 8, 30.0), _scen(110, 9, 60.0)]
    alone = [_state(_alone(sc, t_end)[0]) for sc in scens]

    batch = capi.Batch(nchem=1, ade=True, max_members=4)
    d0 = capi.HipDomain.from_scenario(scens[0], precision=capi.FP32_MIXED, ade=True)
    d1 = capi.HipDomain.from_scenario(scens[1], precision=capi.FP32_MIXED, ade=True)
    s0 = batch.admit(d0, t_end, PRINT)
    s1 = batch.admit(d1, t_end, PRINT)
    next_print = {s0: PRINT, s1: PRINT}
    # run member 1 to its first print, then evict it (its domain keeps its state)
    for _ in range(10000):
        batch.run(8)
        ev = {e[0]: e for e in batch.poll()}
        if s0 in ev:
            next_print[s0] += PRINT
            batch.resume(s0, next_print[s0])
        if s1 in ev:
            break
    batch.evict(s1)
    del next_print[s1]
    # late admission of a third member into the freed slot
    d2 = capi.HipDomain.from_scenario(scens[2], precision=capi.FP32_MIXED, ade=True)
    s2 = batch.admit(d2, t_end, PRINT)
    next_print[s2] = PRINT
    done: set[int] = set()
    _drive(batch, next_print, done)
    for d, a, name in ((d0, alone[0], "first"), (d2, alone[2], "late")):
        b = _state(d)
        for k in a:
            assert np.array_equal(a[k], b[k], equal_nan=True), f"{name} member field {k} differs"
    # the evicted member continues alone from where it left the batch
    pn = PRINT * 2
    while True:
        info = d1.advance(t_end, pn)
        if info.status & 2:
            break
        if info.status & 1:
            pn += PRINT
    b = _state(d1)
    for k in alone[1]:
        assert np.array_equal(alone[1][k], b[k], equal_nan=True), f"evicted member field {k} differs"
    batch.close()


def test_tiling_does_not_change_results():
    """The sweeps' column-block lengths (fx_set_tiling, chosen by the
    autotuner) only regroup the same per-cell arithmetic: the state is
    identical bit for bit across shapes, and the water ledger agrees to
    rounding of its own sums."""
    sc = _scen(140, 11, 55.0)
    states = []
    for ch, ach in ((32, 64), (8, 16), (128, 256), (17, 40)):
        d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, tiling=(ch, ach))
        assert d.tiling() == (ch, ach)
        pn = PRINT
        while True:
            info = d.advance(1500.0, pn)
            if info.status & 2:
                break
            if info.status & 1:
                pn += PRINT
            if info.tim > 700.0 and d.tiling() == (32, 64):
                d.set_tiling(64, 32)                       # a change mid-run keeps the ledger
        st = _state(d)
        st["mb"] = np.array(d.mass_balance())
        states.append(st)
        d.close()
    for st in states[1:]:
        for k in st:
            if k == "mb":
                assert np.allclose(st[k], states[0][k], rtol=1e-12, atol=1e-12)
            else:
                assert np.array_equal(st[k], states[0][k], equal_nan=True), k


def test_members_refuse_changes_until_evicted():
    """The batch's slot holds a member's launch parameters and table
    addresses from admission on, so every call that would change the member
    returns FX_ERR_STATE until it is evicted; reads still work.  The refused
    calls leave the member's run bit-identical to its alone run."""
    t_end = 1300.0
    sc = _scen(48, 21, 45.0)
    ref = _state(_alone(sc, t_end)[0])
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True)
    batch = capi.Batch(nchem=1, ade=True, max_members=4)
    s = batch.admit(d, t_end, PRINT)
    refused = {
        "set_meteo": lambda: d.set_meteo(sc.meteo),
        "set_inflow": lambda: d.set_inflow(sc.meteo, 5, 5),
        "set_field": lambda: d.set("h", np.zeros((sc.nrows + 2, sc.ncols + 2))),
        "set_clock": lambda: d.set_clock(0.0),
        "courant": lambda: d.courant(PRINT),
        "hydrodynamics": lambda: d.hydrodynamics(),
        "advance": lambda: d.advance(t_end, PRINT),
        "set_tiling": lambda: d.set_tiling(8, 16),
    }
    for name, call in refused.items():
        with pytest.raises(capi.FxError) as e:
            call()
        assert e.value.code == -3, name
    assert capi.load().fx_domain_destroy(d.h) == -3
    d.get("h"), d.clock(), d.mass_balance()                     # reads stay allowed
    next_print = {s: PRINT}
    _drive(batch, next_print, set())
    got = _state(d)
    for k in ref:
        assert np.array_equal(ref[k], got[k], equal_nan=True), k
    batch.evict(s)
    d.set_clock(5.0, 1.0, 3)                                    # its own again
    assert d.clock()[:2] == (5.0, 1.0)
    # destroying a batch releases the members still in it
    s = batch.admit(d, t_end, PRINT)
    with pytest.raises(capi.FxError):
        d.close()
    batch.close()
    d.set_clock(0.0)
    d.close()
