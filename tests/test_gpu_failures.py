"""Failure handling of the batches and the serving loop, through the C ABI.

The reference has no detection of NaN or instability and no admission control:
an input error sets errflag and the run exits (main.cpp:119-120,349,370).
Here a scenario whose state turns non-finite is halted and reported alone
(FX_ERR_NONFINITE for a single domain, FX_EV_ERROR for a batch member) while
the others run on unchanged, and the serving loop admits requests by an HBM
budget: a request that does not fit waits, one that can never fit is refused.
The NaN is planted from the host through fx_set_field; every kernel still runs
to its normal end."""
from __future__ import annotations

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen, serve

pytestmark = pytest.mark.gpu
PRINT = 600.0


def _scen(n: int, seed: int, rate: float, t_end: int = 1800):
    sc = gen.scenario_c1(n, seed=seed, t_end=t_end)
    sc.meteo[1:, 1] = np.float32(rate)          # per-scenario melt rate [mm/day]
    return sc


def _alone(sc, to_last_print: bool = False):
    """A scenario run alone with fx_advance, halting at every print as a batch
    does: until the clock passes t_end (main.cpp:271), or, with to_last_print,
    until the print at t_end, which the serving loop returns as the result."""
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True)
    pn = PRINT
    while True:
        info = d.advance(float(sc.t_end), pn)
        if info.status & 2 or (to_last_print and info.tim >= sc.t_end):
            break
        if info.status & 1:
            pn += PRINT
    return d


def _fields(d):
    return {k: d.get(k) for k in ("z", "h", "qx", "qy")} | {"c0": d.get_conc(0)}


def _plant_nan(d: capi.HipDomain) -> None:
    z = d.get("z")
    z[d.nrows // 2, d.ncols // 3] = np.nan
    d.set("z", z)


def test_fp32_domain_reports_nonfinite():
    """fx_advance on an FP32-mixed domain returns FX_ERR_NONFINITE and halts
    within the call (it no longer reads a NaN depth as a dry cell)."""
    sc = _scen(48, 3, 40.0)
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True)
    d.advance(float(sc.t_end), 600.0)
    info = d.advance(float(sc.t_end), 1200.0)     # wet: every iteration runs the solver
    assert info.solver_steps > 1
    _plant_nan(d)
    with pytest.raises(capi.FxError) as e:
        d.advance(float(sc.t_end), 1800.0, 500)
    assert e.value.code == -5
    t, _, _ = d.clock()
    assert t < 1800.0                             # halted on the device, not run to the print
    d.close()


def test_nan_member_evicted_alone():
    """One member of a batch of three gets a NaN in its water surface: only it
    reports FX_EV_ERROR; the other two finish bit-identical to their runs alone."""
    scs = [_scen(48, 11, 30.0), _scen(64, 12, 55.0), _scen(40, 13, 70.0)]
    alone = [_fields(_alone(sc)) for sc in scs]
    batch = capi.Batch(nchem=1, ade=True, max_members=4)
    doms, slots, next_print = [], [], {}
    for k, sc in enumerate(scs):
        d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True)
        if k == 1:
            d.advance(float(sc.t_end), 300.0)     # wet a little first, then plant the NaN
            _plant_nan(d)
        s = batch.admit(d, float(sc.t_end), PRINT)
        doms.append(d)
        slots.append(s)
        next_print[s] = PRINT
    done, errors = set(), set()
    for _ in range(20000):
        if len(done) + len(errors) == 3:
            break
        batch.run(16)
        for slot, kind, tim, it, steps, solver in batch.poll():
            if kind == capi.EV_ERROR:
                errors.add(slot)
                batch.evict(slot)                 # evict that scenario only
            elif kind == capi.EV_DONE:
                done.add(slot)
            else:
                next_print[slot] += PRINT
                batch.resume(slot, next_print[slot])
    assert errors == {slots[1]}
    assert done == {slots[0], slots[2]}
    for k in (0, 2):
        got = _fields(doms[k])
        for f, a in alone[k].items():
            assert np.array_equal(a, got[f], equal_nan=True), (k, f)
    assert np.isnan(doms[1].get("z")).any()
    batch.close()
    for d in doms:
        d.close()


def test_domain_in_two_batches_refused():
    sc = _scen(32, 5, 20.0)
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True)
    a = capi.Batch(nchem=1, ade=True, max_members=2)
    b = capi.Batch(nchem=1, ade=True, max_members=2)
    a.admit(d, float(sc.t_end), PRINT)
    with pytest.raises(capi.FxError) as e:
        b.admit(d, float(sc.t_end), PRINT)
    assert e.value.code == -3
    assert b.info()[0] == 0
    s = next(iter(a.members))
    a.evict(s)
    b.admit(d, float(sc.t_end), PRINT)           # free again once evicted
    b.close()
    a.close()
    d.close()


def _requests(n: int, t_end: int = 1200, rate: float = 0.0):
    """n requests of one size (48^2), each its own DEM and melt rate, all
    arriving at once, or as a seeded Poisson stream of the given rate [1/s]."""
    gaps = np.random.default_rng(7).exponential(1.0 / rate, n) if rate > 0 else np.zeros(n)
    reqs = []
    for i, t in enumerate(np.cumsum(gaps)):
        sc = _scen(48, 20 + i, 25.0 + 10.0 * i, t_end=t_end)
        reqs.append(serve.Request(i, float(t), sc, capi.prepare(sc)))
    return reqs


def test_serve_queues_on_budget_and_results_match_alone():
    """A Poisson stream (500 requests/s, so they arrive while the first ones
    run) under a budget that holds two of the requests at a time: the others
    wait for memory and then run; every result equals the request's run
    alone, whether it ran beside another member or not."""
    reqs = _requests(5, rate=500.0)
    nb = serve.domain_bytes(reqs[0].scenario)
    budget = nb * 5 // 2
    out = serve.serve(reqs, budget_bytes=budget, keep_results=True, steps_per_poll=4)
    assert out["requests"] == 5 and out["failed"] == 0, out
    assert out["queued_for_memory"] >= 2, out
    assert out["peak_members"] == 2, out
    assert out["peak_device_bytes"] <= budget
    assert max(r.admitted - r.arrival for r in reqs) > 0.0
    for r in reqs:
        h = _alone(r.scenario, to_last_print=True).get("h")
        assert np.array_equal(h, r.result, equal_nan=True), r.rid


def test_serve_rejects_a_request_above_the_budget():
    """A request that alone exceeds the budget is refused at once (it could
    never run) and the loop returns; smaller ones still run."""
    reqs = _requests(3, t_end=600)
    big = serve.Request(3, 0.0, _scen(200, 9, 30.0, t_end=600), None)
    nb_small = max(serve.domain_bytes(r.scenario) for r in reqs)
    assert serve.domain_bytes(big.scenario) > nb_small * 2
    big.init = capi.prepare(big.scenario)
    out = serve.serve(reqs + [big], budget_bytes=nb_small * 2)
    assert out["requests"] == 3 and out["failed"] == 1, out
    assert out["failed_ids"] == [3]
    assert "budget" in big.error


def test_serve_evicts_a_nan_request_and_serves_the_rest():
    reqs = _requests(4)

    def hook(req, dom):
        if req.rid == 2:
            dom.advance(float(req.scenario.t_end), 300.0)
            _plant_nan(dom)

    out = serve.serve(reqs, keep_results=True, prepare_hook=hook)
    assert out["requests"] == 3 and out["failed"] == 1, out
    assert out["failed_ids"] == [2]
    assert "NaN" in reqs[2].error
    for r in reqs:
        if r.rid == 2:
            continue
        h = _alone(r.scenario, to_last_print=True).get("h")
        assert np.array_equal(h, r.result, equal_nan=True), r.rid
