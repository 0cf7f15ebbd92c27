"""GPU tests on the benchmark configurations' own generators.

* C1 (the plain-method case) at its full size and length: 256x256, 2
  simulated hours, one solute with the exact ADE.
* A C3 ensemble member (256x256, its own melt rate, 1 simulated hour).
* The C5 generator -- fractal DEM with carved channels, a smooth melt
  multiplier field and four solutes with distinct dispersion coefficients --
  at 192x192 (dx = 1 m) for 40 simulated minutes.

The FP64 parity path must match the oracle bit for bit at every print time:
the oracle's trajectories are recorded in tests/golden/oracle_<case>.json
(written by tests/golden/make_oracle_fixtures.py) as the clock and a SHA-256
of every state grid.  Beside them: six C3 members in one batch, each
bit-identical to its run alone, the ensemble driver's count of their cell
updates, and the C5 generator on the FP32-mixed path with four solutes under
two tile shapes, bit-identical."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen
from tests.golden.make_oracle_fixtures import FIELDS, cases, digest

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
PRINT = 600.0


def _device_grids(d: capi.HipDomain) -> dict:
    g = {k: d.get(k) for k in FIELDS}
    for k in range(d.nchem):
        g[f"conc{k}"] = d.get_conc(k)
    return g


@pytest.mark.parametrize("case", ["c1_256", "c3_04", "c5_192"])
def test_fp64_matches_oracle_fixture_at_every_print(case):
    fx = json.loads((GOLDEN / f"oracle_{case}.json").read_text())
    sc = cases()[case]()
    assert (sc.nrows, sc.ncols, sc.nchem, sc.t_end) == (fx["nrows"], fx["ncols"], fx["nchem"], fx["t_end"])
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP64_PARITY, ade=True)
    print_next = float(sc.print_step)
    steps = 0
    for rec in fx["prints"]:
        info = d.advance(float(sc.t_end), print_next)
        steps += info.steps
        t, _, it = d.clock()
        assert (t, it, steps) == (rec["tim"], rec["it"], rec["steps"]), (case, print_next)
        g = _device_grids(d)
        for k, h in rec["sha256"].items():
            assert digest(g[k]) == h, (case, print_next, k, float(np.nansum(g["h"][1:-1, 1:-1])), rec["sum_h"])
        print_next += sc.print_step
    d.close()


def _run_alone(sc) -> dict:
    d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True)
    pn = PRINT
    while True:
        info = d.advance(float(sc.t_end), pn)
        if info.status & 2:
            break
        if info.status & 1:
            pn += PRINT
    out = {k: d.get(k) for k in ("z", "h", "qx", "qy", "ldry")} | {"c0": d.get_conc(0)}
    out["clock"] = np.array(d.clock())
    d.close()
    return out


def test_c3_members_in_one_batch_match_their_runs_alone():
    """Six members of the 64-scenario ensemble (the smallest: 256-280 cells a
    side, 14.8-73.9 mm/day), each with its own CFL step, in one batch for the
    whole simulated hour."""
    idx = (4, 22, 44, 54, 27, 58)
    scs = [gen.scenario_c3(i) for i in idx]
    alone = [_run_alone(sc) for sc in scs]
    batch = capi.Batch(nchem=1, ade=True, max_members=len(scs))
    doms, slot_of, next_print = [], {}, {}
    for k, sc in enumerate(scs):
        d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True)
        s = batch.admit(d, float(sc.t_end), PRINT)
        doms.append(d)
        slot_of[s] = k
        next_print[s] = PRINT
    done = set()
    for _ in range(100000):
        if len(done) == len(scs):
            break
        batch.run(32)
        for slot, kind, tim, it, steps, solver in batch.poll():
            assert kind != capi.EV_ERROR, slot
            if kind == capi.EV_DONE:
                done.add(slot)
            else:
                next_print[slot] += PRINT
                batch.resume(slot, next_print[slot])
    assert len(done) == len(scs)
    for s, k in slot_of.items():
        d = doms[k]
        got = {f: d.get(f) for f in ("z", "h", "qx", "qy", "ldry")} | {"c0": d.get_conc(0)}
        got["clock"] = np.array(d.clock())
        for f, a in alone[k].items():
            assert np.array_equal(a, got[f], equal_nan=True), (idx[k], f)
    batch.close()
    for d in doms:
        d.close()


def test_run_ensemble_counts_every_member():
    """serve.run_ensemble, the ensemble leg's driver, on the same six members:
    every member runs to its print at t_end (its result; the reference's loop
    takes one more step after it, which nothing prints) and leaves the batch,
    and the reported cell updates are each member's cells times the solver
    steps of its run alone to that print."""
    from fluxos_cpp_amd import serve
    idx = (4, 22, 44, 54, 27, 58)
    scs = [gen.scenario_c3(i) for i in idx]
    inits = [capi.prepare(sc) for sc in scs]
    expect = 0
    for sc, init in zip(scs, inits):
        d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, init=init)
        pn, solver = float(sc.print_step), 0
        while True:
            info = d.advance(float(sc.t_end), pn)
            solver += info.solver_steps
            if info.status & 2 or info.tim >= sc.t_end:
                break
            if info.status & 1:
                pn += float(sc.print_step)
        expect += sc.nrows * sc.ncols * solver
        d.close()
    r = serve.run_ensemble(scs, inits)
    print(f"cell updates {r['cell_updates']} (alone: {expect}), {r['wall_s']:.2f} s", flush=True)
    assert r["failed_slots"] == []
    assert r["scenarios"] == len(scs)
    assert r["cell_updates"] == expect


def test_c5_generator_fp32_four_solutes_independent_of_tiling():
    sc = gen.scenario_c5(192, t_end=2400)
    assert sc.melt_mult is not None and sc.nchem == 4 and len(set(sc.d_coef)) == 4
    states = []
    for tiling in ((8, 16), (32, 64)):
        d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, ade=True, tiling=tiling)
        d.advance(float(sc.t_end), 1200.0)
        d.advance(float(sc.t_end), 1500.0)
        st = {k: d.get(k) for k in ("z", "h", "qx", "qy", "ldry", "twet")}
        for k in range(4):
            st[f"c{k}"] = d.get_conc(k)
        st["clock"] = np.array(d.clock())
        states.append(st)
        d.close()
    for k, a in states[0].items():
        assert np.array_equal(a, states[1][k], equal_nan=True), k
    # the four solutes differ (distinct D), and the channels hold water
    c = [states[0][f"c{k}"][1:-1, 1:-1] for k in range(4)]
    assert all(not np.array_equal(c[0], ck) for ck in c[1:])
    assert (states[0]["h"][1:-1, 1:-1] > 1e-3).sum() > 100
