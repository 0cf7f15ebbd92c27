"""The oracle fixtures of tests/golden/oracle_<case>.json are the oracle's own
output: the first print times of each case are recomputed here on the CPU and
must hash the same (the GPU tests compare the device against the whole
recorded trajectory)."""
from __future__ import annotations

import json
from pathlib import Path

import pytest

from oracle.oracle import OracleDomain
from tests.golden.make_oracle_fixtures import cases, oracle_grids, record

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("case,nprints", [("c1_256", 2), ("c3_04", 2), ("c5_192", 1)])
def test_fixture_prints_reproduce(case, nprints, oracle_lib):
    fx = json.loads((GOLDEN / f"oracle_{case}.json").read_text())
    sc = cases()[case]()
    o = OracleDomain.from_scenario(sc)
    print_next, steps = sc.print_step, 0
    got = []
    while len(got) < nprints:
        o.step(print_next)
        steps += 1
        if o.tim >= print_next:
            got.append(record(oracle_grids(o), o.tim, o.it.value, steps))
            print_next += sc.print_step
    for rec, want in zip(got, fx["prints"][:nprints]):
        assert rec == want, (case, want["tim"])
