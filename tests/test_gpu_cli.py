"""The command-line driver end to end: modset + ESRI DEM + meteo CSV in, the
reference's Results/<t>.txt files out, on the FP64 parity path -- compared
value by value with the oracle running the reference's time loop
(main.cpp:271-470) on the same inputs and writing the same rows
(write_results.cpp:30-102)."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from fluxos_cpp_amd import build, gen
from oracle.oracle import OracleDomain, run_reference_loop

pytestmark = pytest.mark.gpu


def _expected_rows(d: OracleDomain, dxy: int, xll: float, yll: float) -> np.ndarray:
    g = d.g
    rows = []
    for icol in range(1, d.ncols + 1):
        for irow in range(1, d.nrows + 1):
            if g["zb"][irow, icol] != 0.0 and g["h"][irow, icol] > 0.0:     # H_MIN_TO_PRINT 0.005 -> 0
                rows.append([irow, icol, icol * dxy + xll, irow * dxy + yll, g["z"][irow, icol],
                             g["z"][irow, icol] - g["zb"][irow, icol], 0.0, 0.0, g["qx"][irow, icol] * dxy,
                             g["qy"][irow, icol] * dxy, g["us"][irow, icol], d.conc[0][irow, icol], 0.0,
                             g["fe1"][irow, icol], g["fn1"][irow, icol], g["twet"][irow, icol]])
    return np.array(rows)


def test_cli_results_match_oracle(tmp_path):
    if not build.CLI.exists():
        build.build_cli()
    sc = gen.scenario_c1(48, seed=6, t_end=1800)
    modset = gen.write_case(tmp_path, sc)
    print(f"case {modset}", flush=True)
    r = subprocess.run([str(build.CLI), str(modset)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    expected = {}

    def on_print(t, d):
        expected[int(round(t))] = _expected_rows(d, sc.dxy, 0.0, 0.0)

    o = OracleDomain.from_scenario(sc, ade=True)
    run_reference_loop(o, float(sc.t_end), sc.print_step, on_print=on_print)
    assert sorted(expected) == [600, 1200, 1800]
    for t, want in expected.items():
        path = tmp_path / "Results" / f"{t}.txt"
        lines = path.read_text().splitlines()
        got = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])
        assert got.shape == want.shape, (t, got.shape, want.shape)
        assert np.array_equal(got, want), f"Results/{t}.txt differs from the oracle"
