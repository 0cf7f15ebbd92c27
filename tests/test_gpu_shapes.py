"""Grid shapes away from the tile sizes, and more solutes, on both paths.

The FP32-mixed sweeps cut the grid into strips of 61 (step) and 58 (ADE) rows
and blocks of `ch` / `ach` columns, the exact ADE into waves of 64 columns;
the shapes here leave partial strips, partial blocks and partial waves
everywhere (5 x 7 interior is smaller than one strip and one wave).  The
FP64 path must stay bit-identical to the oracle on each; the FP32 path must
give bit-identical results whatever its tile shape."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen
from oracle.oracle import OracleDomain, run_reference_loop
from tests.helpers import matched, max_rel

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scenario(nr, nc, nchem=3, t_end=1200, seed=11, rate=400.0):
    dem = gen.dem_fractal(nr, nc, 2.0, seed, relief=1.0)
    times = np.arange(0, t_end + 1, 300)
    concs = [1.0 + k for k in range(nchem)]
    d_coef = [0.01 * (k + 1) for k in range(nchem)]
    return gen.Scenario(f"shape{nr}x{nc}", dem, 2, gen.meteo_table(times, rate, concs), nchem, d_coef,
                        t_end, seed, print_step=300)


@pytest.mark.parametrize("nr,nc", [(5, 7), (37, 53), (130, 67)])
def test_fp64_odd_shapes_three_solutes(nr, nc):
    sc = _scenario(nr, nc)
    o, d = matched(sc)
    snaps = {}

    def grab(t, od):
        snaps[t] = {"z": od.g["z"].copy(), "qx": od.g["qx"].copy(), "qy": od.g["qy"].copy()}
        for k in range(od.nchem):
            snaps[t][f"c{k}"] = od.conc[k].copy()

    run_reference_loop(o, sc.t_end, sc.print_step, on_print=grab)
    seen = []

    def check(t, dd, info):
        got = {"z": dd.get("z"), "qx": dd.get("qx"), "qy": dd.get("qy")}
        for k in range(sc.nchem):
            got[f"c{k}"] = dd.get_conc(k)
        r = {k: max_rel(snaps[t][k], got[k]) for k in got}
        _say(f"{nr}x{nc} print {t} it {info.it}", r)
        assert all(v == 0.0 for v in r.values()), (t, r)
        seen.append(t)

    capi.run_loop(d, sc.t_end, sc.print_step, on_print=check)
    assert seen == sorted(snaps) and len(seen) == 4
    assert o.wet_count() > 0


@pytest.mark.parametrize("nr,nc", [(37, 53), (130, 67), (200, 41)])
def test_fp32_tile_shape_invariance_odd_shapes(nr, nc):
    sc = _scenario(nr, nc)
    init = capi.prepare(sc)
    out = {}
    for tiling in ((8, 8), (24, 40), (4096, 4096)):
        d = capi.HipDomain.from_scenario(sc, precision=capi.FP32_MIXED, init=init, tiling=tiling)
        capi.run_loop(d, sc.t_end, sc.print_step)
        out[tiling] = [d.get("h"), d.get("qx")] + [d.get_conc(k) for k in range(sc.nchem)]
        info = d.clock()
        d.close()
        _say(f"{nr}x{nc} tiling {tiling}: t {info[0]:.3f} it {info[2]} max h {out[tiling][0].max():.4f}")
    ref = out[(8, 8)]
    for tiling, arrs in out.items():
        for a, b in zip(ref, arrs):
            assert np.array_equal(a, b), tiling
