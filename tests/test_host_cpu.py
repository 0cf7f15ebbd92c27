"""Host-side logic that needs no GPU: the tile-table lookup of the
autotuner, the serving loop's memory admission figure and its arrival
stream."""
from __future__ import annotations

import json

import numpy as np
import pytest

from fluxos_cpp_amd import autotune, capi, gen, serve


def _table(tmp_path, entries):
    p = tmp_path / "tuning.json"
    p.write_text(json.dumps({"entries": entries}))
    return p


def test_tile_table_nearest_size_by_log_distance(tmp_path):
    p = _table(tmp_path, [{"n": 512, "cells": 512 ** 2, "ch": 8, "ach": 16},
                          {"n": 2048, "cells": 2048 ** 2, "ch": 24, "ach": 32},
                          {"n": 4096, "cells": 4096 ** 2, "ch": 8, "ach": 64}])
    assert autotune.choose(512, 512, p) == (8, 16)
    assert autotune.choose(600, 600, p) == (8, 16)
    # 1024^2 is 4x from 512^2 and 4x from 2048^2: the first nearest wins
    assert autotune.choose(1024, 1024, p) == (8, 16)
    assert autotune.choose(1100, 1100, p) == (24, 32)
    assert autotune.choose(3000, 3000, p) == (8, 64)
    assert autotune.choose(16384, 16384, p) == (8, 64)
    assert autotune.choose_cells(0, p) == (8, 16)              # clamped to one cell


def test_tile_table_missing_or_corrupt(tmp_path):
    assert autotune.choose(256, 256, tmp_path / "absent.json") is None
    bad = tmp_path / "bad.json"
    bad.write_text("{not json")
    assert autotune.choose(256, 256, bad) is None


def test_batch_entry_wins_for_batches_else_nearest_total(tmp_path):
    entries = [{"n": 512, "cells": 512 ** 2, "ch": 8, "ach": 16},
               {"n": 4096, "cells": 4096 ** 2, "ch": 24, "ach": 64}]
    p = _table(tmp_path, entries)
    # no batch entry: the batch's total cell count picks the size entry
    assert autotune.choose_batch([256 ** 2] * 4, p) == (8, 16)
    assert autotune.choose_batch([2048 ** 2] * 4, p) == (24, 64)
    p.write_text(json.dumps({"entries": entries, "batch": {"ch": 12, "ach": 32}}))
    assert autotune.choose_batch([2048 ** 2] * 4, p) == (12, 32)
    assert autotune.choose(4096, 4096, p) == (24, 64)         # single domains ignore it


def test_committed_batch_entry_is_the_fastest_measured():
    b = json.loads(autotune.TABLE.read_text())["batch"]
    rates = b["cell_updates_per_s"]
    assert max(rates, key=rates.get) == f"{b['ch']},{b['ach']}"


def test_committed_table_is_well_formed():
    entries = autotune.load_table()
    assert entries, "fluxos_cpp_amd/tuning.json is missing or empty"
    for e in entries:
        assert e["cells"] == e["n"] ** 2
        assert e["ch"] in autotune.CANDIDATES_CH and e["ach"] in autotune.CANDIDATES_ACH
        assert min(e["k32_sweep_us"].values()) == e["k32_sweep_us"][str(e["ch"])]
        assert min(e["k32_ade_us"].values()) == e["k32_ade_us"][str(e["ach"])]


def test_domain_bytes_grows_with_cells_and_solutes(fxlib):
    small = serve.domain_bytes(gen.scenario_c3(0))
    sc = gen.scenario_c3(0)
    big = gen.Scenario("big", np.zeros((2 * sc.nrows, 2 * sc.ncols), dtype=np.float32), sc.dxy, sc.meteo,
                       sc.nchem, sc.d_coef, sc.t_end, sc.seed)
    assert small > 0
    assert 3.5 * small < serve.domain_bytes(big) < 4.5 * small
    assert serve.domain_bytes(sc, ade=False) < small


def test_domain_bytes_refuses_an_fp32_domain_too_large(fxlib):
    sc = gen.scenario_c3(0)
    dem = np.lib.stride_tricks.as_strided(np.zeros(1, dtype=np.float32), shape=(23170, 23170), strides=(0, 0))
    huge = gen.Scenario("huge", dem, sc.dxy, sc.meteo, sc.nchem, sc.d_coef, sc.t_end, sc.seed)
    with pytest.raises(capi.FxError):
        serve.domain_bytes(huge)


def test_poisson_arrivals():
    a = serve.poisson_arrivals(20.0, 50.0, seed=9)
    assert np.all(np.diff(a) > 0) and a[0] >= 0 and a[-1] < 50.0
    assert abs(len(a) - 1000) < 4 * np.sqrt(1000)               # mean rate*seconds, sd sqrt
    gaps = np.diff(np.concatenate([[0.0], a]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.005
    assert np.array_equal(a, serve.poisson_arrivals(20.0, 50.0, seed=9))
