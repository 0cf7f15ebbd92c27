"""The file formats at the boundary (fx_read_dem, fx_read_forcing,
fx_write_results) and the command-line driver's input handling, on the CPU.

Formats follow read_functions.cpp:112-326, common.cpp:110-161 and
write_results.cpp:30-102.  The reference ships no tests for them and one
input file (its Working_example melt forcing, kept under tests/golden/), so
the checks use that file, files written by this package's generator and
hand-made edge cases (header order, integer CELLSIZE, the unsigned-long print
threshold)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from fluxos_cpp_amd import build, capi, gen


DemHeader, ResultsGrids = capi.DemHeader, capi.ResultsGrids
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def lib():
    return capi.load()


def read_dem(lib, path):
    h = DemHeader()
    assert lib.fx_read_dem(str(path).encode(), C.byref(h), None) == 0, lib.fx_last_error()
    dem = np.zeros((h.nrows, h.ncols), dtype=np.float32)
    assert lib.fx_read_dem(str(path).encode(), C.byref(h), dem.ctypes.data_as(C.POINTER(C.c_float))) == 0
    return h, dem


def test_dem_round_trip(lib, tmp_path):
    sc = gen.scenario_c1(40, seed=3)
    gen.write_dem(tmp_path / "a.asc", sc.dem, sc.dxy, xll=425000.5, yll=5785000.0)
    h, dem = read_dem(lib, tmp_path / "a.asc")
    assert (h.nrows, h.ncols, h.cellsize) == (40, 40, 2)
    assert h.nodata == gen.NODATA and h.xllcorner == 425000.5 and h.yllcorner == 5785000.0
    assert np.array_equal(dem, sc.dem)


def test_dem_header_as_the_reference_reads_it(lib, tmp_path):
    """NCOLS/NROWS in either order and case, CELLSIZE and NODATA read as the
    first integer of their line (getIntNumberFromString: '2.7' -> 2)."""
    p = tmp_path / "b.asc"
    vals = np.arange(12, dtype=np.float32).reshape(3, 4) + 500.25
    with open(p, "w") as f:
        f.write("nrows 3\nncols 4\nxllcorner 10.5\nyllcorner -2\ncellsize 2.7\nNODATA_value -9999\n")
        for row in vals:
            f.write("  ".join(f"{v:.9g}" for v in row) + "  \n")
    h, dem = read_dem(lib, p)
    assert (h.nrows, h.ncols, h.cellsize, h.nodata) == (3, 4, 2, -9999.0)
    assert h.xllcorner == 10.5 and h.yllcorner == -2.0
    assert np.array_equal(dem, vals)


def test_dem_errors(lib, tmp_path):
    h = DemHeader()
    rc = lib.fx_read_dem(str(tmp_path / "missing.asc").encode(), C.byref(h), None)
    assert rc == -6 and b"cannot open" in lib.fx_last_error()
    (tmp_path / "short.asc").write_text("NCOLS 4\nNROWS 3\n")
    assert lib.fx_read_dem(str(tmp_path / "short.asc").encode(), C.byref(h), None) == -6


def test_forcing_csv(lib, tmp_path):
    sc = gen.scenario_c2(64)
    gen.write_meteo(tmp_path / "m.csv", sc.meteo)
    n, w = C.c_int32(), C.c_int32()
    path = str(tmp_path / "m.csv").encode()
    assert lib.fx_read_forcing(path, 0, 0, None, C.byref(n), C.byref(w)) == 0
    assert (n.value, w.value) == (sc.meteo.shape[0] - 1, 4)
    tab = np.zeros((n.value, 5), dtype=np.float32)          # one spare column stays 0
    assert lib.fx_read_forcing(path, n.value, 5, tab.ctypes.data_as(C.POINTER(C.c_float)),
                               C.byref(n), C.byref(w)) == 0
    assert np.array_equal(tab[:, :4], sc.meteo[1:]) and not tab[:, 4].any()
    small = np.zeros((2, 4), dtype=np.float32)
    assert lib.fx_read_forcing(path, 2, 4, small.ctypes.data_as(C.POINTER(C.c_float)),
                               C.byref(n), C.byref(w)) == -1


def test_forcing_reference_working_example(lib):
    """The reference's own melt file (Working_example/Qmelt_info_2009): the
    first line is skipped as the header even though it holds data, as
    read_meteo does by starting at row 1 (read_functions.cpp:258); each value
    is the float nearest the decimal text (the meteo table is FP32,
    GlobVar.h:72-73); the two-column layout gives two columns."""
    path = GOLDEN / "Qmelt_info_2009"
    lines = path.read_text().splitlines()
    assert len(lines) == 404 and lines[0].startswith("0,")
    tab = capi.read_forcing_file(path)
    assert tab.shape == (403, 2)
    assert tab[0, 0] == 1800.0 and tab[-1, 0] == 725400.0
    assert tab[0, 1] == np.float32(0.151000252983975)       # second line of the file
    assert tab[-1, 1] == np.float32(0.730419896333635)      # last line
    assert set(np.diff(tab[:, 0]).tolist()) == {1799.0, 1800.0, 1801.0}
    assert tab[:, 1].min() == 0.0 and 89.0 < tab[:, 1].max() < 89.5
    for i in (0, 57, 211, 402):                             # exact text -> float32, line by line
        t, r = lines[i + 1].split(",")
        assert tab[i, 0] == np.float32(float(t)) and tab[i, 1] == np.float32(float(r))
    wide = capi.read_forcing_file(path, ncols=3)             # a concentration column the file lacks
    assert np.array_equal(wide[:, :2], tab) and not wide[:, 2].any()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_results_writer_filter_order_and_columns(lib, tmp_path):
    rng = np.random.default_rng(5)
    nr, nc, dxy = 5, 4, 3
    shape = (nr + 2, nc + 2)
    zb = np.asfortranarray(rng.uniform(500, 501, shape))
    h = np.asfortranarray(np.where(rng.random(shape) < 0.4, 0.0, rng.uniform(0, 2.0, shape)))
    zb[2, 3] = 0.0                      # bed exactly 0 is never printed
    h[2, 3] = 1.0
    z = np.asfortranarray(zb + h + 1e-3)
    f = {k: np.asfortranarray(rng.normal(size=shape)) for k in ("qx", "qy", "conc0", "fe1", "fn1", "twet")}
    for hmin, thresh in ((0.005, 0.0), (1.5, 1.0)):
        g = ResultsGrids(nr, nc, dxy, 0, 100.0, 200.0, hmin, _dp(zb), _dp(z), _dp(h), _dp(f["qx"]), _dp(f["qy"]),
                         None, _dp(f["conc0"]), None, _dp(f["fe1"]), _dp(f["fn1"]), _dp(f["twet"]))
        out = tmp_path / f"r{hmin}.txt"
        assert lib.fx_write_results(str(out).encode(), C.byref(g)) == 0, lib.fx_last_error()
        lines = out.read_text().splitlines()
        assert lines[0].split(",")[0] == "irow [-]" and len(lines[0].split(",")) == 16
        got = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])
        want = []
        for icol in range(1, nc + 1):                   # column-major order, irow fastest
            for irow in range(1, nr + 1):
                if zb[irow, icol] != 0.0 and h[irow, icol] > thresh:
                    want.append([irow, icol, icol * dxy + 100.0, irow * dxy + 200.0, z[irow, icol],
                                 z[irow, icol] - zb[irow, icol], 0, 0, f["qx"][irow, icol] * dxy,
                                 f["qy"][irow, icol] * dxy, 0, f["conc0"][irow, icol], 0, f["fe1"][irow, icol],
                                 f["fn1"][irow, icol], f["twet"][irow, icol]])
        assert np.array_equal(got, np.array(want))     # %.16e round-trips doubles exactly


def _gpu_visible() -> bool:
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU failure path")
def test_cli_reads_inputs_then_fails_loudly_without_gpu(tmp_path):
    """The driver parses the modset (with comments), reads the DEM and the
    forcing, runs read_geo + initiation, and stops at domain creation with the
    library's no-device error -- there is no CPU fallback."""
    cli = build.CLI
    if not cli.exists():
        build.build_cli()
    sc = gen.scenario_c1(32, seed=2, t_end=1200)
    modset = gen.write_case(tmp_path, sc)
    text = modset.read_text()
    modset.write_text(text.replace('"RESTART": false,', '"RESTART": false, /* block comment */'))
    r = subprocess.run([str(cli), str(modset)], capture_output=True, text=True, timeout=60)
    out = r.stdout + r.stderr
    assert r.returncode != 0
    assert "Successful loading of master input file" in out
    assert "Successful loading of DEM file" in out and "Successful loading of METEO file" in out
    assert "Simulation time (days) = 1200" in out
    assert "no HIP device" in out
    assert (tmp_path / "fluxos_run.log").exists()


def test_cli_rejects_unsupported_modules(tmp_path):
    cli = build.CLI
    if not cli.exists():
        build.build_cli()
    sc = gen.scenario_c1(16, seed=2, t_end=600)
    modset = gen.write_case(tmp_path, sc)
    text = modset.read_text()
    modset.write_text(text.replace('"OPENWQ": {\n      "STATUS": false', '"OPENWQ": {\n      "STATUS": true'))
    r = subprocess.run([str(cli), str(modset)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "OPENWQ" in r.stdout
    # a compulsory key missing is the reference's "PROBLEM loading" (read_functions.cpp:98-103)
    modset.write_text(text.replace('"SOIL_RELEASE_RATE"', '"SOIL_RELEASE_RATE_X"'))
    r = subprocess.run([str(cli), str(modset)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "PROBLEM loading of master input file" in r.stdout


# ---------------------------------------------------------------- restart ----
def test_results_reader_round_trips_the_writer(lib, tmp_path):
    rng = np.random.default_rng(11)
    nr, nc, dxy = 6, 5, 2
    shape = (nr + 2, nc + 2)
    zb = np.asfortranarray(rng.uniform(500, 501, shape))
    h = np.asfortranarray(rng.uniform(0.01, 2.0, shape))
    z = np.asfortranarray(zb + h)
    grids = {k: np.asfortranarray(rng.normal(size=shape)) for k in
             ("qx", "qy", "conc0", "soil_mass", "fe1", "fn1", "twet", "ux", "uy")}
    out = tmp_path / "3600.txt"
    capi.write_results_file(out, dxy, 10.0, 20.0, 0.0, zb, z, h, **grids)
    tab = capi.read_results_file(out)
    assert tab.shape == (nr * nc, 16)
    # row of cell (irow, icol) in column-major order
    r = (3 - 1) * nr + (4 - 1)
    assert tab[r, 0] == 4 and tab[r, 1] == 3
    assert tab[r, 4] == z[4, 3] and tab[r, 5] == z[4, 3] - zb[4, 3]
    assert tab[r, 6] == grids["ux"][4, 3] and tab[r, 7] == grids["uy"][4, 3]
    assert tab[r, 8] == grids["qx"][4, 3] * dxy and tab[r, 12] == grids["soil_mass"][4, 3]
    (tmp_path / "empty.txt").write_text("")
    with pytest.raises(capi.FxError):
        capi.read_results_file(tmp_path / "empty.txt")


def test_find_last_step(lib, tmp_path):
    assert capi.find_last_step(tmp_path / "missing") == 0
    for name in ("0.txt", "3600.txt", "86400.txt", "7200.txt", "notes", "1234567.txt"):
        (tmp_path / name).write_text("x")
    # atoi of the first six characters: "1234567.txt" reads as 123456
    assert capi.find_last_step(tmp_path) == 123456
    (tmp_path / "1234567.txt").unlink()
    assert capi.find_last_step(tmp_path) == 86400


@pytest.mark.parametrize("ade", [True, False])
def test_restart_initiation_matches_oracle(ade):
    """fx_initiation_restart (host product code) against the oracle's
    restatement of initiate.cpp's RESTART branch, bit for bit, on a DEM with
    a NODATA ring and a Results table that lists a subset of cells."""
    from oracle.oracle import OracleDomain
    sc = gen.scenario_c1(24, seed=7)
    rng = np.random.default_rng(3)
    nr, nc = sc.nrows, sc.ncols
    cells = [(i, j) for j in range(1, nc + 1) for i in range(1, nr + 1) if rng.random() < 0.3]
    tab = np.zeros((len(cells), 16))
    for a, (i, j) in enumerate(cells):
        tab[a, :2] = i, j
        tab[a, 4:] = rng.normal(size=12)
        tab[a, 4] += 500.0
    zb, ks, weir = capi.read_geo(sc.dem, -99999.0, sc.hdry)
    got = capi.initiation_restart(zb, tab, ade_solver=ade, soil_conc_bckgrd=2.5)
    o = OracleDomain(nr, nc, sc.dxy, sc.hdry, 1, [0.01], ade)
    o.read_geo(sc.dem, sc.hdry)
    o.set_wintra(True, 0.0, 2.5)
    o.initiation_restart(tab)
    pairs = {"zb": o.zb, "z": o.z, "h": o.h, "qx": o.qx, "qy": o.qy, "us": o.us, "ux": o.ux, "uy": o.uy,
             "conc0": o.conc[0], "soil_mass": o.soil, "twet": o.twet, "ldry": o.ldry}
    for k, want in pairs.items():
        assert np.array_equal(got[k], want, equal_nan=True), k
    bad = tab.copy()
    bad[0, 1] = nc + 2
    with pytest.raises(capi.FxError):
        capi.initiation_restart(zb, bad)
