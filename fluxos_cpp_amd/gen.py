"""Synthetic inputs for FLUXOS-OVERLAND runs: DEMs, forcing tables, modset files.

The benchmark workloads are synthetic (no real basin data ships with the
reference: its example DEMs are stripped).  Everything here is seeded and
deterministic (numpy PCG64).  DEMs are returned in ESRI-ASCII *file order*
(first line = northernmost row), as ``float32`` like the reference's reader
(``read_functions.cpp:118,155``).  Every generated DEM has a one-cell ring of
NODATA = -99999 so that the reference's weir boundary is well defined
(``solver_wetdomain.cpp:300-307`` reads an uninitialised ``volrat`` for
non-weir boundary cells) and so that the vertex-to-centre average of
``initiate.cpp:69-84``, which only skips values of magnitude 99999, never mixes
the NODATA value into a bed elevation.  The base elevation is 500 m (an
elevation of exactly 0.0 is NODATA to the reader, ``read_functions.cpp:182``).

Workloads (sizes as stated for each):

* ``c1``: 256x256, dx=2 m, 0.5 % plane + 1.5 m Gaussian depression
  (sigma 40 cells) + U(-0.05, 0.05) noise; 20 mm/day melt, 1 solute at 1 mg/L,
  D=0.01 m2/s, 2 simulated hours.
* ``c2``: 4096x4096, dx=2 m, diamond-square fractal (roughness 0.6, relief 8 m)
  over a 0.3 % slope; melt ramping 0 -> 60 mm/day over 6 h in 60 s rows; two
  solutes (1 and 2 mg/L); 6 simulated hours.
* ``c3``: ensemble member ``i`` of 64: side log-uniform in [256, 2048], fractal
  DEM with seed 1000+i, melt U(5, 80) mm/day, 1 simulated hour.
* ``c5``: 16384x16384, dx=1 m, fractal + carved dendritic channels, smooth melt
  multiplier field, four solutes; 3 simulated hours.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

NODATA = -99999.0
BASE_ELEV = 500.0
HDRY = 1e-4            # ROUGNESS_HEIGHT (also the wet/dry threshold, main.cpp:259)
PRINT_STEP = 600


@dataclass
class Scenario:
    """One run: DEM (file order), forcing table and the run's scalars."""
    name: str
    dem: np.ndarray                  # float32 (NROWS, NCOLS), file order
    dxy: int
    meteo: np.ndarray                # float32 (rows, 2+nchem), row 0 = empty header slot
    nchem: int
    d_coef: list[float]
    t_end: int                       # ntim: last forcing time [s]
    seed: int
    hdry: float = HDRY
    print_step: int = PRINT_STEP
    melt_mult: np.ndarray | None = None   # per-cell multiplier in file order, or None
    manning_n: object = None         # recorded, not consumed by the reference (no friction)
    extra: dict = field(default_factory=dict)

    @property
    def nrows(self) -> int:
        return int(self.dem.shape[0])

    @property
    def ncols(self) -> int:
        return int(self.dem.shape[1])


# ---------------------------------------------------------------------------
# terrain
# ---------------------------------------------------------------------------
def _ring(dem: np.ndarray) -> np.ndarray:
    dem[0, :] = NODATA
    dem[-1, :] = NODATA
    dem[:, 0] = NODATA
    dem[:, -1] = NODATA
    return dem


def diamond_square(n: int, roughness: float, rng: np.random.Generator) -> np.ndarray:
    """Fractal surface on a (2^k+1)^2 lattice covering n x n, scaled to [0, 1].

    Midpoint displacement: the displacement amplitude shrinks by 2^-roughness
    per level (roughness plays the role of the Hurst exponent H).  Each square
    step averages the in-range ones of its four neighbours (up, down, left,
    right, summed in that order) and takes its random displacements in
    row-major order of the points it sets; only those points are touched."""
    k = max(1, math.ceil(math.log2(max(n - 1, 1))))
    m = 2 ** k + 1
    a = np.zeros((m, m), dtype=np.float64)
    a[0, 0], a[0, -1], a[-1, 0], a[-1, -1] = rng.uniform(-1, 1, 4)
    step = m - 1
    amp = 1.0
    scale = 2.0 ** (-roughness)
    while step > 1:
        h = step // 2
        # diamond: centres of squares
        c = 0.25 * (a[0:-1:step, 0:-1:step] + a[0:-1:step, step::step]
                    + a[step::step, 0:-1:step] + a[step::step, step::step])
        a[h::step, h::step] = c + amp * rng.uniform(-1, 1, c.shape)
        # square: edge midpoints.  Type A: rows 0::step, cols h::step; type B:
        # rows h::step, cols 0::step.  Row-major order interleaves them by row.
        ra, ca = np.arange(0, m, step), np.arange(h, m, step)
        rb, cb = np.arange(h, m, step), np.arange(0, m, step)
        centre = a[h::step, h::step]                  # set by this level's diamond step
        # type A: up (rows after the first), down (rows before the last), left, right
        sa = np.zeros((len(ra), len(ca)))
        na_ = np.zeros_like(sa)
        sa[1:] += centre
        na_[1:] += 1
        sa[:-1] += centre
        na_[:-1] += 1
        sa += a[0::step, 0:m - 1:step]
        sa += a[0::step, step::step]
        na_ += 2
        # type B: up, down, left (columns after the first), right (before the last)
        sb = np.zeros((len(rb), len(cb)))
        nb_ = np.full_like(sb, 2.0)
        sb += a[0:m - 1:step, 0::step]
        sb += a[step::step, 0::step]
        sb[:, 1:] += centre
        nb_[:, 1:] += 1
        sb[:, :-1] += centre
        nb_[:, :-1] += 1
        va, vb = sa / na_, sb / nb_
        na, nb = len(ca), len(cb)
        r = rng.uniform(-1, 1, len(ra) * na + len(rb) * nb)
        r = np.concatenate([r, np.zeros(nb)]).reshape(len(ra), na + nb)
        a[0::step, h::step] = va + amp * r[:, :na]
        a[h::step, 0::step] = vb + amp * r[:-1, na:]
        step = h
        amp *= scale
    a = a[:n, :n]
    a -= a.min()
    mx = a.max()
    return a / mx if mx > 0 else a


def dem_c1(n: int = 256, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    irow = np.arange(n, 0, -1, dtype=np.float64)[:, None]        # file line 0 = irow n
    icol = np.arange(1, n + 1, dtype=np.float64)[None, :]
    plane = BASE_ELEV + 0.005 * 2.0 * irow + 0.0 * icol
    cy, cx = (n + 1) / 2.0, (n + 1) / 2.0
    r2 = (irow - cy) ** 2 + (icol - cx) ** 2
    dem = plane - 1.5 * np.exp(-r2 / (2 * 40.0 ** 2)) + rng.uniform(-0.05, 0.05, (n, n))
    return _ring(dem.astype(np.float32))


def dem_fractal(nrows: int, ncols: int, dxy: float, seed: int, roughness: float = 0.6,
                relief: float = 8.0, slope: float = 0.003) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = max(nrows, ncols)
    f = diamond_square(n, roughness, rng)[:nrows, :ncols] * relief
    irow = np.arange(nrows, 0, -1, dtype=np.float64)[:, None]
    dem = BASE_ELEV + f + slope * dxy * irow
    return _ring(dem.astype(np.float32))


def carve_channels(dem: np.ndarray, dxy: float, seed: int, n_roots: int = 16) -> np.ndarray:
    """Dendritic channel network carved 0.5-2 m deep by seeded random walks
    that run downslope (towards the southern edge) and branch."""
    rng = np.random.default_rng(seed)
    nr, nc = dem.shape
    depth = np.zeros(dem.shape, dtype=np.float32)
    walkers = [(int(rng.integers(1, nr // 4)), int(rng.integers(2, nc - 2)), 2.0)
               for _ in range(n_roots)]
    steps = 0
    while walkers and steps < 50 * nr:
        steps += 1
        nxt = []
        for (r, c, d) in walkers:
            if r >= nr - 2 or d < 0.5:
                continue
            depth[r, c] = max(depth[r, c], d)
            r2 = r + 1
            c2 = min(nc - 2, max(1, c + int(rng.integers(-1, 2))))
            nxt.append((r2, c2, d * 0.9995))
            if rng.random() < 0.002 and d > 0.8:
                nxt.append((r2, min(nc - 2, c2 + 2), d * 0.7))
        walkers = nxt[: 4096]
    out = dem.copy()
    inner = out != NODATA
    out[inner] = out[inner] - depth[inner]
    return out


def smooth_field(nrows: int, ncols: int, seed: int, lo: float, hi: float, cells: int = 64) -> np.ndarray:
    """Seeded smooth field in [lo, hi]: bilinear upsampling of a coarse noise grid."""
    rng = np.random.default_rng(seed)
    cr, cc = nrows // cells + 2, ncols // cells + 2
    coarse = rng.uniform(0, 1, (cr, cc))
    y = np.linspace(0, cr - 1.001, nrows)
    x = np.linspace(0, cc - 1.001, ncols)
    y0, x0 = y.astype(int), x.astype(int)
    fy, fx = (y - y0)[:, None], (x - x0)[None, :]
    f = (coarse[y0][:, x0] * (1 - fy) * (1 - fx) + coarse[y0 + 1][:, x0] * fy * (1 - fx)
         + coarse[y0][:, x0 + 1] * (1 - fy) * fx + coarse[y0 + 1][:, x0 + 1] * fy * fx)
    return (lo + (hi - lo) * f).astype(np.float32)


# ---------------------------------------------------------------------------
# forcing tables (read_meteo layout: row 0 empty, data from row 1)
# ---------------------------------------------------------------------------
def meteo_table(times, rates, concs) -> np.ndarray:
    times = np.asarray(times, dtype=np.float64)
    rates = np.broadcast_to(np.asarray(rates, dtype=np.float64), times.shape)
    concs = np.atleast_1d(np.asarray(concs, dtype=np.float64))
    nchem = concs.shape[-1]
    tab = np.zeros((len(times) + 1, 2 + nchem), dtype=np.float32)
    tab[1:, 0] = times
    tab[1:, 1] = rates
    tab[1:, 2:] = np.broadcast_to(concs, (len(times), nchem))
    return tab


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def scenario_c1(n: int = 256, seed: int = 1, t_end: int = 7200) -> Scenario:
    times = np.arange(0, t_end + 1, 600)
    return Scenario("c1", dem_c1(n, seed), 2, meteo_table(times, 20.0, [1.0]), 1, [0.01],
                    t_end, seed, manning_n=0.035)


def scenario_c2(n: int = 4096, seed: int = 2, t_end: int = 21600) -> Scenario:
    times = np.arange(0, t_end + 1, 60)
    rates = 60.0 * times / t_end
    return Scenario("c2", dem_fractal(n, n, 2.0, seed), 2, meteo_table(times, rates, [1.0, 2.0]),
                    2, [0.01, 0.01], t_end, seed, manning_n=("uniform", 0.02, 0.06, seed + 7))


def scenario_c3(i: int, t_end: int = 3600) -> Scenario:
    seed = 1000 + i
    rng = np.random.default_rng(seed)
    side = int(round(math.exp(rng.uniform(math.log(256), math.log(2048)))))
    rate = float(rng.uniform(5.0, 80.0))
    times = np.arange(0, t_end + 1, 600)
    return Scenario(f"c3_{i:02d}", dem_fractal(side, side, 2.0, seed), 2,
                    meteo_table(times, rate, [1.0]), 1, [0.01], t_end, seed)


def scenario_c4_request(seed: int, t_end: int = 1800) -> Scenario:
    rng = np.random.default_rng(seed)
    side = int(rng.choice([512, 1024, 2048], p=[0.6, 0.3, 0.1]))
    rate = float(rng.uniform(5.0, 80.0))
    times = np.arange(0, t_end + 1, 600)
    return Scenario(f"c4_{seed}", dem_fractal(side, side, 2.0, seed), 2,
                    meteo_table(times, rate, [1.0]), 1, [0.01], t_end, seed)


def scenario_c5(n: int = 16384, seed: int = 5, t_end: int = 10800) -> Scenario:
    dem = carve_channels(dem_fractal(n, n, 1.0, seed), 1.0, seed + 1)
    mult = smooth_field(n, n, seed + 2, 10.0, 60.0)      # mm/day, applied as a multiplier of 1 mm/day
    times = np.arange(0, t_end + 1, 600)
    return Scenario("c5", dem, 1, meteo_table(times, 1.0, [1.0, 1.0, 1.0, 1.0]), 4,
                    [0.001, 0.01, 0.03, 0.1], t_end, seed, melt_mult=mult,
                    manning_n=("channel/other", 0.03, 0.05))


# ---------------------------------------------------------------------------
# files in the reference's formats
# ---------------------------------------------------------------------------
def write_dem(path: Path, dem: np.ndarray, dxy: int, xll: float = 0.0, yll: float = 0.0) -> None:
    nr, nc = dem.shape
    with open(path, "w") as f:
        f.write(f"NCOLS {nc}\nNROWS {nr}\nXLLCORNER {xll}\nYLLCORNER {yll}\n"
                f"CELLSIZE {dxy}\nNODATA_VALUE {int(NODATA)}\n")
        for line in dem:
            f.write(" ".join(f"{v:.9g}" for v in line.tolist()))
            f.write("\n")


def write_meteo(path: Path, tab: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write("t (sec), snowmelt (mm/day)" + "".join(f", conc_{k} (mg/l)" for k in range(tab.shape[1] - 2)) + "\n")
        for row in tab[1:]:
            f.write(",".join(f"{v:.9g}" for v in row.tolist()) + "\n")


def write_modset(path: Path, sc: Scenario, dem_file: str, meteo_file: str, out_dir: str,
                 ade: bool = True) -> None:
    m = {
        "COMMNET": f"synthetic {sc.name} seed {sc.seed}",
        "DEM_FILE": dem_file,
        "SIM_DATETIME_START": "2026DEC01 00:00:00",
        "RESTART": False,
        "METEO_FILE": meteo_file,
        "EXTERNAL_MODULES": {
            "ADE_TRANSPORT": {"STATUS": ade, "D_COEF": sc.d_coef[0]},
            "WINTRA": {"STATUS": False, "SWE_STD": 9.5, "SWE_MAX": 9.5},
            "OPENWQ": {"STATUS": False, "MASTERFILE_DIR": ""},
        },
        "OUTPUT": {"OUTPUT_FOLDER": out_dir, "PRINT_STEP": sc.print_step, "H_MIN_TO_PRINT": 0.005},
        "ROUGNESS_HEIGHT": sc.hdry,
        "SOIL_RELEASE_RATE": 0.0,
        "SOIL_CONC_BACKGROUND": 0.0,
    }
    if sc.nchem > 1:
        m["EXTERNAL_MODULES"]["ADE_TRANSPORT"]["D_COEF_PER_SOLUTE"] = list(sc.d_coef)
        m["EXTERNAL_MODULES"]["ADE_TRANSPORT"]["N_SOLUTES"] = sc.nchem
    with open(path, "w") as f:
        f.write("// synthetic FLUXOS modset (comments are allowed, main.cpp:106-109)\n")
        json.dump(m, f, indent=2)
        f.write("\n")


def write_case(directory: Path, sc: Scenario, ade: bool = True) -> Path:
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    write_dem(directory / f"{sc.name}.asc", sc.dem, sc.dxy)
    write_meteo(directory / f"{sc.name}_meteo.csv", sc.meteo)
    modset = directory / f"{sc.name}_modset.json"
    write_modset(modset, sc, str(directory / f"{sc.name}.asc"), str(directory / f"{sc.name}_meteo.csv"),
                 str(directory / "Results"), ade)
    return modset
