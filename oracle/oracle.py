"""ctypes front end of the CPU oracle (liborc_fluxos.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.

``OracleDomain`` owns every grid of the reference's ``GlobVar``
(GlobVar.h:31-73) as a Fortran-ordered numpy array of shape
(NROWS+2, NCOLS+2), i.e. the reference's column-major memory, and exposes the
restated functions one to one (see fluxos_oracle.h for the map).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBPATH = HERE / "liborc_fluxos.so"
MAX_CHEM = 8

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)


class OrcDomain(C.Structure):
    _fields_ = [
        ("nrows", C.c_int32), ("ncols", C.c_int32), ("mrows", C.c_int32), ("mcols", C.c_int32),
        ("dxy", C.c_int64), ("arbase", C.c_int64), ("hdry", C.c_double),
        ("gacc", C.c_float), ("cfl", C.c_float), ("cvdef", C.c_float), ("nuem", C.c_float),
        ("nodata", C.c_double), ("nchem", C.c_int32), ("ade_solver", C.c_int32),
        ("d_coef", C.c_double * MAX_CHEM), ("dtfl", C.c_double), ("tim", C.c_double),
        ("z", _D), ("zb", _D), ("h", _D), ("qx", _D), ("qy", _D), ("us", _D), ("ks", _D),
        ("fe1", _D), ("fe2", _D), ("fe3", _D), ("fn1", _D), ("fn2", _D), ("fn3", _D),
        ("dh", _D), ("dqx", _D), ("dqy", _D), ("qxf", _D), ("qyf", _D), ("twet", _D), ("h0", _D),
        ("ldry", _F), ("weir", _F), ("ldry_prev", _F),
        ("conc", _D * MAX_CHEM),
        ("meteo_nrows", C.c_int32), ("meteo_ncols", C.c_int32), ("meteo", _F), ("melt_mult", _F),
        ("inflow_nrows", C.c_int32), ("inflow_ncols", C.c_int32), ("inflow", _F),
        ("inflow_row", C.c_int64), ("inflow_col", C.c_int64),
        ("ub_volrat", C.c_int64),
        ("soil", _D), ("ux", _D), ("uy", _D),
        ("wintra", C.c_int32), ("pad_", C.c_int32),
        ("soil_rate", C.c_double), ("soil_bckgrd", C.c_double),
    ]


DOUBLE_GRIDS = ("z", "zb", "h", "qx", "qy", "us", "ks", "fe1", "fe2", "fe3", "fn1", "fn2", "fn3",
                "dh", "dqx", "dqy", "qxf", "qyf", "twet", "h0", "soil", "ux", "uy")
FLOAT_GRIDS = ("ldry", "weir", "ldry_prev")


def build(lib_path: Path | None = None, plant: str | None = None) -> Path:
    """Compile the oracle (optionally with a planted mistake ORC_PLANT_<plant>)."""
    if plant is None:
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        return LIBPATH
    subprocess.run(["make", "-s", "-C", str(HERE), "plant", f"P={plant}"], check=True)
    return HERE / "_plants" / f"liborc_{plant}.so"


_LIBS: dict[str, C.CDLL] = {}


def load(path: Path | None = None) -> C.CDLL:
    path = Path(path or LIBPATH)
    key = str(path)
    if key in _LIBS:
        return _LIBS[key]
    if not path.exists():
        build()
    lib = C.CDLL(str(path))
    P = C.POINTER(OrcDomain)
    lib.orc_courant.argtypes = [P, C.c_double]
    lib.orc_courant.restype = C.c_double
    lib.orc_add_meteo.argtypes = [P]
    lib.orc_add_inflow.argtypes = [P]
    lib.orc_hydrodynamics.argtypes = [P]
    lib.orc_hydrodynamics.restype = None
    lib.orc_solver_dry.argtypes = [P, C.c_int, C.c_int]
    lib.orc_solver_dry.restype = None
    lib.orc_solver_wet.argtypes = [P, C.c_int, C.c_int]
    lib.orc_solver_wet.restype = None
    lib.orc_adesolver.argtypes = [P, C.c_int, C.c_int, _D]
    lib.orc_adesolver.restype = None
    lib.orc_read_geo.argtypes = [P, _F, C.c_double]
    lib.orc_read_geo.restype = None
    lib.orc_initiation.argtypes = [P]
    lib.orc_initiation.restype = None
    lib.orc_initiation_restart.argtypes = [P, _D, C.c_int64]
    lib.orc_initiation_restart.restype = C.c_int
    lib.orc_wintra.argtypes = [P, C.c_int]
    lib.orc_wintra.restype = None
    lib.orc_step.argtypes = [P, C.c_double, C.POINTER(C.c_int32), _D]
    lib.orc_num_threads.restype = C.c_int
    lib.orc_step_timed.argtypes = [P, C.c_double, C.POINTER(C.c_int32), _D, _D]
    lib.orc_set_threads.argtypes = [C.c_int]
    lib.orc_set_threads.restype = None
    lib.orc_proc_bind.restype = C.c_int
    _LIBS[key] = lib
    return lib


class OracleDomain:
    """All grids of one domain plus the reference's scalars."""

    def __init__(self, nrows: int, ncols: int, dxy: int, hdry: float, nchem: int = 1,
                 d_coef=(0.01,), ade: bool = True, nodata: float = -99999.0,
                 lib_path: Path | None = None):
        self.lib = load(lib_path)
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.mrows, self.mcols = self.nrows + 2, self.ncols + 2
        shape = (self.mrows, self.mcols)
        self.g = {k: np.zeros(shape, dtype=np.float64, order="F") for k in DOUBLE_GRIDS}
        self.g.update({k: np.zeros(shape, dtype=np.float32, order="F") for k in FLOAT_GRIDS})
        self.nchem = int(nchem)
        self.conc = [np.zeros(shape, dtype=np.float64, order="F") for _ in range(self.nchem)]
        self.s = OrcDomain()
        s = self.s
        s.nrows, s.ncols, s.mrows, s.mcols = self.nrows, self.ncols, self.mrows, self.mcols
        s.dxy = int(dxy)
        s.arbase = int(dxy) * int(dxy)
        s.hdry = float(hdry)
        s.gacc, s.cfl, s.cvdef, s.nuem = 9.80665, 1.0, 0.07, 1.793e-6
        s.nodata = float(nodata)
        s.nchem = self.nchem
        s.ade_solver = 1 if ade else 0
        for k, v in enumerate(d_coef):
            s.d_coef[k] = float(v)
        s.dtfl, s.tim = 0.0, 0.0
        for k in DOUBLE_GRIDS:
            setattr(s, k, self.g[k].ctypes.data_as(_D))
        for k in FLOAT_GRIDS:
            setattr(s, k, self.g[k].ctypes.data_as(_F))
        for k in range(self.nchem):
            s.conc[k] = self.conc[k].ctypes.data_as(_D)
        self._meteo = None
        self._mult = None
        self._inflow = None
        self.it = C.c_int32(0)
        self._scratch = np.zeros(3 * self.mrows * self.mcols, dtype=np.float64)

    # -- grids as attributes -------------------------------------------------
    def __getattr__(self, name):
        g = self.__dict__.get("g")
        if g is not None and name in g:
            return g[name]
        raise AttributeError(name)

    @property
    def tim(self) -> float:
        return self.s.tim

    @tim.setter
    def tim(self, v: float) -> None:
        self.s.tim = float(v)

    @property
    def dtfl(self) -> float:
        return self.s.dtfl

    @dtfl.setter
    def dtfl(self, v: float) -> None:
        self.s.dtfl = float(v)

    # -- forcing -----------------------------------------------------------------
    def set_meteo(self, table: np.ndarray | None, melt_mult: np.ndarray | None = None) -> None:
        if table is None:
            self._meteo = None
            self.s.meteo = _F()
            self.s.meteo_nrows = 0
        else:
            self._meteo = np.ascontiguousarray(table, dtype=np.float32)
            self.s.meteo = self._meteo.ctypes.data_as(_F)
            self.s.meteo_nrows, self.s.meteo_ncols = self._meteo.shape
        if melt_mult is None:
            self._mult = None
            self.s.melt_mult = _F()
        else:
            self._mult = np.asfortranarray(melt_mult, dtype=np.float32)
            self.s.melt_mult = self._mult.ctypes.data_as(_F)

    def set_inflow(self, table: np.ndarray | None, irow: int = 0, icol: int = 0) -> None:
        if table is None:
            self._inflow = None
            self.s.inflow = _F()
            return
        self._inflow = np.ascontiguousarray(table, dtype=np.float32)
        self.s.inflow = self._inflow.ctypes.data_as(_F)
        self.s.inflow_nrows, self.s.inflow_ncols = self._inflow.shape
        self.s.inflow_row, self.s.inflow_col = int(irow), int(icol)

    # -- restated reference functions ----------------------------------------------
    def read_geo(self, dem: np.ndarray, ks_input: float) -> None:
        dem = np.ascontiguousarray(dem, dtype=np.float32)
        assert dem.shape == (self.nrows, self.ncols)
        self.lib.orc_read_geo(C.byref(self.s), dem.ctypes.data_as(_F), float(ks_input))

    def initiation(self) -> None:
        self.lib.orc_initiation(C.byref(self.s))

    def initiation_restart(self, table: np.ndarray) -> None:
        """RESTART branch of initiation; table = the Results file's data rows (n x 16)."""
        t = np.ascontiguousarray(table, dtype=np.float64).reshape(-1, 16)
        if self.lib.orc_initiation_restart(C.byref(self.s), t.ctypes.data_as(_D), t.shape[0]) != 0:
            raise ValueError("restart row outside the grid")

    def set_wintra(self, on: bool, soil_rate: float, soil_bckgrd: float = 0.0) -> None:
        self.s.wintra = 1 if on else 0
        self.s.soil_rate = float(soil_rate)
        self.s.soil_bckgrd = float(soil_bckgrd)

    def wintra(self, ichem: int = 0) -> None:
        self.lib.orc_wintra(C.byref(self.s), int(ichem))

    def courant(self, print_next: float) -> float:
        return self.lib.orc_courant(C.byref(self.s), float(print_next))

    def add_meteo(self) -> None:
        self.lib.orc_add_meteo(C.byref(self.s))

    def add_inflow(self) -> int:
        return self.lib.orc_add_inflow(C.byref(self.s))

    def hydrodynamics(self) -> None:
        self.lib.orc_hydrodynamics(C.byref(self.s))

    def solver_dry(self, irow: int, icol: int) -> None:
        self.lib.orc_solver_dry(C.byref(self.s), irow, icol)

    def solver_wet(self, irow: int, icol: int) -> None:
        self.lib.orc_solver_wet(C.byref(self.s), irow, icol)

    def adesolver(self, it: int, ichem: int) -> None:
        self.lib.orc_adesolver(C.byref(self.s), int(it), int(ichem), self._scratch.ctypes.data_as(_D))

    def step(self, print_next: float) -> tuple[int, float]:
        hp = C.c_double(0.0)
        flags = self.lib.orc_step(C.byref(self.s), float(print_next), C.byref(self.it), C.byref(hp))
        return flags, hp.value

    def step_timed(self, print_next: float, phase_s: np.ndarray) -> tuple[int, float]:
        """step() with the wall time of CFL, forcing, hydrodynamics and ADE
        added to phase_s[0..3] (a float64 array of 4)."""
        assert phase_s.dtype == np.float64 and phase_s.size == 4 and phase_s.flags.c_contiguous
        hp = C.c_double(0.0)
        flags = self.lib.orc_step_timed(C.byref(self.s), float(print_next), C.byref(self.it), C.byref(hp),
                                        phase_s.ctypes.data_as(_D))
        return flags, hp.value

    @property
    def ub_volrat(self) -> int:
        return int(self.s.ub_volrat)

    # -- convenience ---------------------------------------------------------------
    @classmethod
    def from_scenario(cls, sc, ade: bool = True, lib_path: Path | None = None) -> "OracleDomain":
        """read_geo + initiation of a gen.Scenario (main.cpp:145-262 for the parts that
        touch the grids); forcing installed."""
        d = cls(sc.nrows, sc.ncols, sc.dxy, sc.hdry, sc.nchem, sc.d_coef, ade, lib_path=lib_path)
        d.read_geo(sc.dem, sc.hdry)
        d.initiation()
        mult = None
        if sc.melt_mult is not None:
            mult = np.zeros((d.mrows, d.mcols), dtype=np.float32, order="F")
            mult[1:-1, 1:-1] = np.flipud(sc.melt_mult)
        d.set_meteo(sc.meteo, mult)
        return d

    def wet_count(self) -> int:
        return int((self.g["ldry"][1:-1, 1:-1] == 0).sum())

    def copy_state_from(self, other: "OracleDomain") -> None:
        for k in self.g:
            self.g[k][...] = other.g[k]
        for k in range(self.nchem):
            self.conc[k][...] = other.conc[k]
        self.s.tim, self.s.dtfl = other.s.tim, other.s.dtfl
        self.it.value = other.it.value


def run_reference_loop(d: OracleDomain, t_end: float, print_step: int, timstart: int = 0,
                       on_print=None, max_steps: int | None = None):
    """The time loop of main.cpp:271-470 over the oracle; returns the dt sequence
    of every iteration and the solver flag of each."""
    print_next = timstart + print_step                # main.cpp:260-261
    dts, solver = [], []
    n = 0
    while d.tim <= t_end:                             # main.cpp:271
        flags, _ = d.step(print_next)
        dts.append(d.dtfl)
        solver.append(flags & 1)
        if d.tim >= print_next:                       # main.cpp:444
            if on_print:
                on_print(print_next, d)
            print_next = print_next + print_step
        n += 1
        if max_steps is not None and n >= max_steps:
            break
    return np.array(dts), np.array(solver)
