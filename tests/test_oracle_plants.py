"""Each pin of the oracle catches a plausible mistake.

The oracle is rebuilt with one planted mistake at a time (the ORC_PLANT_*
switches in oracle/fluxos_oracle.c: a halved or doubled coefficient, a flipped
sign, a shifted index, a wrong table row, a dropped term) and the pin that
covers that part of the step must fail on it.  A pin that passes on a planted
mistake pins nothing.
"""
from __future__ import annotations

import pytest

from oracle import oracle
import tests.test_oracle_pins as pins

PLANTS = [
    ("PRESSURE_HALF", pins.check_lake_at_rest),   # 1/2 g h^2 -> g h^2
    ("SMOOTH", pins.check_lake_at_rest),          # 0.8/0.1 face smoothing -> 0.9
    ("RITTER", pins.check_ritter),                # 8/27 -> 0.3
    ("DIFF", pins.check_ade_diffusion),           # sigc 0.5 -> 1.0
    ("CFL_V", pins.check_cfl),                    # |v| replaced by |u|
    ("MELT_ROW", pins.check_forcing),             # "first row > t" -> ">= t"
    ("MELT_ROW", pins.check_inflow),              # the same lookup serves the inflow table
    ("TURB", pins.check_turbulent_shear),         # sign of dv/dx in the stress
    ("ROE_SIGN", pins.check_roe_step),            # sign of the Roe mass dissipation
    ("LOOP3_INDEX", pins.check_water_conservation),  # west neighbour read as east
    ("ADE_CARRY", pins.check_ade_conservation),   # south-face solute flux carried between columns dropped
    ("QUICK", pins.check_ade_advection),          # upwind-biased face weights 3/4, 3/8 -> 7/8, 1/4
    ("DRY_BED", pins.check_dry_face_pressure),    # face bed = the dry cell's own bed, not the mean
    ("CONV", pins.check_uniform_flow),            # x-momentum carried with v instead of u
    ("WINTRA_AREA", pins.check_wintra),           # released mass divided by h only, not h dxy^2
    ("RESTART_QX", pins.check_restart),           # q*dxy read back divided by dxy
    ("ROE_ABS", pins.check_roe_supercritical),    # |c2| - |c3| in the mass dissipation taken without the absolute values
    ("ROE_SHEAR", pins.check_roe_shear_wave),     # the shear wave's speed |u| taken as u
    ("DZ_LIMIT", pins.check_roe_dz_limiter),      # east-face step not capped at the depth difference
    ("STALE", pins.check_stale_fluxes_persist),   # fluxes of an all-dry neighbourhood zeroed
    ("ADE_CLAMP", pins.check_ade_clamp),          # new concentration not clamped to the neighbour bounds
    ("ROE_A21", pins.check_roe_matrix),           # (c3a - c2a) written as (c2a - c3a) in the momentum row
    ("ADE_LIMIT", pins.check_ade_outflow_split),  # outflow of both faces not limited to the held solute
    ("ADE_LIMIT1", pins.check_ade_single_outflow),  # outflow of a single face not limited
    ("ADE_BILAN", pins.check_ade_emptied_cell),   # outflows of an emptied cell not cancelled
    ("ADE_PASS", pins.check_ade_pass_through),    # an emptied cell's east outflow not limited to its inflow
    ("ADE_NEG", pins.check_ade_overdrained_cell),  # an overdrained cell's east outflow not cancelled
]


@pytest.fixture(scope="module")
def plant_libs():
    return {name: str(oracle.build(plant=name)) for name, _ in PLANTS}


@pytest.mark.parametrize("plant,check", PLANTS, ids=[p for p, _ in PLANTS])
def test_pin_fails_on_planted_mistake(plant_libs, plant, check):
    check()                                   # the real oracle passes
    with pytest.raises(AssertionError):
        check(lib=plant_libs[plant])
