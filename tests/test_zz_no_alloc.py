"""fx_mass_balance_noalloc: the water ledger of fx_mass_balance computed
without allocating, in a device array of fixed size.  It shares the volume
kernel with fx_mass_balance and sums the partials in the same order with the
same Kahan steps, so the two must agree bit for bit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen

pytestmark = pytest.mark.gpu


def _noalloc(d: capi.HipDomain):
    """fx_mass_balance_noalloc through the C ABI."""
    lib = capi.load()
    out = (C.c_double * 3)()
    rc = lib.fx_mass_balance_noalloc(d.h, out)
    assert rc == 0, lib.fx_last_error().decode()
    return out[0], out[1], out[2]


def _same_bits(a, b) -> bool:
    return np.array_equal(np.array(a, dtype=np.float64).view(np.uint64), np.array(b, dtype=np.float64).view(np.uint64))


def _wet_domain(n: int, seed: int, precision: int, t: float) -> capi.HipDomain:
    sc = gen.scenario_c1(n, seed=seed)
    d = capi.HipDomain.from_scenario(sc, precision=precision, ade=False)
    d.advance(float(sc.t_end), t)
    return d


@pytest.mark.parametrize("precision", [capi.FP32_MIXED, capi.FP64_PARITY], ids=["fp32", "fp64"])
def test_noalloc_matches_mass_balance_two_calls(precision):
    """Two calls in a row, the second on a smaller domain (the array keeps the
    first call's larger set of partials), each bit-identical to fx_mass_balance;
    the ledger is non-trivial (melt in, water on the ring's weir faces)."""
    big = _wet_domain(120, 3, precision, 2400.0)
    small = _wet_domain(40, 5, precision, 1200.0)
    for d in (big, small):
        ref = d.mass_balance()
        got = _noalloc(d)
        assert ref[0] > 0.0, ref
        if precision == capi.FP32_MIXED:          # the ledger is kept by FP32-mixed domains
            assert ref[1] > 0.0, ref
        assert _same_bits(ref, got), (ref, got)
    # the same domain again after more steps: the array's old contents are not read
    big.advance(1e30, 1e30, 25)
    assert _same_bits(big.mass_balance(), _noalloc(big))
    big.close()
    small.close()


def test_noalloc_fp32_tiling_and_pending_weir():
    """An FP32-mixed domain whose last step left weir-face volume in the sweep
    partials (not yet folded by the clock), under two tile shapes."""
    d = _wet_domain(96, 7, capi.FP32_MIXED, 3000.0)
    for ch, ach in ((8, 16), (32, 64)):
        d.set_tiling(ch, ach)
        d.advance(1e30, 1e30, 7)
        ref = d.mass_balance()
        assert ref[2] != 0.0, ref
        assert _same_bits(ref, _noalloc(d)), (ch, ach)
    d.close()
