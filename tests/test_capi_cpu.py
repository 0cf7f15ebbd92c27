"""CPU-side checks of the C ABI library: it loads, exports every entry point
that include/fluxos_hip.h declares, and its host input layer (read_geo +
initiation) is bit-identical to the oracle's restatement."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

from fluxos_cpp_amd import capi, gen
from oracle.oracle import OracleDomain

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols() -> list[str]:
    text = (ROOT / "include" / "fluxos_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fx_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol(fxlib):
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(fxlib, s), s
    assert set(syms) == set(capi.EXPORTS)


def test_version_and_error_text(fxlib):
    assert b"gfx950" in fxlib.fx_version()
    assert isinstance(fxlib.fx_last_error(), bytes)


@pytest.mark.parametrize("n,seed", [(16, 1), (40, 7), (64, 2)])
def test_read_geo_and_initiation_match_oracle(n, seed):
    sc = gen.scenario_c1(n, seed=seed)
    o = OracleDomain.from_scenario(sc)
    st = capi.prepare(sc)
    for name, a, b in (("zb", o.zb, st.zb), ("ks", o.ks, st.ks), ("weir", o.weir, st.weir),
                       ("z", o.z, st.z), ("h", o.h, st.h), ("ldry", o.ldry, st.ldry),
                       ("conc0", o.conc[0], st.conc0)):
        same = (a == b) | (np.isnan(a) & np.isnan(b))
        assert same.all(), name


def test_read_geo_rejects_bad_arguments(fxlib):
    with pytest.raises(capi.FxError):
        capi.read_geo(np.zeros((2, 2), dtype=np.float32), -99999.0, 1e-4)


def test_domain_create_without_gpu_fails_loudly():
    """No silent CPU path: without a device, creation reports FX_ERR_NOGPU."""
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.FxError) as e:
        capi.HipDomain(16, 16, 2, 1e-4)
    assert e.value.code == -8


def _desc(n, precision):
    d = capi.DomainDesc()
    d.nrows = d.ncols = n
    d.dxy, d.nchem = 1, 1
    d.hdry, d.nodata = 1e-3, -99999.0
    d.gacc, d.cfl, d.cvdef, d.nuem = 9.80665, 1.0, 0.07, 1.793e-6
    d.d_coef[0] = 0.01
    d.ade_solver = 1
    d.precision = precision
    d.ade_mode = capi.ADE_FAST if precision == capi.FP32_MIXED else capi.ADE_EXACT
    return d


def test_fp32_domain_size_limit(fxlib):
    """The FP32-mixed kernels address cells by 32-bit byte offsets: a domain of
    2^29 cells or more (ghost rows included) is refused before any allocation;
    the FP64 path is not limited by it.  (NROWS + 2)^2 crosses 2^29 between
    NROWS = 23168 and 23169."""
    import ctypes as C
    below, above = _desc(23168, capi.FP32_MIXED), _desc(23169, capi.FP32_MIXED)
    assert (23168 + 2) ** 2 < 1 << 29 <= (23169 + 2) ** 2
    assert fxlib.fx_domain_bytes(C.byref(below)) > 0
    assert fxlib.fx_domain_bytes(C.byref(above)) == -1
    assert b"2^29" in fxlib.fx_last_error()
    assert fxlib.fx_domain_bytes(C.byref(_desc(23169, capi.FP64_PARITY))) > 0
