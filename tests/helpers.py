"""Shared set-up for parity tests: matched oracle / device domains."""
from __future__ import annotations

import numpy as np

from fluxos_cpp_amd import capi, gen
from oracle.oracle import OracleDomain

STATE_D = ("z", "h", "qx", "qy", "fe1", "fe2", "fe3", "fn1", "fn2", "fn3", "twet", "h0")
STATE_F = ("ldry", "ldry_prev")


def matched(sc, ade=True, precision=capi.FP64_PARITY, ade_mode=None):
    """Oracle and device domain started from the same read_geo + initiation."""
    o = OracleDomain.from_scenario(sc, ade=ade)
    init = capi.prepare(sc, ade=ade)
    d = capi.HipDomain.from_scenario(sc, precision=precision, ade=ade, ade_mode=ade_mode, init=init)
    return o, d


def push_oracle_state(o: OracleDomain, d: capi.HipDomain) -> None:
    """Copy every oracle grid the device keeps into the device domain."""
    for k in STATE_D:
        d.set(k, o.g[k])
    for k in STATE_F:
        d.set(k, o.g[k])
    for k in range(o.nchem):
        d.set_conc(k, o.conc[k])
    d.set_clock(o.tim, o.dtfl, o.it.value)


def max_rel(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    den = np.maximum(np.abs(a), np.abs(b))
    diff = np.abs(a - b)
    r = np.where(den > 0, diff / np.where(den > 0, den, 1), 0.0)
    r = np.where(both_nan, 0.0, r)
    r = np.where(np.isnan(a) ^ np.isnan(b), np.inf, r)
    return float(r.max()) if r.size else 0.0


def compare(o: OracleDomain, d: capi.HipDomain, fields=("z", "h", "qx", "qy", "ldry"), conc=True):
    """Max relative difference per field (0.0 means bit-identical up to NaN placement)."""
    out = {}
    for k in fields:
        out[k] = max_rel(o.g[k], d.get(k))
    if conc:
        for k in range(o.nchem):
            out[f"conc{k}"] = max_rel(o.conc[k], d.get_conc(k))
    return out


def random_wet_state(sc, seed: int, wet_frac: float = 0.6, hmax: float = 0.05):
    """An oracle domain with a random mix of wet, shallow and dry cells, random
    discharges and stale fluxes, to exercise every branch of the solvers."""
    rng = np.random.default_rng(seed)
    o = OracleDomain.from_scenario(sc)
    inner = (slice(1, -1), slice(1, -1))
    zb = o.zb
    h = np.where(rng.random(zb.shape) < wet_frac, rng.uniform(0, hmax, zb.shape), 0.0)
    h = np.where(rng.random(zb.shape) < 0.1, rng.uniform(0, 2 * sc.hdry, zb.shape), h)
    z = zb + h
    o.z[inner] = z[inner]
    o.h[inner] = np.maximum(0.0, o.z[inner] - zb[inner])
    wet = o.h > sc.hdry
    o.qx[inner] = np.where(wet[inner], rng.normal(0, 0.01, h[inner].shape), 0.0)
    o.qy[inner] = np.where(wet[inner], rng.normal(0, 0.01, h[inner].shape), 0.0)
    for k in ("fe1", "fe2", "fe3", "fn1", "fn2", "fn3"):
        o.g[k][inner] = np.where(o.weir[inner] == 0, rng.normal(0, 1e-4, h[inner].shape), 0.0)
    o.ldry[inner] = np.where(wet[inner], 0.0, 1.0).astype(np.float32)
    o.ldry_prev[inner] = np.where(rng.random(h[inner].shape) < 0.9, o.ldry[inner], 1 - o.ldry[inner])
    o.h0[inner] = np.maximum(0.0, o.h[inner] * rng.uniform(0.8, 1.2, h[inner].shape))
    for k in range(o.nchem):
        o.conc[k][inner] = rng.uniform(0, 2, h[inner].shape)
    o.dtfl = 0.5
    o.tim = 1000.0
    return o
