"""Pins of the CPU oracle that do not restate its formulas.

The reference ships no tests or golden vectors, only one input file (a melt
record, used by the reader and device tests), and it cannot be built here
(Armadillo is absent), so the oracle is pinned against what the physics and
the reference's documented behaviour fix independently:

* CFL: dt = cfl*dxy/(max(|u|,|v|,1e-6) + sqrt(g h)) -- textbook bound;
* forcing: the "first row later than t" lookup, unit conversion mm/day -> m;
* a dry basin only accumulates melt (no fluxes exist);
* lake at rest: zero mass flux, and the only momentum change is the
  hydrostatic thrust 1/2 g H^2 against the W/S walls (whose faces carry no
  flux) -- times dt/dxy, then the 0.8 centre weight of the face smoothing;
* Ritter's dam-break solution q = 8/27 h sqrt(g h), depth 4/9 h at the
  front, for a lone wet cell and for a dry cell next to a wet one;
* Roe's upwind mass flux -1/2 c dz across a still free-surface step, and
  the eddy-viscosity shear stress nu H dv/dx in a linear shear flow;
* Roe's scheme with moving water: the dissipation is -1/2 |A| dU with the
  flux Jacobian at the face's mean state (|A| by numpy's eigendecomposition),
  pure upwinding of the mass flux in supercritical flow, the shear wave
  damped by 1/2 |u| times its jump, and the cap of the dissipated step at
  the depth difference over a bed step;
* stale fluxes of an all-dry neighbourhood applied again at every step;
* the ADE's new values within their neighbours' bounds under strong
  transport, and its outflow limiter sending out exactly the held solute
  through one face, or split a : b between two, nothing from a cell another
  face has emptied, and from such a cell exactly what it takes in (nothing,
  if it was drained of more than it holds);
* conservation: water in a walled basin, solute under pure transport;
* diffusion: one ADE sweep with still water is the explicit 5-point scheme
  c + 2 D dt / dx^2 * Laplacian(c) (sigc = 0.5), exact on a quadratic;
* read_geo + initiation on hand-worked DEMs, including the two quirks the
  reference has (the corner with all-NODATA vertices becomes NaN, and the
  first-column copy taken before the ghost fill, initiate.cpp:44-51).

tests/test_oracle_plants.py rebuilds the oracle with planted mistakes and
shows each of these checks fails on its mistake.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.oracle import OracleDomain

G = float(np.float32(9.80665))          # gacc is a float (GlobVar.h:78)


def flat_domain(n=12, dxy=2, h=0.0, hdry=1e-4, ade=True, nchem=1, lib=None, bed=500.0,
                ring_bed=None):
    """n x n interior, weir ring, flat bed, depth h everywhere (ring included)."""
    d = OracleDomain(n, n, dxy, hdry, nchem, [0.01] * nchem, ade, lib_path=lib)
    d.zb[...] = bed
    d.weir[...] = 0.0
    d.weir[1, 1:-1] = d.weir[-2, 1:-1] = d.weir[1:-1, 1] = d.weir[1:-1, -2] = 1.0
    if ring_bed is not None:
        ring = d.weir == 1.0
        d.zb[ring] = ring_bed
    d.ks[1:-1, 1:-1] = hdry
    d.z[...] = d.zb + h
    d.h[...] = np.maximum(d.z - d.zb, 0.0)
    d.ldry[1:-1, 1:-1] = np.where(d.h[1:-1, 1:-1] > hdry, 0.0, 1.0)
    return d


# ------------------------------------------------- Roe with moving water ----
def check_roe_supercritical(lib=None):
    """Uniform depth over a flat bed, flowing faster than the wave speed
    (|u| > sqrt(g H)), with a jump in discharge across one face: every
    characteristic speed has the sign of u, so Roe's scheme is pure upwind and
    the face carries the upwind cell's discharge as its mass flux
    (solver_wetdomain.cpp:213-285), along both axes and in both directions."""
    H, n = 0.2, 12
    c = math.sqrt(G * H)
    for axis in (0, 1):
        for sgn in (1.0, -1.0):
            u0, u1 = sgn * 2.6 * c, sgn * 2.9 * c
            d = flat_domain(n=n, h=H, lib=lib, bed=0.0)
            q = d.qx if axis == 0 else d.qy
            q[...] = H * u0
            if axis == 0:
                q[6:, :] = H * u1
            else:
                q[:, 6:] = H * u1
            d.dtfl = 1e-3
            d.hydrodynamics()
            got = d.fe1[5, 3:-3] if axis == 0 else d.fn1[3:-3, 5]
            want = H * (u0 if sgn > 0 else u1)
            assert np.allclose(got, want, rtol=1e-13, atol=0), (axis, sgn, got[:3], want)


def check_roe_shear_wave(lib=None):
    """Uniform depth and normal velocity u0 over a flat bed, with a jump in the
    transverse velocity across one face (a shear wave, speed u0): Roe's
    dissipation of the transverse-momentum flux is 1/2 |u0| times the jump, so
    the face carries H u0 v of the upwind side, less the eddy-viscosity stress
    nu H dv/dx (solver_wetdomain.cpp:177-285), along both axes and in both
    directions."""
    H, n, dxy, v0, v1 = 0.2, 12, 2.0, 0.1, -0.05
    nu = float(np.float32(1.793e-6))
    for axis in (0, 1):
        for u0 in (0.3, -0.3):
            d = flat_domain(n=n, h=H, lib=lib, bed=0.0)
            qn, qt = (d.qx, d.qy) if axis == 0 else (d.qy, d.qx)
            qn[...] = H * u0
            qt[...] = H * v0
            if axis == 0:
                qt[6:, :] = H * v1
            else:
                qt[:, 6:] = H * v1
            d.dtfl = 1e-3
            d.hydrodynamics()
            want = H * u0 * (v0 if u0 > 0 else v1) - nu * H * (v1 - v0) / dxy
            got = d.fe3[5, 3:-3] if axis == 0 else d.fn2[3:-3, 5]
            assert np.allclose(got, want, rtol=1e-12, atol=0), (axis, u0, got[:3], want)


def _abs_matrix(a):
    """|A| = R |Lambda| R^-1 by eigendecomposition (A has real eigenvalues)."""
    w, r = np.linalg.eig(a)
    return (r @ np.diag(np.abs(w)) @ np.linalg.inv(r)).real


def check_roe_matrix(lib=None):
    """Random subcritical and supercritical states that vary along one axis
    over a flat bed, with a uniform transverse velocity (so no eddy stress):
    at every face the Roe dissipation, the face flux less its convective and
    pressure parts, is -1/2 |A| dU, with A the Jacobian of the shallow-water
    flux at the face's mean state (h, q, r averaged) and dU the jump of
    (h, q, r) across the face (solver_wetdomain.cpp:213-285)."""
    n, H = 12, 0.3
    rng = np.random.default_rng(7)
    c = math.sqrt(G * H)
    for axis in (0, 1):
        d = flat_domain(n=n, h=H, lib=lib, bed=0.0)
        hv = H * (1.0 + 0.08 * rng.uniform(-1, 1, n + 2))        # |dh| < h/2: no step cap
        un = c * rng.uniform(-1.6, 1.6, n + 2)                    # normal velocity
        vt = 0.07                                                 # transverse velocity
        hh = hv[:, None] if axis == 0 else hv[None, :]
        uu = un[:, None] if axis == 0 else un[None, :]
        d.h[...] = hh + 0.0 * d.h
        d.z[...] = d.zb + d.h
        (d.qx if axis == 0 else d.qy)[...] = d.h * uu
        (d.qy if axis == 0 else d.qx)[...] = d.h * vt
        d.dtfl = 1e-4
        h0, qn0, qt0 = d.h.copy(), (d.qx if axis == 0 else d.qy).copy(), (d.qy if axis == 0 else d.qx).copy()
        d.hydrodynamics()
        for k in range(3, n - 2):
            p, e = ((k, 5), (k + 1, 5)) if axis == 0 else ((5, k), (5, k + 1))
            hm = 0.5 * (h0[p] + h0[e])
            qm, rm = 0.5 * (qn0[p] + qn0[e]), 0.5 * (qt0[p] + qt0[e])
            um, vm, cm = qm / hm, rm / hm, math.sqrt(G * hm)
            jac = np.array([[0.0, 1.0, 0.0], [cm * cm - um * um, 2 * um, 0.0], [-um * vm, vm, um]])
            want = -0.5 * _abs_matrix(jac) @ np.array([h0[e] - h0[p], qn0[e] - qn0[p], qt0[e] - qt0[p]])
            if axis == 0:
                got = (d.fe1[p] - qm, d.fe2[p] - qm * um - 0.5 * G * hm * hm, d.fe3[p] - qm * vm)
            else:
                got = (d.fn1[p] - qm, d.fn3[p] - qm * um - 0.5 * G * hm * hm, d.fn2[p] - qm * vm)
            assert np.allclose(got, want, rtol=1e-9, atol=1e-14), (axis, k, got, want)


def check_roe_dz_limiter(lib=None):
    """Still water of equal depth H either side of a bed step: the free surface
    steps by the bed step dz.  Roe's still-step mass flux -1/2 c dz applies on
    an east face while |dz| <= H/2; beyond that the reference caps the step it
    dissipates at the depth difference, here zero, and on north faces it always
    does (solver_wetdomain.cpp:216-225, 249-255).  A capped face carries no
    mass and only the hydrostatic thrust 1/2 g H^2."""
    H, n = 0.3, 10
    for axis, dz, capped in ((0, 0.1, False), (0, 0.2, True), (1, 0.1, True)):
        d = flat_domain(n=n, h=H, lib=lib, bed=0.0)
        if axis == 0:
            d.zb[6:-1, :] = dz
        else:
            d.zb[:, 6:-1] = dz
        d.z[...] = d.zb + H
        d.h[...] = H
        d.dtfl = 1e-3
        d.hydrodynamics()
        m = d.fe1[5, 3:-3] if axis == 0 else d.fn1[3:-3, 5]
        p = d.fe2[5, 3:-3] if axis == 0 else d.fn3[3:-3, 5]
        want = 0.0 if capped else -0.5 * math.sqrt(G * H) * dz
        assert np.allclose(m, want, rtol=1e-12, atol=1e-16), (axis, dz, m[:3], want)
        if capped:
            assert np.allclose(p, 0.5 * G * H * H, rtol=1e-12, atol=0), (axis, dz, p[:3])


def check_stale_fluxes_persist(lib=None):
    """A dry cell whose four neighbours are dry keeps the face fluxes it last
    had (solver_drydomain.cpp:60-76), and the update applies them again at
    every step: the water surface of the cell and of its east neighbour drift
    linearly, by -/+ dt fe / dx per step, while the neighbourhood stays dry."""
    n, dxy, F, dt, steps = 10, 2, 1e-6, 0.5, 6
    d = flat_domain(n=n, h=0.0, lib=lib, bed=0.0)
    d.fe1[5, 5] = F
    d.dtfl = dt
    z0 = d.z.copy()
    for k in range(1, steps + 1):
        d.hydrodynamics()
        step = k * dt * F / dxy
        assert d.z[5, 5] == pytest.approx(z0[5, 5] - step, rel=1e-12), k
        assert d.z[6, 5] == pytest.approx(z0[6, 5] + step, rel=1e-12), k
    moved = np.argwhere(d.z != z0)
    assert sorted(map(tuple, moved)) == [(5, 5), (6, 5)]


def check_ade_outflow_split(lib=None):
    """One cell holding solute in still, solute-free water, with east and north
    face discharges a, b that would carry out more than it holds (no
    diffusion): the outflow limiter sends out exactly what it holds, split
    a : b between the two faces (ADEsolver_calc.cpp:244-253), so the east and
    north neighbours end at a/(a+b) and b/(a+b) of its concentration and the
    cell itself at zero."""
    n, cP, a, b = 14, 0.8, 0.3, 0.2
    d = _ade_box(lib, lambda i, j: 0.0 * i, D=0.0, n=n)
    d.conc[0][6, 6] = cP
    d.fe1[6, 6], d.fn1[6, 6] = a, b
    d.ldry_prev[...] = 1.0                     # no rescale: depths do not change
    d.adesolver(2, 0)
    c = d.conc[0]
    assert c[7, 6] == pytest.approx(a / (a + b) * cP, rel=1e-12)
    assert c[6, 7] == pytest.approx(b / (a + b) * cP, rel=1e-12)
    assert abs(c[6, 6]) < 1e-15


def check_ade_single_outflow(lib=None):
    """One cell holding solute in still water (no diffusion) whose east (or
    north) face discharge would carry out more than it holds: the limiter lets
    the face carry exactly the held solute (ADEsolver_calc.cpp:254-257), so the
    receiving neighbour gains the whole of it, c_Q + c_P at equal depths, and
    the cell ends at zero.  A cell beyond the receiver holds more solute, so
    the receiver's own bounds do not cap its value."""
    n, cP, cQ, a = 14, 0.5, 0.1, 0.3
    for axis in (0, 1):
        d = _ade_box(lib, lambda i, j: 0.0 * i, D=0.0, n=n)
        q, qq = ((7, 6), (8, 6)) if axis == 0 else ((6, 7), (6, 8))
        d.conc[0][6, 6], d.conc[0][q], d.conc[0][qq] = cP, cQ, 1.0
        (d.fe1 if axis == 0 else d.fn1)[6, 6] = a
        d.ldry_prev[...] = 1.0                 # no rescale: depths do not change
        d.adesolver(2, 0)
        c = d.conc[0]
        assert c[q] == pytest.approx(cQ + cP, rel=1e-12), axis
        assert abs(c[6, 6]) < 1e-15, axis


def check_ade_emptied_cell(lib=None):
    """Still water (no diffusion); a cell P holding solute, whose west face
    discharge flows west strongly enough to take all of it (the west cell's
    face flux is capped at what P holds), while P's own east face discharge
    points east: P has nothing left to send, so the reference cancels its
    east and north outflows (ADEsolver_calc.cpp:259, 274-276).  The west cell
    ends with P's concentration, P and its east neighbour at zero."""
    n, cP = 14, 0.5
    d = _ade_box(lib, lambda i, j: 0.0 * i, D=0.0, n=n)
    d.conc[0][6, 6] = cP
    d.fe1[5, 6] = -0.5                         # west face of P: flow from P to the west
    d.fe1[6, 6] = 0.3                          # east face of P: flow to the east
    d.ldry_prev[...] = 1.0                     # no rescale: depths do not change
    d.adesolver(2, 0)
    c = d.conc[0]
    assert c[5, 6] == pytest.approx(cP, rel=1e-12)
    assert abs(c[6, 6]) < 1e-15
    assert c[7, 6] == 0.0


def check_ade_pass_through(lib=None):
    """Still water (no diffusion); a cell P holding solute is emptied through
    its west face (or its south face), takes solute in through its north face
    (or its east face) from a neighbour holding some, and its discharge through
    the remaining face points out: P can pass on only what it takes in
    (ADEsolver_calc.cpp:259-273).  So P ends at zero, the neighbour that
    emptied it at P's concentration, and at equal depths the cell downstream of
    P gains exactly what the supplying neighbour lost."""
    n, cP, cS = 14, 0.5, 0.4
    for axis in (0, 1):
        d = _ade_box(lib, lambda i, j: 0.0 * i, D=0.0, n=n)
        d.conc[0][6, 6] = cP
        if axis == 0:      # emptied west, supplied from the north, sends east
            drain, supply, out = (5, 6), (6, 7), (7, 6)
            d.fe1[5, 6], d.fn1[6, 6], d.fe1[6, 6] = -0.5, -0.05, 0.3
        else:              # emptied south, supplied from the east, sends north
            drain, supply, out = (6, 5), (7, 6), (6, 7)
            d.fn1[6, 5], d.fe1[6, 6], d.fn1[6, 6] = -0.5, -0.05, 0.3
        d.conc[0][supply] = cS
        d.ldry_prev[...] = 1.0                 # no rescale: depths do not change
        d.adesolver(2, 0)
        c = d.conc[0]
        assert abs(c[6, 6]) < 1e-15, axis
        assert c[drain] == pytest.approx(cP, rel=1e-12), axis
        assert 0.0 < c[out] == pytest.approx(cS - c[supply], rel=1e-12), (axis, c[out], cS - c[supply])


def check_ade_overdrained_cell(lib=None):
    """Still water (no diffusion); a cell P drained through both its west and
    its south face, each taking all that P holds (each face's flux is capped
    at P's content), so P's balance is negative; a smaller inflow through its
    north face (or its east face) does not make up for it, and P's discharge
    through the remaining face points out: P sends nothing through that face
    (ADEsolver_calc.cpp:264-266, 271-273), so the cell downstream of it keeps
    its zero."""
    n, cP, cS = 14, 0.5, 0.4
    for axis in (0, 1):
        d = _ade_box(lib, lambda i, j: 0.0 * i, D=0.0, n=n)
        d.conc[0][6, 6] = cP
        d.fe1[5, 6], d.fn1[6, 5] = -0.5, -0.5          # west and south faces drain P
        if axis == 0:      # supplied from the north, sends east
            supply, out = (6, 7), (7, 6)
            d.fn1[6, 6], d.fe1[6, 6] = -0.05, 0.3
        else:              # supplied from the east, sends north
            supply, out = (7, 6), (6, 7)
            d.fe1[6, 6], d.fn1[6, 6] = -0.05, 0.3
        d.conc[0][supply] = cS
        d.ldry_prev[...] = 1.0                 # no rescale: depths do not change
        d.adesolver(2, 0)
        assert d.conc[0][out] == 0.0, axis


def check_ade_clamp(lib=None):
    """Strong transport of random sign (face Courant numbers up to 0.75)
    through uniform water carrying a random field: whatever the fluxes, each
    new concentration lies within the smallest and largest of its four
    neighbours' values (ADEsolver_calc.cpp:47-62, 284-285).  No cell changes
    depth (the cells count as dry before the step, so pass 1 leaves them as
    they are), so those values are the step's inputs."""
    n = 14
    rng = np.random.default_rng(5)
    d = _ade_box(lib, lambda i, j: rng.random(i.shape),
                 fe=lambda i, j: rng.uniform(-0.15, 0.15, i.shape),
                 fn=lambda i, j: rng.uniform(-0.15, 0.15, i.shape), n=n)
    d.ldry_prev[...] = 1.0
    c0 = d.conc[0].copy()
    d.adesolver(2, 0)
    c1 = d.conc[0]
    changed = 0
    for iy in range(1, n + 1):
        for ix in range(1, n + 1):
            if d.ldry[ix, iy] != 0.0 or (ix, iy) == (n, n):
                continue                       # dry cells zero; the last cell is not visited
            nb = (c0[ix - 1, iy], c0[ix + 1, iy], c0[ix, iy - 1], c0[ix, iy + 1])
            assert min(nb) <= c1[ix, iy] <= max(nb), (ix, iy, c1[ix, iy], nb)
            changed += c1[ix, iy] != c0[ix, iy]
    assert changed > (n - 2) ** 2 // 2


# ---------------------------------------------------------------- CFL ----
def check_cfl(lib=None):
    d = flat_domain(h=0.3, lib=lib)
    d.qx[5, 6] = 0.12
    d.qy[7, 4] = -0.21
    d.tim = 100.0
    hp = d.courant(print_next=1e6)
    dt_still = 2.0 / (1e-6 + math.sqrt(G * 0.3))
    dt_u = 2.0 / (0.12 / 0.3 + math.sqrt(G * 0.3))
    dt_v = 2.0 / (0.21 / 0.3 + math.sqrt(G * 0.3))
    assert hp == pytest.approx(0.3, rel=1e-14)
    assert d.dtfl == pytest.approx(min(dt_still, dt_u, dt_v), rel=1e-14)
    assert d.tim == pytest.approx(100.0 + d.dtfl, rel=1e-15)
    # print clip (main.cpp:321)
    d2 = flat_domain(h=0.3, lib=lib)
    d2.tim = 599.5
    d2.courant(print_next=600.0)
    assert d2.dtfl == 0.5 and d2.tim == 600.0


def test_cfl_closed_form():
    check_cfl()


def test_cfl_flags_and_copies():
    d = flat_domain(h=0.0)
    d.h[3, 3] = 0.05
    d.ldry[3, 3] = 1.0
    hp = d.courant(600.0)
    assert hp == 0.05
    assert d.ldry[3, 3] == 0.0 and d.ldry_prev[3, 3] == 1.0 and d.h0[3, 3] == 0.05
    assert (d.ldry[1:-1, 1:-1].sum()) == d.nrows * d.ncols - 1


# ------------------------------------------------------------ forcing ----
def check_forcing(lib=None):
    d = flat_domain(h=0.0, lib=lib)
    tab = np.array([[0, 0, 0], [0, 10, 1], [600, 20, 2], [1200, 30, 3]], dtype=np.float32)
    d.set_meteo(tab)
    d.zb[4, 4] = -99999.0                      # a NODATA cell receives nothing
    z0 = d.z.copy()
    d.tim, d.dtfl = 600.0, 60.0                # exactly on a row time: the NEXT row applies
    d.add_meteo()
    melt = 30.0 / (1000.0 * 3600.0 * 24.0) * 60.0
    inner = np.zeros_like(d.z, dtype=bool)
    inner[1:-1, 1:-1] = True
    inner[4, 4] = False
    assert np.allclose(d.z[inner] - z0[inner], melt, rtol=0, atol=1e-13)
    assert d.z[4, 4] == z0[4, 4]
    # past the end of the table: no melt
    d.tim = 5000.0
    z1 = d.z.copy()
    d.add_meteo()
    assert np.array_equal(d.z, z1)


def test_forcing_lookup_and_units():
    check_forcing()


def test_melt_mixing_of_concentration():
    d = flat_domain(h=0.0)
    d.set_meteo(np.array([[0, 0, 0], [3600, 86.4, 5.0]], dtype=np.float32))
    d.z[5, 5] += 0.002                        # one cell already holds 2 mm at c = 1
    d.conc[0][5, 5] = 1.0
    d.tim, d.dtfl = 10.0, 1000.0
    d.add_meteo()
    m = float(np.float32(86.4)) / 86.4e6 * 1000.0   # ~1 mm; the table is FP32 (GlobVar.h:72-73)
    h = 0.002 + m
    assert d.conc[0][5, 5] == pytest.approx((1.0 * 0.002 + m * 5.0) / h, rel=1e-12)
    # a cell that was dry gets water but no solute (common.cpp:226)
    assert d.conc[0][6, 6] == 0.0


def check_inflow(lib=None):
    """Point inflow (common.cpp:277-390): the discharge of the first table row
    later than t, q*dt/dxy^2 of depth into one cell, solute mixed by volume;
    past the end of the table nothing enters and concentrations stay."""
    d = flat_domain(n=8, h=0.1, lib=lib, bed=0.0)
    d.conc[0][...] = 2.0
    tab = np.array([[0.0, 0.0, 0.0], [100.0, 0.5, 5.0], [200.0, 1.0, 7.0]], dtype=np.float32)
    d.set_inflow(tab, 4, 5)
    d.tim, d.dtfl = 150.0, 0.4
    z0 = d.z.copy()
    d.add_inflow()
    dz = 1.0 * 0.4 / 4.0
    assert d.z[4, 5] == pytest.approx(0.1 + dz, rel=1e-15)
    assert d.h[4, 5] == pytest.approx(0.1 + dz, rel=1e-15)
    assert d.conc[0][4, 5] == pytest.approx((2.0 * 0.1 + dz * 7.0) / (0.1 + dz), rel=1e-14)
    other = np.ones_like(z0, dtype=bool)
    other[4, 5] = False
    assert np.array_equal(d.z[other], z0[other])
    d.tim = 100.0                              # row 1 is not later than t: row 2 again
    d.add_inflow()
    assert d.z[4, 5] == pytest.approx(0.1 + 2 * dz, rel=1e-14)
    d.tim = 250.0                              # past the table
    c = d.conc[0][4, 5]
    d.add_inflow()
    assert d.z[4, 5] == pytest.approx(0.1 + 2 * dz, rel=1e-14)
    assert d.conc[0][4, 5] == pytest.approx(c, rel=1e-15)


def test_inflow_point_source_closed_form():
    check_inflow()


def test_dry_basin_only_accumulates_melt():
    """No cell ever exceeds hdry: the solver never runs, dt is the print step,
    and the surface rises by exactly the melt."""
    d = flat_domain(h=0.0, hdry=1.0)
    d.set_meteo(np.array([[0, 0, 0], [3600, 10, 1], [7200, 40, 1]], dtype=np.float32))
    z0 = d.z.copy()
    pn = 600.0
    total = 0.0
    while d.tim <= 7200:
        flags, hp = d.step(pn)
        assert flags & 1 == 0 and hp == 0.0 and d.dtfl == pytest.approx(600.0)
        total += (10.0 if d.tim < 3600 else 40.0 if d.tim < 7200 else 0.0) / 86.4e6 * 600.0
        if d.tim >= pn:
            pn += 600.0
    assert np.allclose(d.z[1:-1, 1:-1] - z0[1:-1, 1:-1], total, atol=1e-12)
    assert np.count_nonzero(d.fe1) == 0


# -------------------------------------------------------- lake at rest ----
def check_lake_at_rest(lib=None):
    H = 0.25
    d = flat_domain(n=10, h=H, lib=lib)
    d.dtfl = 0.3
    d.hydrodynamics()
    dt, dxy = 0.3, 2.0
    assert np.array_equal(d.z, flat_domain(n=10, h=H).z), "still water must not move"
    thrust = -(0.5 * G * H * H) / dxy * dt * 0.8
    qx, qy = d.qx[1:-1, 1:-1], d.qy[1:-1, 1:-1]
    inner = slice(1, -1)          # non-weir rows/cols 2..n-1 -> indices 1..n-2 of the interior
    # first non-weir row (irow = 2) feels the wall to its west; every other cell is balanced
    assert np.allclose(qx[1, 1:-1], thrust, rtol=1e-12, atol=0)
    assert np.allclose(qx[2:-1, 1:-1], 0.0, atol=1e-15)
    assert np.allclose(qy[1:-1, 1], thrust, rtol=1e-12, atol=0)
    assert np.allclose(qy[1:-1, 2:-1], 0.0, atol=1e-15)
    assert np.count_nonzero(d.fe1) == 0 and np.count_nonzero(d.fn1) == 0


def test_lake_at_rest_wall_thrust():
    check_lake_at_rest()


# -------------------------------------------------------------- Ritter ----
def check_ritter(lib=None):
    H, dt, dxy = 0.1, 0.5, 2.0
    d = flat_domain(n=10, h=0.0, lib=lib)
    d.z[5, 5] += H
    d.h[5, 5] = H
    d.dtfl = dt
    d.hydrodynamics()
    q = 0.296 * H * math.sqrt(G * H)               # 8/27 h sqrt(g h) to 3 digits
    assert 0.296 == pytest.approx(8 / 27, abs=5e-4)
    assert d.fe1[5, 5] == pytest.approx(q, rel=1e-13)
    assert d.fn1[5, 5] == pytest.approx(q, rel=1e-13)
    # the dry cell west of the wet one draws the same Ritter discharge
    # inward, with depth 4/9 H at the front (solver_drydomain.cpp:127-138)
    hf = 0.444 * H
    assert d.fe1[4, 5] == pytest.approx(-q, rel=1e-13)
    assert d.fe2[4, 5] == pytest.approx(q * q / hf + 0.5 * G * hf * hf, rel=1e-12)
    assert d.fn1[5, 4] == pytest.approx(-q, rel=1e-13)


def test_ritter_dam_break():
    check_ritter()


def test_outflow_limited_to_available_water():
    """A long step drains the lone wet cell exactly: the limiter caps the
    outflow at arbase*h/dt (solver_wetdomain.cpp:309-332)."""
    H = 0.1
    d = flat_domain(n=10, h=0.0)
    d.z[5, 5] += H
    d.h[5, 5] = H
    d.dtfl = 50.0
    hp = d.z[5, 5] - d.zb[5, 5]               # H as the cell holds it at a 500 m datum
    d.hydrodynamics()
    vol_rate = 4.0 * hp / 50.0
    out = (d.fe1[5, 5] + d.fn1[5, 5]) * 2.0
    assert out == pytest.approx(vol_rate, rel=1e-14)
    assert d.fe1[5, 5] > 0 and d.fn1[5, 5] > 0


# ------------------------------------------------- Roe and turbulence ----
def check_roe_step(lib=None):
    """Still water with a small step in the free surface: with u = 0 the Roe
    matrix |A| is c*I, so the mass flux through the step's face is the
    textbook -1/2 c (z_e - z_p), c = sqrt(g h_face), h_face the mean depth
    (solver_wetdomain.cpp:213-285), along both axes."""
    H1, H2, n = 0.30, 0.28, 10
    for axis in (0, 1):
        d = flat_domain(n=n, h=H1, lib=lib, bed=0.0)
        if axis == 0:
            d.z[6:-1, :] = H2
        else:
            d.z[:, 6:-1] = H2
        d.h[...] = d.z - d.zb
        d.dtfl = 0.01
        d.hydrodynamics()
        hf = 0.5 * (H1 + H2)
        want = 0.5 * math.sqrt(G * hf) * (H1 - H2)
        got = d.fe1[5, 2:-2] if axis == 0 else d.fn1[2:-2, 5]
        assert np.allclose(got, want, rtol=1e-12, atol=0), (axis, got, want)
        # one cell further from the step nothing moves
        away = d.fe1[3, 2:-2] if axis == 0 else d.fn1[2:-2, 3]
        assert np.allclose(away, 0.0, atol=1e-16)


def check_turbulent_shear(lib=None):
    """A lake of depth H at rest in x with a y-velocity that grows linearly
    with x (v = a + s x): the eddy-viscosity stress is nu H dv/dx, carried
    by the y-momentum flux through east faces and, through the symmetric
    off-diagonal term, by the x-momentum flux through north faces
    (solver_wetdomain.cpp:177-211), with the opposite sign of the gradient."""
    H, a, s, n, dxy = 0.2, 0.1, 0.02, 12, 2.0
    d = flat_domain(n=n, h=H, lib=lib, bed=0.0)
    ii = np.arange(n + 2, dtype=float)[:, None] + np.zeros((1, n + 2))
    d.qy[...] = H * (a + s * ii)
    d.dtfl = 0.01
    d.hydrodynamics()
    nu = float(np.float32(1.793e-6))
    want = -nu * H * s / dxy
    assert np.allclose(d.fe3[2:-2, 2:-2], want, rtol=1e-9, atol=0)
    assert np.allclose(d.fn2[2:-2, 2:-2], want, rtol=1e-9, atol=0)


def check_dry_face_pressure(lib=None):
    """A dry cell whose bed stands above the water surface of its wet east /
    north neighbour: no water moves, the face only carries the hydrostatic
    thrust 1/2 g h_f^2 of the water standing against it, h_f = z_e - the mean
    bed of the two cells (solver_drydomain.cpp:121-126, 146-153)."""
    d = flat_domain(n=10, h=0.0, lib=lib, bed=0.0)
    d.zb[5, 5] = 0.2                       # dry cell P
    d.zb[6, 5] = 0.0                       # east neighbour, wet
    d.zb[5, 6] = 0.0                       # north neighbour, wet
    d.z[...] = d.zb
    d.z[6, 5] = 0.15
    d.z[5, 6] = 0.13
    d.h[...] = d.z - d.zb
    d.dtfl = 0.1
    d.hydrodynamics()
    he, hn = 0.15 - 0.1, 0.13 - 0.1
    assert d.fe1[5, 5] == 0.0 and d.fn1[5, 5] == 0.0
    assert d.fe2[5, 5] == pytest.approx(0.5 * G * he * he, rel=1e-12)
    assert d.fn3[5, 5] == pytest.approx(0.5 * G * hn * hn, rel=1e-12)


def check_uniform_flow(lib=None):
    """Uniform depth H moving at u0 over a flat bed: every interior east face
    carries the physical flux (H u0, H u0^2 + 1/2 g H^2, 0), with no Roe
    dissipation (no jumps) and no shear; likewise along y for v0."""
    H, u0, n = 0.2, 0.3, 12
    for axis in (0, 1):
        d = flat_domain(n=n, h=H, lib=lib, bed=0.0)
        (d.qx if axis == 0 else d.qy)[...] = H * u0
        d.dtfl = 0.01
        d.hydrodynamics()
        sl = (slice(2, -3), slice(2, -3))
        if axis == 0:
            m, p, t = d.fe1[sl], d.fe2[sl], d.fe3[sl]
        else:
            m, p, t = d.fn1[sl], d.fn3[sl], d.fn2[sl]
        assert np.allclose(m, H * u0, rtol=1e-13, atol=0)
        assert np.allclose(p, H * u0 * u0 + 0.5 * G * H * H, rtol=1e-12, atol=0)
        assert np.allclose(t, 0.0, atol=1e-15)


def check_ade_advection(lib=None):
    """Uniform east-going discharge q through water of depth H carrying a
    quadratic profile c = a + b i + e i^2: the upwind-biased face value
    -1/8 c(i-1) + 3/4 c(i) + 3/8 c(i+1) is exactly c(i+1/2) for it, so the
    advective change of cell i is -Cr (c(i+1/2) - c(i-1/2)) = -Cr (b + 2 e i)
    with the cell Courant number Cr = q dt / (dx H), plus the diffusion step
    2 D dt / dx^2 * 2e (sigc = 0.5).  The sequential sweep only adds a term of
    second order in Cr; face weights that still sum to one but are wrong leave
    a first-order residue e Cr * O(1)."""
    n, H, q, dt, dx, D, a, b, e = 14, 0.05, 1e-3, 0.5, 2.0, 0.01, 1.0, 0.0, 0.02
    d = _ade_box(lib, lambda i, j: a + b * i + e * i * i + 0.0 * j, H=H, n=n, D=D,
                 fe=lambda i, j: np.where((i >= 1) & (j >= 2) & (j <= n - 1), q, 0.0) + 0.0 * j,
                 fn=lambda i, j: 0.0 * i)
    d.conc[0][d.weir == 1.0] = 100.0       # no wet cell is an extreme of its neighbourhood
    c0 = d.conc[0].copy()
    d.adesolver(2, 0)
    cr = q * dt / (dx * H)
    i = np.arange(n + 2, dtype=float)[:, None] + np.zeros((1, n + 2))
    want = -cr * (b + 2 * e * i) + 2 * D * dt / dx ** 2 * (2 * e)
    sl = (slice(5, -5), slice(4, -4))
    got = (d.conc[0] - c0)[sl]
    # tolerance Cr e / 16: ~50x the sequential sweep's second-order term here,
    # 4x below the residue Cr e / 4 of face weights off by 1/8
    assert np.allclose(got, want[sl], rtol=0, atol=cr * e / 16)


def test_dry_cell_face_pressure():
    check_dry_face_pressure()


def test_uniform_flow_physical_flux():
    check_uniform_flow()


def test_ade_advects_quadratic_profile():
    check_ade_advection()


def test_roe_flux_at_a_free_surface_step():
    check_roe_step()


def test_turbulent_shear_stress():
    check_turbulent_shear()


# -------------------------------------------------------- conservation ----
def check_water_conservation(lib=None, steps=60):
    rng = np.random.default_rng(7)
    n = 16
    d = flat_domain(n=n, h=0.0, lib=lib, ring_bed=600.0)
    d.zb[2:-2, 2:-2] += rng.uniform(-0.2, 0.2, (n - 2, n - 2))
    d.z[...] = d.zb
    d.z[5:9, 5:9] += 0.3
    d.h[...] = np.maximum(d.z - d.zb, 0.0)
    nonweir = (d.weir == 0.0)
    nonweir[0, :] = nonweir[-1, :] = nonweir[:, 0] = nonweir[:, -1] = False
    v0 = d.z[nonweir].sum()
    for _ in range(steps):
        d.courant(1e9)
        d.hydrodynamics()
    v1 = d.z[nonweir].sum()
    assert abs(v1 - v0) <= 1e-11 * abs(v0)


def test_water_conserved_in_walled_basin():
    check_water_conservation()


def _ade_box(lib, pattern, fe=0.0, fn=0.0, H=0.05, n=14, D=0.01, seed=3):
    d = flat_domain(n=n, h=H, lib=lib)
    d.s.d_coef[0] = D
    ring = d.weir == 1.0
    d.ldry[1:-1, 1:-1] = 0.0
    d.ldry[ring] = 1.0
    d.ldry_prev[...] = d.ldry
    d.h0[...] = d.h
    ii, jj = np.meshgrid(np.arange(n + 2), np.arange(n + 2), indexing="ij")
    d.conc[0][...] = pattern(ii.astype(float), jj.astype(float))
    d.conc[0][ring] = 0.0
    d.conc[0][0, :] = d.conc[0][-1, :] = d.conc[0][:, 0] = d.conc[0][:, -1] = 0.0
    if callable(fe):
        d.fe1[...] = fe(ii, jj)
        d.fn1[...] = fn(ii, jj)
    d.dtfl = 0.5
    return d


def check_ade_diffusion(lib=None):
    """Still water, uniform depth: one sweep is the explicit 5-point diffusion
    step with coefficient 2D (sigc = 0.5), exact on a quadratic profile."""
    e, D, dt, dx = 0.01, 0.01, 0.5, 2.0
    d = _ade_box(lib, lambda i, j: 1.0 + 0.1 * i + e * i * i, D=D)
    c0 = d.conc[0].copy()
    d.adesolver(2, 0)
    got = d.conc[0][4:-4, 4:-4]
    want = c0[4:-4, 4:-4] + 2 * D * dt / dx ** 2 * (2 * e)
    assert np.allclose(got, want, rtol=0, atol=1e-13)


def test_ade_diffusion_closed_form():
    check_ade_diffusion()


def test_ade_zero_stays_zero_and_uniform_stays_uniform():
    d = _ade_box(None, lambda i, j: 0.0 * i)
    d.adesolver(2, 0)
    assert np.count_nonzero(d.conc[0]) == 0
    d = _ade_box(None, lambda i, j: 0.7 + 0.0 * i)
    d.adesolver(2, 0)
    assert np.allclose(d.conc[0][3:-3, 3:-3], 0.7, atol=1e-15)


def check_ade_conservation(lib=None):
    """Advection + diffusion in a closed box of still-depth water: the sweep
    moves solute between cells without creating or losing any.  The field is
    monotone and the transport small, so no cell leaves the range of its
    neighbours (the reference clamps to that range, ADEsolver_calc.cpp:284-285,
    which is not conservative at interior extrema)."""
    n = 14

    def ramp(i, j):
        return 1.0 + 0.02 * i + 0.015 * j

    def fe(i, j):
        f = 2e-4 * (1.0 + 0.5 * np.cos(j / 5.0)) + 0.0 * i
        return np.where((i >= 2) & (i <= n - 2) & (j >= 2) & (j <= n - 1), f, 0.0)

    def fn(i, j):
        f = -1.5e-4 * (1.0 + 0.5 * np.sin(i / 4.0)) + 0.0 * j
        return np.where((i >= 2) & (i <= n - 1) & (j >= 2) & (j <= n - 2), f, 0.0)

    d = _ade_box(lib, ramp, fe=fe, fn=fn, n=n)
    # the dry ring holds a high raw value: the E/N bound terms read it before
    # pass 1 zeroes it, so no wet cell is an extreme of its neighbourhood
    d.conc[0][d.weir == 1.0] = 100.0
    wet = d.ldry == 0.0
    m0 = (d.conc[0] * d.h)[wet].sum()
    c0 = d.conc[0].copy()
    d.adesolver(2, 0)
    m1 = (d.conc[0] * d.h)[wet].sum()
    assert np.abs(d.conc[0] - c0)[wet].max() > 1e-6          # something did move
    assert abs(m1 - m0) <= 1e-12 * m0


def test_ade_conserves_solute():
    check_ade_conservation()


def test_ade_skipped_on_first_solver_step():
    d = _ade_box(None, lambda i, j: 1.0 + 0.1 * i)
    c0 = d.conc[0].copy()
    d.adesolver(1, 0)
    assert np.array_equal(d.conc[0], c0)


# ------------------------------------------------ read_geo + initiation ----
def _dem_domain(dem, nodata=-99999.0):
    nr, nc = dem.shape
    d = OracleDomain(nr, nc, 2, 1e-4, 1, [0.01], True, nodata=nodata)
    d.read_geo(dem.astype(np.float32), 1e-4)
    d.initiation()
    return d


def test_read_geo_weir_ring_and_means():
    n = 6
    dem = np.full((n, n), 510.0)
    dem[0, :] = dem[-1, :] = dem[:, 0] = dem[:, -1] = -99999.0
    d = _dem_domain(dem)
    w = d.weir[1:-1, 1:-1]
    assert w[0, :].all() and w[-1, :].all() and w[:, 0].all() and w[:, -1].all()
    assert not w[1:-1, 1:-1].any()
    # interior constant bed survives the vertex averaging
    assert np.all(d.zb[2:-2, 2:-2] == 510.0)
    # the corner whose four vertices are all NODATA: 0/0 (initiate.cpp:85)
    assert math.isnan(d.zb[n, n])
    # (NROWS, 1): its SE vertex is the first-column copy taken before the ghost
    # fill, i.e. 0 (initiate.cpp:44-51): (0 + 510 + 510) / 3
    assert d.zb[n, 1] == pytest.approx(340.0, rel=1e-15)
    # cold start: still water at the bed, everything dry
    assert np.array_equal(d.z[1:-1, 1:-1], d.zb[1:-1, 1:-1], equal_nan=True)
    assert (d.ldry[1:-1, 1:-1] == 1.0).all()


def test_initiation_vertex_to_centre_on_a_ramp():
    """A bed rising by 1 m per row: cell centres sit half a row up."""
    n = 8
    irow = np.arange(n, 0, -1, dtype=float)[:, None]     # file line 0 is irow = n
    dem = np.broadcast_to(500.0 + irow, (n, n)).copy()
    dem[0, :] = dem[-1, :] = dem[:, 0] = dem[:, -1] = -99999.0
    d = _dem_domain(dem)
    for r in range(2, n - 2):
        assert np.allclose(d.zb[r, 2:n - 2], 500.0 + r + 0.5, rtol=0, atol=1e-12)
    # ghosts are linear extrapolations of the first two rows (initiate.cpp:193-198)
    assert np.allclose(d.zb[0, 3:n - 2], 1.5 * d.zb[1, 3:n - 2] - 0.5 * d.zb[2, 3:n - 2])


# ------------------------------------------------------------- WINTRA ----
def check_wintra(lib=None):
    """Soil release in a still basin of fixed depth: the soil mass decays
    geometrically, m_n = m_0 (1 - r dt / 3600)^n (r per hour, dt in s), and
    what leaves the soil arrives as solute mass c h dxy^2 in the same cell.
    Dry cells and cells whose bed is NODATA keep both."""
    n, dxy, H, r, m0, dt, steps = 10, 3, 0.25, 0.8, 7.0, 1.7, 25
    d = flat_domain(n=n, dxy=dxy, h=H, lib=lib)
    d.set_wintra(True, r, 0.0)
    d.soil[1:-1, 1:-1] = m0
    d.h[4, 5] = 0.5 * d.s.hdry                 # a dry cell
    d.zb[6, 3] = d.s.nodata                    # a NODATA bed (h stays H)
    d.dtfl = dt
    for _ in range(steps):
        d.wintra(0)
    expect_soil = m0 * (1.0 - r * dt / 3600.0) ** steps
    wet = np.ones((n + 2, n + 2), dtype=bool)
    wet[0, :] = wet[-1, :] = wet[:, 0] = wet[:, -1] = False
    wet[4, 5] = wet[6, 3] = False
    assert np.allclose(d.soil[wet], expect_soil, rtol=1e-13, atol=0)
    assert np.allclose(d.conc[0][wet], (m0 - expect_soil) / (H * dxy * dxy), rtol=1e-12, atol=0)
    for (i, j) in ((4, 5), (6, 3)):
        assert d.soil[i, j] == m0 and d.conc[0][i, j] == 0.0
    # ghosts are never touched
    assert (d.soil[0, :] == 0).all() and (d.conc[0][:, -1] == 0).all()


def test_wintra_geometric_release():
    check_wintra()


def test_wintra_runs_inside_the_solver_gate_only():
    """main.cpp:385-404: after the ADE, only on solver steps, and switched
    off without the ADE solver (main.cpp:207-213)."""
    for ade, h, expect_release in ((True, 0.3, True), (False, 0.3, False), (True, 0.0, False)):
        d = flat_domain(n=10, h=h, ade=ade)
        d.set_wintra(True, 1.5, 0.0)
        d.soil[1:-1, 1:-1] = 4.0
        d.conc[0][...] = 0.0
        d.step(1e9)
        centre = d.soil[5, 5]
        if expect_release:
            # the step's own dt, one release
            assert centre == pytest.approx(4.0 * (1.0 - 1.5 * d.dtfl / 3600.0), rel=1e-14)
            assert d.conc[0][5, 5] == pytest.approx((4.0 - centre) / (d.h[5, 5] * 4), rel=1e-12)
        else:
            assert centre == 4.0


# ------------------------------------------------------------ restart ----
def _restart_rows(dxy):
    # irow, icol, X, Y, z, h, ux, uy, qx*dxy, qy*dxy, us, conc, soil, fe1, fn1, twet
    return np.array([
        [3, 4, 4 * dxy, 3 * dxy, 512.5, 2.0, 0.1, -0.2, 0.6, -0.8, 0.0, -3.0, 9.0, 0.01, 0.02, 77.0],
        [2, 2, 2 * dxy, 2 * dxy, 511.0, 1.0, 0.0, 0.0, 0.3, 0.4, 0.0, 2.5, 1.5, 0.0, 0.0, 12.0],
    ])


def check_restart(lib=None):
    """Results rows as initial conditions (initiate.cpp:109-156): listed cells
    take the file's z, h, velocities and discharges (written as q*dxy and
    read back without the division), the solute clamped at 0, the soil mass
    and the wet-time tracer, and are marked wet; every other cell keeps the
    pre-load state: water at the bed, background soil mass, solute 0."""
    n, dxy = 6, 2
    dem = np.full((n, n), 510.0)
    dem[0, :] = dem[-1, :] = dem[:, 0] = dem[:, -1] = -99999.0
    d = OracleDomain(n, n, dxy, 1e-4, 1, [0.01], True, lib_path=lib)
    d.read_geo(dem.astype(np.float32), 1e-4)
    d.set_wintra(True, 0.5, 5.0)
    d.initiation_restart(_restart_rows(dxy))
    assert (d.z[3, 4], d.h[3, 4], d.ux[3, 4], d.uy[3, 4]) == (512.5, 2.0, 0.1, -0.2)
    assert (d.qx[3, 4], d.qy[3, 4]) == (0.6, -0.8)
    assert d.conc[0][3, 4] == 0.0 and d.conc[0][2, 2] == 2.5
    assert (d.soil[3, 4], d.soil[2, 2], d.twet[3, 4]) == (9.0, 1.5, 77.0)
    assert d.ldry[3, 4] == 0.0 and d.ldry[2, 2] == 0.0
    # an unlisted interior cell
    assert d.z[4, 3] == d.zb[4, 3] and d.h[4, 3] == 0.0
    assert d.soil[4, 3] == 5.0 and d.conc[0][4, 3] == 0.0 and d.ldry[4, 3] == 0.0
    # ghost rows follow the loaded interior (initiate.cpp:193-199)
    assert d.z[0, 4] == 1.5 * d.z[1, 4] - 0.5 * d.z[2, 4]


def test_restart_overlay():
    check_restart()


def test_restart_without_ade_marks_solute_absent():
    n, dxy = 6, 2
    dem = np.full((n, n), 510.0)
    d = OracleDomain(n, n, dxy, 1e-4, 1, [0.01], False)
    d.read_geo(dem.astype(np.float32), 1e-4)
    d.initiation_restart(_restart_rows(dxy))
    assert d.conc[0][3, 4] == -1.0 and d.conc[0][4, 3] == 0.0


def test_restart_row_outside_grid_is_refused():
    d = OracleDomain(6, 6, 2, 1e-4, 1, [0.01], True)
    d.read_geo(np.full((6, 6), 510.0, dtype=np.float32), 1e-4)
    rows = _restart_rows(2)
    rows[1, 0] = 9
    with pytest.raises(ValueError):
        d.initiation_restart(rows)


def test_roe_upwind_mass_flux_in_supercritical_flow():
    check_roe_supercritical()


def test_roe_shear_wave_dissipation():
    check_roe_shear_wave()


def test_roe_step_limiter_at_a_bed_step():
    check_roe_dz_limiter()


def test_stale_fluxes_persist_in_a_dry_neighbourhood():
    check_stale_fluxes_persist()


def test_ade_new_values_within_neighbour_bounds():
    check_ade_clamp()


def test_roe_dissipation_is_half_the_jacobian_modulus():
    check_roe_matrix()


def test_ade_outflow_limiter_splits_the_held_solute():
    check_ade_outflow_split()


def test_ade_outflow_limiter_single_face():
    check_ade_single_outflow()


def test_ade_emptied_cell_sends_nothing_on():
    check_ade_emptied_cell()


def test_ade_emptied_cell_passes_on_what_it_takes_in():
    check_ade_pass_through()


def test_ade_overdrained_cell_sends_nothing_on():
    check_ade_overdrained_cell()
