/*
 * fluxos_oracle.h -- CPU restatement of the FLUXOS-OVERLAND explicit
 * finite-volume time step, used ONLY as test infrastructure (the checker for
 * the HIP library and the CPU baseline timed by bench.py).  Nothing in the
 * product links or calls this code.
 *
 * It follows the reference's double-precision OpenMP path function by
 * function, with identical expression order and dtypes so that results can
 * be compared bit for bit:
 *
 *   orc_courant          <- src/fluxos/main.cpp:300-325   (CFL block, tim += dt)
 *   orc_add_meteo        <- src/fluxos/common.cpp:164-273
 *   orc_add_inflow       <- src/fluxos/common.cpp:277-390
 *   orc_hydrodynamics    <- src/fluxos/hydrodynamics_calc.cpp:27-157
 *   orc_solver_dry       <- src/fluxos/solver_drydomain.cpp:27-194
 *   orc_solver_wet       <- src/fluxos/solver_wetdomain.cpp:27-342
 *   orc_adesolver        <- src/fluxos/ADEsolver_calc.cpp:27-291
 *   orc_read_geo         <- src/fluxos/read_functions.cpp:112-235 (from a parsed DEM)
 *   orc_initiation       <- src/fluxos/initiate.cpp:27-231 (cold start branch)
 *   orc_initiation_restart <- src/fluxos/initiate.cpp:27-231 (RESTART branch, :109-156)
 *   orc_wintra           <- src/fluxos/WINTRAsolver_calc.cpp:26-57
 *   orc_step             <- src/fluxos/main.cpp:271-406 (one loop iteration)
 *
 * Layout: every grid is (nrows+2) x (ncols+2), column-major, index
 * irow + icol*mrows, exactly like the reference's arma::Mat storage.
 * Memory the reference never writes is zero here (the reference relies on
 * Armadillo zero-filling its matrices).
 *
 * Extensions (off unless set): several solutes with their own dispersion
 * coefficient (meteo column 2+k carries solute k), and a per-cell melt
 * multiplier.  With nchem=1 and melt_mult=NULL the reference is reproduced.
 *
 * Pinning.  The reference ships no tests or golden vectors and cannot be
 * built here (Armadillo is absent), so each function is checked against what
 * the physics fixes on its own (tests/test_oracle_pins.py), and each check is
 * shown to fail on a planted mistake (tests/test_oracle_plants.py):
 *   orc_courant        dt = cfl dxy / (max(|u|,|v|,1e-6) + sqrt(g h)), flags
 *   orc_add_meteo      table lookup, mm/day -> m, solute mixing, dry basin
 *   orc_add_inflow     table lookup, q dt / dxy^2 into one cell, mixing
 *   orc_solver_dry     Ritter inflow front (8/27, 4/9), hydrostatic face thrust
 *   orc_solver_wet     lake at rest, uniform-flow physical flux, Roe mass flux
 *                      at a still step, eddy-viscosity shear, Ritter outflow,
 *                      outflow limiter
 *   orc_hydrodynamics  water conservation in a walled basin, 0.8 smoothing
 *   orc_adesolver      diffusion and advection closed forms, conservation,
 *                      zero / uniform fields, first-step skip
 *   orc_read_geo, orc_initiation   hand-worked DEMs
 *   orc_initiation_restart  hand-written Results rows: listed cells take the
 *                      file's values (qx/qy unscaled, conc clamped at 0,
 *                      ldry 0), the rest the background soil mass
 *   orc_wintra         geometric decay m (1 - r dt/3600)^n of the soil mass
 *                      in a still wet basin, soil + solute mass conserved,
 *                      dry and NODATA cells untouched
 *   orc_solver_wet     also with moving water: the Roe dissipation is
 *                      -1/2 |A| dU (A the flux Jacobian at the face's mean
 *                      state, |A| by eigendecomposition) on random
 *                      subcritical and supercritical states; pure upwind
 *                      mass flux in supercritical flow; a shear wave damped
 *                      by 1/2 |u| times its jump; the |dz| cap at a bed step
 *                      (east faces above half the depth, north faces always)
 *   orc_solver_dry     stale fluxes of an all-dry neighbourhood re-applied
 *                      at every step
 *   orc_adesolver      new values within the neighbours' bounds under
 *                      strong transport; the outflow limiter sending out
 *                      exactly the held solute through one face (east or
 *                      north), or split a : b between two; nothing from a
 *                      cell another face has emptied, and from such a cell
 *                      exactly what it takes in through a third face
 *                      (nothing, if it was drained of more than it holds)
 * Parity unpinned: the weir discharge sqrt(g) h^1.5 of outer-ring cells that
 * are not weir cells (solver_drydomain.cpp:178-185, solver_wetdomain.cpp:
 * 300-307).  Supported inputs never reach it: the library refuses a domain
 * whose outer interior ring is not all weir cells.
 */
#ifndef FLUXOS_ORACLE_H
#define FLUXOS_ORACLE_H

#include <stdint.h>

#define ORC_MAX_CHEM 8

typedef struct orc_domain {
    /* sizes */
    int32_t nrows, ncols;   /* interior NROWS, NCOLS */
    int32_t mrows, mcols;   /* padded = n + 2 */
    int64_t dxy;            /* size_t in GlobVar.h:76 */
    int64_t arbase;         /* dxy*dxy, main.cpp:176 */
    double hdry;            /* = ks(1,1) = ROUGNESS_HEIGHT, main.cpp:259 */
    float gacc, cfl, cvdef, nuem;   /* GlobVar.h:78-80, main.cpp:132-135 */
    double nodata;          /* NODATA_VALUE */
    int32_t nchem;
    int32_t ade_solver;     /* EXTERNAL_MODULES.ADE_TRANSPORT.STATUS */
    double d_coef[ORC_MAX_CHEM];
    double dtfl, tim;

    /* grids (caller-owned, mrows*mcols each) */
    double *z, *zb, *h, *qx, *qy, *us, *ks;
    double *fe1, *fe2, *fe3, *fn1, *fn2, *fn3;
    double *dh, *dqx, *dqy, *qxf, *qyf, *twet, *h0;
    float *ldry, *weir, *ldry_prev;
    double *conc[ORC_MAX_CHEM];

    /* meteo forcing: meteo_nrows rows of meteo_ncols floats (row-major);
     * row 0 is the empty header slot (all zero), data from row 1
     * (read_functions.cpp:259-266).  meteo_ncols = 2 + nchem. */
    int32_t meteo_nrows, meteo_ncols;
    const float *meteo;
    const float *melt_mult;     /* optional per-cell multiplier, NULL = 1 */

    /* inflow point source (read_functions.cpp:285-326, rows from 0) */
    int32_t inflow_nrows, inflow_ncols;
    const float *inflow;        /* NULL = no INFLOW_FILE */
    int64_t inflow_row, inflow_col;

    /* diagnostics */
    int64_t ub_volrat;          /* weir branch reached with volrat unset */

    /* WINTRA and restart (appended; NULL / 0 = off) */
    double *soil;               /* soil_mass grid (GlobVar.h:68) */
    double *ux, *uy;            /* ux, uy: only carried from a Results file to the next */
    int32_t wintra;             /* EXTERNAL_MODULES.WINTRA.STATUS */
    int32_t pad_;
    double soil_rate;           /* SOIL_RELEASE_RATE [1/h] */
    double soil_bckgrd;         /* SOIL_CONC_BACKGROUND */
} orc_domain;

#ifdef __cplusplus
extern "C" {
#endif

/* CFL block of main.cpp:300-323 followed by tim += dtfl (main.cpp:325).
 * Returns hpall (the solver gate, main.cpp:385). */
double orc_courant(orc_domain *d, double print_next);
/* common.cpp:164-273.  Returns 0. */
int orc_add_meteo(orc_domain *d);
/* common.cpp:277-390.  Returns 0, or 1 when the cell is NODATA. */
int orc_add_inflow(orc_domain *d);
/* hydrodynamics_calc.cpp:27-157 (all four OpenMP loops). */
void orc_hydrodynamics(orc_domain *d);
void orc_solver_dry(orc_domain *d, int irow, int icol);
void orc_solver_wet(orc_domain *d, int irow, int icol);
/* ADEsolver_calc.cpp:27-291; scratch needs 3*mrows*mcols doubles or NULL. */
void orc_adesolver(orc_domain *d, int it, int ichem, double *scratch);
/* read_functions.cpp:145-225 applied to a parsed DEM: dem holds NROWS lines
 * of NCOLS values in file order (first line = northernmost row). */
void orc_read_geo(orc_domain *d, const float *dem, double ks_input);
/* initiate.cpp:27-231, cold start (RESTART false). */
void orc_initiation(orc_domain *d);
/* initiate.cpp:27-231 with RESTART true and a loaded Results file.  table
 * holds the file's data rows (the line after the header first), 16 doubles
 * each (irow, icol, X, Y, z, h, ux, uy, qx*dxy, qy*dxy, us, conc, soil_mass,
 * fe_1, fn_1, twet); the reference's filedata row a is table row a-1.  The
 * grids must be zero on entry, like GlobVar's new matrices.  Returns 0, or
 * -1 if a row names a cell outside the padded grid (the reference writes
 * out of bounds there). */
int orc_initiation_restart(orc_domain *d, const double *table, int64_t trows);
/* WINTRAsolver_calc.cpp:26-57 for solute ichem. */
void orc_wintra(orc_domain *d, int ichem);

/* One iteration of main.cpp:271-406.  *it is the solver-step counter.
 * Returns bit0 = solver ran, bit1 = inflow error. */
int orc_step(orc_domain *d, double print_next, int32_t *it, double *hpall_out);

int orc_num_threads(void);
/* timing helpers of the CPU baseline (not part of the restatement) */
int orc_step_timed(orc_domain *d, double print_next, int32_t *it, double *hpall_out, double *phase_s);
void orc_set_threads(int n);
int orc_proc_bind(void);

#ifdef __cplusplus
}
#endif

#endif
