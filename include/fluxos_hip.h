/*
 * fluxos_hip.h -- C ABI of libfluxos_hip.so, the MI355X (gfx950) time step of
 * FLUXOS-OVERLAND.
 *
 * The reference has no plugin or FFI layer: its hot path is a set of C++ free
 * functions taking `GlobVar&`.  Each entry point below names the reference
 * seam it replaces.  Host buffers use the reference's own memory layout
 * (arma::Mat, column-major, (NROWS+2) x (NCOLS+2), index irow + icol*MROWS,
 * GlobVar.h:31-73), so a driver can hand over Mat::memptr() directly: double
 * for state grids, float for the flag grids (ldry, innerNeumannBCWeir,
 * ldry_prev).  Buffers are copied at the call and never retained.
 *
 * Errors: every call returns FX_OK (0) or a negative FX_ERR_* code and sets a
 * thread-local message readable with fx_last_error().  Nothing calls exit().
 *
 * Threading: one host thread per domain at a time.  Each domain owns one HIP
 * stream on its device; several domains may share a device.
 */
#ifndef FLUXOS_HIP_H
#define FLUXOS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FX_MAX_CHEM 8

/* status codes */
#define FX_OK               0
#define FX_ERR_ARG         -1   /* invalid argument                         */
#define FX_ERR_HIP         -2   /* HIP runtime error                        */
#define FX_ERR_STATE       -3   /* call not valid in the domain's state     */
#define FX_ERR_UNSUPPORTED -4   /* input the reference leaves undefined     */
#define FX_ERR_NONFINITE   -5   /* NaN/Inf found in the state               */
#define FX_ERR_IO          -6   /* file could not be read or written        */
#define FX_ERR_NOMEM       -7   /* device memory exhausted                  */
#define FX_ERR_NOGPU       -8   /* no HIP device available                  */

/* precision of the device state */
#define FX_FP64_PARITY 0    /* the reference's FP64 arithmetic, bit for bit   */
#define FX_FP32_MIXED  1    /* FP32 fluxes, FP64 depth/mass/dt                */

/* advection-dispersion mode */
#define FX_ADE_OFF   0
#define FX_ADE_EXACT 1      /* the reference's sequential sweep (wavefront)  */
#define FX_ADE_FAST  2      /* fully parallel (Jacobi) face-flux form        */

typedef struct fx_domain_s *fx_domain;

/* Replaces the scalars of GlobVar (GlobVar.h:75-104) plus main.cpp:132-135. */
typedef struct fx_domain_desc {
    int32_t nrows, ncols;        /* NROWS, NCOLS of the DEM                   */
    int32_t dxy;                 /* CELLSIZE, an integer (read_functions.cpp:137) */
    int32_t nchem;               /* solutes (1 without OpenWQ, main.cpp:244-249) */
    double hdry;                 /* ROUGNESS_HEIGHT = ks(1,1) (main.cpp:259)  */
    double nodata;               /* NODATA_VALUE of the DEM header            */
    float gacc, cfl, cvdef, nuem;/* 9.80665, 1, 0.07, 1.793e-6 in the reference */
    double d_coef[FX_MAX_CHEM];  /* D_COEF per solute                         */
    int32_t ade_solver;          /* ADE_TRANSPORT.STATUS                      */
    int32_t precision;           /* FX_FP64_PARITY | FX_FP32_MIXED            */
    int32_t ade_mode;            /* FX_ADE_*                                  */
    int32_t device;              /* HIP device ordinal                        */
} fx_domain_desc;

/* state fields for fx_set_field / fx_get_field (GlobVar.h:81-92).  An
 * FX_FP32_MIXED domain keeps d = z - zb, qx, qy, the stored fluxes, the
 * wet-time tracer, ldry (as the reference sets it) and the solutes; it derives
 * h, h0 and ldry_prev from the current depth on reading and ignores them on
 * writing. */
#define FX_F_Z          0   /* double  water surface elevation z          */
#define FX_F_H          1   /* double  depth h                             */
#define FX_F_QX         2   /* double  qx                                  */
#define FX_F_QY         3   /* double  qy                                  */
#define FX_F_US         4   /* double  us (always zero in the reference)   */
#define FX_F_LDRY       5   /* float   ldry                                */
#define FX_F_LDRY_PREV  6   /* float   ldry_prev                           */
#define FX_F_FE1        7   /* double  fe_1 .. fe_3, fn_1 .. fn_3          */
#define FX_F_FE2        8
#define FX_F_FE3        9
#define FX_F_FN1       10
#define FX_F_FN2       11
#define FX_F_FN3       12
#define FX_F_TWET      13   /* double  twetimetracer                      */
#define FX_F_H0        14   /* double  h0                                  */
#define FX_F_ZB        15   /* double  zb (read only; set by fx_set_static)*/
#define FX_F_SOIL      16   /* double  soil_mass (WINTRA; GlobVar.h:68)    */
#define FX_F_CONC0     32   /* double  conc_SW[k] is FX_F_CONC0 + k        */

/* Per-call step summary of fx_advance. */
typedef struct fx_step_info {
    double tim;          /* ds.tim after the last step                      */
    double dtfl;         /* ds.dtfl of the last step                        */
    double hpall;        /* hpall of the last step (solver gate)            */
    double print_next;   /* print time the call was given                   */
    int32_t it;          /* solver-step counter (main.cpp:387)              */
    int32_t steps;       /* loop iterations executed by this call           */
    int32_t solver_steps;/* of which ran the solver (hpall != 0)            */
    int32_t status;      /* bit0 print due, bit1 tim > t_end, bit2 NaN/Inf  */
} fx_step_info;

/* ---- lifetime ----------------------------------------------------------
 * Replaces `GlobVar ds(NROWS+2, NCOLS+2)` (main.cpp:125-127, GlobVar.h:31-74):
 * allocates the domain's device state.  (NROWS+2)*(NCOLS+2) is at most 2^34
 * cells, and below 2^29 for FX_FP32_MIXED (its kernels address a cell by a
 * 32-bit byte offset); larger descriptors give FX_ERR_ARG.                 */
int fx_domain_create(const fx_domain_desc *desc, fx_domain *out);
int fx_domain_destroy(fx_domain dom);
/* Device bytes the domain holds (for admission by memory budget). */
int64_t fx_domain_bytes(const fx_domain_desc *desc);

/* ---- state transfer -----------------------------------------------------
 * Static grids after read_geo + initiation (read_functions.cpp:221-223,
 * initiate.cpp:43-90,192-228): zb and ks double, weir = innerNeumannBCWeir
 * float.  Every cell of the outer interior ring must be a weir cell: the
 * reference's weir branch reads an uninitialised `volrat` otherwise
 * (solver_drydomain.cpp:36,178-185; solver_wetdomain.cpp:37,300-307). */
int fx_set_static(fx_domain dom, const double *zb, const double *ks, const float *weir);
int fx_set_field(fx_domain dom, int32_t field, const void *host);
int fx_get_field(fx_domain dom, int32_t field, void *host);
/* ds.tim, ds.dtfl and the solver-step counter `it` (main.cpp:72,271-325). */
int fx_set_clock(fx_domain dom, double tim, double dtfl, int32_t it);
int fx_get_clock(fx_domain dom, double *tim, double *dtfl, int32_t *it);

/* ---- forcing ---------------------------------------------------------------
 * Meteo table as read by read_meteo (read_functions.cpp:237-282): nrows rows
 * of ncols = 2+nchem floats, row-major (t [s], rate [mm/day], conc_k), row 0
 * being the empty header slot.  melt_mult (optional, padded grid of floats)
 * scales the melt per cell; NULL = uniform as in the reference.           */
int fx_set_meteo(fx_domain dom, int32_t nrows, int32_t ncols, const float *table,
                 const float *melt_mult);
/* Inflow table as read by read_inflow (read_functions.cpp:285-326): rows
 * from 0, and the discharge cell (read_functions.cpp:228-231), an interior
 * cell.  fx_advance returns FX_ERR_ARG if the cell's bed is NODATA_VALUE
 * (common.cpp:377-380).                                                   */
int fx_set_inflow(fx_domain dom, int32_t nrows, int32_t ncols, const float *table,
                  int64_t irow, int64_t icol);

/* ---- granular seams (FX_FP64_PARITY domains) --------------------------------
 * fx_courant_condition: the inline CFL block main.cpp:300-323 plus
 * `ds.tim += ds.dtfl` (main.cpp:325).  Returns dtfl and hpall.            */
int fx_courant_condition(fx_domain dom, double print_next, double *dtfl, double *hpall);
/* add_meteo, common.cpp:164-273 (common.h:50). */
int fx_add_meteo(fx_domain dom);
/* add_inflow, common.cpp:277-390 (common.h:69). */
int fx_add_inflow(fx_domain dom);
/* hydrodynamics_calc, hydrodynamics_calc.cpp:27-157 (hydrodynamics_calc.h:24),
 * including solver_dry / solver_wet (solver_drydomain.h:25, solver_wetdomain.h:25). */
int fx_hydrodynamics_calc(fx_domain dom);
/* adesolver_calc, ADEsolver_calc.cpp:27-291 (ADEsolver_calc.h:25). */
int fx_adesolver_calc(fx_domain dom, int32_t it, int32_t ichem);
/* wintrasolver_calc, WINTRAsolver_calc.cpp:26-57 (WINTRAsolver_calc.h:25):
 * release of soil_mass into solute ichem in every wet cell whose bed is not
 * NODATA.  Needs fx_set_wintra(dom, 1, ...). */
int fx_wintrasolver_calc(fx_domain dom, int32_t ichem);

/* ---- WINTRA module ------------------------------------------------------------
 * EXTERNAL_MODULES.WINTRA.STATUS and SOIL_RELEASE_RATE [1/h]
 * (read_functions.cpp:46,65).  When on, fx_advance runs the release after the
 * ADE of every solver step (main.cpp:399-404) and the domain holds a
 * soil_mass grid (FX_F_SOIL, zero until set).  The reference switches the
 * module off without the ADE solver (main.cpp:207-213): status 1 on such a
 * domain is FX_ERR_STATE.  Batches do not take WINTRA domains. */
int fx_set_wintra(fx_domain dom, int32_t status, double soil_release_rate);

/* ---- device-resident loop ----------------------------------------------------
 * Iterations of main.cpp:271-406 without host round trips: each iteration is
 * CFL -> tim += dt -> forcing -> (if hpall != 0) hydrodynamics + ADE.  Runs
 * while tim <= t_end and stops after the iteration that reaches print_next
 * (main.cpp:444) or after max_steps iterations.                           */
int fx_advance(fx_domain dom, double t_end, double print_next, int32_t max_steps,
               fx_step_info *info);
/* Water volume of the non-weir cells and the running ledger of forcing input
 * and weir-face exchange, all accumulated in FP64 (extension; no counterpart
 * in the reference).  out[0]=volume, out[1]=melt and inflow in (with the
 * inflow cell's deficit filled up to the bed, common.cpp:324-325),
 * out[2]=net weir inflow.  The ledger is kept by FP32-mixed domains. */
int fx_mass_balance(fx_domain dom, double out[3]);
/* fx_mass_balance without any allocation (extension; no counterpart in the
 * reference).  For callers that cannot allocate device or pinned memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It runs fx_mass_balance's volume kernel into a fixed device
 * array of the library (2^20 block partials, 8 MiB per device, held until the
 * process ends) and sums the partials on the device in the same order, so
 * out[] is bit-identical to fx_mass_balance's.  A domain whose volume kernel
 * has more blocks than that (FP32-mixed: ceil(NROWS/61) * ceil(NCOLS/ch);
 * FP64: ceil(NROWS/64) * ceil(NCOLS/4)) is FX_ERR_ARG.  Calls on one device
 * share the array: the library serializes them within a process.          */
int fx_mass_balance_noalloc(fx_domain dom, double out[3]);
/* Solute ichem's mass in the non-weir cells, sum of c h dxy^2 in FP64
 * (out[0]), and the running FP64 ledger of the solute the melt and the
 * inflow brought into water (out[1], FP32-mixed domains; 0 for FP64
 * domains) (extension; no counterpart in the reference).  The reference's
 * scheme does not conserve solute: it is zeroed in cells that dry and in weir
 * cells, and the ADE's min/max clamp adds or removes it, so out[0] and
 * out[1] differ by those sinks.                                            */
int fx_solute_balance(fx_domain dom, int32_t ichem, double out[2]);
/* Tile shape of the FP32-mixed sweeps (extension; no counterpart in the
 * reference): columns walked per block by the step sweep (ch) and by the
 * ADE sweep (ach), 8..4096; defaults 32 and 64.  Results do not depend on it.
 * Chosen per grid size by the autotuner (fluxos_cpp_amd/autotune.py).     */
int fx_set_tiling(fx_domain dom, int32_t ch, int32_t ach);
int fx_get_tiling(fx_domain dom, int32_t *ch, int32_t *ach);
int fx_synchronize(fx_domain dom);

/* ---- batches of scenarios (extension; no counterpart in the reference) ------
 * Many independent FP32-mixed domains advance together: one batched step runs
 * one iteration of main.cpp:271-406 for every member, each with its own CFL
 * dt, clock, forcing and print time, in one launch per kernel over all
 * members' blocks (replayed from a captured hipGraph).  Members of a batch
 * carry the same number of solutes and the same ADE setting.  A member halts
 * when its clock reaches its print time or passes t_end; fx_batch_poll
 * reports it, the host may then read its fields with fx_get_field and resume
 * or evict it.  Between fx_batch_run and the next fx_batch_poll the members'
 * state belongs to the batch: no other call may touch a member domain.  At
 * any time, calls that change a member's state, clock, tables, tiling or
 * WINTRA setting, the granular seams, fx_advance and fx_domain_destroy
 * return FX_ERR_STATE until it is evicted (the batch holds the member's
 * launch parameters and table addresses from admission on).  While the
 * batch's fast-ADE blocks are too few to fill the GPU (few members left),
 * it runs the members' ADE on narrower tiles than their own; results do not
 * depend on the tiling.                                                      */
typedef struct fx_batch_s *fx_batch;
#define FX_EV_PRINT 1       /* member reached its print time (tim <= t_end)   */
#define FX_EV_DONE  2       /* member's clock passed t_end (main.cpp:271)     */
#define FX_EV_ERROR 3       /* NaN/Inf in the member's state                  */
typedef struct fx_batch_event {
    int32_t slot;        /* member slot returned by fx_batch_admit          */
    int32_t kind;        /* FX_EV_*                                         */
    double tim;          /* member's clock                                  */
    int32_t it;          /* its solver-step counter                         */
    int32_t steps;       /* loop iterations it ran since admission          */
    int32_t solver_steps;/* of which ran the solver (hpall != 0)            */
    int32_t pad;
} fx_batch_event;
int fx_batch_create(int32_t device, int32_t nchem, int32_t ade_solver, int32_t max_members, fx_batch *out);
/* Waits for the batch's queued steps; members still in it leave it with
 * their state, as if evicted.                                             */
int fx_batch_destroy(fx_batch batch);
/* Admit a prepared domain (static fields, state and forcing set) that runs
 * from its current clock until tim > t_end, halting at print_next.        */
int fx_batch_admit(fx_batch batch, fx_domain dom, double t_end, double print_next, int32_t *slot);
/* Continue a member halted at a print time, with its next print time. */
int fx_batch_resume(fx_batch batch, int32_t slot, double print_next);
/* Remove a member; the domain stays valid and keeps its state. */
int fx_batch_evict(fx_batch batch, int32_t slot);
/* Enqueue nsteps batched steps (asynchronous). */
int fx_batch_run(fx_batch batch, int32_t nsteps);
/* Wait for the queued steps; list members that halted since last listed. */
int fx_batch_poll(fx_batch batch, fx_batch_event *events, int32_t max_events, int32_t *n_events);
/* Members, their interior cells, and cell-updates done (as of the last poll). */
int fx_batch_info(fx_batch batch, int32_t *n_members, int64_t *cells, int64_t *cell_updates);

/* ---- host input layer (no GPU needed) -----------------------------------------
 * read_geo on a parsed DEM (read_functions.cpp:145-225): dem has NROWS lines
 * of NCOLS values in file order.  Outputs are padded grids.               */
int fx_read_geo(int32_t nrows, int32_t ncols, double nodata, const float *dem,
                double ks_input, double *zb, double *ks, float *weir);
/* initiation, cold start (initiate.cpp:27-231): zb in place (vertex to cell
 * centre, ghost extrapolation) and the initial z, h, ldry, conc grids.    */
int fx_initiation(int32_t nrows, int32_t ncols, int32_t ade_solver, double *zb,
                  double *z, double *h, float *ldry, double *conc0);
/* initiation with RESTART true and a loaded Results file (initiate.cpp:27-231,
 * the load at :109-156).  table: the file's data rows, 16 doubles each, as
 * fx_read_results returns them.  Cells the file lists take its z, h, ux, uy,
 * qx, qy (the file holds q*dxy; the reference reads it back as the
 * discharge, initiate.cpp:144-145), us, conc_SW[0] clamped at 0 (-1 without
 * the ADE solver), soil_mass and twetimetracer, and ldry = 0; every other
 * cell keeps z = zb, h = 0, zero velocities, ldry = 0 and the background soil
 * mass (initiate.cpp:93-107).  Every output grid is overwritten.  A row
 * naming a cell outside the padded grid is FX_ERR_ARG.                    */
typedef struct fx_init_grids {
    double *zb;                  /* in: read_geo's bed; out: the initial bed  */
    double *z, *h, *qx, *qy, *us, *ux, *uy, *conc0, *soil_mass, *twet;
    float *ldry;
} fx_init_grids;
int fx_initiation_restart(int32_t nrows, int32_t ncols, int32_t ade_solver, double soil_conc_bckgrd,
                          const double *table, int64_t table_rows, fx_init_grids *grids);

/* ---- files in the reference's formats (no GPU needed) ----------------------------
 * ESRI-ASCII DEM: NROWS/NCOLS from the first two header lines in either
 * order (get_domain_size, common.cpp:110-161), then XLLCORNER, YLLCORNER,
 * CELLSIZE (an integer) and NODATA_VALUE (an integer) (read_functions.cpp:
 * 129-141).  dem (NROWS*NCOLS floats, file order; NULL = header only).     */
typedef struct fx_dem_header {
    int32_t nrows, ncols, cellsize, pad;
    double xllcorner, yllcorner, nodata;
} fx_dem_header;
int fx_read_dem(const char *path, fx_dem_header *hdr, float *dem);
/* Forcing CSV (read_meteo / read_inflow, read_functions.cpp:237-326): the
 * first line is a header; each later line is one row of comma-separated
 * numbers.  Returns the row and column counts; with table != NULL also the
 * rows, row-major in a max_rows x max_cols float buffer (missing cells 0). */
int fx_read_forcing(const char *path, int32_t max_rows, int32_t max_cols, float *table, int32_t *nrows,
                    int32_t *ncols);
/* Results/<t>.txt (write_results, write_results.cpp:30-102): one line per
 * interior cell with zb != 0 and h > H_MIN_TO_PRINT (truncated to an integer
 * as the reference stores it), 16 columns: irow, icol, X, Y, z, z-zb, ux,
 * uy, qx*dxy, qy*dxy, us, conc_SW[0], soil_mass, fe_1, fn_1, twetimetracer.
 * Numbers are written as %.16e (the reference's Armadillo CSV formatting
 * is not pinned by any fixture).  NULL optional grids print as 0.         */
typedef struct fx_results_grids {
    int32_t nrows, ncols, dxy, pad;
    double xllcorner, yllcorner, h_min_print;
    const double *zb, *z, *h;                      /* required */
    const double *qx, *qy, *us, *conc0, *soil_mass, *fe1, *fn1, *twet;
    const double *ux, *uy;                         /* carried from a restart file */
} fx_results_grids;
int fx_write_results(const char *path, const fx_results_grids *grids);
/* A Results file read back as initial conditions (initiate.cpp:118-120):
 * the lines after the header, 16 numbers each.  Returns the row count; with
 * table != NULL also the rows (max_rows x 16 doubles, row-major).          */
int fx_read_results(const char *path, int64_t max_rows, double *table, int64_t *nrows);
/* findLastStep (common.cpp:69-106): the largest time among the folder's file
 * names, each read as atoi of its first six characters ("3600.txt" ->
 * 3600; the reference copies sizeof(char*)-2 bytes).  The reference skips
 * the first two directory entries, assuming they are "." and ".."; here
 * those two names are skipped wherever they come.  0 for an empty or
 * missing folder. */
int fx_find_last_step(const char *folder, int32_t *timstart);

/* ---- kernel timing (extension; no counterpart in the reference) --------------
 * When enabled (on >= 1), kernels the domain launches are bracketed by HIP
 * events recorded on the domain's own stream: every launch with on = 1, one
 * launch in `on` of each kernel class otherwise (less event overhead).
 * fx_profile_read returns the summed device time [ms] and count of the timed
 * launches of one kernel class since the last reset. */
#define FX_K_CFL        0   /* FP64 CFL reduction                          */
#define FX_K_MELT       1   /* FP64 forcing + hydro loop 1                 */
#define FX_K_FLUX       2   /* FP64 solver_dry / solver_wet                */
#define FX_K_UPDATE     3   /* FP64 hydro loops 3+4                        */
#define FX_K_ADE_PASS1  4   /* FP64 ADE pass 1                             */
#define FX_K_ADE_SWEEP  5   /* FP64 ADE wavefront sweep                    */
#define FX_K_STEP32     6   /* FP32-mixed fused step                       */
#define FX_K_ADE32      7   /* FP32-mixed fast ADE                         */
#define FX_K_CFL32      8   /* FP32-mixed stand-alone CFL                  */
#define FX_K_COUNT     16
int fx_profile(fx_domain dom, int32_t on);
int fx_profile_read(fx_domain dom, int32_t kernel, double *total_ms, int64_t *launches);

const char *fx_last_error(void);
const char *fx_version(void);

#ifdef __cplusplus
}
#endif

#endif
