// fx_internal.h -- device-side data structures shared by the HIP sources of
// libfluxos_hip.so.  Not part of the public ABI (include/fluxos_hip.h is).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/fluxos_hip.h"

// Step control block, one per domain, resident in device memory.  The CFL
// kernel's last block writes it; every later kernel of the step reads it, so
// a step needs no host round trip.
struct fx_ctl {
    double tim;          // ds.tim after the current step's CFL
    double dtfl;         // ds.dtfl of the current step
    double hpall;        // max wet depth seen by the CFL (solver gate)
    double print_next;   // next print time
    double t_end;        // stop once tim > t_end (main.cpp:271)
    double meteoi;       // melt depth of this step [m] (before multiplier)
    double inflowi;      // inflow depth of this step [m]
    double cmelt[FX_MAX_CHEM];
    double cin[FX_MAX_CHEM];
    // FP64 water ledger (extension): melt volume added to non-weir cells and
    // net volume entering non-weir cells through faces shared with weir cells
    double led_melt, led_melt_c;      // Kahan sum and compensation
    double led_weir, led_weir_c;
    // FP64 solute ledger (extension): solute the melt and the inflow brought
    // into water, per solute (Kahan sum and compensation)
    double sled[FX_MAX_CHEM], sled_c[FX_MAX_CHEM];
    int32_t it;          // solver steps so far
    int32_t gate;        // hpall != 0 for the current step
    int32_t halt;        // steps are no-ops while set
    int32_t pending_halt;
    int32_t flip_pending;// FP32 device loop: the step's buffer flip is left to the next clock
    int32_t steps;       // iterations run since the host last cleared them
    int32_t solver_steps;
    int32_t status;      // bit2: non-finite value seen
    uint32_t blocks_done;// last-block-done counter of the CFL reduction
    uint32_t ade_ticket; // block tickets of the ADE wavefront
    uint32_t ade_error;  // the wavefront's bounded wait expired
    // FP32-mixed path: the step kernel leaves the next step's CFL inputs here
    int32_t cur;         // current ping-pong buffer of d/qx/qy
    int32_t nblk32;      // number of partials written by the last step kernel
    double dt_raw;       // min over wet cells of the CFL bound (no print clip)
    double s_wet;        // wet-time prefix sum [h] over solver steps
    double s_wet_prev;
    double melt_area;    // sum of the melt multiplier over non-weir bed cells
    // per-step FP32 constants of the FP32-mixed kernels (set by k32_clock)
    float k_dt_dxy, k_dt, k_half_dx_dt, k_arbase_dt, k_dxy_dt, k_pad;
};

// launch mode bits
#define FXM_GATED 1u     // honour ctl->halt / ctl->gate (device-resident loop)
#define FXM_MELT  2u     // apply add_meteo in the melt kernel
#define FXM_LOOP1 4u     // apply hydro loop 1 in the melt kernel
#define FXM_INFLOW 8u    // apply add_inflow in the melt kernel

// FP64 parity state (reference dtypes; flags as bytes on the device)
struct fx_state64 {
    double *z, *zb, *h, *qx, *qy, *ks;
    double *fe1, *fe2, *fe3, *fn1, *fn2, *fn3;
    double *twet, *h0;
    uint8_t *ldry, *ldry_prev, *weir;
    double *conc[FX_MAX_CHEM];
    double *conc_tmp;          // pass-1 copy used by the exact ADE
};

// FP32-mixed state: depth relative to the bed in FP64 (ping-pong), discharges
// and fluxes in FP32, bed in FP32, per-cell flag byte.
struct fx_state32 {
    double *d[2];              // z - zb, signed (the reference keeps deficits)
    float2 *q[2];              // (qx, qy) of each cell, one 8-byte load
    double *zb;                // bed elevation (FP64: bed differences to ~1e-13 m)
    float *f6;                 // persistent fluxes fe1,fe2,fe3,fn1,fn2,fn3 (S1), six per cell
    float2 *ff;                // (fe1, fn1) of the step for the fast ADE
    float *mass[2][FX_MAX_CHEM];  // solute mass per unit area (c*h), ping-pong
    uint8_t *flags;            // FXF_* bits
    float2 *tw;                // wet-time tracer: (closed runs, start of the open run)
    double *partials;          // 3 per step-kernel block (dt, hpall, weir volume)
    double *part2;             // 3 per clock block (second level of the fold)
    uint32_t *ticket;          // last-block counter of the clock
    double *rec;               // FX32_REC folds of the partials, by the fast ADE launch
    int64_t npart;
};

#define FXF_WEIR    1u   // innerNeumannBCWeir
#define FXF_STALE   2u   // f6 holds nonzero fluxes for this cell
#define FXF_NODATA  4u   // bed equals NODATA: no melt
#define FXF_WETPREV 8u   // wet in loop 1 of the last solver step (open wet-time run)
#define FXF_INNER  16u   // interior cell (1..NROWS x 1..NCOLS)
#define FXF_DRY4   32u   // dry at the end of the last solver step: the reference's ldry

// per-kernel HIP-event timing on the domain's stream (fx_profile)
struct fx_prof {
    bool on = false;
    int stride = 1;            // time one launch in `stride` of each kernel class
    bool skipping = false;     // the open begin/end pair is not timed
    int64_t seq[FX_K_COUNT] = {0};
    struct rec { int id; hipEvent_t a, b; };
    std::vector<rec> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[FX_K_COUNT] = {0};
    int64_t count[FX_K_COUNT] = {0};
};

struct fx_domain_s {
    fx_domain_desc desc;
    int32_t mrows, mcols;
    int64_t ncell;             // mrows*mcols
    hipStream_t stream;
    fx_ctl *ctl;               // device
    fx_ctl *ctl_host;          // pinned mirror
    double *partials;          // CFL block partials (2 per block)
    int32_t npartials;
    unsigned long long *ade_edges;   // exact ADE: last-column edge words of every block
    int32_t ade_blocks;
    float *meteo; int32_t meteo_nrows, meteo_ncols;
    float *melt_mult;
    float *inflow; int32_t inflow_nrows, inflow_ncols;
    int64_t inflow_irow, inflow_icol;
    bool has_static;
    double ks0;                // interior roughness (uniform in FP32-mixed domains)
    bool melt_area_ok;         // ctl->melt_area is current
    bool meteo_nonneg;         // no negative melt rate or multiplier in the current tables
    bool inflow_nonneg;        // no negative rate in the current inflow table
    fx_state64 s64;
    fx_state32 s32;
    std::vector<double> zb_host;   // FP64 bed (FP32 mode output conversion)
    hipGraphExec_t graph;      // one captured iteration of the loop
    bool graph_ok;
    int32_t ch, ach;           // FP32 sweep column-block lengths (fx_set_tiling)
    bool in_batch;             // member of a batch: tiling is fixed
    bool cfl_fresh;            // FP32: the CFL partials are the last sweep's, of the current state
    bool wintra;               // WINTRA soil release on (fx_set_wintra)
    double soil_rate;          // SOIL_RELEASE_RATE [1/h]
    double *soil;              // soil_mass grid (FP64, both precisions); NULL until needed
    int64_t bytes;
    fx_prof prof;
};

// bracket a launch with timing events when profiling is on
void fx_prof_begin(fx_domain d, int id);
void fx_prof_end(fx_domain d, int id);
int fx_prof_drain(fx_domain d);

// error helpers (fx_capi.hip)
int fx_fail(int code, const std::string &msg);
int fx_hip_check(hipError_t e, const char *what);
#define FX_HIP(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return fx_hip_check(e_, #call); } while (0)

// launchers implemented in the kernel sources
int fx64_launch_cfl(fx_domain d, unsigned mode);
int fx64_launch_melt(fx_domain d, unsigned mode);
int fx64_launch_flux(fx_domain d, unsigned mode);
int fx64_launch_update(fx_domain d, unsigned mode);
int fx64_launch_ade_exact(fx_domain d, unsigned mode, int k0, int nsol);   // solutes k0 .. k0+nsol-1
int fx_launch_forcing(fx_domain d);
int fx_launch_step_end(fx_domain d);
int fx64_launch_volume(fx_domain d, double *vol_out);
int fx64_launch_wintra(fx_domain d, unsigned mode, int k_first, int k_count);
int fx_inflow_cell_ok(fx_domain d);
int fx32_launch_wintra(fx_domain d);

int fx32_init_state(fx_domain d);
int fx32_launch_step(fx_domain d, bool first);
int fx32_launch_step_close(fx_domain d);
int fx32_launch_cfl_initial(fx_domain d);
int fx32_get_field(fx_domain d, int field, void *host);
int fx32_set_field(fx_domain d, int field, const void *host);
int fx32_launch_volume(fx_domain d, double *vol_out);
int fx32_launch_ade(fx_domain d, bool fold);
int fx32_pending_weir(fx_domain d, double *vol);
int fx32_launch_solute(fx_domain d, int k, double *mass_out);
// water-volume block partials of fx_mass_balance into a caller's buffer
int64_t fx32_volume_nparts(fx_domain d);
int fx32_volume_partials(fx_domain d, double *part);
int64_t fx64_volume_nparts(fx_domain d);
int fx64_volume_partials(fx_domain d, double *part);
int fx64_launch_solute(fx_domain d, int k, double *mass_out);
