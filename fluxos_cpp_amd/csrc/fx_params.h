// fx_params.h -- launch parameters shared by the kernels (passed by value).
#pragma once

#include "fx_internal.h"

struct fx_params {
    int32_t nrows, ncols;     // interior size
    int32_t ld;               // leading dimension = mrows
    int32_t nchem;
    int32_t ade;              // ADE_TRANSPORT.STATUS
    int32_t meteo_nrows, meteo_ncols;
    int32_t inflow_nrows, inflow_ncols;
    int32_t rates_nonneg;
    int32_t ch, ach;          // column-block lengths of the FP32 step and ADE sweeps
    int32_t wintra;           // WINTRA module on (soil release after the ADE)
    double soil_rate;         // SOIL_RELEASE_RATE [1/h]
    int64_t inflow_p;         // linear index of the inflow cell, -1 = none
    double dxy;               // CELLSIZE as double
    double arbase;            // dxy*dxy as double
    uint32_t dxy_u;           // CELLSIZE as unsigned (solver_wet's dx, dy)
    double hdry;
    double ks0;               // interior roughness (FP32-mixed path)
    double nodata;
    double g;                 // (double)gacc
    float cfl_dxy;            // cfl*dxy evaluated in float (main.cpp:316)
    double cvdef, nuem;       // float constants promoted to double
    double D[FX_MAX_CHEM];
    const float *meteo;
    const float *melt_mult;
    const float *inflow;
    const double *d32[2];     // FP32-mixed depth planes (the clock's inflow ledger), else NULL
};

fx_params fx_make_params(fx_domain d);

// Forcing lookup (common.cpp:195-203, 307-317): first row whose time exceeds
// tim; past the end of the table rate and concentrations are zero.
__device__ inline int fx_forcing_row(const float *tab, int nrows, int ncols, double tim)
{
    for (int a = 0; a < nrows; a++)
        if ((double)tab[(size_t)a * ncols] > tim) return a;
    return -1;
}

// The same lookup by every thread of a block (all must call it): the smallest
// row index whose time exceeds tim, -1 if none.
__device__ inline int fx_forcing_row_block(const float *tab, int nrows, int ncols, double tim)
{
    __shared__ int best;
    if (threadIdx.x == 0) best = 0x7fffffff;
    __syncthreads();
    int mine = 0x7fffffff;
    for (int a = threadIdx.x; a < nrows; a += blockDim.x)
        if ((double)tab[(size_t)a * ncols] > tim) { mine = a; break; }
    if (mine != 0x7fffffff) atomicMin(&best, mine);
    __syncthreads();
    const int r = best;
    __syncthreads();
    return r == 0x7fffffff ? -1 : r;
}

// Evaluate both forcings for the clock held in ctl (one thread); mrow / irow
// are the table rows if already looked up (-2 = look up here).
__device__ inline void fx_forcing_eval(const fx_params &P, fx_ctl *ctl, int mrow = -2, int irow = -2)
{
    double rate = 0.0;
    for (int k = 0; k < FX_MAX_CHEM; k++) { ctl->cmelt[k] = 0.0; ctl->cin[k] = 0.0; }
    if (P.meteo) {
        int row = mrow != -2 ? mrow : fx_forcing_row(P.meteo, P.meteo_nrows, P.meteo_ncols, ctl->tim);
        if (row >= 0) {
            rate = (double)P.meteo[(size_t)row * P.meteo_ncols + 1];
            for (int k = 0; k < P.nchem && 2 + k < P.meteo_ncols; k++)
                ctl->cmelt[k] = (double)P.meteo[(size_t)row * P.meteo_ncols + 2 + k];
        }
    }
    ctl->meteoi = rate / (1000. * 3600. * 24.) * ctl->dtfl;
    double q = 0.0;
    if (P.inflow) {
        int row = irow != -2 ? irow : fx_forcing_row(P.inflow, P.inflow_nrows, P.inflow_ncols, ctl->tim);
        if (row >= 0) {
            q = (double)P.inflow[(size_t)row * P.inflow_ncols + 1];
            for (int k = 0; k < P.nchem && 2 + k < P.inflow_ncols; k++)
                ctl->cin[k] = (double)P.inflow[(size_t)row * P.inflow_ncols + 2 + k];
        }
    }
    ctl->inflowi = q * ctl->dtfl / (P.dxy * P.dxy);
}
