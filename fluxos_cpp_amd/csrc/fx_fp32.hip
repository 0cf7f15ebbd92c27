// fx_fp32.hip -- FP32-mixed path for gfx950: one fused, LDS-tiled stencil
// kernel per time step.
//
// State in HBM (per cell): depth relative to the bed d = z - zb in FP64
// (ping-pong), qx, qy in FP32 (ping-pong), bed zb in FP64, one flag byte, and
// -- touched only where needed -- the persistent fluxes of cells that may be
// "all dry" next step (the reference keeps stale fluxes there,
// solver_drydomain.cpp:60-76) and the wet-time tracer (written only when a
// cell changes between wet and dry).
//
// k32_sweep (one per step; described at the kernel): one wave per 63-row
// strip walks along the columns, so every cell is read from HBM once and its
// fluxes are computed once.  Per column it applies the melt of this step
// (common.cpp:210-263) and hydro loop 1 (hydrodynamics_calc.cpp:53-73) as the
// cells arrive, computes the face fluxes (solver_dry / solver_wet in FP32;
// solver_drydomain.cpp:27-194, solver_wetdomain.cpp:27-342), updates the
// state with the 0.1/0.8/0.1 face smoothing (hydrodynamics_calc.cpp:93-154)
// and folds the CFL bound of the new state for the next step
// (main.cpp:300-323) into per-block FP64 partials that k32_clock reduces.
// Face volume terms are FP32 products summed exactly in FP64, so every face
// moves the same volume out of one cell and into the other and the FP64
// depth conserves water.
// The bed is held in FP64 and only its differences between neighbours enter
// the FP32 arithmetic, so the water-surface gradient of millimetre films over
// a 500 m datum keeps its accuracy (~1e-13 m), and the front faces, which
// decide in FP64, see the bed as the FP64 path does.
#include "fx_kern32.h"

#include <cstdlib>

// multiply-adds are fused within one source expression only: the column
// routines are inlined several times (block prologue, loop body, tail), and
// fusing across statements can differ between copies, which would make a
// cell's rounding depend on its place in the tile and results on the tiling
#pragma clang fp contract(on)


namespace {

// the clock of one loop iteration (main.cpp:271-325, 385-387) from the
// folded CFL partials of the previous step sweep: used by k32_clock (one
// domain) and k32b_clock (batch members)

// the CFL time step from the largest wave speed of the domain (9e10 s, the
// reference's initial bound, where no cell is above hdry)
__device__ inline double cfl_dt(const fx_params &P, double smax)
{
    return smax > 0.0 ? (double)(P.cfl_dxy * fx_rcp((float)smax)) : 9.e10;
}

__device__ inline void kahan_add(double &sum, double &comp, double v)
{
    const double y = v - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

__device__ inline void clock_advance(const fx_params &P, fx_ctl *ctl, double a, double b, double c, double m)
{
    // Kahan-compensated ledger of the weir-face exchange of the previous step
    kahan_add(ctl->led_weir, ctl->led_weir_c, c * P.arbase);
    // solute the previous step's melt brought into water (m: melt area of
    // the cells that held water; meteoi and cmelt are still that step's)
    if (P.meteo && P.ade)
        for (int k = 0; k < P.nchem; k++)
            kahan_add(ctl->sled[k], ctl->sled_c[k], ctl->meteoi * m * P.arbase * ctl->cmelt[k]);
    const double dtc = cfl_dt(P, a);                   // a: the largest wave speed
    ctl->dt_raw = dtc;
    const double dt = fmin(ctl->print_next - ctl->tim, dtc);
    ctl->dtfl = dt;
    ctl->hpall = b;
    ctl->tim = ctl->tim + dt;
    const int gate = (b != 0.0);
    ctl->gate = gate;
    ctl->it += gate;
    ctl->steps += 1;
    ctl->solver_steps += gate;
    ctl->k_dt_dxy = (float)(dt / P.dxy);
    ctl->k_dt = (float)dt;
    ctl->k_half_dx_dt = (float)(0.5 * P.dxy / dt);
    ctl->k_arbase_dt = (float)(P.arbase / dt);
    ctl->k_dxy_dt = (float)(P.dxy / dt);
    ctl->s_wet_prev = ctl->s_wet;
    if (gate) ctl->s_wet = ctl->s_wet + dt / 3600;
}

__device__ inline void clock_forcing(const fx_params &P, fx_ctl *ctl, int mrow, int irow)
{
    fx_forcing_eval(P, ctl, mrow, irow);
    if (P.inflow) {                  // the inflow cell's bed is never NODATA here (checked by the host)
        // water in: the source, plus a deficit of the cell filled up to the
        // bed first (common.cpp:324-325), judged after this step's melt
        double vin = ctl->inflowi;
        if (P.d32[0]) {
            const size_t p = (size_t)P.inflow_p;
            double dm = P.d32[ctl->cur][p];
            if (P.meteo) dm += ctl->meteoi * (P.melt_mult ? (double)P.melt_mult[p] : 1.0);
            if (dm <= 0.0) vin -= dm;
            // its solute enters only water (common.cpp:350-361)
            if (dm > 0.0 && P.ade)
                for (int k = 0; k < P.nchem; k++)
                    kahan_add(ctl->sled[k], ctl->sled_c[k], ctl->inflowi * P.arbase * ctl->cin[k]);
        }
        const double y = vin * P.arbase - ctl->led_melt_c;
        const double t = ctl->led_melt + y;
        ctl->led_melt_c = (t - ctl->led_melt) - y;
        ctl->led_melt = t;
    }
    if (P.meteo) {
        const double y = ctl->meteoi * ctl->melt_area * P.arbase - ctl->led_melt_c;
        const double t = ctl->led_melt + y;
        ctl->led_melt_c = (t - ctl->led_melt) - y;
        ctl->led_melt = t;
    }
    ctl->pending_halt = (ctl->tim >= ctl->print_next);
}

// clock_finish: the step's clock from the folded (wave speed, hpall, weir).  Every
// thread of the block calls it (the forcing-table lookup is shared); thread 0
// holds the folded values and writes the control block.
__device__ inline void clock_finish(const fx_params &P, fx_ctl *ctl, double a, double b, double c, double m)
{
    __shared__ double tim_new;
    if (threadIdx.x == 0) clock_advance(P, ctl, a, b, c, m);
    if (threadIdx.x == 0) tim_new = ctl->tim;
    __syncthreads();
    const int mrow = P.meteo ? fx_forcing_row_block(P.meteo, P.meteo_nrows, P.meteo_ncols, tim_new) : -1;
    const int irow = P.inflow ? fx_forcing_row_block(P.inflow, P.inflow_nrows, P.inflow_ncols, tim_new) : -1;
    if (threadIdx.x == 0) clock_forcing(P, ctl, mrow, irow);
}

// The forcing row for tim with thread t holding the time of row t (tables of
// at most blockDim.x rows, loaded before the clock knows tim): the smallest
// row whose time exceeds tim, -1 if none.  All threads call it.
__device__ inline int forcing_row_held(float t_row, int nrows, double tim)
{
    __shared__ int best;
    if (threadIdx.x == 0) best = 0x7fffffff;
    __syncthreads();
    if ((int)threadIdx.x < nrows && (double)t_row > tim) atomicMin(&best, (int)threadIdx.x);
    __syncthreads();
    const int r = best;
    __syncthreads();
    return r == 0x7fffffff ? -1 : r;
}

// clock_finish with the forcing tables' times already held (mt, it: row
// threadIdx.x of the melt and inflow tables; held = both tables fit)
__device__ inline void clock_finish_held(const fx_params &P, fx_ctl *ctl, double a, double b, double c, double m,
                                         bool held, float mt, float it)
{
    if (!held) {
        clock_finish(P, ctl, a, b, c, m);
        return;
    }
    __shared__ double tim_new2;
    if (threadIdx.x == 0) clock_advance(P, ctl, a, b, c, m);
    if (threadIdx.x == 0) tim_new2 = ctl->tim;
    __syncthreads();
    const int mrow = P.meteo ? forcing_row_held(mt, P.meteo_nrows, tim_new2) : -1;
    const int irow = P.inflow ? forcing_row_held(it, P.inflow_nrows, tim_new2) : -1;
    if (threadIdx.x == 0) clock_forcing(P, ctl, mrow, irow);
}

// Lane masks of the three ring columns (c-1, c, c+1, by ring slot), one bit
// per lane of the wave, in scalar registers: w = wet in loop 1, a = an
// interior cell that is not a weir cell (it computes fluxes).  A neighbour's
// bit is the mask shifted by one lane (rows) or another slot's mask
// (columns), so neighbour tests cost no vector instructions.
struct cmask {
    uint64_t w[3], a[3];
};
#define FLUX_LANES 0x7ffffffffffffffeull   // lanes 1..62 compute fluxes

struct consts32 {
    float hdry, g, nuem, ks, dx, idx;
    float sqrt_g, q_isg;               // sqrt(g), 1/(4 sqrt(g))
    float nuem_idx;                    // nuem / dx
    float half_dx_over_dt, arbase_over_dt;
    double hdry64;
};

// A wet/dry front in FP64: the depths after melt of P, E and N and the bed
// differences E-P and N-P, from which a front face forms its free-surface
// difference (ze - zp = (de - dp) + dzb) and its mean depth over the face's
// bed (ze - zbpe = de + dzb/2), as the reference does on z in FP64.  The
// branch tests of the front faces (wet above hdry, inflow or pressure only)
// and the Ritter depth then see what the FP64 path sees; the flux magnitudes
// follow in FP32.
struct front64 {
    double dp, de, dn, dzbE, dzbN;
};

// Ritter inflow face state from a free-surface drop dz (solver_drydomain.cpp:
// 127-138, solver_wetdomain.cpp:131-138): q = 0.296 dz sqrt(g dz) (8/27),
// depth 0.444 dz (4/9); q^2/h is 0 where dz rounds to 0 in FP32
__device__ inline float ritter_q(float g, float dz) { return 0.296f * dz * fx_sqrt(g * dz); }
__device__ inline float q2_over_h(float q, float h) { return h > 0.f ? q * (q * fx_rcp(h)) : 0.f; }

// solver_dry (solver_drydomain.cpp:78-193): pressure-only or Ritter inflow
// from a wet E / N neighbour, decided in FP64 on the front
__device__ inline void solver_dry32(const consts32 &K, const front64 &fr, bool wetE, bool wetN,
                                    float dxy_over_dt, float *f)
{
    const float g = K.g;
    float fe1 = 0.f, fe2 = 0.f, fn1 = 0.f, fn3 = 0.f;
    if (wetE) {
        const double hme = fr.de + 0.5 * fr.dzbE;                      // ze - zbpe (its max with 0
        if (hme > K.hdry64) {                                            // against hdry > 0)
            const double dze = (fr.de - fr.dp) + fr.dzbE;              // ze - zp
            if (dze <= 0.0) {
                const float hm = (float)hme;
                fe2 = .5f * g * hm * hm;
            } else {
                const float dz = fx_min_nn((float)dze, (float)hme);
                const float he = fx_max0((float)fr.de);
                const float hm = 0.444f * dz;
                const float q = -fx_min_nn(ritter_q(g, dz), dxy_over_dt * he);
                fe1 = q;
                fe2 = q2_over_h(q, hm) + .5f * g * hm * hm;
            }
        }
    }
    if (wetN) {
        const double hmn = fr.dn + 0.5 * fr.dzbN;
        if (hmn > K.hdry64) {
            const double dzn = (fr.dn - fr.dp) + fr.dzbN;
            if (dzn <= 0.0) {
                const float hm = (float)hmn;
                fn3 = .5f * g * hm * hm;
            } else {
                const float dz = fx_min_nn((float)dzn, (float)hmn);
                const float hn = fx_max0((float)fr.dn);
                const float hm = 0.444f * dz;
                const float r = -fx_min_nn(ritter_q(g, dz), dxy_over_dt * hn);
                fn1 = r;
                fn3 = q2_over_h(r, hm) + .5f * g * hm * hm;
            }
        }
    }
    f[0] = fe1; f[1] = fe2; f[2] = 0.f; f[3] = fn1; f[4] = 0.f; f[5] = fn3;
}

// solver_wet in FP32 (solver_wetdomain.cpp:27-342), same branch structure.
// The eight neighbour velocities q/max(h, hdry, ks) of the eddy-viscosity
// term (:191-200) are per-cell quantities, computed once when a cell is
// loaded (u, v below); likewise sqrt(h) for the eddy coefficients (:188-189).
struct nb32 {
    float dp, de, dn;                  // depths after melt (may be negative)
    float dzbE, dzbN;                  // bed differences E-P, N-P
    float qp, qe, qn;                  // qx at P, E, N
    float rp, re, rn;                  // qy at P, E, N
    float uP, vP, uN, uS, uSE, uNE;    // cell velocities q/max(h,hdry,ks)
    float vE, vW, vNW, vNE;
    float sqP, sqE, sqN;               // sqrt(max(0, d))
    bool wetE, wetN;
};

__device__ inline void solver_wet32(const consts32 &K, const nb32 &n, const front64 &fr, float *f)
{
    const float hdry = K.hdry, g = K.g, dx = K.dx;
    const float hp = fx_max0(n.dp);
    float he = fx_max0(n.de);
    float hn = fx_max0(n.dn);
    float qp = n.qp, qe = n.qe, qn = n.qn;
    float rp = n.rp, re = n.re, rn = n.rn;
    float hme = .5f * (hp + he), qme = .5f * (qp + qe), rme = .5f * (rp + re);
    float hmn = .5f * (hp + hn), qmn = .5f * (qp + qn), rmn = .5f * (rp + rn);
    float dze = (n.de - n.dp) + n.dzbE;
    const float dqe = qe - qp, dre = re - rp;
    float dzn = (n.dn - n.dp) + n.dzbN;
    const float drn = rn - rp, dqn = qn - qp;
    bool lroe = true;
    // a dry E / N neighbour: the front face, decided in FP64 (:121-175)
    if (!n.wetE) {
        const double hm64 = fr.dp - 0.5 * fr.dzbE;                     // zp - zbpe
        he = 0.f; qe = 0.f; re = 0.f;
        qme = 0.f; rme = 0.f; hme = 0.f;
        if (hm64 > K.hdry64) {
            hme = (float)hm64;
            const double dz64 = (fr.de - fr.dp) + fr.dzbE;               // ze - zp
            if (dz64 < 0.0) {
                const float dz = fx_min_nn((float)(-dz64), hme);
                qme = fx_min_nn(ritter_q(g, dz), K.half_dx_over_dt * hp);
            }
        }
        lroe = false;
    }
    if (!n.wetN) {
        const double hm64 = fr.dp - 0.5 * fr.dzbN;
        hn = 0.f; qn = 0.f; rn = 0.f;
        qmn = 0.f; rmn = 0.f; hmn = 0.f;
        if (hm64 > K.hdry64) {
            hmn = (float)hm64;
            const double dz64 = (fr.dn - fr.dp) + fr.dzbN;
            if (dz64 < 0.0) {
                const float dz = fx_min_nn((float)(-dz64), hmn);
                rmn = fx_min_nn(ritter_q(g, dz), K.half_dx_over_dt * hp);
            }
        }
        lroe = false;
    }
    // one reciprocal square root per face, r = 1/sqrt(max(hm, hdry)), gives
    // 1/max(hm, hdry) = r^2, the celerity sqrt(g hm) = sqrt(g) hm r and the
    // Roe scale 1/(4c) = r/(4 sqrt(g)) where hm > hdry; where hm <= hdry the
    // celerity stays finite and meets a zero scale
    const float rse = fx_rsq(fx_max_nn(hme, hdry)), rsn = fx_rsq(fx_max_nn(hmn, hdry));
    const float cme = K.sqrt_g * (hme * rse), cmn = K.sqrt_g * (hmn * rsn);
    const float ihme = rse * rse, ihmn = rsn * rsn;
    const float ume = qme * ihme, vme = rme * ihme, umn = qmn * ihmn, vmn = rmn * ihmn;
    // eddy coefficients: us is identically zero, so cvdef*us*h + nuem = nuem;
    // a dry neighbour contributes h = 0 and velocity 0 (:124-126,151-153)
    const float hne_t = K.nuem_idx * n.sqP * (n.wetE ? n.sqE : 0.f);       // with the 1/dx of the gradients
    const float hnn_t = K.nuem_idx * n.sqP * (n.wetN ? n.sqN : 0.f);
    const float un = n.wetN ? n.uN : 0.f;
    const float ve = n.wetE ? n.vE : 0.f;
    const float idx = K.idx;
    const float txye = hne_t * ((ve - n.vP) + .25f * (un + n.uNE - n.uS - n.uSE));
    const float txyn = hnn_t * ((un - n.uP) + .25f * (ve + n.vNE - n.vW - n.vNW));
    float fe1 = qme, fe2 = qme * ume, fe3 = qme * vme - txye;
    float fn1 = rmn, fn2 = rmn * umn - txyn, fn3 = rmn * vmn;
    if (lroe) {
        // Roe dissipation -1/2 |A| dU (solver_wetdomain.cpp:218-277) in
        // characteristic form: the strengths w1, w3 of the two gravity waves
        // (speeds c3 = u - c and c2 = u + c, halved through cc = 1/(4c)) and of
        // the shear wave (speed u, halved through hs) along the eigenvectors
        // (1, c3, v), (1, c2, v) and (0, 0, 1).  It equals the reference's
        // matrix form term by term, in about half the operations.
        float cc = 0.f, hs = 0.f;
        if (hme > hdry) {
            float dzea = fabsf(dze);
            if (dzea > 0.5f * hme) dze = copysignf(fx_min_nn(dzea, fabsf(he - hp)), dze);
            cc = K.q_isg * rse;
            hs = .5f;
        }
        {
            const float c2 = ume + cme, c3 = ume - cme;
            const float w1 = fabsf(c3) * ((c2 * dze - dqe) * cc);
            const float w3 = fabsf(c2) * ((dqe - c3 * dze) * cc);
            const float m1 = w1 + w3;
            fe1 -= m1;
            fe2 -= c3 * w1 + c2 * w3;
            fe3 -= vme * m1 + hs * fabsf(ume) * (dre - vme * dze);
        }
        cc = 0.f;
        hs = 0.f;
        if (hmn > hdry) {
            dzn = copysignf(fx_min_nn(fabsf(dzn), fabsf(hn - hp)), dzn);
            cc = K.q_isg * rsn;
            hs = .5f;
        }
        {
            const float c2 = vmn + cmn, c3 = vmn - cmn;
            const float w1 = fabsf(c3) * ((c2 * dzn - drn) * cc);
            const float w3 = fabsf(c2) * ((drn - c3 * dzn) * cc);
            const float m1 = w1 + w3;
            fn1 -= m1;
            fn2 -= umn * m1 + hs * fabsf(vmn) * (dqn - umn * dzn);
            fn3 -= c3 * w1 + c2 * w3;
        }
    }
    fe2 += .5f * g * (hme * hme);
    fn3 += .5f * g * (hmn * hmn);
    // outflow limited to the water in the cell (solver_wetdomain.cpp:309-332)
    const float volrat = K.arbase_over_dt * hp;
    if (volrat > 0.f) {
        if (fe1 > 0.f && fn1 > 0.f) {
            const float tot = (fe1 + fn1) * dx;
            if (tot > volrat) {
                const float cf = fn1 * fx_rcp(fe1 + fn1);
                fe1 = (1.f - cf) * volrat * idx;
                fn1 = cf * volrat * idx;
            }
        } else if (fe1 > 0.f) {
            fe1 = fx_min_nn(fe1 * dx, volrat) * idx;
        } else if (fn1 > 0.f) {
            fn1 = fx_min_nn(fn1 * dx, volrat) * idx;
        }
    } else {
        fe1 = 0.f; fn1 = 0.f;
    }
    f[0] = fe1; f[1] = fe2; f[2] = fe3; f[3] = fn1; f[4] = fn2; f[5] = fn3;
}

// CFL bound of one updated cell (main.cpp:308-316).  The cell's bound is
// cfl*dxy/(max(u0, v0) + c0), since min(a/(u0+c0), a/(v0+c0)) = a/(max(u0,v0)+c0)
// (division is monotonic); the wave speeds max(u0, v0) + c0 are reduced by
// their maximum, in FP32 (exact in any order, so the same for every tiling),
// and the clock divides once (cfl_dt).  h is the depth or the signed
// d = z - zb (the same above hdry > 0); the host holds hdry >= 1e-30, so hf is
// a normal FP32 number.  The largest depth (hpall, main.cpp:312) is kept in
// FP32 as well: the loop only tests it against zero (the solver gate), and
// the FP32 maximum is exact in any order.
__device__ inline void cfl_cell(const fx_params &P, float sqrt_g, double h, float qx, float qy, float &smax,
                                float &hmx)
{
    if (h > P.hdry) {
        const float hf = (float)h;
        hmx = fx_max_nn(hmx, hf);
        const float r = fx_rsq(hf);                          // 1/h = r^2, sqrt(g h) = sqrt(g) h r
        const float c0 = sqrt_g * (hf * r);
        const float u = fx_max_nn(1e-6f, fx_max_nn(fabsf(qx), fabsf(qy)) * (r * r));
        smax = fx_max_nn(smax, u + c0);
    }
}

// ---------------------------------------------------------------------------
// The sweep.  Lane k of the block's single wave holds row r0-2+k of every
// column it visits; lanes 1..62 compute fluxes (rows r0-1..r0+60) and lanes
// 2..62 own and update rows r0..r0+60 (TH = 61).  The wave walks the columns
// c0..c0+CH-1; per column:
//   L(c+1): cells of column c+1 from the prefetch registers into the LDS ring
//           (melt, loop 1, cell velocities), then prefetch column c+2;
//   F(c):   fluxes of column c (needs columns c-1..c+1);
//   U(c):   update of the own rows of column c (needs fluxes of c-1 and c).
// Every lane does the same work in every phase (no divergent halo passes);
// each cell is read from HBM once per block and only three halo rows are
// shared with the neighbouring block.
// ---------------------------------------------------------------------------
#define TH FX32_TH
#define NS 64                 // rows held per column: r0-2 .. r0+61

struct pref {                 // one cell's HBM values in flight
    double d;
    double zb;
    float qx, qy, mult;
    uint8_t f;                // flag byte (kept a byte: a widening after the
                              // load would wait for the load there)
};

// Three columns (c-1, c, c+1) of what the flux stencil reads from the S
// neighbour (u; the wet flags are lane masks, cmask); columns c and c+1 of the rest, and the
// fluxes only where the update reads them: the E face of column c and the N
// faces of columns c-1 and c.  What two columns keep sits in one record per
// column parity.  7808 bytes per wave.
struct ring2 {
    float dm[NS], qx[NS], qy[NS], v[NS], sq[NS];
    double zb[NS];            // FP64 bed
    double dm64[NS];          // FP64 depth after melt
    float fn[3][NS];          // fn_1..fn_3 of this column
    uint8_t raw[NS];          // flag bytes
};
struct ring {
    float u[3][NS];
    float fe[3][NS];          // fe_1..fe_3 of column c
    ring2 p[2];               // by column parity
};

// the FP64 front of row i (columns of parity pc and pn = pc ^ 1): depths
// after melt and bed differences
__device__ inline front64 front_of(const ring &R, int pc, int pn, int i)
{
    front64 fr;
    fr.dp = R.p[pc].dm64[i];
    fr.de = R.p[pc].dm64[i + 1];
    fr.dn = R.p[pn].dm64[i];
    fr.dzbE = R.p[pc].zb[i + 1] - R.p[pc].zb[i];
    fr.dzbN = R.p[pn].zb[i] - R.p[pc].zb[i];
    return fr;
}

// flux m (0..5 = fe_1..fe_3, fn_1..fn_3) of row i of the column of parity par
__device__ inline float &rflux(ring &R, int par, int m, int i)
{
    return m < 3 ? R.fe[m][i] : R.p[par].fn[m - 3][i];
}

// per-block constants of the sweep
struct sweep_ctx {
    const fx_params *P;
    const s32p *S;
    const double *d_in;
    double *d_out;
    const float2 *q_in;
    float2 *q_out;
    consts32 K;
    double meteoi, inflowi, s_wet_prev;
    uint32_t inflow_p;        // linear index of the inflow cell, ~0u = none
    int inflow_c;             // the inflow cell's column if it is in this lane's row, else -1
    float dxy_over_dt, kf, dtf, dtf1, hdryf;     // dtf1 = dt / 10
    int gate, ade_on, k, row, mcols;
    bool row_ok, own;
    uint32_t ld;
};

__device__ inline void load_col(const sweep_ctx &X, int c, pref &a)
{
    a.f = 0;
    if (!X.row_ok || c < 0 || c >= X.mcols) return;      // column test is uniform
    const uint32_t p = (uint32_t)X.row + (uint32_t)c * X.ld;
    const uint32_t o4 = p * 4u;
    const s32p &S = *X.S;
    a.f = at_off(S.flags, p);
    a.d = at_off(X.d_in, p * 8u);
    const float2 q = at_off(X.q_in, p * 8u);
    a.zb = at_off(S.zb, p * 8u);
    a.qx = q.x;
    a.qy = q.y;
    a.mult = X.P->melt_mult ? at_off(X.P->melt_mult, o4) : 1.0f;
}

// L: this lane's cell of column col into ring slots SL (three-column arrays)
// and col & 1 (two-column record): melt of this step, hydro loop 1, cell
// velocities.  The inflow cell is found by its column, not by a bit of the
// prefetched flags: a test on a loaded value right after the load would wait
// for it there.
template <int SL>
__device__ inline void store_col(const sweep_ctx &X, ring &R, cmask &M, const pref &a, int col, bool own_col,
                                 double &ms)
{
    const int i = X.k, par = col & 1;
    const bool inner = a.f & FXF_INNER;
    // melt area that brings solute (cells holding water before the melt,
    // common.cpp:214-233), summed over the cells this lane owns
    if (own_col && X.own && inner && !(a.f & FXF_NODATA) && a.d > 0.0) ms += (double)a.mult;
    double d = inner ? a.d : 0.0;
    if (inner && !(a.f & FXF_NODATA)) d += X.meteoi * (double)a.mult;
    if (col == X.inflow_c) {                            // add_inflow after add_meteo (common.cpp:324-330)
        if (d <= 0.0) d = 0.0;
        d += X.inflowi;
    }
    const bool wet = d > X.P->hdry;                     // loop 1: h = max(0,d) > hdry
    float qx = inner ? a.qx : 0.f, qy = inner ? a.qy : 0.f;
    if (X.gate && !wet) { qx = 0.f; qy = 0.f; }
    const float dm = (float)d;
    const float h = fx_max0(dm);
    const float den = fx_rcp(fx_max_nn(fx_max_nn(h, X.hdryf), X.K.ks));
    R.p[par].dm[i] = dm;
    R.p[par].qx[i] = qx;
    R.p[par].qy[i] = qy;
    R.u[SL][i] = qx * den;
    R.p[par].v[i] = qy * den;
    R.p[par].sq[i] = fx_sqrt(h);
    R.p[par].zb[i] = inner ? a.zb : 0.0;
    M.w[SL] = __builtin_amdgcn_ballot_w64(wet);
    M.a[SL] = __builtin_amdgcn_ballot_w64(inner && !(a.f & FXF_WEIR));
    R.p[par].raw[i] = (uint8_t)a.f;
    R.p[par].dm64[i] = d;
}

// F(c): fluxes of this lane's row; column c in slots S0 (three-column arrays)
// and pc = c & 1 (two-column record), c-1 in SS, c+1 in SN and pc ^ 1
template <int S0, int SS, int SN>
__device__ inline void flux_col(const sweep_ctx &X, ring &R, const cmask &M, int c)
{
    const int i = X.k;
    float f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (X.gate && lane_bit(M.a[S0] & FLUX_LANES)) {
        const bool wetE = lane_bit(M.w[S0] >> 1), wetN = lane_bit(M.w[SN]);
        const int pc = c & 1, pn = pc ^ 1;     // slots of columns c and c+1
        if (!lane_bit(M.w[S0])) {
            if (!lane_bit((M.w[S0] << 1) | (M.w[S0] >> 1) | M.w[SS] | M.w[SN])) {
                if (R.p[pc].raw[i] & FXF_STALE) {  // stale fluxes of an all-dry neighbourhood
                    const uint32_t p = (uint32_t)X.row + (uint32_t)c * X.ld;
                    const FX_GLOBAL float *s6 = (const FX_GLOBAL float *)(X.S->f6 + (size_t)p * 6);
                    for (int m = 0; m < 6; m++) f[m] = s6[m];
                }
            } else {
                solver_dry32(X.K, front_of(R, pc, pn, i), wetE, wetN, X.dxy_over_dt, f);
            }
        } else {
            const float dzbE = (float)(R.p[pc].zb[i + 1] - R.p[pc].zb[i]);
            const float dzbN = (float)(R.p[pn].zb[i] - R.p[pc].zb[i]);
            front64 fr;
            if (!wetE || !wetN) fr = front_of(R, pc, pn, i);
            nb32 n;
            n.dp = R.p[pc].dm[i]; n.de = R.p[pc].dm[i + 1]; n.dn = R.p[pn].dm[i];
            n.dzbE = dzbE; n.dzbN = dzbN;
            n.qp = R.p[pc].qx[i]; n.qe = R.p[pc].qx[i + 1]; n.qn = R.p[pn].qx[i];
            n.rp = R.p[pc].qy[i]; n.re = R.p[pc].qy[i + 1]; n.rn = R.p[pn].qy[i];
            n.uP = R.u[S0][i]; n.vP = R.p[pc].v[i];
            n.uN = R.u[SN][i]; n.uS = R.u[SS][i];
            n.uSE = R.u[SS][i + 1]; n.uNE = R.u[SN][i + 1];
            n.vE = R.p[pc].v[i + 1]; n.vW = R.p[pc].v[i - 1];
            n.vNW = R.p[pn].v[i - 1]; n.vNE = R.p[pn].v[i + 1];
            n.sqP = R.p[pc].sq[i]; n.sqE = R.p[pc].sq[i + 1]; n.sqN = R.p[pn].sq[i];
            n.wetE = wetE; n.wetN = wetN;
            solver_wet32(X.K, n, fr, f);
        }
    }
    const int fs = c & 1;
    for (int m = 0; m < 6; m++) rflux(R, fs, m, i) = f[m];
}

// U(c): hydro loops 3 and 4 for this lane's own cell, and its wave speed
template <int S0, int SS, int SN>
__device__ inline void update_col(const sweep_ctx &X, ring &R, const cmask &M, int c, float &smax, float &hmx,
                                  double &wv, bool &bad)
{
    if (!X.own) return;
    const fx_params &P = *X.P;
    const s32p &S = *X.S;
    const int i = X.k, fs = c & 1, fsS = fs ^ 1;
    const uint32_t p = (uint32_t)X.row + (uint32_t)c * X.ld;
    const uint32_t o4 = p * 4u;
    const uint8_t f = R.p[fs].raw[i];
    const bool weir = !lane_bit(M.a[S0]);             // own cells are interior: weir = no fluxes
    double d = R.p[fs].dm64[i];
    float qxn, qyn;
    uint8_t nf = f;
    float fe1P = 0.f, fn1P = 0.f;
    if (!X.gate) {
        qxn = R.p[fs].qx[i];
        qyn = R.p[fs].qy[i];
        // no solver step, so every cell was at or below hdry at the CFL block,
        // which marks it dry (main.cpp:300-325); add_meteo, when there is a
        // melt table, then marks wet each cell its melt leaves at or below
        // hdry (common.cpp:219-222), and add_inflow sets its cell from its own
        // depth (common.cpp:333-337)
        const double hm = fmax(0.0, d);
        const bool dry = (p == X.inflow_p) ? hm <= P.hdry
                                           : (!P.meteo || (f & FXF_NODATA) || hm > P.hdry);
        nf = dry ? (uint8_t)(nf | FXF_DRY4) : (uint8_t)(nf & ~FXF_DRY4);
    } else {
        const float kf = X.kf, dtf = X.dtf;
        const float fe1W = rflux(R, fs, 0, i - 1), fn1S = rflux(R, fsS, 3, i);
        fe1P = rflux(R, fs, 0, i);
        fn1P = rflux(R, fs, 3, i);
        float dqx = 0.f, dqy = 0.f;
        if (!weir) {
            const float a1 = fe1W * kf, b1 = fe1P * kf, c1 = fn1S * kf, e1 = fn1P * kf;
            d += ((double)a1 - (double)b1) + ((double)c1 - (double)e1);
            dqx = ((rflux(R, fs, 1, i - 1) - rflux(R, fs, 1, i)) + (rflux(R, fsS, 4, i) - rflux(R, fs, 4, i))) * kf;
            dqy = ((rflux(R, fs, 2, i - 1) - rflux(R, fs, 2, i)) + (rflux(R, fsS, 5, i) - rflux(R, fs, 5, i))) * kf;
            // into / out of a weir cell: the neighbours of a non-weir cell
            // are interior cells (the outer interior ring is weir), so a
            // neighbour that computes no fluxes is a weir cell
            if (!lane_bit(M.a[S0] >> 1)) wv -= (double)b1;
            if (!lane_bit(M.a[SN])) wv -= (double)e1;
        }
        // h = max(0, d) against hdry > 0 is d against hdry
        // the reference's loop-4 wet flag (hydrodynamics_calc.cpp:123-154)
        const bool dry4 = d < P.hdry || weir;
        nf = dry4 ? (uint8_t)(nf | FXF_DRY4) : (uint8_t)(nf & ~FXF_DRY4);
        if (dry4) {
            qxn = 0.f;
            qyn = 0.f;
        } else {
            // 0.1 of each face's discharge and 0.8 of the updated cell's
            // (hydrodynamics_calc.cpp:145-150), the two faces summed first
            qxn = (fe1W + fe1P) * X.dtf1 + .8f * (R.p[fs].qx[i] + dqx);
            qyn = (fn1S + fn1P) * X.dtf1 + .8f * (R.p[fs].qy[i] + dqy);
        }
        // keep the fluxes of a cell that may be all-dry next step (S1)
        const bool wet1 = lane_bit(M.w[S0]);
        if (!weir && (d <= P.hdry || !P.rates_nonneg)) {
            const bool alldry =
                !lane_bit(M.w[S0] | (M.w[S0] << 1) | (M.w[S0] >> 1) | M.w[SS] | M.w[SN]);
            if (!alldry) {
                bool nz = false;
                for (int m = 0; m < 6; m++) nz |= rflux(R, fs, m, i) != 0.f;
                if (nz) {
                    FX_GLOBAL float *s6 = (FX_GLOBAL float *)(S.f6 + (size_t)p * 6);
                    for (int m = 0; m < 6; m++) s6[m] = rflux(R, fs, m, i);
                    nf |= FXF_STALE;
                } else {
                    nf &= ~FXF_STALE;
                }
            }
        }
        // wet-time tracer: runs are opened / closed on transitions only
        if (wet1 != (bool)(f & FXF_WETPREV)) {
            if (wet1) {
                at_off(reinterpret_cast<float *>(S.tw), p * 8u + 4u) = (float)X.s_wet_prev;
                nf |= FXF_WETPREV;
            } else {
                const float2 w = at_off(S.tw, p * 8u);
                at_off(reinterpret_cast<float *>(S.tw), p * 8u) = w.x + (float)(X.s_wet_prev - (double)w.y);
                nf &= ~FXF_WETPREV;
            }
        }
    }
    // a non-finite depth or discharge is reported (ctl->status bit 2): the
    // wet/dry tests below would otherwise read NaN as a dry cell.  Weir cells
    // are left out: they never compute, and initiation leaves the bed of the
    // far corner NaN (0/0 of four NODATA vertices, initiate.cpp:85), so its
    // water surface is NaN in the reference too
    bad |= !weir && (!fx_finite(d) || !fx_finitef(qxn) || !fx_finitef(qyn));
    at_off(X.d_out, p * 8u) = d;
    at_off(X.q_out, p * 8u) = make_float2(qxn, qyn);
    if (X.ade_on) {
        at_off(S.ff, p * 8u) = make_float2(fe1P, fn1P);
    }
    if (nf != f) at_off(S.flags, p) = nf;
    cfl_cell(P, X.K.sqrt_g, d, qxn, qyn, smax, hmx);
}

// one column of the sweep; SC = ring slot of column c in the three-column
// arrays (column c0 + j uses j % 3); its two-column record is c & 1
template <int SC>
__device__ inline void column_step(const sweep_ctx &X, ring &R, cmask &M, int c, int c_last, pref &a,
                                   float &smax, float &hmx, double &wv, double &ms, bool &bad)
{
    constexpr int SN = (SC + 1) % 3, SS = (SC + 2) % 3;
    __syncthreads();
    store_col<SN>(X, R, M, a, c + 1, c + 1 <= c_last, ms);   // column c+1 (prefetched last step)
    if (c + 2 <= c_last + 1) load_col(X, c + 2, a);
    __syncthreads();
    flux_col<SC, SS, SN>(X, R, M, c);
    __syncthreads();
    update_col<SC, SS, SN>(X, R, M, c, smax, hmx, wv, bad);
}

__device__ inline float wmaxf_nn(float v)            // v >= 0
{
    for (int o = 32; o > 0; o >>= 1) v = fx_max_nn(v, __shfl_xor(v, o, 64));
    return v;
}

// one block of the sweep: strip bx (rows), column block by; gx = strips per
// column block (the block's partials go to slot bx + by * gx)
__device__ inline void sweep_tile(const fx_params &P, const s32p &S, fx_ctl *ctl, int ade_on, int bx, int by,
                                  int gx, ring &R)
{
    if (ctl->halt) return;
    sweep_ctx X;
    const int cur = ctl->cur;
    X.P = &P;
    X.S = &S;
    X.d_in = S.d_in[cur];
    X.d_out = const_cast<double *>(S.d_in[cur ^ 1]);
    X.q_in = S.q_in[cur];
    X.q_out = const_cast<float2 *>(S.q_in[cur ^ 1]);
    X.meteoi = P.meteo ? ctl->meteoi : 0.0;
    X.inflowi = ctl->inflowi;
    X.inflow_p = P.inflow ? (uint32_t)P.inflow_p : ~0u;
    X.gate = ctl->gate;
    X.ade_on = ade_on;
    X.s_wet_prev = ctl->s_wet_prev;
    X.k = threadIdx.x;
    const int r0 = 1 + bx * TH;
    const int c0 = 1 + by * P.ch;
    const int c_last = min(c0 + P.ch - 1, P.ncols);
    X.row = r0 - 2 + X.k;                             // this lane's row
    X.row_ok = X.row >= 0 && X.row < P.ld;
    X.own = X.k >= 2 && X.k <= NS - 2 && X.row <= P.nrows;
    X.inflow_c = -1;                                  // the inflow cell is interior (host-checked)
    if (P.inflow && X.row == (int)(X.inflow_p % (uint32_t)P.ld)) X.inflow_c = (int)(X.inflow_p / (uint32_t)P.ld);
    X.mcols = P.ncols + 2;
    X.ld = (uint32_t)P.ld;
    X.hdryf = (float)P.hdry;
    X.K.hdry = (float)P.hdry;
    X.K.hdry64 = P.hdry;
    X.K.g = (float)P.g;
    X.K.sqrt_g = fx_sqrt(X.K.g);
    X.K.q_isg = .25f * fx_rcp(X.K.sqrt_g);
    X.K.nuem = (float)P.nuem;
    X.K.ks = (float)P.ks0;
    X.K.dx = (float)P.dxy;
    X.K.idx = 1.f / X.K.dx;
    X.K.nuem_idx = X.K.nuem * X.K.idx;
    X.K.half_dx_over_dt = ctl->k_half_dx_dt;
    X.K.arbase_over_dt = ctl->k_arbase_dt;
    X.dxy_over_dt = ctl->k_dxy_dt;
    X.kf = ctl->k_dt_dxy;
    X.dtf = ctl->k_dt;
    X.dtf1 = .1f * X.dtf;

    pref a;
    float smax = 0.f;
    float hmx = 0.f;
    double wv = 0.0, ms = 0.0;
    bool bad = false;
    // prologue: columns c0-2 (slot 1), c0-1 (slot 2), c0 (slot 0), fluxes of c0-1
    cmask M;
    load_col(X, c0 - 2, a); store_col<1>(X, R, M, a, c0 - 2, false, ms);
    load_col(X, c0 - 1, a); store_col<2>(X, R, M, a, c0 - 1, false, ms);
    load_col(X, c0, a); store_col<0>(X, R, M, a, c0, true, ms);
    load_col(X, c0 + 1, a);
    __syncthreads();
    flux_col<2, 1, 0>(X, R, M, c0 - 1);
    // column c uses ring slot (c - c0) % 3, unrolled by three
    int c = c0;
    for (; c + 2 <= c_last; c += 3) {
        column_step<0>(X, R, M, c, c_last, a, smax, hmx, wv, ms, bad);
        column_step<1>(X, R, M, c + 1, c_last, a, smax, hmx, wv, ms, bad);
        column_step<2>(X, R, M, c + 2, c_last, a, smax, hmx, wv, ms, bad);
    }
    if (c <= c_last) { column_step<0>(X, R, M, c, c_last, a, smax, hmx, wv, ms, bad); c++; }
    if (c <= c_last) { column_step<1>(X, R, M, c, c_last, a, smax, hmx, wv, ms, bad); c++; }
    smax = wmaxf_nn(smax);
    hmx = wmaxf_nn(hmx);
    wv = wsum(wv);
    ms = wsum(ms);
    if (__any(bad) && X.k == 0) atomicOr(&ctl->status, 4);
    if (X.k == 0) {
        const size_t bid = bx + (size_t)by * gx;
        S.partials[FX32_NPV * bid] = (double)smax;
        S.partials[FX32_NPV * bid + 1] = (double)hmx;
        S.partials[FX32_NPV * bid + 2] = wv;
        S.partials[FX32_NPV * bid + 3] = ms;
    }
}

__global__ __launch_bounds__(64) void k32_sweep(fx_params P, s32p S, fx_ctl *ctl, int ade_on, int gx, int n,
                                                int chunk)
{
    __shared__ ring R;
    const int t = fx_xcd_tile((int)blockIdx.x, n, chunk);
    sweep_tile(P, S, ctl, ade_on, t % gx, t / gx, gx, R);
}

// batched: block b runs block FX_BMAP_BLK(tmap[b]) of member FX_BMAP_SLOT(tmap[b])
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k32b_sweep(const fx_bslot *__restrict__ slots, const uint32_t *__restrict__ tmap)
{
    __shared__ ring R;
    const uint32_t m = tmap[blockIdx.x];
    const fx_bslot &B = slots[FX_BMAP_SLOT(m)];
    const int blk = (int)FX_BMAP_BLK(m);
    sweep_tile(B.P, B.S, B.ctl, B.ade_on, blk % B.gx, blk / B.gx, B.gx, R);
}

// CFL partials of the current state (start of a device-resident run), same
// tiling as k32_sweep so the partials line up
__global__ __launch_bounds__(64) void k32_cfl(fx_params P, s32p S, fx_ctl *ctl)
{
    const int cur = ctl->cur;
    const double *d_in = S.d_in[cur];
    const float2 *q_in = S.q_in[cur];
    const int k = threadIdx.x;
    const int r0 = 1 + blockIdx.x * TH, c0 = 1 + blockIdx.y * P.ch;
    const int row = r0 - 2 + k;
    const float sqrt_g = fx_sqrt((float)P.g);
    float smax = 0.f, hmx = 0.f;
    if (k >= 2 && k <= NS - 2 && row <= P.nrows) {
        for (int c = c0; c <= min(c0 + P.ch - 1, P.ncols); c++) {
            const uint32_t p = (uint32_t)row + (uint32_t)c * (uint32_t)P.ld;
            cfl_cell(P, sqrt_g, d_in[p], q_in[p].x, q_in[p].y, smax, hmx);
        }
    }
    smax = wmaxf_nn(smax);
    hmx = wmaxf_nn(hmx);
    if (k == 0) {
        const size_t bid = blockIdx.x + (size_t)blockIdx.y * gridDim.x;
        S.partials[FX32_NPV * bid] = (double)smax;
        S.partials[FX32_NPV * bid + 1] = (double)hmx;
    }
}

// Clock of one loop iteration (main.cpp:271-325, 385-387): fold the CFL
// partials left by the previous step, clip to the print time, advance tim,
// evaluate the forcing, gate the solver, update the FP64 water ledger.
//
// clock_fold: one block's max wave speed / max hpall / weir-volume sum over partials
// [lo, hi), in a fixed order (the weir entries are consumed: set to 0).  All
// loads come before the stores, so a thread's loads are in flight together.
__device__ inline void clock_fold(double *partials, int lo, int hi, double &a, double &b, double &c, double &m)
{
    const int tid = threadIdx.x;
    a = 0.0; b = 0.0; c = 0.0; m = 0.0;
#pragma unroll 4
    for (int i = lo + tid; i < hi; i += blockDim.x) {
        const double2 *q = reinterpret_cast<const double2 *>(partials + FX32_NPV * (size_t)i);
        const double2 x = q[0], y = q[1];
        a = fmax(a, x.x);
        b = fmax(b, x.y);
        c += y.x;
        m += y.y;
    }
    for (int i = lo + tid; i < hi; i += blockDim.x)
        reinterpret_cast<double2 *>(partials + FX32_NPV * (size_t)i)[1] = make_double2(0.0, 0.0);
    __shared__ double sm[4][16];
    a = wmax(a); b = wmax(b); c = wsum(c); m = wsum(m);
    if ((tid & 63) == 0) { sm[0][tid >> 6] = a; sm[1][tid >> 6] = b; sm[2][tid >> 6] = c; sm[3][tid >> 6] = m; }
    __syncthreads();
    if (tid == 0)
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            a = fmax(a, sm[0][w]); b = fmax(b, sm[1][w]); c += sm[2][w]; m += sm[3][w];
        }
}

// the end of a batch member's step (buffer flip, print / end / NaN halt: a
// NaN/Inf member halts alone), done by the next step's clock or, after a
// run's last step, by k32b_end
__device__ inline void batch_step_close(fx_ctl *ctl)
{
    if (ctl->flip_pending) {
        ctl->flip_pending = 0;
        ctl->cur ^= 1;
        ctl->halt = ctl->pending_halt | ((ctl->status >> 2) & 1);
    }
}

// The end of a step (the buffer flip and the print / end / NaN halt), as the
// next clock finds it: the device loop leaves it to the step after, one
// launch fewer per step.  Returns whether the loop goes on; it also halts at
// the end time (main.cpp:271).
__device__ inline bool step_close_and_check(fx_ctl &C)
{
    if (C.flip_pending) {
        C.flip_pending = 0;
        C.cur ^= 1;
        C.halt = C.pending_halt | ((C.status >> 2) & 1);      // halt on NaN/Inf
    }
    if (C.halt) return false;
    if (C.tim > C.t_end) {
        C.halt = 1;
        return false;
    }
    return true;
}

// the control block into / out of a block's shared copy, word by word
__device__ inline void ctl_load(fx_ctl &C, const fx_ctl *ctl)
{
    static_assert(sizeof(fx_ctl) % 8 == 0, "fx_ctl is copied in 8-byte words");
    const uint64_t *g = reinterpret_cast<const uint64_t *>(ctl);
    uint64_t *l = reinterpret_cast<uint64_t *>(&C);
    for (int k = threadIdx.x; k < (int)(sizeof(fx_ctl) / 8); k += blockDim.x) l[k] = g[k];
}
__device__ inline void ctl_store(fx_ctl *ctl, const fx_ctl &C)
{
    uint64_t *g = reinterpret_cast<uint64_t *>(ctl);
    const uint64_t *l = reinterpret_cast<const uint64_t *>(&C);
    for (int k = threadIdx.x; k < (int)(sizeof(fx_ctl) / 8); k += blockDim.x) g[k] = l[k];
}

// One domain.  Every block reads the control block and decides the same way
// whether the loop goes on, before any partial is consumed; nothing writes
// the control block before every block has read it.  With nrec > 0 (one
// block) the fast ADE launch of the previous step has folded the partials
// into nrec records, and the block folds those.  Otherwise one block folds
// every partial, or several each fold a fixed chunk into part2 and the last
// block to finish (ticket) folds part2 in block order.  Either way the result
// does not depend on block timing.  The block that runs the clock works on a
// shared copy of the control block and writes it back; the forcing tables'
// times are loaded with it when they fit one entry per thread.
__global__ __launch_bounds__(1024) void k32_clock(fx_params P, double *partials, int nblk, double *part2,
                                                  uint32_t *ticket, const double *rec, int nrec, fx_ctl *ctl)
{
    __shared__ fx_ctl C;
    __shared__ int go_s;
    const bool held = (!P.meteo || P.meteo_nrows <= (int)blockDim.x) &&
                      (!P.inflow || P.inflow_nrows <= (int)blockDim.x);
    float mt = 0.f, it = 0.f;
    if (held && P.meteo && (int)threadIdx.x < P.meteo_nrows) mt = P.meteo[(size_t)threadIdx.x * P.meteo_ncols];
    if (held && P.inflow && (int)threadIdx.x < P.inflow_nrows) it = P.inflow[(size_t)threadIdx.x * P.inflow_ncols];
    ctl_load(C, ctl);
    __syncthreads();
    if (threadIdx.x == 0) go_s = step_close_and_check(C);   // the previous step closed
    __syncthreads();
    if (!go_s) {
        // written back once every block has read the control block: by the
        // only block, or by the last to take a ticket
        if (gridDim.x > 1) {
            __shared__ int last0;
            if (threadIdx.x == 0) last0 = atomicAdd(ticket, 1u) == gridDim.x - 1;
            __syncthreads();
            if (!last0) return;
            if (threadIdx.x == 0) *ticket = 0;
        }
        ctl_store(ctl, C);
        return;
    }
    double a, b, c, m;
    if (nrec > 0) {
        // the records (their partials are rewritten by the next sweep, so
        // none needs clearing)
        a = 0.0; b = 0.0; c = 0.0; m = 0.0;
        if ((int)threadIdx.x < nrec) {
            const double2 *q = reinterpret_cast<const double2 *>(rec + FX32_NPV * (size_t)threadIdx.x);
            const double2 x = q[0], y = q[1];
            a = x.x; b = x.y; c = y.x; m = y.y;
        }
        a = wmax(a); b = wmax(b); c = wsum(c); m = wsum(m);  // nrec <= 64: the first wave
        __shared__ double r4[4];
        if (threadIdx.x == 0) { r4[0] = a; r4[1] = b; r4[2] = c; r4[3] = m; }
        __syncthreads();
        a = r4[0]; b = r4[1]; c = r4[2]; m = r4[3];
    } else {
        const int chunk = (nblk + (int)gridDim.x - 1) / (int)gridDim.x;
        const int lo = min(nblk, (int)blockIdx.x * chunk), hi = min(nblk, lo + chunk);
        clock_fold(partials, lo, hi, a, b, c, m);
    }
    if (gridDim.x > 1) {
        __shared__ int last;
        if (threadIdx.x == 0) {
            part2[FX32_NPV * blockIdx.x] = a;
            part2[FX32_NPV * blockIdx.x + 1] = b;
            part2[FX32_NPV * blockIdx.x + 2] = c;
            part2[FX32_NPV * blockIdx.x + 3] = m;
            __threadfence();
            last = atomicAdd(ticket, 1u) == gridDim.x - 1;
        }
        __syncthreads();
        if (!last) return;
        // second level, by the whole block: the other blocks' part2 entries are
        // read past this CU's cache (volatile), one entry per thread
        __threadfence();
        double A = 0.0, B = 0.0, Cw = 0.0, Mw = 0.0;
        for (int k = threadIdx.x; k < (int)gridDim.x; k += blockDim.x) {
            const volatile double *q = part2 + FX32_NPV * (size_t)k;
            A = fmax(A, q[0]);
            B = fmax(B, q[1]);
            Cw += q[2];
            Mw += q[3];
        }
        __shared__ double sm2[4][16];
        A = wmax(A); B = wmax(B); Cw = wsum(Cw); Mw = wsum(Mw);
        const int tid = threadIdx.x;
        if ((tid & 63) == 0) { sm2[0][tid >> 6] = A; sm2[1][tid >> 6] = B; sm2[2][tid >> 6] = Cw; sm2[3][tid >> 6] = Mw; }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
                A = fmax(A, sm2[0][w]); B = fmax(B, sm2[1][w]); Cw += sm2[2][w]; Mw += sm2[3][w];
            }
            *ticket = 0;
        }
        a = A; b = B; c = Cw; m = Mw;
    }
    clock_finish_held(P, &C, a, b, c, m, held, mt, it);   // thread 0 holds the folded values
    if (threadIdx.x == 0) C.flip_pending = 1;          // this step's flip: the next clock's
    __syncthreads();
    ctl_store(ctl, C);
}

// the flip and halt of a device loop's last step, which the clock of a
// following step would do: once at the end of fx_advance
__global__ void k32_flip_end(fx_ctl *ctl)
{
    if (ctl->flip_pending) {
        ctl->flip_pending = 0;
        ctl->cur ^= 1;
        ctl->halt = ctl->pending_halt | ((ctl->status >> 2) & 1);     // halt on NaN/Inf
    }
}

// batched: one block per member
// One block per member, on a shared copy of its control block: the forcing
// tables' times are loaded with it (when they fit one entry per thread), and
// the copy is written back only if it changed (free slots share one halted
// control block)
__global__ __launch_bounds__(1024) void k32b_clock(const fx_bslot *__restrict__ slots)
{
    __shared__ fx_ctl C;
    __shared__ int go_s, dirty_s;
    const fx_bslot &B = slots[blockIdx.x];
    const fx_params &P = B.P;
    const bool held = (!P.meteo || P.meteo_nrows <= (int)blockDim.x) &&
                      (!P.inflow || P.inflow_nrows <= (int)blockDim.x);
    float mt = 0.f, it = 0.f;
    if (held && P.meteo && (int)threadIdx.x < P.meteo_nrows) mt = P.meteo[(size_t)threadIdx.x * P.meteo_ncols];
    if (held && P.inflow && (int)threadIdx.x < P.inflow_nrows) it = P.inflow[(size_t)threadIdx.x * P.inflow_ncols];
    ctl_load(C, B.ctl);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int was_pending = C.flip_pending, was_halt = C.halt;
        batch_step_close(&C);                          // the previous step's end
        int go = !C.halt;
        if (go && C.tim > C.t_end) {                   // loop condition, main.cpp:271
            C.halt = 1;
            go = 0;
        }
        go_s = go;
        dirty_s = go || was_pending || C.halt != was_halt;
    }
    __syncthreads();
    if (go_s) {
        double a, b, c, m;
        clock_fold(B.S.partials, 0, B.npart, a, b, c, m);
        clock_finish_held(P, &C, a, b, c, m, held, mt, it);
        if (threadIdx.x == 0) C.flip_pending = 1;      // this step's end: the next clock's
    }
    __syncthreads();
    if (dirty_s) ctl_store(B.ctl, C);
}

// WINTRAsolver_calc.cpp:26-57 on the step's new state (before the flip):
// a wet cell whose bed is not NODATA releases soil_rate/3600 of its soil mass
// per second into each solute in turn; the solute is held as mass per unit
// area, so c += dm / (h dxy^2) becomes m += dm / dxy^2
__global__ __launch_bounds__(256) void k32_wintra(fx_params P, a32p A, double *soil, fx_ctl *ctl)
{
    if (ctl->halt || !ctl->gate) return;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= A.mm) return;
    const uint8_t f = A.flags[i];
    if (!(f & FXF_INNER) || (f & FXF_NODATA)) return;
    const int nxt = ctl->cur ^ 1;
    const double hp = A.d[nxt][i];
    if (!(hp > P.hdry)) return;
    const double dtl = ctl->dtfl;
    double m = soil[i];
    for (int k = 0; k < P.nchem; k++) {
        const double deltam = m * P.soil_rate / 3600 * dtl;
        m = m - deltam;
        A.mass[nxt][k][i] += (float)(deltam / P.arbase);
    }
    soil[i] = m;
}

// batched end of a run: the last step's buffer flip and print/end halt of
// every member (k32_flip + k_step_end), and the compact status the host polls
__global__ void k32b_end(const fx_bslot *slots, int nslot, fx_bstat *stat)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslot) return;
    fx_ctl *ctl = slots[i].ctl;
    batch_step_close(ctl);                             // the last step's end
    fx_bstat st;
    st.tim = ctl->tim;
    st.dtfl = ctl->dtfl;
    st.it = ctl->it;
    st.steps = ctl->steps;
    st.solver_steps = ctl->solver_steps;
    st.halt = ctl->halt;
    st.status = ctl->status;
    st.pad = 0;
    stat[i] = st;
}

// water volume of non-weir interior cells (FP64, block partials)
__global__ __launch_bounds__(64) void k32_volume(fx_params P, s32p S, fx_ctl *ctl, double *part)
{
    const double *d_in = S.d_in[ctl->cur];
    const int k = threadIdx.x;
    const int r0 = 1 + blockIdx.x * TH, c0 = 1 + blockIdx.y * P.ch;
    const int row = r0 - 2 + k;
    double v = 0.0;
    if (k >= 2 && k <= NS - 2 && row <= P.nrows) {
        for (int c = c0; c <= min(c0 + P.ch - 1, P.ncols); c++) {
            const uint32_t p = (uint32_t)row + (uint32_t)c * (uint32_t)P.ld;
            if (!(S.flags[p] & FXF_WEIR)) v += d_in[p];
        }
    }
    v = wsum(v);
    if (k == 0) part[blockIdx.x + (size_t)blockIdx.y * gridDim.x] = v;
}

inline dim3 grid32(const fx_params &P)
{
    return dim3((P.nrows + TH - 1) / TH, (P.ncols + P.ch - 1) / P.ch);
}



} // namespace

s32p make_s32p(fx_domain d)
{
    fx_state32 &t = d->s32;
    s32p S;
    S.d_in[0] = t.d[0]; S.d_in[1] = t.d[1];
    S.q_in[0] = t.q[0]; S.q_in[1] = t.q[1];
    S.zb = t.zb;
    S.f6 = t.f6;
    S.ff = t.ff;
    S.flags = t.flags;
    S.tw = t.tw;
    S.partials = t.partials;
    S.mm = d->ncell;
    return S;
}

int fx32_npart(fx_domain d)
{
    fx_params P = fx_make_params(d);
    const dim3 g = grid32(P);
    return (int)(g.x * g.y);
}

int fx32b_launch_clock(hipStream_t st, const fx_bslot *slots, int nslot)
{
    hipLaunchKernelGGL(k32b_clock, dim3(nslot), dim3(1024), 0, st, slots);
    return fx_hip_check(hipGetLastError(), "k32b_clock");
}

int fx32b_launch_sweep(hipStream_t st, const fx_bslot *slots, const uint32_t *tmap, int nblocks)
{
    hipLaunchKernelGGL(k32b_sweep, dim3(nblocks), dim3(64), 0, st, slots, tmap);
    return fx_hip_check(hipGetLastError(), "k32b_sweep");
}

int fx32b_launch_end(hipStream_t st, const fx_bslot *slots, int nslot, fx_bstat *stat)
{
    hipLaunchKernelGGL(k32b_end, dim3((nslot + 63) / 64), dim3(64), 0, st, slots, nslot, stat);
    return fx_hip_check(hipGetLastError(), "k32b_end");
}

namespace {

template <typename T>
int dalloc32(fx_domain d, T **p, int64_t n)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)std::max<int64_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return fx_hip_check(e, "hipMalloc");
    e = hipMemsetAsync(q, 0, (size_t)std::max<int64_t>(n, 1) * sizeof(T), d->stream);
    if (e != hipSuccess) return fx_hip_check(e, "hipMemsetAsync");
    *p = (T *)q;
    return FX_OK;
}

int get_cur(fx_domain d, int *cur)
{
    FX_HIP(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(fx_ctl), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    *cur = d->ctl_host->cur;
    return FX_OK;
}

} // namespace

int fx32_init_state(fx_domain d)
{
    fx_state32 &t = d->s32;
    const int64_t mm = d->ncell;
    int rc;
    for (int k = 0; k < 2; k++) {
        if ((rc = dalloc32(d, &t.d[k], mm))) return rc;
        if ((rc = dalloc32(d, &t.q[k], mm))) return rc;
    }
    if ((rc = dalloc32(d, &t.zb, mm))) return rc;
    if ((rc = dalloc32(d, &t.f6, 6 * mm))) return rc;
    if ((rc = dalloc32(d, &t.ff, mm))) return rc;
    if ((rc = dalloc32(d, &t.flags, mm))) return rc;
    if ((rc = dalloc32(d, &t.tw, mm))) return rc;
    if (d->desc.ade_solver)
        for (int b = 0; b < 2; b++)
            for (int k = 0; k < d->desc.nchem; k++)
                if ((rc = dalloc32(d, &t.mass[b][k], mm))) return rc;
    fx_params P = fx_make_params(d);
    dim3 g = grid32(P);
    t.npart = (int64_t)g.x * g.y;
    // room for the partials of the smallest column block fx_set_tiling allows
    const int64_t cap = (int64_t)g.x * ((P.ncols + FX32_CH_MIN - 1) / FX32_CH_MIN);
    if ((rc = dalloc32(d, &t.partials, FX32_NPV * cap))) return rc;
    if ((rc = dalloc32(d, &t.part2, FX32_NPV * FX32_FOLD_MAX))) return rc;
    if ((rc = dalloc32(d, &t.ticket, 1))) return rc;
    if ((rc = dalloc32(d, &t.rec, FX32_NPV * FX32_REC))) return rc;
    return FX_OK;
}

// host-side conversions between the reference layout and the FP32-mixed state
int fx32_set_field(fx_domain d, int field, const void *host)
{
    fx_state32 &t = d->s32;
    const int64_t mm = d->ncell;
    int cur = 0, rc;
    if ((rc = get_cur(d, &cur))) return rc;
    if (field == FX_F_ZB) {          // from fx_set_static: FP32 bed, flags base
        const double *zb = (const double *)host;
        FX_HIP(hipMemcpy(t.zb, zb, mm * 8, hipMemcpyHostToDevice));
        return FX_OK;
    }
    if (field == -1) {               // weir mask (uint8 0/1) -> flag bytes
        const uint8_t *w = (const uint8_t *)host;
        std::vector<uint8_t> fl(mm, 0);
        double melt_area = 0.0;
        const double nod = d->desc.nodata;
        for (int icol = 1; icol <= d->desc.ncols; icol++)
            for (int irow = 1; irow <= d->desc.nrows; irow++) {
                const int64_t p = irow + (int64_t)icol * d->mrows;
                uint8_t f = FXF_INNER | FXF_DRY4;
                if (w[p]) f |= FXF_WEIR;
                if (d->zb_host[p] == nod) f |= FXF_NODATA;
                fl[p] = f;
            }
        FX_HIP(hipMemcpy(t.flags, fl.data(), mm, hipMemcpyHostToDevice));
        (void)melt_area;
        return FX_OK;
    }
    if (!d->has_static && field != FX_F_ZB) return fx_fail(FX_ERR_STATE, "fx_set_static first");
    switch (field) {
    case FX_F_Z: {
        const double *z = (const double *)host;
        std::vector<double> dv(mm, 0.0);
        for (int64_t i = 0; i < mm; i++) dv[i] = z[i] - d->zb_host[i];
        for (int k = 0; k < 2; k++)
            FX_HIP(hipMemcpy(t.d[k], dv.data(), mm * 8, hipMemcpyHostToDevice));
        return FX_OK;
    }
    case FX_F_QX:
    case FX_F_QY: {
        const double *q = (const double *)host;
        std::vector<float2> q2(mm);
        FX_HIP(hipMemcpy(q2.data(), t.q[cur], mm * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < mm; i++) {
            if (field == FX_F_QX) q2[i].x = (float)q[i];
            else q2[i].y = (float)q[i];
        }
        FX_HIP(hipMemcpy(t.q[cur], q2.data(), mm * 8, hipMemcpyHostToDevice));
        return FX_OK;
    }
    case FX_F_FE1: case FX_F_FE2: case FX_F_FE3: case FX_F_FN1: case FX_F_FN2: case FX_F_FN3: {
        const double *q = (const double *)host;
        const int m = field - FX_F_FE1;
        std::vector<float> f6(6 * mm), qf(mm);
        FX_HIP(hipMemcpy(f6.data(), t.f6, 6 * mm * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < mm; i++) f6[6 * i + m] = qf[i] = (float)q[i];
        FX_HIP(hipMemcpy(t.f6, f6.data(), 6 * mm * 4, hipMemcpyHostToDevice));
        // stored-flux flag bit
        std::vector<uint8_t> fl(mm);
        FX_HIP(hipMemcpy(fl.data(), t.flags, mm, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < mm; i++)
            if (qf[i] != 0.0f && (fl[i] & FXF_INNER) && !(fl[i] & FXF_WEIR)) fl[i] |= FXF_STALE;
        FX_HIP(hipMemcpy(t.flags, fl.data(), mm, hipMemcpyHostToDevice));
        return FX_OK;
    }
    case FX_F_TWET: {
        const double *q = (const double *)host;
        std::vector<float2> w(mm);
        for (int64_t i = 0; i < mm; i++) w[i] = make_float2((float)q[i], 0.0f);
        FX_HIP(hipMemcpy(t.tw, w.data(), mm * 8, hipMemcpyHostToDevice));
        std::vector<uint8_t> fl(mm);
        FX_HIP(hipMemcpy(fl.data(), t.flags, mm, hipMemcpyDeviceToHost));
        for (auto &f : fl) f &= ~FXF_WETPREV;
        FX_HIP(hipMemcpy(t.flags, fl.data(), mm, hipMemcpyHostToDevice));
        return FX_OK;
    }
    case FX_F_LDRY: {                // the flag bit the solver steps keep
        const float *ld = (const float *)host;
        std::vector<uint8_t> fl(mm);
        FX_HIP(hipMemcpy(fl.data(), t.flags, mm, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < mm; i++) {
            if (!(fl[i] & FXF_INNER)) continue;
            fl[i] = (uint8_t)(ld[i] == 1.0f ? (fl[i] | FXF_DRY4) : (fl[i] & ~FXF_DRY4));
        }
        FX_HIP(hipMemcpy(t.flags, fl.data(), mm, hipMemcpyHostToDevice));
        return FX_OK;
    }
    case FX_F_H: case FX_F_LDRY_PREV: case FX_F_H0:
        return FX_OK;                // derived from d in this state form
    default:
        if (field >= FX_F_CONC0 && field < FX_F_CONC0 + d->desc.nchem) {
            if (!d->desc.ade_solver) return FX_OK;
            const int k = field - FX_F_CONC0;
            const double *c = (const double *)host;
            std::vector<double> dv(mm);
            FX_HIP(hipMemcpy(dv.data(), t.d[cur], mm * 8, hipMemcpyDeviceToHost));
            std::vector<float> m(mm, 0.0f);
            for (int64_t i = 0; i < mm; i++) m[i] = (float)(c[i] * std::max(0.0, dv[i]));
            FX_HIP(hipMemcpy(t.mass[cur][k], m.data(), mm * 4, hipMemcpyHostToDevice));
            return FX_OK;
        }
        return fx_fail(FX_ERR_ARG, "unknown field");
    }
}

int fx32_get_field(fx_domain d, int field, void *host)
{
    fx_state32 &t = d->s32;
    const int64_t mm = d->ncell;
    int cur = 0, rc;
    if ((rc = get_cur(d, &cur))) return rc;
    std::vector<double> dv(mm);
    std::vector<uint8_t> fl(mm);
    auto load_d = [&]() -> int {
        FX_HIP(hipMemcpy(dv.data(), t.d[cur], mm * 8, hipMemcpyDeviceToHost));
        return FX_OK;
    };
    auto load_fl = [&]() -> int {
        FX_HIP(hipMemcpy(fl.data(), t.flags, mm, hipMemcpyDeviceToHost));
        return FX_OK;
    };
    double *out = (double *)host;
    switch (field) {
    case FX_F_ZB:
        std::copy(d->zb_host.begin(), d->zb_host.end(), out);
        return FX_OK;
    case FX_F_Z:
    case FX_F_H: {
        if ((rc = load_d())) return rc;
        for (int64_t i = 0; i < mm; i++) {
            const bool inner = (i % d->mrows) >= 1 && (i % d->mrows) <= d->desc.nrows &&
                               (i / d->mrows) >= 1 && (i / d->mrows) <= d->desc.ncols;
            if (field == FX_F_Z) out[i] = inner ? d->zb_host[i] + dv[i] : d->zb_host[i];
            else out[i] = inner ? std::max(0.0, dv[i]) : 0.0;
        }
        return FX_OK;
    }
    case FX_F_QX:
    case FX_F_QY: {
        std::vector<float2> q2(mm);
        FX_HIP(hipMemcpy(q2.data(), t.q[cur], mm * 8, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < mm; i++) out[i] = field == FX_F_QX ? q2[i].x : q2[i].y;
        return FX_OK;
    }
    case FX_F_LDRY: {                // as the last solver step left it
        if ((rc = load_fl())) return rc;
        float *o = (float *)host;
        for (int64_t i = 0; i < mm; i++)
            o[i] = (fl[i] & FXF_INNER) && (fl[i] & FXF_DRY4) ? 1.0f : 0.0f;
        return FX_OK;
    }
    case FX_F_LDRY_PREV: {           // derived from the current depth
        if ((rc = load_d())) return rc;
        if ((rc = load_fl())) return rc;
        float *o = (float *)host;
        for (int64_t i = 0; i < mm; i++) {
            if (!(fl[i] & FXF_INNER)) { o[i] = 0.0f; continue; }
            const double h = std::max(0.0, dv[i]);
            o[i] = (h < d->desc.hdry || (fl[i] & FXF_WEIR)) ? 1.0f : 0.0f;
        }
        return FX_OK;
    }
    case FX_F_H0: {
        if ((rc = load_d())) return rc;
        for (int64_t i = 0; i < mm; i++) out[i] = std::max(0.0, dv[i]);
        return FX_OK;
    }
    case FX_F_FE1: case FX_F_FE2: case FX_F_FE3: case FX_F_FN1: case FX_F_FN2: case FX_F_FN3: {
        const int m = field - FX_F_FE1;
        std::vector<float> f6(6 * mm);
        FX_HIP(hipMemcpy(f6.data(), t.f6, 6 * mm * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < mm; i++) out[i] = f6[6 * i + m];
        return FX_OK;
    }
    case FX_F_TWET: {
        if ((rc = load_fl())) return rc;
        std::vector<float2> w(mm);
        FX_HIP(hipMemcpy(w.data(), t.tw, mm * 8, hipMemcpyDeviceToHost));
        const double s = d->ctl_host->s_wet;
        for (int64_t i = 0; i < mm; i++)
            out[i] = w[i].x + ((fl[i] & FXF_WETPREV) ? s - (double)w[i].y : 0.0);
        return FX_OK;
    }
    default:
        if (field >= FX_F_CONC0 && field < FX_F_CONC0 + d->desc.nchem) {
            const int k = field - FX_F_CONC0;
            if (!d->desc.ade_solver) { std::fill(out, out + mm, -1.0); return FX_OK; }
            if ((rc = load_d())) return rc;
            std::vector<float> m(mm);
            FX_HIP(hipMemcpy(m.data(), t.mass[cur][k], mm * 4, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < mm; i++) {
                const double h = std::max(0.0, dv[i]);
                out[i] = h > 0.0 ? (double)m[i] / h : 0.0;
            }
            return FX_OK;
        }
        return fx_fail(FX_ERR_ARG, "unknown field");
    }
}

int fx32_launch_cfl_initial(fx_domain d)
{
    fx_params P = fx_make_params(d);
    s32p S = make_s32p(d);
    // melt_area for the ledger: sum of the multiplier over non-weir bed cells
    if (d->meteo && !d->melt_area_ok) {
        const int64_t mm = d->ncell;
        std::vector<uint8_t> fl(mm);
        FX_HIP(hipMemcpy(fl.data(), d->s32.flags, mm, hipMemcpyDeviceToHost));
        std::vector<float> mult;
        if (d->melt_mult) {
            mult.resize(mm);
            FX_HIP(hipMemcpy(mult.data(), d->melt_mult, mm * 4, hipMemcpyDeviceToHost));
        }
        double s = 0.0;
        for (int64_t i = 0; i < mm; i++)
            if ((fl[i] & FXF_INNER) && !(fl[i] & FXF_WEIR) && !(fl[i] & FXF_NODATA))
                s += mult.empty() ? 1.0 : (double)mult[i];
        FX_HIP(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(fx_ctl), hipMemcpyDeviceToHost, d->stream));
        FX_HIP(hipStreamSynchronize(d->stream));
        d->ctl_host->melt_area = s;
        FX_HIP(hipMemcpyAsync(d->ctl, d->ctl_host, sizeof(fx_ctl), hipMemcpyHostToDevice, d->stream));
        d->melt_area_ok = true;
    }
    fx_prof_begin(d, FX_K_CFL32);
    hipLaunchKernelGGL(k32_cfl, grid32(P), dim3(64), 0, d->stream, P, S, d->ctl);
    fx_prof_end(d, FX_K_CFL32);
    return fx_hip_check(hipGetLastError(), "k32_cfl");
}

namespace {
int env_int(const char *name, int dflt)
{
    const char *e = std::getenv(name);
    return (e && *e) ? std::atoi(e) : dflt;
}
} // namespace

// runs of four column blocks: as fast as one contiguous run per XCD at
// 4096^2 (0.603 against 0.605 ms per step; no remap: 0.652), and finer dealt
int fx32_xcd_chunk(int gx, int n)
{
    static const int cols = env_int("FLUXOS_XCD_COLS", 4);
    if (cols == 0) return 0;
    if (cols < 0) return (n + 7) / 8;                // one contiguous run per XCD
    return cols * gx;
}

// batches are dealt in runs of 32 tiles (FLUXOS_XCD_BATCH; 0 = round-robin):
// short enough that members differing in wet area load the XCDs evenly, long
// enough that a tile's strip neighbours, whose halo rows it re-reads, are
// mostly on its XCD (64 scenarios of 256^2-2048^2, 32.1 s with runs of 16, 32
// or 64 tiles, 32.5 s with 128, 33.8 s with 512, 33.1 s round-robin)
int fx32_xcd_chunk_batch()
{
    static const int t = env_int("FLUXOS_XCD_BATCH", 32);
    return t > 0 ? t : 0;
}

int fx32_launch_wintra(fx_domain d)
{
    fx_params P = fx_make_params(d);
    a32p A = make_a32p(d);
    const unsigned nb = (unsigned)((d->ncell + 255) / 256);
    hipLaunchKernelGGL(k32_wintra, dim3(nb), dim3(256), 0, d->stream, P, A, d->soil, d->ctl);
    return fx_hip_check(hipGetLastError(), "k32_wintra");
}

// One loop iteration of the device loop: its clock (which also closes the
// previous iteration), the step sweep, the fast ADE and WINTRA.  The
// iteration's own flip and halt are left to the next clock, or to
// fx32_launch_step_close at the end of the loop.
int fx32_launch_step(fx_domain d, bool first)
{
    fx_params P = fx_make_params(d);
    s32p S = make_s32p(d);
    const int ade_on = (d->desc.ade_solver && d->desc.ade_mode == FX_ADE_FAST) ? 1 : 0;
    // after an iteration of this loop the fast ADE has folded the partials
    // into records; else one block folds up to 16 partials per thread, and
    // past that blocks of 1024 fold chunks
    const int nrec = (!first && ade_on && P.nchem >= 1) ? std::min(FX32_REC, fx32_ade_blocks(d)) : 0;
    const int nfold = (nrec > 0 || d->s32.npart <= 16 * 1024)
                          ? 1
                          : (int)std::min<int64_t>(FX32_FOLD_MAX, (d->s32.npart + 1023) / 1024);
    hipLaunchKernelGGL(k32_clock, dim3(nfold), dim3(1024), 0, d->stream, P, S.partials,
                       (int)d->s32.npart, d->s32.part2, d->s32.ticket, (const double *)d->s32.rec, nrec, d->ctl);
    FX_HIP(hipGetLastError());
    fx_prof_begin(d, FX_K_STEP32);
    const dim3 g = grid32(P);
    const int n = (int)(g.x * g.y);
    hipLaunchKernelGGL(k32_sweep, dim3(n), dim3(64), 0, d->stream, P, S, d->ctl, ade_on, (int)g.x, n,
                       fx32_xcd_chunk((int)g.x, n));
    fx_prof_end(d, FX_K_STEP32);
    FX_HIP(hipGetLastError());
    if (ade_on) {
        int rc = fx32_launch_ade(d, true);
        if (rc) return rc;
    }
    if (d->wintra) {
        int rc = fx32_launch_wintra(d);
        if (rc) return rc;
    }
    return FX_OK;
}

// the flip and halt of the device loop's last iteration
int fx32_launch_step_close(fx_domain d)
{
    hipLaunchKernelGGL(k32_flip_end, dim3(1), dim3(1), 0, d->stream, d->ctl);
    return fx_hip_check(hipGetLastError(), "k32_flip_end");
}

int64_t fx32_volume_nparts(fx_domain d)
{
    const dim3 g = grid32(fx_make_params(d));
    return (int64_t)g.x * g.y;
}

// block partials of the water volume into part[0 .. fx32_volume_nparts)
int fx32_volume_partials(fx_domain d, double *part)
{
    fx_params P = fx_make_params(d);
    s32p S = make_s32p(d);
    hipLaunchKernelGGL(k32_volume, grid32(P), dim3(64), 0, d->stream, P, S, d->ctl, part);
    return fx_hip_check(hipGetLastError(), "k32_volume");
}

int fx32_launch_volume(fx_domain d, double *vol_out)
{
    fx_params P = fx_make_params(d);
    const size_t nb = (size_t)fx32_volume_nparts(d);
    double *part = nullptr;
    FX_HIP(hipMalloc((void **)&part, nb * sizeof(double)));
    int rc = fx32_volume_partials(d, part);
    if (rc) {
        (void)hipFree(part);
        return rc;
    }
    std::vector<double> h(nb);
    FX_HIP(hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    FX_HIP(hipFree(part));
    double s = 0.0, c = 0.0;
    for (double v : h) { double y = v - c; double t = s + y; c = (t - s) - y; s = t; }
    *vol_out = s * P.arbase;
    return FX_OK;
}

// weir-face exchange of the last step, not yet folded by k32_clock
namespace {
// solute mass per unit area (c h) summed over interior non-weir cells, block partials
__global__ __launch_bounds__(256) void k32_solute(const float *mass, const uint8_t *flags, int64_t mm, double *part)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < mm && (flags[i] & FXF_INNER) && !(flags[i] & FXF_WEIR)) v = (double)mass[i];
    v = wsum(v);
    __shared__ double sm[4];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}
} // namespace

int fx32_launch_solute(fx_domain d, int k, double *mass_out)
{
    FX_HIP(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(fx_ctl), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    const float *m = d->s32.mass[d->ctl_host->cur][k];
    const size_t nb = (size_t)((d->ncell + 255) / 256);
    double *part = nullptr;
    FX_HIP(hipMalloc((void **)&part, nb * sizeof(double)));
    hipLaunchKernelGGL(k32_solute, dim3((unsigned)nb), dim3(256), 0, d->stream, m, d->s32.flags, d->ncell, part);
    std::vector<double> h(nb);
    FX_HIP(hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    FX_HIP(hipFree(part));
    double s = 0.0, c = 0.0;
    for (double v : h) { double y = v - c; double t = s + y; c = (t - s) - y; s = t; }
    *mass_out = s * (double)d->desc.dxy * (double)d->desc.dxy;
    return FX_OK;
}

int fx32_pending(fx_domain d, double *weir_vol, double *solute_area)
{
    const size_t nb = (size_t)d->s32.npart;
    std::vector<double> h(FX32_NPV * nb);
    FX_HIP(hipMemcpyAsync(h.data(), d->s32.partials, FX32_NPV * nb * sizeof(double), hipMemcpyDeviceToHost,
                          d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    double w = 0.0, m = 0.0;
    for (size_t i = 0; i < nb; i++) {
        w += h[FX32_NPV * i + 2];
        m += h[FX32_NPV * i + 3];
    }
    if (weir_vol) *weir_vol = w * (double)d->desc.dxy * (double)d->desc.dxy;
    if (solute_area) *solute_area = m;
    return FX_OK;
}

int fx32_pending_weir(fx_domain d, double *vol)
{
    return fx32_pending(d, vol, nullptr);
}

// change the column-block lengths of the FP32 sweeps.  The pending weir-face
// exchange of the last step is held in the old tiling's partials: it is folded
// into the FP64 ledger first, and the partials are cleared.
int fx32_set_tiling(fx_domain d, int ch, int ach)
{
    if (ch == d->ch && ach == d->ach) return FX_OK;
    double pend = 0.0, area = 0.0;
    int rc;
    if ((rc = fx32_pending(d, &pend, &area))) return rc;
    FX_HIP(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(fx_ctl), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    fx_ctl *c = d->ctl_host;
    // the partials are cleared: fold what they hold into the ledgers first
    auto kahan = [](double &sum, double &comp, double v) {
        const double y = v - comp;
        const double t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    };
    kahan(c->led_weir, c->led_weir_c, pend);
    if (d->meteo && d->desc.ade_solver)
        for (int k = 0; k < d->desc.nchem; k++)
            kahan(c->sled[k], c->sled_c[k], c->meteoi * area * (double)d->desc.dxy * d->desc.dxy * c->cmelt[k]);
    FX_HIP(hipMemcpyAsync(d->ctl, c, sizeof(fx_ctl), hipMemcpyHostToDevice, d->stream));
    FX_HIP(hipMemsetAsync(d->s32.partials, 0, FX32_NPV * (size_t)d->s32.npart * sizeof(double), d->stream));
    d->ch = ch;
    d->ach = ach;
    fx_params P = fx_make_params(d);
    const dim3 g = grid32(P);
    d->s32.npart = (int64_t)g.x * g.y;
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}
