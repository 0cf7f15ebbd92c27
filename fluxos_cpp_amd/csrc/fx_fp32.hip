// fx_fp32.hip -- FP32-mixed path for gfx950: one fused, LDS-tiled stencil
// kernel per time step.
//
// State in HBM (per cell): depth relative to the bed d = z - zb in FP64
// (ping-pong), qx, qy in FP32 (ping-pong), bed zb in FP32, one flag byte, and
// -- touched only where needed -- the persistent fluxes of cells that may be
// "all dry" next step (the reference keeps stale fluxes there,
// solver_drydomain.cpp:60-76) and the wet-time tracer (written only when a
// cell changes between wet and dry).
//
// k32_step, one 64x8 tile per 256-thread block:
//   1. load the tile plus a 2-cell W/S and 1-cell E/N halo into LDS, applying
//      the melt of this step (common.cpp:210-263) and hydro loop 1
//      (hydrodynamics_calc.cpp:53-73) on the fly;
//   2. face fluxes of every non-weir cell of the tile and its W/S halo row and
//      column (solver_dry / solver_wet in FP32; solver_drydomain.cpp:27-194,
//      solver_wetdomain.cpp:27-342);
//   3. flux divergence + update with 0.1/0.8/0.1 face smoothing
//      (hydrodynamics_calc.cpp:93-154).  Face volume terms are FP32 products
//      summed exactly in FP64, so every face moves the same volume out of one
//      cell and into the other and the FP64 depth conserves water;
//   4. CFL bound of the updated state for the next step (main.cpp:300-323),
//      reduced per block in FP64; k32_clock folds the partials.
// The bed is held as an FP32 high part plus an FP16 low part: differences of
// neighbouring high parts are exact (Sterbenz) and the low parts restore the
// FP64 bed to ~6e-8 m, so the water-surface gradient of millimetre films over
// a 500 m datum keeps its accuracy.
#include "fx_params.h"

#pragma clang fp contract(fast)

#define TX 64
#define TY 8
#define LR (TX + 3)
#define LC (TY + 3)
#define FR (TX + 1)
#define FC (TY + 1)
#define NT 256

namespace {

struct s32p {
    const double *d_in[2];
    const float *qx_in[2], *qy_in[2];
    const float *zb;
    const _Float16 *zb_lo;
    float *f6;
    float *ff;
    uint8_t *flags;
    float *tw_acc, *tw_t0;
    double *partials;
    int64_t mm;
};

// LDS flag bits
#define LW 1u   // wet in loop 1
#define LX 2u   // weir
#define LI 4u   // interior
#define LS 8u   // stored fluxes nonzero

struct consts32 {
    float hdry, g, nuem, ks, dx;
    float half_dx_over_dt, arbase_over_dt;
};

// solver_dry in FP32 (solver_drydomain.cpp:78-193); pressure-only or Ritter
// inflow from a wet E / N neighbour.  dzbE = zb(E) - zb(P), dzbN likewise.
__device__ inline void solver_dry32(const consts32 &K, float dp, float de, float dn, float dzbE,
                                    float dzbN, bool wetE, bool wetN, float dxy_over_dt,
                                    float *f)
{
    const float g = K.g;
    float fe1 = 0.f, fe2 = 0.f, fn1 = 0.f, fn3 = 0.f;
    if (wetE) {
        // hme = max(0, ze - zbpe); ze <= zp test on the free surface
        float hme = fmaxf(0.f, de + 0.5f * dzbE);
        if (hme > K.hdry) {
            const float he = fmaxf(0.f, de);
            const float dze = (de - dp) + dzbE;
            if (dze <= 0.f) {
                fe2 = .5f * g * hme * hme;
            } else {
                const float dz = fminf(fabsf(dze), hme);
                float q = 0.296f * dz * sqrtf(g * dz);
                hme = 0.444f * dz;
                q = -fminf(q, dxy_over_dt * he);
                fe1 = q;
                fe2 = q * (q / hme) + .5f * g * hme * hme;
            }
        }
    }
    if (wetN) {
        float hmn = fmaxf(0.f, dn + 0.5f * dzbN);
        if (hmn > K.hdry) {
            const float hn = fmaxf(0.f, dn);
            const float dzn = (dn - dp) + dzbN;
            if (dzn <= 0.f) {
                fn3 = .5f * g * hmn * hmn;
            } else {
                const float dz = fminf(fabsf(dzn), hmn);
                float r = 0.296f * dz * sqrtf(g * dz);
                hmn = 0.444f * dz;
                r = -fminf(r, dxy_over_dt * hn);
                fn1 = r;
                fn3 = r * (r / hmn) + .5f * g * hmn * hmn;
            }
        }
    }
    f[0] = fe1; f[1] = fe2; f[2] = 0.f; f[3] = fn1; f[4] = 0.f; f[5] = fn3;
}

// solver_wet in FP32 (solver_wetdomain.cpp:27-342), same branch structure.
struct nb32 {
    float dp, de, dn, dw, ds;          // depths (may be negative)
    float dzbE, dzbN;                  // bed differences E-P, N-P
    float qp, qe, qn, qs, qse, qne;    // qx at P, E, N, S, SE, NE
    float rp, re, rn, rw, rwn, rne;    // qy at P, E, N, W, NW, NE
    float hse, hne, hnw;               // loop-1 depths at SE, NE, NW
    bool wetE, wetN;
};

__device__ inline void solver_wet32(const consts32 &K, const nb32 &n, float *f)
{
    const float hdry = K.hdry, g = K.g, ks = K.ks, dx = K.dx;
    const float hp = fmaxf(0.f, n.dp);
    const float hp0 = fmaxf(fmaxf(hdry, hp), ks);
    const float hw = fmaxf(0.f, n.dw);
    float he = fmaxf(0.f, n.de);
    const float hs = fmaxf(0.f, n.ds);
    float hn = fmaxf(0.f, n.dn);
    float qp = n.qp, qe = n.qe, qn = n.qn;
    float rp = n.rp, re = n.re, rn = n.rn;
    float hme = .5f * (hp + he), qme = .5f * (qp + qe), rme = .5f * (rp + re);
    float hmn = .5f * (hp + hn), qmn = .5f * (qp + qn), rmn = .5f * (rp + rn);
    float dze = (n.de - n.dp) + n.dzbE;
    const float dqe = qe - qp, dre = re - rp;
    float dzn = (n.dn - n.dp) + n.dzbN;
    const float drn = rn - rp, dqn = qn - qp;
    bool lroe = true;
    if (!n.wetE) {
        hme = fmaxf(0.f, n.dp - 0.5f * n.dzbE);
        he = 0.f; qe = 0.f; re = 0.f;
        if (hme > hdry) {
            if (dze >= 0.f) {
                qme = 0.f; rme = 0.f;
            } else {
                dze = fminf(fabsf(dze), hme);
                qme = fminf(0.296f * dze * sqrtf(g * dze), K.half_dx_over_dt * hp);
                rme = 0.f;
            }
        } else {
            qme = 0.f; rme = 0.f; hme = 0.f;
        }
        lroe = false;
    }
    if (!n.wetN) {
        hmn = fmaxf(0.f, n.dp - 0.5f * n.dzbN);
        hn = 0.f; qn = 0.f; rn = 0.f;
        if (hmn > hdry) {
            if (dzn >= 0.f) {
                qmn = 0.f; rmn = 0.f;
            } else {
                dzn = fminf(fabsf(dzn), hmn);
                rmn = fminf(0.296f * dzn * sqrtf(g * dzn), K.half_dx_over_dt * hp);
                qmn = 0.f;
            }
        } else {
            qmn = 0.f; rmn = 0.f; hmn = 0.f;
        }
        lroe = false;
    }
    const float cme = sqrtf(g * hme), cmn = sqrtf(g * hmn);
    const float ihme = 1.f / fmaxf(hme, hdry), ihmn = 1.f / fmaxf(hmn, hdry);
    const float ume = qme * ihme, vme = rme * ihme, umn = qmn * ihmn, vmn = rmn * ihmn;
    // eddy coefficients: us is identically zero, so cvdef*us*h + nuem = nuem
    const float hne_t = K.nuem * sqrtf(hp * he);
    const float hnn_t = K.nuem * sqrtf(hp * hn);
    const float ihp0 = 1.f / hp0;
    const float up = qp * ihp0, vp = rp * ihp0;
    const float un = qn / fmaxf(fmaxf(hn, hdry), ks);
    const float us0 = n.qs / fmaxf(fmaxf(hs, hdry), ks);
    const float use0 = n.qse / fmaxf(fmaxf(n.hse, hdry), ks);
    const float une = n.qne / fmaxf(fmaxf(n.hne, hdry), ks);
    const float ve = re / fmaxf(fmaxf(he, hdry), ks);
    const float vw = n.rw / fmaxf(fmaxf(hw, hdry), ks);
    const float vwn = n.rwn / fmaxf(fmaxf(n.hnw, hdry), ks);
    const float ven = n.rne / fmaxf(fmaxf(n.hne, hdry), ks);
    const float idx = 1.f / dx;
    const float txye = hne_t * ((ve - vp) * idx + .25f * (un + une - us0 - use0) * idx);
    const float txyn = hnn_t * ((un - up) * idx + .25f * (ve + ven - vw - vwn) * idx);
    float fe1 = qme, fe2 = qme * ume, fe3 = qme * vme - txye;
    float fn1 = rmn, fn2 = rmn * umn - txyn, fn3 = rmn * vmn;
    if (lroe) {
        float cc = 0.f;
        if (hme > hdry) {
            float dzea = fabsf(dze);
            if (dzea > 0.5f * hme) dze = copysignf(fminf(dzea, fabsf(he - hp)), dze);
            cc = .25f / cme;
        }
        {
            const float c1 = ume, c2 = ume + cme, c3 = ume - cme;
            const float c1a = fabsf(c1), c2a = fabsf(c2), c3a = fabsf(c3);
            const float a11 = c2 * c3a - c2a * c3, a12 = c2a - c3a;
            const float a21 = c2 * c3 * (c3a - c2a), a22 = c2a * c2 - c3a * c3;
            const float a31 = vme * (c2 * c3a - 2.f * cme * c1a - c2a * c3);
            const float a32 = vme * (c2a - c3a), a33 = 2.f * cme * c1a;
            fe1 -= (a11 * dze + a12 * dqe) * cc;
            fe2 -= (a21 * dze + a22 * dqe) * cc;
            fe3 -= (a31 * dze + a32 * dqe + a33 * dre) * cc;
        }
        cc = 0.f;
        if (hmn > hdry) {
            dzn = copysignf(fminf(fabsf(dzn), fabsf(hn - hp)), dzn);
            cc = .25f / cmn;
        }
        {
            const float c1 = vmn, c2 = vmn + cmn, c3 = vmn - cmn;
            const float c1a = fabsf(c1), c2a = fabsf(c2), c3a = fabsf(c3);
            const float a11 = c2 * c3a - c2a * c3, a13 = c2a - c3a;
            const float a21 = umn * (c2 * c3a - 2.f * cmn * c1a - c2a * c3);
            const float a22 = 2.f * cmn * c1a, a23 = umn * (c2a - c3a);
            const float a31 = c2 * c3 * (c3a - c2a), a33 = c2a * c2 - c3a * c3;
            fn1 -= (a11 * dzn + a13 * drn) * cc;
            fn2 -= (a21 * dzn + a22 * dqn + a23 * drn) * cc;
            fn3 -= (a31 * dzn + a33 * drn) * cc;
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
                const float cf = fn1 / (fe1 + fn1);
                fe1 = (1.f - cf) * volrat * idx;
                fn1 = cf * volrat * idx;
            }
        } else if (fe1 > 0.f) {
            fe1 = fminf(fe1 * dx, volrat) * idx;
        } else if (fn1 > 0.f) {
            fn1 = fminf(fn1 * dx, volrat) * idx;
        }
    } else {
        fe1 = 0.f; fn1 = 0.f;
    }
    f[0] = fe1; f[1] = fe2; f[2] = fe3; f[3] = fn1; f[4] = fn2; f[5] = fn3;
}

__device__ inline double wmin(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wmax(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// block partials: 3 doubles per block (dt bound, max wet depth, weir-face volume)
__device__ inline void block_partials(double dtm, double hmx, double wv, double *partials,
                                      unsigned bid, int tid, bool write_weir)
{
    __shared__ double sm[3][NT / 64];
    dtm = wmin(dtm);
    hmx = wmax(hmx);
    wv = wsum(wv);
    if ((tid & 63) == 0) { sm[0][tid >> 6] = dtm; sm[1][tid >> 6] = hmx; sm[2][tid >> 6] = wv; }
    __syncthreads();
    if (tid == 0) {
        double a = sm[0][0], b = sm[1][0], c = sm[2][0];
        for (int w = 1; w < NT / 64; w++) { a = fmin(a, sm[0][w]); b = fmax(b, sm[1][w]); c += sm[2][w]; }
        partials[3 * (size_t)bid] = a;
        partials[3 * (size_t)bid + 1] = b;
        if (write_weir) partials[3 * (size_t)bid + 2] = c;
    }
}

// CFL bound of one updated cell (main.cpp:308-316), FP64
__device__ inline void cfl_cell(const fx_params &P, double h, float qx, float qy, double &dtm, double &hmx)
{
    if (h > P.hdry) {
        hmx = fmax(hmx, h);
        const double c0 = sqrt(P.g * h);
        const double u0 = fmax(1e-6, fabs((double)qx / h));
        const double v0 = fmax(1e-6, fabs((double)qy / h));
        dtm = fmin(dtm, fmin((double)P.cfl_dxy / (u0 + c0), (double)P.cfl_dxy / (v0 + c0)));
    }
}

__global__ __launch_bounds__(NT) void k32_step(fx_params P, s32p S, fx_ctl *ctl, int ade_on)
{
    if (ctl->halt) return;
    const int cur = ctl->cur;
    const double *d_in = S.d_in[cur];
    double *d_out = const_cast<double *>(S.d_in[cur ^ 1]);
    const float *qx_in = S.qx_in[cur], *qy_in = S.qy_in[cur];
    float *qx_out = const_cast<float *>(S.qx_in[cur ^ 1]);
    float *qy_out = const_cast<float *>(S.qy_in[cur ^ 1]);
    const int gate = ctl->gate;
    const double dt = ctl->dtfl;
    const double meteoi = P.meteo ? ctl->meteoi : 0.0;
    const int ld = P.ld, mcols = P.ncols + 2;
    const int r0 = 1 + blockIdx.x * TX, c0 = 1 + blockIdx.y * TY;
    const int tid = threadIdx.x + threadIdx.y * 64;
    const unsigned bid = blockIdx.x + blockIdx.y * gridDim.x;

    __shared__ float s_d[LR * LC], s_qx[LR * LC], s_qy[LR * LC], s_zb[LR * LC], s_zl[LR * LC];
    __shared__ uint8_t s_fl[LR * LC];
    __shared__ float s_f[6][FR * FC];

    // 1. load tile + halo, melt and loop 1 on the fly
    for (int i = tid; i < LR * LC; i += NT) {
        const int lr = i % LR, lc = i / LR;
        const int r = r0 - 2 + lr, c = c0 - 2 + lc;
        float dv = 0.f, qxv = 0.f, qyv = 0.f, zbv = 0.f, zlv = 0.f;
        uint8_t fl = 0;
        if (r >= 0 && r < ld && c >= 0 && c < mcols) {
            const size_t p = (size_t)r + (size_t)c * ld;
            const uint8_t f = S.flags[p];
            if (f & FXF_INNER) {
                double d = d_in[p];
                if (!(f & FXF_NODATA)) d += P.melt_mult ? meteoi * (double)P.melt_mult[p] : meteoi;
                const bool wet = d > P.hdry;           // h = max(0,d) > hdry
                qxv = qx_in[p];
                qyv = qy_in[p];
                if (gate && !wet) { qxv = 0.f; qyv = 0.f; }
                dv = (float)d;
                zbv = S.zb[p];
                zlv = (float)S.zb_lo[p];
                fl = (wet ? LW : 0) | ((f & FXF_WEIR) ? LX : 0) | LI | ((f & FXF_STALE) ? LS : 0);
            }
        }
        s_d[i] = dv; s_qx[i] = qxv; s_qy[i] = qyv; s_zb[i] = zbv; s_zl[i] = zlv; s_fl[i] = fl;
    }
    __syncthreads();

    consts32 K;
    K.hdry = (float)P.hdry;
    K.g = (float)P.g;
    K.nuem = (float)P.nuem;
    K.ks = (float)P.ks0;           // uniform interior roughness (checked at fx_set_static)
    K.dx = (float)P.dxy;
    K.half_dx_over_dt = (float)(0.5 * P.dxy / dt);
    K.arbase_over_dt = (float)(P.arbase / dt);
    const float dxy_over_dt = (float)(P.dxy / dt);

    // 2. fluxes of the tile and its W/S halo
    if (gate) {
        for (int i = tid; i < FR * FC; i += NT) {
            const int fr = i % FR, fc = i / FR;
            const int li = (fr + 1) + (fc + 1) * LR;
            float f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            const uint8_t fl = s_fl[li];
            if ((fl & LI) && !(fl & LX)) {
                const bool wetE = s_fl[li + 1] & LW, wetN = s_fl[li + LR] & LW;
                if (!(fl & LW)) {
                    const bool wetW = s_fl[li - 1] & LW, wetS = s_fl[li - LR] & LW;
                    if (!wetW && !wetE && !wetS && !wetN) {
                        if (fl & LS) {         // stale fluxes of an all-dry neighbourhood
                            const int r = r0 - 1 + fr, c = c0 - 1 + fc;
                            const size_t p = (size_t)r + (size_t)c * ld;
                            for (int k = 0; k < 6; k++) f[k] = S.f6[k * S.mm + p];
                        }
                    } else {
                        // neighbour bed differences: exact high parts plus the low parts
                        const float dzbE = (s_zb[li + 1] - s_zb[li]) + (s_zl[li + 1] - s_zl[li]);
                        const float dzbN = (s_zb[li + LR] - s_zb[li]) + (s_zl[li + LR] - s_zl[li]);
                        solver_dry32(K, s_d[li], s_d[li + 1], s_d[li + LR], dzbE, dzbN, wetE, wetN,
                                     dxy_over_dt, f);
                    }
                } else {
                    nb32 n;
                    n.dp = s_d[li]; n.de = s_d[li + 1]; n.dn = s_d[li + LR];
                    n.dw = s_d[li - 1]; n.ds = s_d[li - LR];
                    n.dzbE = (s_zb[li + 1] - s_zb[li]) + (s_zl[li + 1] - s_zl[li]);
                    n.dzbN = (s_zb[li + LR] - s_zb[li]) + (s_zl[li + LR] - s_zl[li]);
                    n.qp = s_qx[li]; n.qe = s_qx[li + 1]; n.qn = s_qx[li + LR]; n.qs = s_qx[li - LR];
                    n.qse = s_qx[li + 1 - LR]; n.qne = s_qx[li + 1 + LR];
                    n.rp = s_qy[li]; n.re = s_qy[li + 1]; n.rn = s_qy[li + LR]; n.rw = s_qy[li - 1];
                    n.rwn = s_qy[li - 1 + LR]; n.rne = s_qy[li + 1 + LR];
                    n.hse = fmaxf(0.f, s_d[li + 1 - LR]);
                    n.hne = fmaxf(0.f, s_d[li + 1 + LR]);
                    n.hnw = fmaxf(0.f, s_d[li - 1 + LR]);
                    n.wetE = wetE; n.wetN = wetN;
                    solver_wet32(K, n, f);
                }
            }
            for (int k = 0; k < 6; k++) s_f[k][i] = f[k];
        }
    }
    __syncthreads();

    // 3. update own cells, 4. CFL bound of the new state
    double dtm = 9.e10, hmx = 0.0, wv = 0.0;
    const float kf = (float)(dt / P.dxy);
    const float dtf = (float)dt;
    for (int j = 0; j < TY / 4; j++) {
        const int lr = threadIdx.x, lc = threadIdx.y + 4 * j;
        const int r = r0 + lr, c = c0 + lc;
        if (r > P.nrows || c > P.ncols) continue;
        const size_t p = (size_t)r + (size_t)c * ld;
        const int li = (lr + 2) + (lc + 2) * LR;
        const int fi = (lr + 1) + (lc + 1) * FR;
        const uint8_t f = S.flags[p];
        const bool weir = f & FXF_WEIR;
        double d = d_in[p];
        if (!(f & FXF_NODATA)) d += P.melt_mult ? meteoi * (double)P.melt_mult[p] : meteoi;
        float qxn, qyn;
        uint8_t nf = f;
        float fe1P = 0.f, fn1P = 0.f;
        if (!gate) {
            qxn = qx_in[p];
            qyn = qy_in[p];
        } else {
            const float fe1W = s_f[0][fi - 1], fn1S = s_f[3][fi - FR];
            fe1P = s_f[0][fi];
            fn1P = s_f[3][fi];
            float dqx = 0.f, dqy = 0.f;
            if (!weir) {
                const float a = fe1W * kf, b = fe1P * kf, cc = fn1S * kf, e = fn1P * kf;
                d += ((double)a - (double)b) + ((double)cc - (double)e);
                dqx = ((s_f[1][fi - 1] - s_f[1][fi]) + (s_f[4][fi - FR] - s_f[4][fi])) * kf;
                dqy = ((s_f[2][fi - 1] - s_f[2][fi]) + (s_f[5][fi - FR] - s_f[5][fi])) * kf;
                if (s_fl[li + 1] & LX) wv -= (double)b;        // into / out of a weir cell
                if (s_fl[li + LR] & LX) wv -= (double)e;
            }
            const double h = fmax(0.0, d);
            if (h < P.hdry || weir) {
                qxn = 0.f;
                qyn = 0.f;
            } else {
                qxn = .1f * (fe1W * dtf) + .8f * (s_qx[li] + dqx) + .1f * (fe1P * dtf);
                qyn = .1f * (fn1S * dtf) + .8f * (s_qy[li] + dqy) + .1f * (fn1P * dtf);
            }
            // keep the fluxes of a cell that may be all-dry next step (S1)
            if (!weir) {
                const bool wet1 = s_fl[li] & LW;
                const bool alldry = !wet1 && !(s_fl[li - 1] & LW) && !(s_fl[li + 1] & LW) &&
                                    !(s_fl[li - LR] & LW) && !(s_fl[li + LR] & LW);
                if (!alldry && (h <= P.hdry || !P.rates_nonneg)) {
                    bool nz = false;
                    for (int k = 0; k < 6; k++) nz |= s_f[k][fi] != 0.f;
                    if (nz) {
                        for (int k = 0; k < 6; k++) S.f6[k * S.mm + p] = s_f[k][fi];
                        nf |= FXF_STALE;
                    } else {
                        nf &= ~FXF_STALE;
                    }
                }
            }
            // wet-time tracer: close / open runs on transitions only
            const bool wet1 = s_fl[li] & LW;
            if (wet1 && !(f & FXF_WETPREV)) {
                S.tw_t0[p] = (float)ctl->s_wet_prev;
                nf |= FXF_WETPREV;
            } else if (!wet1 && (f & FXF_WETPREV)) {
                S.tw_acc[p] += (float)(ctl->s_wet_prev - (double)S.tw_t0[p]);
                nf &= ~FXF_WETPREV;
            }
        }
        d_out[p] = d;
        qx_out[p] = qxn;
        qy_out[p] = qyn;
        if (ade_on) {
            S.ff[p] = fe1P;
            S.ff[S.mm + p] = fn1P;
        }
        if (nf != f) S.flags[p] = nf;
        cfl_cell(P, fmax(0.0, d), qxn, qyn, dtm, hmx);
    }
    block_partials(dtm, hmx, wv, S.partials, bid, tid, true);
}

// CFL partials of the current state (start of a device-resident run)
__global__ __launch_bounds__(NT) void k32_cfl(fx_params P, s32p S, fx_ctl *ctl)
{
    const int cur = ctl->cur;
    const double *d_in = S.d_in[cur];
    const float *qx_in = S.qx_in[cur], *qy_in = S.qy_in[cur];
    const int r0 = 1 + blockIdx.x * TX, c0 = 1 + blockIdx.y * TY;
    const int tid = threadIdx.x + threadIdx.y * 64;
    const unsigned bid = blockIdx.x + blockIdx.y * gridDim.x;
    double dtm = 9.e10, hmx = 0.0;
    for (int j = 0; j < TY / 4; j++) {
        const int r = r0 + threadIdx.x, c = c0 + threadIdx.y + 4 * j;
        if (r > P.nrows || c > P.ncols) continue;
        const size_t p = (size_t)r + (size_t)c * P.ld;
        cfl_cell(P, fmax(0.0, d_in[p]), qx_in[p], qy_in[p], dtm, hmx);
    }
    block_partials(dtm, hmx, 0.0, S.partials, bid, tid, false);
}

// Clock of one loop iteration (main.cpp:271-325, 385-387): fold the CFL
// partials left by the previous step, clip to the print time, advance tim,
// evaluate the forcing, gate the solver, update the FP64 water ledger.
__global__ __launch_bounds__(1024) void k32_clock(fx_params P, double *partials, int nblk, fx_ctl *ctl)
{
    if (ctl->halt) return;
    const int tid = threadIdx.x;
    if (ctl->tim > ctl->t_end) {                       // loop condition, main.cpp:271
        if (tid == 0) ctl->halt = 1;
        return;
    }
    double a = 9.e10, b = 0.0, c = 0.0;
    for (int i = tid; i < nblk; i += 1024) {
        a = fmin(a, partials[3 * (size_t)i]);
        b = fmax(b, partials[3 * (size_t)i + 1]);
        c += partials[3 * (size_t)i + 2];
        partials[3 * (size_t)i + 2] = 0.0;
    }
    __shared__ double sm[3][16];
    a = wmin(a); b = wmax(b); c = wsum(c);
    if ((tid & 63) == 0) { sm[0][tid >> 6] = a; sm[1][tid >> 6] = b; sm[2][tid >> 6] = c; }
    __syncthreads();
    if (tid != 0) return;
    for (int w = 1; w < 16; w++) { a = fmin(a, sm[0][w]); b = fmax(b, sm[1][w]); c += sm[2][w]; }
    // Kahan-compensated ledger of the weir-face exchange of the previous step
    {
        const double y = c * P.arbase - ctl->led_weir_c;
        const double t = ctl->led_weir + y;
        ctl->led_weir_c = (t - ctl->led_weir) - y;
        ctl->led_weir = t;
    }
    ctl->dt_raw = a;
    const double dt = fmin(ctl->print_next - ctl->tim, a);
    ctl->dtfl = dt;
    ctl->hpall = b;
    ctl->tim = ctl->tim + dt;
    const int gate = (b != 0.0);
    ctl->gate = gate;
    ctl->it += gate;
    ctl->steps += 1;
    ctl->solver_steps += gate;
    ctl->s_wet_prev = ctl->s_wet;
    if (gate) ctl->s_wet = ctl->s_wet + dt / 3600;
    fx_forcing_eval(P, ctl);
    if (P.meteo) {
        const double y = ctl->meteoi * ctl->melt_area * P.arbase - ctl->led_melt_c;
        const double t = ctl->led_melt + y;
        ctl->led_melt_c = (t - ctl->led_melt) - y;
        ctl->led_melt = t;
    }
    ctl->pending_halt = (ctl->tim >= ctl->print_next);
}

__global__ void k32_flip(fx_ctl *ctl)
{
    if (!ctl->halt) ctl->cur ^= 1;
}

// water volume of non-weir interior cells (FP64, block partials)
__global__ __launch_bounds__(NT) void k32_volume(fx_params P, s32p S, fx_ctl *ctl, double *part)
{
    const double *d_in = S.d_in[ctl->cur];
    const int r0 = 1 + blockIdx.x * TX, c0 = 1 + blockIdx.y * TY;
    const int tid = threadIdx.x + threadIdx.y * 64;
    double v = 0.0;
    for (int j = 0; j < TY / 4; j++) {
        const int r = r0 + threadIdx.x, c = c0 + threadIdx.y + 4 * j;
        if (r > P.nrows || c > P.ncols) continue;
        const size_t p = (size_t)r + (size_t)c * P.ld;
        if (!(S.flags[p] & FXF_WEIR)) v += d_in[p];
    }
    v = wsum(v);
    __shared__ double sm[NT / 64];
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < NT / 64; w++) s += sm[w];
        part[blockIdx.x + blockIdx.y * gridDim.x] = s;
    }
}

inline dim3 grid32(const fx_params &P)
{
    return dim3((P.nrows + TX - 1) / TX, (P.ncols + TY - 1) / TY);
}

s32p make_s32p(fx_domain d)
{
    fx_state32 &t = d->s32;
    s32p S;
    S.d_in[0] = t.d[0]; S.d_in[1] = t.d[1];
    S.qx_in[0] = t.qx[0]; S.qx_in[1] = t.qx[1];
    S.qy_in[0] = t.qy[0]; S.qy_in[1] = t.qy[1];
    S.zb = t.zb;
    S.zb_lo = t.zb_lo;
    S.f6 = t.f6;
    S.ff = t.ff;
    S.flags = t.flags;
    S.tw_acc = t.tw_acc;
    S.tw_t0 = t.tw_t0;
    S.partials = t.partials;
    S.mm = d->ncell;
    return S;
}

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
        if ((rc = dalloc32(d, &t.qx[k], mm))) return rc;
        if ((rc = dalloc32(d, &t.qy[k], mm))) return rc;
    }
    if ((rc = dalloc32(d, &t.zb, mm))) return rc;
    if ((rc = dalloc32(d, &t.zb_lo, mm))) return rc;
    if ((rc = dalloc32(d, &t.f6, 6 * mm))) return rc;
    if ((rc = dalloc32(d, &t.ff, 2 * mm))) return rc;
    if ((rc = dalloc32(d, &t.flags, mm))) return rc;
    if ((rc = dalloc32(d, &t.tw_acc, mm))) return rc;
    if ((rc = dalloc32(d, &t.tw_t0, mm))) return rc;
    if (d->desc.ade_solver)
        for (int b = 0; b < 2; b++)
            for (int k = 0; k < d->desc.nchem; k++)
                if ((rc = dalloc32(d, &t.mass[b][k], mm))) return rc;
    fx_params P = fx_make_params(d);
    dim3 g = grid32(P);
    t.npart = (int64_t)g.x * g.y;
    if ((rc = dalloc32(d, &t.partials, 3 * t.npart))) return rc;
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
        std::vector<float> zf(mm);
        std::vector<_Float16> zl(mm);
        for (int64_t i = 0; i < mm; i++) {
            zf[i] = (float)zb[i];
            zl[i] = (_Float16)(zb[i] - (double)zf[i]);
        }
        FX_HIP(hipMemcpy(t.zb, zf.data(), mm * 4, hipMemcpyHostToDevice));
        FX_HIP(hipMemcpy(t.zb_lo, zl.data(), mm * 2, hipMemcpyHostToDevice));
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
                uint8_t f = FXF_INNER;
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
        std::vector<float> qf(mm);
        for (int64_t i = 0; i < mm; i++) qf[i] = (float)q[i];
        float *dst = field == FX_F_QX ? t.qx[cur] : t.qy[cur];
        FX_HIP(hipMemcpy(dst, qf.data(), mm * 4, hipMemcpyHostToDevice));
        return FX_OK;
    }
    case FX_F_FE1: case FX_F_FE2: case FX_F_FE3: case FX_F_FN1: case FX_F_FN2: case FX_F_FN3: {
        const double *q = (const double *)host;
        std::vector<float> qf(mm);
        for (int64_t i = 0; i < mm; i++) qf[i] = (float)q[i];
        FX_HIP(hipMemcpy(t.f6 + (int64_t)(field - FX_F_FE1) * mm, qf.data(), mm * 4, hipMemcpyHostToDevice));
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
        std::vector<float> qf(mm);
        for (int64_t i = 0; i < mm; i++) qf[i] = (float)q[i];
        FX_HIP(hipMemcpy(t.tw_acc, qf.data(), mm * 4, hipMemcpyHostToDevice));
        std::vector<uint8_t> fl(mm);
        FX_HIP(hipMemcpy(fl.data(), t.flags, mm, hipMemcpyDeviceToHost));
        for (auto &f : fl) f &= ~FXF_WETPREV;
        FX_HIP(hipMemcpy(t.flags, fl.data(), mm, hipMemcpyHostToDevice));
        return FX_OK;
    }
    case FX_F_H: case FX_F_LDRY: case FX_F_LDRY_PREV: case FX_F_H0:
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
        std::vector<float> q(mm);
        FX_HIP(hipMemcpy(q.data(), field == FX_F_QX ? t.qx[cur] : t.qy[cur], mm * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < mm; i++) out[i] = q[i];
        return FX_OK;
    }
    case FX_F_LDRY: case FX_F_LDRY_PREV: {
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
        std::vector<float> q(mm);
        FX_HIP(hipMemcpy(q.data(), t.f6 + (int64_t)(field - FX_F_FE1) * mm, mm * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < mm; i++) out[i] = q[i];
        return FX_OK;
    }
    case FX_F_TWET: {
        if ((rc = load_fl())) return rc;
        std::vector<float> acc(mm), t0(mm);
        FX_HIP(hipMemcpy(acc.data(), t.tw_acc, mm * 4, hipMemcpyDeviceToHost));
        FX_HIP(hipMemcpy(t0.data(), t.tw_t0, mm * 4, hipMemcpyDeviceToHost));
        const double s = d->ctl_host->s_wet;
        for (int64_t i = 0; i < mm; i++)
            out[i] = acc[i] + ((fl[i] & FXF_WETPREV) ? s - (double)t0[i] : 0.0);
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
    hipLaunchKernelGGL(k32_cfl, grid32(P), dim3(64, 4), 0, d->stream, P, S, d->ctl);
    fx_prof_end(d, FX_K_CFL32);
    return fx_hip_check(hipGetLastError(), "k32_cfl");
}

int fx32_launch_step(fx_domain d)
{
    fx_params P = fx_make_params(d);
    s32p S = make_s32p(d);
    hipLaunchKernelGGL(k32_clock, dim3(1), dim3(1024), 0, d->stream, P, S.partials,
                       (int)d->s32.npart, d->ctl);
    FX_HIP(hipGetLastError());
    const int ade_on = (d->desc.ade_solver && d->desc.ade_mode == FX_ADE_FAST) ? 1 : 0;
    fx_prof_begin(d, FX_K_STEP32);
    hipLaunchKernelGGL(k32_step, grid32(P), dim3(64, 4), 0, d->stream, P, S, d->ctl, ade_on);
    fx_prof_end(d, FX_K_STEP32);
    FX_HIP(hipGetLastError());
    if (ade_on) {
        int rc = fx32_launch_ade(d);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k32_flip, dim3(1), dim3(1), 0, d->stream, d->ctl);
    return fx_hip_check(hipGetLastError(), "k32_flip");
}

int fx32_launch_volume(fx_domain d, double *vol_out)
{
    fx_params P = fx_make_params(d);
    s32p S = make_s32p(d);
    dim3 g = grid32(P);
    const size_t nb = (size_t)g.x * g.y;
    double *part = nullptr;
    FX_HIP(hipMalloc((void **)&part, nb * sizeof(double)));
    hipLaunchKernelGGL(k32_volume, g, dim3(64, 4), 0, d->stream, P, S, d->ctl, part);
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
int fx32_pending_weir(fx_domain d, double *vol)
{
    const size_t nb = (size_t)d->s32.npart;
    std::vector<double> h(3 * nb);
    FX_HIP(hipMemcpyAsync(h.data(), d->s32.partials, 3 * nb * sizeof(double), hipMemcpyDeviceToHost,
                          d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    double s = 0.0;
    for (size_t i = 0; i < nb; i++) s += h[3 * i + 2];
    *vol = s * (double)d->desc.dxy * (double)d->desc.dxy;
    return FX_OK;
}
