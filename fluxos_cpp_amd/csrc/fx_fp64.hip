// fx_fp64.hip -- FP64 parity kernels for gfx950.
//
// This path reproduces the reference's double-precision arithmetic bit for
// bit: same expression order, same float->double promotions, IEEE division
// and square root, and no FP contraction (the library is compiled with
// -ffp-contract=off).  It is the plain multi-kernel form of the reference's
// loops, one kernel per loop, each a coalesced pass over the column-major
// grid (irow is the fast, contiguous axis and the 64-lane axis of a block):
//
//   k64_cfl     <- main.cpp:300-325      (block min/max + last-block finalise)
//   k64_melt    <- common.cpp:164-390 and hydrodynamics_calc.cpp:53-73
//   k64_flux    <- hydrodynamics_calc.cpp:76-90 + solver_drydomain.cpp:27-194
//                  + solver_wetdomain.cpp:27-342
//   k64_update  <- hydrodynamics_calc.cpp:93-154 (loops 3 and 4 fused: qxf/qyf
//                  are recomputed from the neighbour's fe_1/fn_1)
#include <algorithm>

#include "fx_params.h"

#define BX 64
#define BY 4

namespace {

__device__ inline bool fx_interior(const fx_params &P, int &irow, int &icol)
{
    irow = 1 + blockIdx.x * BX + threadIdx.x;
    icol = 1 + blockIdx.y * BY + threadIdx.y;
    return irow <= P.nrows && icol <= P.ncols;
}

// wave-wide min/max of doubles (64 lanes)
__device__ inline double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline double ld_coherent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// CFL (main.cpp:300-325).  Every interior cell: h0 = h, ldry_prev = ldry,
// ldry = !(h > hdry); wet cells give dt candidates.  dt and hpall are a
// min/max, so the reduction order does not change the result.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k64_cfl(fx_params P, fx_state64 s, fx_ctl *ctl,
                                               double *partials, unsigned mode)
{
    if ((mode & FXM_GATED) && ctl->halt) return;
    // grid-stride over columns (a bounded number of blocks keeps the
    // last-block-done counter cheap), threads along the contiguous irow axis
    double dtmin = 9.e10, hmax = 0.0;
    const int tid = threadIdx.x;
    for (int icol = 1 + blockIdx.x; icol <= P.ncols; icol += gridDim.x) {
        for (int irow = 1 + tid; irow <= P.nrows; irow += 256) {
            const size_t p = (size_t)irow + (size_t)icol * P.ld;
            const double hc = s.h[p];
            s.h0[p] = hc;
            s.ldry_prev[p] = s.ldry[p];
            if (hc > P.hdry) {
                s.ldry[p] = 0;
                const double hp = fmax(hc, P.hdry);
                hmax = fmax(hmax, hc);
                const double c0 = sqrt(P.g * hc);
                const double u0 = fmax(.000001, fabs(s.qx[p] / hp));
                const double v0 = fmax(.000001, fabs(s.qy[p] / hp));
                dtmin = fmin(fmin((double)P.cfl_dxy / (u0 + c0), (double)P.cfl_dxy / (v0 + c0)), dtmin);
            } else {
                s.ldry[p] = 1;
            }
        }
    }
    // block reduction
    __shared__ double sm_dt[256 / 64], sm_h[256 / 64];
    __shared__ bool am_last;
    dtmin = wave_min(dtmin);
    hmax = wave_max(hmax);
    if ((tid & 63) == 0) { sm_dt[tid >> 6] = dtmin; sm_h[tid >> 6] = hmax; }
    __syncthreads();
    const unsigned nblk = gridDim.x;
    const unsigned bid = blockIdx.x;
    if (tid == 0) {
        double a = sm_dt[0], b = sm_h[0];
        for (int w = 1; w < 4; w++) { a = fmin(a, sm_dt[w]); b = fmax(b, sm_h[w]); }
        partials[2 * bid] = a;
        partials[2 * bid + 1] = b;
        __threadfence();
        unsigned prev = atomicAdd(&ctl->blocks_done, 1u);
        am_last = (prev == nblk - 1);
    }
    __syncthreads();
    if (!am_last) return;
    __threadfence();
    double a = 9.e10, b = 0.0;
    for (unsigned i = tid; i < nblk; i += 256) {
        a = fmin(a, ld_coherent(&partials[2 * i]));
        b = fmax(b, ld_coherent(&partials[2 * i + 1]));
    }
    a = wave_min(a);
    b = wave_max(b);
    if ((tid & 63) == 0) { sm_dt[tid >> 6] = a; sm_h[tid >> 6] = b; }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < 4; w++) { a = fmin(a, sm_dt[w]); b = fmax(b, sm_h[w]); }
        double dt = fmin(ctl->print_next - ctl->tim, a);   // main.cpp:321
        ctl->dtfl = dt;
        ctl->hpall = b;
        ctl->tim = ctl->tim + dt;                         // main.cpp:325
        const int gate = (b != 0.0);                      // main.cpp:385
        ctl->gate = gate;
        if (mode & FXM_GATED) {
            ctl->it += gate;
            ctl->steps += 1;
            ctl->solver_steps += gate;
            ctl->pending_halt = (ctl->tim >= ctl->print_next) || (ctl->tim > ctl->t_end);
        }
        fx_forcing_eval(P, ctl);
        ctl->blocks_done = 0;
    }
}

// ---------------------------------------------------------------------------
// Forcing (common.cpp:164-390) and hydro loop 1 (hydrodynamics_calc.cpp:53-73)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k64_melt(fx_params P, fx_state64 s, fx_ctl *ctl,
                                                unsigned mode)
{
    if ((mode & FXM_GATED) && ctl->halt) return;
    int irow, icol;
    if (!fx_interior(P, irow, icol)) return;
    const size_t p = (size_t)irow + (size_t)icol * P.ld;
    const double dtl = ctl->dtfl;
    double z = s.z[p];
    const double zb = s.zb[p];
    if ((mode & FXM_MELT) && P.meteo && zb != P.nodata) {
        const double meteoi = P.melt_mult ? ctl->meteoi * (double)P.melt_mult[p] : ctl->meteoi;
        const double hp = fmax(z - zb, 0.0);
        z = z + meteoi;
        const double h = fmax(z - zb, 0.0);
        s.z[p] = z;
        s.h[p] = h;
        if (h <= P.hdry) s.ldry[p] = 0;                   // common.cpp:219-222
        s.h0[p] = h;
        if (P.ade && hp != 0.0) {
            for (int k = 0; k < P.nchem; k++)
                s.conc[k][p] = (s.conc[k][p] * hp + (meteoi * ctl->cmelt[k])) / (h);
        }
    }
    if ((mode & FXM_INFLOW) && P.inflow && (int64_t)p == P.inflow_p) {
        if (zb != P.nodata) {                             // common.cpp:324-376
            if (z <= zb) z = zb;
            const double hp = fmax(z - zb, 0.0);
            z = z + ctl->inflowi;
            const double h = fmax(z - zb, 0.0);
            s.z[p] = z;
            s.h[p] = h;
            s.ldry[p] = (h <= P.hdry) ? 1 : 0;
            s.h0[p] = h;
            if (P.ade && hp != 0.0) {
                for (int k = 0; k < P.nchem; k++)
                    s.conc[k][p] = (s.conc[k][p] * hp + (ctl->inflowi * ctl->cin[k])) / (h);
            }
        } else {
            atomicOr(&ctl->status, 8);
        }
    }
    if ((mode & FXM_LOOP1) && (!(mode & FXM_GATED) || ctl->gate)) {
        const double hp = fmax(0.0, z - zb);
        s.h[p] = hp;
        if (hp <= P.hdry) {
            s.qx[p] = 0.0;
            s.qy[p] = 0.0;
            s.ldry[p] = 1;
        } else {
            s.ldry[p] = 0;
            s.twet[p] += dtl / 3600;
        }
    }
}

// ---------------------------------------------------------------------------
// solver_dry (solver_drydomain.cpp:27-194).  Weir branch: unreachable because
// the outer ring is required to be weir cells (checked at fx_set_static).
// ---------------------------------------------------------------------------
__device__ inline void solver_dry64(const fx_params &P, const fx_state64 &s, double dtl,
                                    int irow, int icol)
{
    const size_t ld = P.ld;
    const size_t p = (size_t)irow + (size_t)icol * ld;
    const size_t pw = p - 1, pe = p + 1, ps = p - ld, pn = p + ld;
    const uint8_t ldw = s.ldry[pw], ldp = s.ldry[p], lde = s.ldry[pe];
    const uint8_t lds = s.ldry[ps], ldn = s.ldry[pn];
    if (ldw == 1 && ldp == 1 && lde == 1 && lds == 1 && ldn == 1) return;   // :60-76
    const double g = P.g;
    const double zbp = s.zb[p], zbe = s.zb[pe], zbn = s.zb[pn];
    const double zp = s.z[p], ze = s.z[pe], zn = s.z[pn];
    const double hp = fmax(0.0, zp - zp);                                  // :85
    const double he = fmax(0.0, ze - zbe);
    const double hn = fmax(0.0, zn - zbn);
    const double zbpe = .5 * (zbe + zbp);
    const double zbpn = .5 * (zbn + zbp);
    double fe1 = 0.0, fe2 = 0.0, fe3 = 0.0, fn1 = 0.0, fn2 = 0.0, fn3 = 0.0;
    (void)hp;
    if (lde == 0) {
        double hme = fmax(0.0, ze - zbpe);
        if (hme > P.hdry) {
            if (ze <= zp) {
                fe2 = .5 * g * hme * hme;
            } else {
                double dze = fmin(fabs(ze - zp), hme);
                double qme = 0.296 * dze * sqrt(g * dze);
                hme = 0.444 * dze;
                const double volrat = P.dxy * he / dtl;
                qme = -fmin(qme, volrat);
                const double ume = qme / hme;
                fe1 = qme;
                fe2 = qme * ume + .5 * g * hme * hme;
            }
        }
    }
    if (ldn == 0) {
        double hmn = fmax(0.0, zn - zbpn);
        if (hmn > P.hdry) {
            if (zn <= zp) {
                fn3 = .5 * g * hmn * hmn;
            } else {
                double dzn = fmin(fabs(zn - zp), hmn);
                double rmn = 0.296 * dzn * sqrt(g * dzn);
                hmn = 0.444 * dzn;
                const double volrat = P.dxy * hn / dtl;
                rmn = -fmin(rmn, volrat);
                const double vmn = rmn / hmn;
                fn1 = rmn;
                fn3 = rmn * vmn + .5 * g * hmn * hmn;
            }
        }
    }
    s.fn1[p] = fn1; s.fn2[p] = fn2; s.fn3[p] = fn3;
    s.fe1[p] = fe1; s.fe2[p] = fe2; s.fe3[p] = fe3;
}

// ---------------------------------------------------------------------------
// solver_wet (solver_wetdomain.cpp:27-342).  us is identically zero in the
// reference (hydrodynamics_calc.cpp:65,137; initiate.cpp:176), so the eddy
// coefficients cvdef*us*h + nuem equal nuem exactly and us is not stored.
// ---------------------------------------------------------------------------
__device__ inline void solver_wet64(const fx_params &P, const fx_state64 &s, double dtl,
                                    int irow, int icol)
{
    const size_t ld = P.ld;
    const double dx = (double)P.dxy_u, dy = (double)P.dxy_u;
    const double hdryl = P.hdry, gaccl = P.g, nueml = P.nuem, cvdefl = P.cvdef;
    const size_t p = (size_t)irow + (size_t)icol * ld;
    const size_t pw = p - 1, pe = p + 1, ps = p - ld, pn = p + ld;
    const size_t pes = pe - ld, pen = pe + ld, pwn = pw + ld;
    const double kspl = s.ks[p];
    const uint8_t ldp = s.ldry[p], lde = s.ldry[pe], ldn = s.ldry[pn];
    bool lroe = true;

    const double zbw = s.zb[pw], zbp = s.zb[p], zbe = s.zb[pe], zbs = s.zb[ps], zbn = s.zb[pn];
    const double zw = s.z[pw], zp = s.z[p], ze = s.z[pe], zs = s.z[ps], zn = s.z[pn];
    double qp = s.qx[p], qe = s.qx[pe], qn = s.qx[pn];
    double rw = s.qy[pw], rp = s.qy[p], re = s.qy[pe], rn = s.qy[pn];

    const double zbpe = .5 * (zbe + zbp);
    const double zbpn = .5 * (zbn + zbp);
    const double hp = fmax(0.0, zp - zbp);
    const double hp0 = fmax(fmax(hdryl, hp), kspl);
    const double hw = fmax(0.0, zw - zbw);
    double he = fmax(0.0, ze - zbe);
    const double hs = fmax(0.0, zs - zbs);
    double hn = fmax(0.0, zn - zbn);

    double hme = .5 * (hp + he), qme = .5 * (qp + qe), rme = .5 * (rp + re);
    double hmn = .5 * (hp + hn), qmn = .5 * (qp + qn), rmn = .5 * (rp + rn);
    double dze = ze - zp;
    const double dqe = qe - qp, dre = re - rp;
    double dzn = zn - zp;
    const double drn = rn - rp, dqn = qn - qp;

    if (lde == 1) {
        hme = fmax(0.0, zp - zbpe);
        he = 0.0; qe = 0.0; re = 0.0;
        if (hme > hdryl) {
            if (ze >= zp) {
                qme = 0.0; rme = 0.0;
            } else {
                dze = fmin(fabs(dze), hme);
                qme = 0.296 * dze * sqrt(gaccl * dze);
                const double volrat = 0.5 * dx * hp / dtl;
                qme = fmin(qme, volrat);
                rme = 0.0;
            }
        } else {
            qme = 0.0; rme = 0.0; hme = 0.0;
        }
        lroe = false;
    }
    if (ldn == 1) {
        hmn = fmax(0.0, zp - zbpn);
        hn = 0.0; qn = 0.0; rn = 0.0;
        if (hmn > hdryl) {
            if (zn >= zp) {
                qmn = 0.0; rmn = 0.0;
            } else {
                dzn = fmin(fabs(dzn), hmn);
                rmn = 0.296 * dzn * sqrt(gaccl * dzn);
                qmn = 0.0;
                const double volrat = 0.5 * dy * hp / dtl;
                rmn = fmin(rmn, volrat);
            }
        } else {
            qmn = 0.0; rmn = 0.0; hmn = 0.0;
        }
        lroe = false;
    }

    const double cme = sqrt(gaccl * hme);
    const double cmn = sqrt(gaccl * hmn);
    const double ume = qme / fmax(hme, hdryl);
    const double vme = rme / fmax(hme, hdryl);
    const double umn = qmn / fmax(hmn, hdryl);
    const double vmn = rmn / fmax(hmn, hdryl);
    const double cnp = cvdefl * 0.0 * hp + nueml;
    const double cne = cvdefl * 0.0 * he + nueml;
    const double cnn = cvdefl * 0.0 * hn + nueml;
    const double hne = .5 * (cnp + cne) * sqrt(hp * he);
    const double hnn = .5 * (cnp + cnn) * sqrt(hp * hn);
    const double up = qp / hp0;
    const double un = qn / fmax(fmax(hn, hdryl), s.ks[pn]);
    const double us0 = s.qx[ps] / fmax(fmax(hs, hdryl), s.ks[ps]);
    const double use0 = s.qx[pes] / fmax(fmax(s.h[pes], hdryl), s.ks[pes]);
    const double une = s.qx[pen] / fmax(fmax(s.h[pen], hdryl), s.ks[pen]);
    const double vp = rp / hp0;
    const double ve = re / fmax(fmax(he, hdryl), s.ks[pe]);
    const double vw = rw / fmax(fmax(hw, hdryl), s.ks[pw]);
    const double vwn = s.qy[pwn] / fmax(fmax(s.h[pwn], hdryl), s.ks[pwn]);
    const double ven = s.qy[pen] / fmax(fmax(s.h[pen], hdryl), s.ks[pen]);
    const double txye = hne * ((ve - vp) / fabs(dx) + .25 * (un + une - us0 - use0) / dy);
    const double txyn = hnn * ((un - up) / fabs(dy) + .25 * (ve + ven - vw - vwn) / dx);

    const double fe1c = qme, fe2c = qme * ume, fe3c = qme * vme - txye;
    const double fn1c = rmn, fn2c = rmn * umn - txyn, fn3c = rmn * vmn;

    double fe1r = 0.0, fe2r = 0.0, fe3r = 0.0, fn1r = 0.0, fn2r = 0.0, fn3r = 0.0;
    if (lroe) {
        double cc;
        if (hme > P.hdry) {
            double dzea = fabs(dze);
            if (dzea > 0.5 * hme) {
                const double dhea = fabs(he - hp);
                dzea = fmin(dzea, dhea);
                dze = copysign(dzea, dze);
            }
            cc = .25 / cme;
        } else {
            cc = 0.0;
        }
        {
            const double c1 = ume, c2 = ume + cme, c3 = ume - cme;
            const double c1a = fabs(c1), c2a = fabs(c2), c3a = fabs(c3);
            const double a11 = c2 * c3a - c2a * c3;
            const double a12 = c2a - c3a;
            const double a21 = c2 * c3 * (c3a - c2a);
            const double a22 = c2a * c2 - c3a * c3;
            const double a31 = vme * (c2 * c3a - 2. * cme * c1a - c2a * c3);
            const double a32 = vme * (c2a - c3a);
            const double a33 = 2. * cme * c1a;
            fe1r = -(a11 * dze + a12 * dqe) * cc;
            fe2r = -(a21 * dze + a22 * dqe) * cc;
            fe3r = -(a31 * dze + a32 * dqe + a33 * dre) * cc;
        }
        if (ldp == 0 && hmn > P.hdry) {
            double dzna = fabs(dzn);
            const double dhna = fabs(hn - hp);
            dzna = fmin(dzna, dhna);
            dzn = copysign(dzna, dzn);
            cc = .25 / cmn;
        } else {
            cc = 0.0;
        }
        {
            const double c1 = vmn, c2 = vmn + cmn, c3 = vmn - cmn;
            const double c1a = fabs(c1), c2a = fabs(c2), c3a = fabs(c3);
            const double a11 = c2 * c3a - c2a * c3;
            const double a13 = c2a - c3a;
            const double a21 = umn * (c2 * c3a - 2. * cmn * c1a - c2a * c3);
            const double a22 = 2. * cmn * c1a;
            const double a23 = umn * (c2a - c3a);
            const double a31 = c2 * c3 * (c3a - c2a);
            const double a33 = c2a * c2 - c3a * c3;
            fn1r = -(a11 * dzn + a13 * drn) * cc;
            fn2r = -(a21 * dzn + a22 * dqn + a23 * drn) * cc;
            fn3r = -(a31 * dzn + a33 * drn) * cc;
        }
    }
    const double fe2p = .5 * gaccl * (hme * hme);
    const double fn3p = .5 * gaccl * (hmn * hmn);
    double fe1 = fe1c + fe1r;
    const double fe2 = fe2c + fe2r + fe2p;
    const double fe3 = fe3c + fe3r;
    double fn1 = fn1c + fn1r;
    const double fn2 = fn2c + fn2r;
    const double fn3 = fn3c + fn3r + fn3p;

    const double volpot = P.arbase * hp;
    const double volrat = volpot / dtl;
    if (volrat > 0.0) {
        if (fe1 > 0.0 && fn1 > 0.0) {
            if (fe1 * dy + fn1 * dx > volrat) {
                const double cf = fn1 * dx / (fe1 * dy + fn1 * dx);
                fe1 = (1. - cf) * volrat / dy;
                fn1 = cf * volrat / dx;
            }
        } else if (fe1 > 0.0) {
            fe1 = fmin(fe1 * dy, volrat) / dy;
        } else if (fn1 > 0.0) {
            fn1 = fmin(fn1 * dx, volrat) / dx;
        }
    } else {
        fe1 = 0.0; fn1 = 0.0;
    }
    s.fn1[p] = fn1; s.fn2[p] = fn2; s.fn3[p] = fn3;
    s.fe1[p] = fe1; s.fe2[p] = fe2; s.fe3[p] = fe3;
}

__global__ __launch_bounds__(256) void k64_flux(fx_params P, fx_state64 s, fx_ctl *ctl,
                                                unsigned mode)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate)) return;
    int irow, icol;
    if (!fx_interior(P, irow, icol)) return;
    const size_t p = (size_t)irow + (size_t)icol * P.ld;
    if (s.weir[p]) return;                               // :81-88
    const double dtl = ctl->dtfl;
    if (s.ldry[p] == 1) solver_dry64(P, s, dtl, irow, icol);
    else solver_wet64(P, s, dtl, irow, icol);
}

// loops 3 and 4 (hydrodynamics_calc.cpp:93-154)
__global__ __launch_bounds__(256) void k64_update(fx_params P, fx_state64 s, fx_ctl *ctl,
                                                  unsigned mode)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate)) return;
    int irow, icol;
    if (!fx_interior(P, irow, icol)) return;
    const size_t ld = P.ld;
    const size_t p = (size_t)irow + (size_t)icol * ld;
    const size_t pw = p - 1, ps = p - ld;
    const double dtl = ctl->dtfl, dxy = P.dxy;
    const bool weir = s.weir[p] != 0;
    double dh = 0.0, dqx = 0.0, dqy = 0.0, qxf = 0.0, qyf = 0.0;
    if (!weir) {
        const double fe1p = s.fe1[p], fn1p = s.fn1[p];
        dh = ((s.fe1[pw] - fe1p) / dxy + (s.fn1[ps] - fn1p) / dxy) * dtl;
        dqx = ((s.fe2[pw] - s.fe2[p]) / dxy + (s.fn2[ps] - s.fn2[p]) / dxy) * dtl;
        dqy = ((s.fe3[pw] - s.fe3[p]) / dxy + (s.fn3[ps] - s.fn3[p]) / dxy) * dtl;
        qxf = fe1p * dtl;
        qyf = fn1p * dtl;
    }
    // face discharges of the W and S neighbours as loop 3 left them
    const double qxfw = (irow > 1 && !s.weir[pw]) ? s.fe1[pw] * dtl : 0.0;
    const double qyfs = (icol > 1 && !s.weir[ps]) ? s.fn1[ps] * dtl : 0.0;
    const double z = s.z[p] + dh;
    s.z[p] = z;
    const double hp = fmax(0.0, z - s.zb[p]);
    s.h[p] = hp;
    if (hp < P.hdry || weir) {
        s.qx[p] = 0.0;
        s.qy[p] = 0.0;
        s.ldry[p] = 1;
    } else {
        double qx = s.qx[p] + dqx;
        double qy = s.qy[p] + dqy;
        qx = .1 * qxfw + .8 * qx + .1 * qxf;
        qy = .1 * qyfs + .8 * qy + .1 * qyf;
        s.qx[p] = qx;
        s.qy[p] = qy;
        s.ldry[p] = 0;
        if (!isfinite(z) || !isfinite(qx) || !isfinite(qy)) atomicOr(&ctl->status, 4);
    }
}

// WINTRAsolver_calc.cpp:26-57: a wet cell whose bed is not NODATA releases
// soil_rate/3600 of its soil mass per second of the step into the solute,
// once per solute in solute order (main.cpp:400-403)
__global__ __launch_bounds__(256) void k64_wintra(fx_params P, fx_state64 s, double *soil, fx_ctl *ctl,
                                                  unsigned mode, int k_first, int k_count)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate)) return;
    int irow, icol;
    if (!fx_interior(P, irow, icol)) return;
    const size_t p = (size_t)irow + (size_t)icol * P.ld;
    const double hp = s.h[p];
    const double zbp = s.zb[p];
    if (!(hp > P.hdry && zbp != P.nodata)) return;
    const double dtl = ctl->dtfl;
    double m = soil[p];
    for (int k = k_first; k < k_first + k_count; k++) {
        const double deltam = m * P.soil_rate / 3600 * dtl;
        m = m - deltam;
        s.conc[k][p] = s.conc[k][p] + deltam / (hp * P.arbase);
    }
    soil[p] = m;
}

__global__ void k_forcing(fx_params P, fx_ctl *ctl) { fx_forcing_eval(P, ctl); }

__global__ void k_step_end(fx_ctl *ctl)
{
    if (!ctl->halt) ctl->halt = ctl->pending_halt | ((ctl->status >> 2) & 1);   // halt on NaN/Inf
}

// water volume of non-weir cells (diagnostic, block partials + host sum)
__global__ __launch_bounds__(256) void k64_volume(fx_params P, fx_state64 s, double *part)
{
    int irow, icol;
    double v = 0.0;
    if (fx_interior(P, irow, icol)) {
        const size_t p = (size_t)irow + (size_t)icol * P.ld;
        if (!s.weir[p]) v = (s.z[p] - s.zb[p]);
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __shared__ double sm[4];
    const int tid = threadIdx.x + threadIdx.y * BX;
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) part[blockIdx.x + blockIdx.y * gridDim.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

inline dim3 grid_of(const fx_params &P)
{
    return dim3((P.nrows + BX - 1) / BX, (P.ncols + BY - 1) / BY);
}

} // namespace

int fx64_launch_cfl(fx_domain d, unsigned mode)
{
    fx_params P = fx_make_params(d);
    const int nb = std::min(std::min(P.ncols, 2048), d->npartials);
    fx_prof_begin(d, FX_K_CFL);
    hipLaunchKernelGGL(k64_cfl, dim3(nb), dim3(256), 0, d->stream, P, d->s64, d->ctl,
                       d->partials, mode);
    fx_prof_end(d, FX_K_CFL);
    return fx_hip_check(hipGetLastError(), "k64_cfl");
}

int fx64_launch_melt(fx_domain d, unsigned mode)
{
    fx_params P = fx_make_params(d);
    fx_prof_begin(d, FX_K_MELT);
    hipLaunchKernelGGL(k64_melt, grid_of(P), dim3(BX, BY), 0, d->stream, P, d->s64, d->ctl, mode);
    fx_prof_end(d, FX_K_MELT);
    return fx_hip_check(hipGetLastError(), "k64_melt");
}

int fx64_launch_flux(fx_domain d, unsigned mode)
{
    fx_params P = fx_make_params(d);
    fx_prof_begin(d, FX_K_FLUX);
    hipLaunchKernelGGL(k64_flux, grid_of(P), dim3(BX, BY), 0, d->stream, P, d->s64, d->ctl, mode);
    fx_prof_end(d, FX_K_FLUX);
    return fx_hip_check(hipGetLastError(), "k64_flux");
}

int fx64_launch_wintra(fx_domain d, unsigned mode, int k_first, int k_count)
{
    fx_params P = fx_make_params(d);
    hipLaunchKernelGGL(k64_wintra, grid_of(P), dim3(BX, BY), 0, d->stream, P, d->s64, d->soil, d->ctl, mode,
                       k_first, k_count);
    return fx_hip_check(hipGetLastError(), "k64_wintra");
}

int fx64_launch_update(fx_domain d, unsigned mode)
{
    fx_params P = fx_make_params(d);
    fx_prof_begin(d, FX_K_UPDATE);
    hipLaunchKernelGGL(k64_update, grid_of(P), dim3(BX, BY), 0, d->stream, P, d->s64, d->ctl, mode);
    fx_prof_end(d, FX_K_UPDATE);
    return fx_hip_check(hipGetLastError(), "k64_update");
}

int fx_launch_forcing(fx_domain d)
{
    fx_params P = fx_make_params(d);
    hipLaunchKernelGGL(k_forcing, dim3(1), dim3(1), 0, d->stream, P, d->ctl);
    return fx_hip_check(hipGetLastError(), "k_forcing");
}

int fx_launch_step_end(fx_domain d)
{
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, d->stream, d->ctl);
    return fx_hip_check(hipGetLastError(), "k_step_end");
}

namespace {
// solute mass per unit area (c h) of interior non-weir cells, block partials
__global__ __launch_bounds__(256) void k64_solute(fx_params P, fx_state64 s, int k, double *part)
{
    int irow, icol;
    double v = 0.0;
    if (fx_interior(P, irow, icol)) {
        const size_t p = (size_t)irow + (size_t)icol * P.ld;
        if (!s.weir[p]) v = s.conc[k][p] * s.h[p];
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __shared__ double sm[4];
    const int tid = threadIdx.x + threadIdx.y * BX;
    if ((tid & 63) == 0) sm[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) part[blockIdx.x + blockIdx.y * gridDim.x] = sm[0] + sm[1] + sm[2] + sm[3];
}
} // namespace

int fx64_launch_solute(fx_domain d, int k, double *mass_out)
{
    fx_params P = fx_make_params(d);
    dim3 g = grid_of(P);
    const size_t nb = (size_t)g.x * g.y;
    double *part = nullptr;
    FX_HIP(hipMallocAsync((void **)&part, nb * sizeof(double), d->stream));
    hipLaunchKernelGGL(k64_solute, g, dim3(BX, BY), 0, d->stream, P, d->s64, k, part);
    std::vector<double> h(nb);
    FX_HIP(hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipFreeAsync(part, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    double s = 0.0, c = 0.0;
    for (double v : h) { double y = v - c; double t = s + y; c = (t - s) - y; s = t; }
    *mass_out = s * P.arbase;
    return FX_OK;
}

int64_t fx64_volume_nparts(fx_domain d)
{
    const dim3 g = grid_of(fx_make_params(d));
    return (int64_t)g.x * g.y;
}

// block partials of the water volume into part[0 .. fx64_volume_nparts)
int fx64_volume_partials(fx_domain d, double *part)
{
    fx_params P = fx_make_params(d);
    hipLaunchKernelGGL(k64_volume, grid_of(P), dim3(BX, BY), 0, d->stream, P, d->s64, part);
    return fx_hip_check(hipGetLastError(), "k64_volume");
}

int fx64_launch_volume(fx_domain d, double *vol_out)
{
    fx_params P = fx_make_params(d);
    const size_t nb = (size_t)fx64_volume_nparts(d);
    double *part = nullptr;
    FX_HIP(hipMallocAsync((void **)&part, nb * sizeof(double), d->stream));
    int rc = fx64_volume_partials(d, part);
    if (rc) {
        (void)hipFreeAsync(part, d->stream);
        return rc;
    }
    std::vector<double> h(nb);
    FX_HIP(hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipFreeAsync(part, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    // Kahan over block partials
    double s = 0.0, c = 0.0;
    for (double v : h) { double y = v - c; double t = s + y; c = (t - s) - y; s = t; }
    *vol_out = s * P.arbase;
    return FX_OK;
}
