/*
 * fluxos_oracle.c -- TEST INFRASTRUCTURE ONLY.  CPU restatement of the
 * FLUXOS-OVERLAND time step (see fluxos_oracle.h for the function map).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, as the checker; the product never links it.
 *
 * Expression order, operand types and float/double promotions follow the
 * reference so that a correct implementation agrees bit for bit when built
 * without FP contraction (-ffp-contract=off).  Each block cites the
 * reference lines it restates.  ORC_PLANT_* macros plant deliberate
 * mistakes; tests/test_oracle_plants.py builds such variants to show that
 * the pinning tests catch them.  A normal build defines none of them.
 */
#include "fluxos_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IX(d, r, c) ((size_t)(r) + (size_t)(c) * (size_t)(d)->mrows)

#ifdef ORC_PLANT_PRESSURE_HALF
#define K_PRESS 1.0
#else
#define K_PRESS .5
#endif
#ifdef ORC_PLANT_SMOOTH
#define K_SMOOTH_C .9
#else
#define K_SMOOTH_C .8
#endif
#ifdef ORC_PLANT_RITTER
#define K_RITTER 0.3
#else
#define K_RITTER 0.296
#endif
#ifdef ORC_PLANT_DIFF
#define K_SIGC 1.0
#else
#define K_SIGC 0.5
#endif

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* thread count of the hydro loops' parallel region (timing only) */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

/* the OpenMP thread-binding policy in effect (omp_proc_bind_t: 0 false,
 * 1 true, 2 master/primary, 3 close, 4 spread) */
int orc_proc_bind(void)
{
#ifdef _OPENMP
    return (int)omp_get_proc_bind();
#else
    return 0;
#endif
}

/* ------------------------------------------------------------------ */
/* CFL condition: main.cpp:300-325                                      */
/* ------------------------------------------------------------------ */
double orc_courant(orc_domain *d, double print_next)
{
    double dtfl = 9.e10;              /* main.cpp:274 */
    double hpall = 0.0;               /* main.cpp:275 */
    /* ds.cfl*ds.dxy: float times size_t is evaluated in float */
    const float cfl_dxy = d->cfl * (float)d->dxy;
    const double g = (double)d->gacc;
    for (int icol = 1; icol <= d->ncols; icol++) {
        for (int irow = 1; irow <= d->nrows; irow++) {
            const size_t p = IX(d, irow, icol);
            double hp = d->h[p];
            d->h0[p] = hp;
            d->ldry_prev[p] = d->ldry[p];
            if (hp > d->hdry) {
                d->ldry[p] = 0.0f;
                hp = fmax(d->h[p], d->hdry);
                hpall = fmax(hpall, d->h[p]);
                double c0 = sqrt(g * d->h[p]);
                double u0 = fmax(.000001, fabs(d->qx[p] / hp));
#ifdef ORC_PLANT_CFL_V
                double v0 = u0;
#else
                double v0 = fmax(.000001, fabs(d->qy[p] / hp));
#endif
                dtfl = fmin(fmin((double)cfl_dxy / (u0 + c0),
                                 (double)cfl_dxy / (v0 + c0)), dtfl);
            } else {
                d->ldry[p] = 1.0f;
            }
            dtfl = fmin(print_next - d->tim, dtfl);
        }
    }
    d->dtfl = dtfl;
    d->tim = d->tim + dtfl;           /* main.cpp:325 */
    return hpall;
}

/* ------------------------------------------------------------------ */
/* Forcing row lookup shared by add_meteo/add_inflow                    */
/* (common.cpp:195-200, 307-312): first row whose time exceeds tim.     */
/* Past the end of the table the rate and concentrations are zero.      */
/* ------------------------------------------------------------------ */
static int forcing_row(const float *tab, int nrows, int ncols, double tim)
{
    for (int a = 0; a < nrows; a++) {
#ifdef ORC_PLANT_MELT_ROW
        if ((double)tab[(size_t)a * ncols] >= tim) return a;
#else
        if ((double)tab[(size_t)a * ncols] > tim) return a;
#endif
    }
    return -1;
}

/* common.cpp:164-273 */
int orc_add_meteo(orc_domain *d)
{
    if (d->meteo == NULL) return 0;              /* common.cpp:190-191 */
    const int row = forcing_row(d->meteo, d->meteo_nrows, d->meteo_ncols, d->tim);
    double rate = 0.0;
    double cmelt[ORC_MAX_CHEM] = {0};
    if (row >= 0) {
        rate = (double)d->meteo[(size_t)row * d->meteo_ncols + 1];
        for (int k = 0; k < d->nchem; k++)
            cmelt[k] = (double)d->meteo[(size_t)row * d->meteo_ncols + 2 + k];
    }
    /* mm/day -> m per step: common.cpp:203 */
    const double meteoi0 = rate / (1000. * 3600. * 24.) * d->dtfl;
    for (int icol = 1; icol <= d->ncols; icol++) {
        for (int irow = 1; irow <= d->nrows; irow++) {
            const size_t p = IX(d, irow, icol);
            if (d->zb[p] == d->nodata) continue;
            const double meteoi = d->melt_mult ? meteoi0 * (double)d->melt_mult[p] : meteoi0;
            double hp = fmax(d->z[p] - d->zb[p], 0.0);
            d->z[p] = d->z[p] + meteoi;
            d->h[p] = fmax(d->z[p] - d->zb[p], 0.0);
            if (d->h[p] <= d->hdry) d->ldry[p] = 0.0f;   /* common.cpp:219-222 */
            d->h0[p] = d->h[p];
            if (d->ade_solver && hp != 0.0) {
                for (int k = 0; k < d->nchem; k++)
                    d->conc[k][p] = (d->conc[k][p] * hp + (meteoi * cmelt[k])) / (d->h[p]);
            }
        }
    }
    return 0;
}

/* common.cpp:277-390 */
int orc_add_inflow(orc_domain *d)
{
    if (d->inflow == NULL) return 0;
    const int row = forcing_row(d->inflow, d->inflow_nrows, d->inflow_ncols, d->tim);
    double q = 0.0;
    double cin[ORC_MAX_CHEM] = {0};
    if (row >= 0) {
        q = (double)d->inflow[(size_t)row * d->inflow_ncols + 1];
        for (int k = 0; k < d->nchem; k++)
            cin[k] = (double)d->inflow[(size_t)row * d->inflow_ncols + 2 + k];
    }
    const double dxy = (double)d->dxy;
    const double inflowi = q * d->dtfl / (pow(dxy, 2));     /* common.cpp:317 */
    const size_t p = IX(d, d->inflow_row, d->inflow_col);
    if (d->zb[p] == d->nodata) return 1;
    if (d->z[p] <= d->zb[p]) d->z[p] = d->zb[p];
    double hp = fmax(d->z[p] - d->zb[p], 0.0);
    d->z[p] = d->z[p] + inflowi;
    d->h[p] = fmax(d->z[p] - d->zb[p], 0.0);
    d->ldry[p] = (d->h[p] <= d->hdry) ? 1.0f : 0.0f;
    d->h0[p] = d->h[p];
    if (d->ade_solver && hp != 0.0) {
        for (int k = 0; k < d->nchem; k++)
            d->conc[k][p] = (d->conc[k][p] * hp + (inflowi * cin[k])) / (d->h[p]);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* solver_dry: solver_drydomain.cpp:27-194                              */
/* ------------------------------------------------------------------ */
void orc_solver_dry(orc_domain *d, int irow, int icol)
{
    const size_t p = IX(d, irow, icol);
    const size_t pw = IX(d, irow - 1, icol), pe = IX(d, irow + 1, icol);
    const size_t ps = IX(d, irow, icol - 1), pn = IX(d, irow, icol + 1);
    const float ldw = d->ldry[pw], ldp = d->ldry[p], lde = d->ldry[pe];
    const float lds = d->ldry[ps], ldn = d->ldry[pn];
    const double g = (double)d->gacc;

    /* all-dry neighbourhood: nothing stored, fluxes keep their old value
     * (solver_drydomain.cpp:60-76) */
    if (ldw == 1 && ldp == 1 && lde == 1 && lds == 1 && ldn == 1) {
        d->dh[p] = 0.0; d->dqx[p] = 0.0; d->dqy[p] = 0.0;
        d->qxf[p] = 0.0; d->qyf[p] = 0.0; d->qx[p] = 0.0; d->qy[p] = 0.0;
#ifdef ORC_PLANT_STALE
        d->fe1[p] = d->fe2[p] = d->fe3[p] = d->fn1[p] = d->fn2[p] = d->fn3[p] = 0.0;
#endif
        return;
    }
    const double zbp = d->zb[p], zbe = d->zb[pe], zbn = d->zb[pn];
    const double zp = d->z[p], ze = d->z[pe], zn = d->z[pn];
    /* hp is z(P)-z(P), i.e. zero (solver_drydomain.cpp:85) */
    const double hp = fmax(0.0, d->z[p] - d->z[p]);
    const double he = fmax(0.0, ze - zbe);
    const double hn = fmax(0.0, zn - zbn);
#ifdef ORC_PLANT_DRY_BED
    const double zbpe = zbp;
#else
    const double zbpe = .5 * (zbe + zbp);
#endif
    const double zbpn = .5 * (zbn + zbp);
    double hme, qme, hmn, rmn, dze = ze - zp, dzn = zn - zp;
    double volrat = 0.0;
    int volrat_set = 0;
    const double dtl = d->dtfl;
    double fe1 = 0.0, fe2 = 0.0, fe3 = 0.0, fn1 = 0.0, fn2 = 0.0, fn3 = 0.0;

    if (lde == 0.0f) {                                   /* :115-145 */
        hme = fmax(0.0, ze - zbpe);
        if (hme > d->hdry) {
            if (ze <= zp) {
                double fe2p = K_PRESS * g * hme * hme;
                fe1 = 0.0; fe2 = fe2p; fe3 = 0.0;
            } else {
                dze = fmin(fabs(dze), hme);
                qme = K_RITTER * dze * sqrt(g * dze);
                hme = 0.444 * dze;
                volrat = (double)d->dxy * he / dtl; volrat_set = 1;
                qme = -fmin(qme, volrat);
                double ume = qme / hme;
                fe1 = qme;
                fe2 = qme * ume + K_PRESS * g * hme * hme;
                fe3 = 0.0;
            }
        } else {
            fe1 = 0.0; fe2 = 0.0; fe3 = 0.0;
        }
    }
    if (ldn == 0.0f) {                                   /* :146-175 */
        hmn = fmax(0.0, zn - zbpn);
        if (hmn > d->hdry) {
            if (zn <= zp) {
                double fn3p = K_PRESS * g * hmn * hmn;
                fn1 = 0.0; fn2 = 0.0; fn3 = fn3p;
            } else {
                dzn = fmin(fabs(dzn), hmn);
                rmn = K_RITTER * dzn * sqrt(g * dzn);
                hmn = 0.444 * dzn;
                volrat = (double)d->dxy * hn / dtl; volrat_set = 1;
                rmn = -fmin(rmn, volrat);
                double vmn = rmn / hmn;
                fn1 = rmn; fn2 = 0.0;
                fn3 = rmn * vmn + K_PRESS * g * hmn * hmn;
            }
        } else {
            fn1 = 0.0; fn2 = 0.0; fn3 = 0.0;
        }
    }
    /* weir discharge at the outer interior ring (:178-185); the reference
     * reads volrat even when no branch above set it */
    if (icol == 1 || icol == d->ncols) {
        if (!volrat_set) { d->ub_volrat++; volrat = INFINITY; }
        fn1 = fmin(volrat, sqrt(g) * pow(fmax(hp, 0.0), 1.5));
    }
    if (irow == 1 || irow == d->nrows) {
        if (!volrat_set) { d->ub_volrat++; volrat = INFINITY; }
        fe1 = fmin(volrat, sqrt(g) * pow(fmax(hp, 0.0), 1.5));
    }
    d->fn1[p] = fn1; d->fn2[p] = fn2; d->fn3[p] = fn3;
    d->fe1[p] = fe1; d->fe2[p] = fe2; d->fe3[p] = fe3;
}

/* ------------------------------------------------------------------ */
/* solver_wet: solver_wetdomain.cpp:27-342                              */
/* ------------------------------------------------------------------ */
void orc_solver_wet(orc_domain *d, int irow, int icol)
{
    const unsigned dx = (unsigned)d->dxy, dy = (unsigned)d->dxy;
    const double hdryl = d->hdry;
    const double gaccl = (double)d->gacc;
    const double nueml = (double)d->nuem, cvdefl = (double)d->cvdef;
    const size_t p = IX(d, irow, icol);
    const size_t pw = IX(d, irow - 1, icol), pe = IX(d, irow + 1, icol);
    const size_t ps = IX(d, irow, icol - 1), pn = IX(d, irow, icol + 1);
    const size_t pes = IX(d, irow + 1, icol - 1), pen = IX(d, irow + 1, icol + 1);
    const size_t pwn = IX(d, irow - 1, icol + 1);
    const double kspl = d->ks[p];
    const float ldp = d->ldry[p], lde = d->ldry[pe], ldn = d->ldry[pn];
    int lroe = 1;
    double volrat = 0.0;
    int volrat_set = 0;

    const double zbw = d->zb[pw], zbp = d->zb[p], zbe = d->zb[pe];
    const double zbs = d->zb[ps], zbn = d->zb[pn];
    const double zw = d->z[pw], zp = d->z[p], ze = d->z[pe], zs = d->z[ps], zn = d->z[pn];
    double qp = d->qx[p], qe = d->qx[pe], qn = d->qx[pn];
    double rw = d->qy[pw], rp = d->qy[p], re = d->qy[pe], rn = d->qy[pn];

    const double zbpe = .5 * (zbe + zbp);
    const double zbpn = .5 * (zbn + zbp);
    const double hp = fmax(0.0, d->z[p] - zbp);
    const double hp0 = fmax(fmax(hdryl, hp), kspl);
    const double hw = fmax(0.0, zw - zbw);
    double he = fmax(0.0, ze - zbe);
    const double hs = fmax(0.0, zs - zbs);
    double hn = fmax(0.0, zn - zbn);
    /* the (irow, icol+2) depth is overwritten before use (:100,189) */

    double hme = .5 * (hp + he), qme = .5 * (qp + qe), rme = .5 * (rp + re);
    double hmn = .5 * (hp + hn), qmn = .5 * (qp + qn), rmn = .5 * (rp + rn);
    double dze = ze - zp;
    const double dqe = qe - qp, dre = re - rp;
    double dzn = zn - zp;
    const double drn = rn - rp, dqn = qn - qp;
    const double dtl = d->dtfl;

    if (lde == 1.0f) {                                   /* :121-147 */
        hme = fmax(0.0, zp - zbpe);
        he = 0.0; qe = 0.0; re = 0.0;
        if (hme > hdryl) {
            if (ze >= zp) {
                qme = 0.0; rme = 0.0;
            } else {
                dze = fmin(fabs(dze), hme);
                qme = K_RITTER * dze * sqrt(gaccl * dze);
                volrat = 0.5 * dx * hp / dtl; volrat_set = 1;
                qme = fmin(qme, volrat);
                rme = 0.0;
            }
        } else {
            qme = 0.0; rme = 0.0; hme = 0.0;
        }
        lroe = 0;
    }
    if (ldn == 1.0f) {                                   /* :148-175 */
        hmn = fmax(0.0, zp - zbpn);
        hn = 0.0; qn = 0.0; rn = 0.0;
        if (hmn > hdryl) {
            if (zn >= zp) {
                qmn = 0.0; rmn = 0.0;
            } else {
                dzn = fmin(fabs(dzn), hmn);
                rmn = K_RITTER * dzn * sqrt(gaccl * dzn);
                qmn = 0.0;
                volrat = 0.5 * dy * hp / dtl; volrat_set = 1;
                rmn = fmin(rmn, volrat);
            }
        } else {
            qmn = 0.0; rmn = 0.0; hmn = 0.0;
        }
        lroe = 0;
    }

    /* turbulent stress (:177-203) */
    const double cme = sqrt(gaccl * hme);
    const double cmn = sqrt(gaccl * hmn);
    const double ume = qme / fmax(hme, hdryl);
    const double vme = rme / fmax(hme, hdryl);
    const double umn = qmn / fmax(hmn, hdryl);
    const double vmn = rmn / fmax(hmn, hdryl);
    const double cnp = cvdefl * d->us[p] * hp + nueml;
    const double cne = cvdefl * d->us[pe] * he + nueml;
    const double cnn = cvdefl * d->us[pn] * hn + nueml;
    const double hne = .5 * (cnp + cne) * sqrt(hp * he);
    const double hnn = .5 * (cnp + cnn) * sqrt(hp * hn);
    const double up = qp / hp0;
    const double un = qn / fmax(fmax(hn, hdryl), d->ks[pn]);
    const double us0 = d->qx[ps] / fmax(fmax(hs, hdryl), d->ks[ps]);
    const double use0 = d->qx[pes] / fmax(fmax(d->h[pes], hdryl), d->ks[pes]);
    const double une = d->qx[pen] / fmax(fmax(d->h[pen], hdryl), d->ks[pen]);
    const double vp = rp / hp0;
    const double ve = re / fmax(fmax(he, hdryl), d->ks[pe]);
    const double vw = rw / fmax(fmax(hw, hdryl), d->ks[pw]);
    const double vwn = d->qy[pwn] / fmax(fmax(d->h[pwn], hdryl), d->ks[pwn]);
    const double ven = d->qy[pen] / fmax(fmax(d->h[pen], hdryl), d->ks[pen]);
#ifdef ORC_PLANT_TURB
    const double txye = hne * ((vp - ve) / fabs((double)dx) + .25 * (un + une - us0 - use0) / dy);
#else
    const double txye = hne * ((ve - vp) / fabs((double)dx) + .25 * (un + une - us0 - use0) / dy);
#endif
    const double txyn = hnn * ((un - up) / fabs((double)dy) + .25 * (ve + ven - vw - vwn) / dx);

    /* convective fluxes (:205-211) */
#ifdef ORC_PLANT_CONV
    const double fe1c = qme, fe2c = qme * vme, fe3c = qme * vme - txye;
#else
    const double fe1c = qme, fe2c = qme * ume, fe3c = qme * vme - txye;
#endif
    const double fn1c = rmn, fn2c = rmn * umn - txyn, fn3c = rmn * vmn;

    /* Roe dissipation (:213-285) */
    double fe1r, fe2r, fe3r, fn1r, fn2r, fn3r;
    if (lroe) {
        double cc, c1, c2, c3, c1a, c2a, c3a;
        double a11, a12, a13, a21, a22, a23, a31, a32, a33;
        if (hme > d->hdry) {
            double dzea = fabs(dze);
            if (dzea > 0.5 * hme) {
#ifndef ORC_PLANT_DZ_LIMIT
                double dhea = fabs(he - hp);
                dzea = fmin(dzea, dhea);
                dze = copysign(dzea, dze);
#endif
            }
            cc = .25 / cme;
        } else {
            cc = 0.0;
        }
        c1 = ume; c2 = ume + cme; c3 = ume - cme;
        c1a = fabs(c1); c2a = fabs(c2); c3a = fabs(c3);
        a11 = c2 * c3a - c2a * c3;
#ifdef ORC_PLANT_ROE_ABS
        a12 = c2 - c3;
#else
        a12 = c2a - c3a;
#endif
#ifdef ORC_PLANT_ROE_A21
        a21 = c2 * c3 * (c2a - c3a);
#else
        a21 = c2 * c3 * (c3a - c2a);
#endif
        a22 = c2a * c2 - c3a * c3;
        a31 = vme * (c2 * c3a - 2. * cme * c1a - c2a * c3);
        a32 = vme * (c2a - c3a);
#ifdef ORC_PLANT_ROE_SHEAR
        a33 = 2. * cme * c1;
#else
        a33 = 2. * cme * c1a;
#endif
#ifdef ORC_PLANT_ROE_SIGN
        fe1r = (a11 * dze + a12 * dqe) * cc;
#else
        fe1r = -(a11 * dze + a12 * dqe) * cc;
#endif
        fe2r = -(a21 * dze + a22 * dqe) * cc;
        fe3r = -(a31 * dze + a32 * dqe + a33 * dre) * cc;

        if (ldp == 0.0f && hmn > d->hdry) {
            double dzna = fabs(dzn);
            double dhna = fabs(hn - hp);
            dzna = fmin(dzna, dhna);
            dzn = copysign(dzna, dzn);
            cc = .25 / cmn;
        } else {
            cc = 0.0;
        }
        c1 = vmn; c2 = vmn + cmn; c3 = vmn - cmn;
        c1a = fabs(c1); c2a = fabs(c2); c3a = fabs(c3);
        a11 = c2 * c3a - c2a * c3;
        a13 = c2a - c3a;
        a21 = umn * (c2 * c3a - 2. * cmn * c1a - c2a * c3);
        a22 = 2. * cmn * c1a;
        a23 = umn * (c2a - c3a);
        a31 = c2 * c3 * (c3a - c2a);
        a33 = c2a * c2 - c3a * c3;
        fn1r = -(a11 * dzn + a13 * drn) * cc;
        fn2r = -(a21 * dzn + a22 * dqn + a23 * drn) * cc;
        fn3r = -(a31 * dzn + a33 * drn) * cc;
    } else {
        fe1r = 0.0; fe2r = 0.0; fe3r = 0.0;
        fn1r = 0.0; fn2r = 0.0; fn3r = 0.0;
    }

    /* pressure at the cell side (:288-289) */
    const double fe2p = K_PRESS * gaccl * (hme * hme);
    const double fn3p = K_PRESS * gaccl * (hmn * hmn);

    double fe1 = fe1c + fe1r;
    const double fe2 = fe2c + fe2r + fe2p;
    const double fe3 = fe3c + fe3r;
    double fn1 = fn1c + fn1r;
    const double fn2 = fn2c + fn2r;
    const double fn3 = fn3c + fn3r + fn3p;

    /* weir discharge (:300-307) */
    if (icol == 1 || icol == d->ncols) {
        if (!volrat_set) { d->ub_volrat++; volrat = INFINITY; }
        fn1 = fmin(volrat, sqrt(gaccl) * pow(fmax(hp, 0.0), 1.5));
    }
    if (irow == 1 || irow == d->nrows) {
        if (!volrat_set) { d->ub_volrat++; volrat = INFINITY; }
        fe1 = fmin(volrat, sqrt(gaccl) * pow(fmax(hp, 0.0), 1.5));
    }

    /* outflow limited to the water in the cell (:309-332) */
    const double volpot = (double)d->arbase * hp;
    volrat = volpot / dtl;
    if (volrat > 0.0) {
        if (fe1 > 0.0 && fn1 > 0.0) {
            if (fe1 * dy + fn1 * dx > volrat) {
                double cf = fn1 * dx / (fe1 * dy + fn1 * dx);
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
    d->fn1[p] = fn1; d->fn2[p] = fn2; d->fn3[p] = fn3;
    d->fe1[p] = fe1; d->fe2[p] = fe2; d->fe3[p] = fe3;
}

/* ------------------------------------------------------------------ */
/* hydrodynamics_calc.cpp:27-157                                        */
/* ------------------------------------------------------------------ */
void orc_hydrodynamics(orc_domain *d)
{
    const double dtl = d->dtfl;
    const double dxy = (double)d->dxy;
    const int nr = d->nrows, nc = d->ncols;
#pragma omp parallel
    {
        /* loop 1: depth, wet/dry flag, wet-time tracer (:53-73) */
#pragma omp for schedule(static)
        for (int icol = 1; icol <= nc; icol++) {
            for (int irow = 1; irow <= nr; irow++) {
                const size_t p = IX(d, irow, icol);
                double hp = fmax(0.0, d->z[p] - d->zb[p]);
                d->h[p] = hp;
                if (hp <= d->hdry) {
                    d->qx[p] = 0.0; d->qy[p] = 0.0; d->us[p] = 0.0;
                    d->ldry[p] = 1.0f;
                } else {
                    d->ldry[p] = 0.0f;
                    d->twet[p] += dtl / 3600;
                }
            }
        }
        /* loop 2: face fluxes (:76-90) */
#pragma omp for schedule(static)
        for (int icol = 1; icol <= nc; icol++) {
            for (int irow = 1; irow <= nr; irow++) {
                const size_t p = IX(d, irow, icol);
                const float cell_neumann = d->weir[p];
                if (d->ldry[p] == 1.0f && cell_neumann == 0.0f)
                    orc_solver_dry(d, irow, icol);
                else if (cell_neumann == 0.0f)
                    orc_solver_wet(d, irow, icol);
            }
        }
        /* loop 3: flux divergence (:93-120) */
#pragma omp for schedule(static)
        for (int icol = 1; icol <= nc; icol++) {
            for (int irow = 1; irow <= nr; irow++) {
                const size_t p = IX(d, irow, icol);
#ifdef ORC_PLANT_LOOP3_INDEX
                const size_t pw = IX(d, irow + 1, icol);
#else
                const size_t pw = IX(d, irow - 1, icol);
#endif
                const size_t ps = IX(d, irow, icol - 1);
                if (d->weir[p] == 0.0f) {
                    d->dh[p] = ((d->fe1[pw] - d->fe1[p]) / dxy
                              + (d->fn1[ps] - d->fn1[p]) / dxy) * dtl;
                    d->dqx[p] = ((d->fe2[pw] - d->fe2[p]) / dxy
                               + (d->fn2[ps] - d->fn2[p]) / dxy) * dtl;
                    d->dqy[p] = ((d->fe3[pw] - d->fe3[p]) / dxy
                               + (d->fn3[ps] - d->fn3[p]) / dxy) * dtl;
                    d->qxf[p] = d->fe1[p] * dtl;
                    d->qyf[p] = d->fn1[p] * dtl;
                } else {
                    d->dh[p] = 0.0; d->dqx[p] = 0.0; d->dqy[p] = 0.0;
                    d->qxf[p] = 0.0; d->qyf[p] = 0.0;
                }
            }
        }
        /* loop 4: new state with 0.1/0.8/0.1 face smoothing (:123-154) */
#pragma omp for schedule(static)
        for (int icol = 1; icol <= nc; icol++) {
            for (int irow = 1; irow <= nr; irow++) {
                const size_t p = IX(d, irow, icol);
                const size_t pw = IX(d, irow - 1, icol), ps = IX(d, irow, icol - 1);
                d->z[p] = d->z[p] + d->dh[p];
                double hp = fmax(0.0, d->z[p] - d->zb[p]);
                d->h[p] = hp;
                if (hp < d->hdry || d->weir[p] == 1.0f) {
                    d->qx[p] = 0.0; d->qy[p] = 0.0; d->us[p] = 0.0;
                    d->ldry[p] = 1.0f;
                } else {
                    d->qx[p] = d->qx[p] + d->dqx[p];
                    d->qy[p] = d->qy[p] + d->dqy[p];
                    d->qx[p] = .1 * d->qxf[pw] + K_SMOOTH_C * d->qx[p] + .1 * d->qxf[p];
                    d->qy[p] = .1 * d->qyf[ps] + K_SMOOTH_C * d->qy[p] + .1 * d->qyf[p];
                    d->ldry[p] = 0.0f;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* adesolver_calc: ADEsolver_calc.cpp:27-291                            */
/* ------------------------------------------------------------------ */
void orc_adesolver(orc_domain *d, int it, int ichem, double *scratch)
{
    if (it <= 1) return;                                 /* :45,74-77 */
    const size_t mm = (size_t)d->mrows * d->mcols;
    double *own = NULL;
    if (!scratch) { own = (double *)calloc(mm * 2 + (size_t)d->mrows * d->mcols, sizeof(double)); scratch = own; }
    double *cmaxr = scratch, *cminr = scratch + mm;
    double *qfcds = scratch + 2 * mm;                   /* mrows*mcols x 1, zeroed */
    memset(qfcds, 0, mm * sizeof(double));
    double *c = d->conc[ichem];
    const int nr = d->nrows, nc = d->ncols;
    const double nt = 1, sigc = K_SIGC;

    /* pass 1: neighbour bounds and rescale to the new depth (:47-73) */
    for (int iy = 1; iy <= nc; iy++) {
        for (int ix = 1; ix <= nr; ix++) {
            const size_t p = IX(d, ix, iy);
            cmaxr[p] = fmax(c[IX(d, ix - 1, iy)],
                       fmax(c[IX(d, ix + 1, iy)],
                       fmax(c[IX(d, ix, iy - 1)], c[IX(d, ix, iy + 1)])));
            cminr[p] = fmin(c[IX(d, ix - 1, iy)],
                       fmin(c[IX(d, ix + 1, iy)],
                       fmin(c[IX(d, ix, iy - 1)], c[IX(d, ix, iy + 1)])));
            double hnew = d->h[p];
            if (d->ldry[p] == 0 && d->ldry_prev[p] == 0)
                c[p] = c[p] * d->h0[p] / hnew;
            else if (d->ldry[p] == 1)
                c[p] = 0.0;
        }
    }

    const double dx = (double)d->dxy, dy = (double)d->dxy;
    const double area = (double)d->arbase, areae = area, arean = area;
    const double dtfl = d->dtfl, D = d->d_coef[ichem];
    double pfe = 0.0;
    /* pass 2: sequential sweep, irow fastest; the last cell is not visited
     * (:93-290) */
    const long long ncell = (long long)nr * nc;
    for (long long a = 1; a < ncell; a++) {
        const int iy = (int)((a - 1) / nr) + 1;
        const int ix = (int)(a - (long long)nr * (iy - 1));
        const int is = iy - 1, in = iy + 1, inn = (iy + 2 < nc + 1) ? iy + 2 : nc + 1;
        const int iw = ix - 1, ie = ix + 1, iee = (ix + 2 < nr + 1) ? ix + 2 : nr + 1;
        const size_t p = IX(d, ix, iy);

        if (ix == 1) {                                   /* :106-116 */
            pfe = c[IX(d, 0, iy)] * d->fe1[IX(d, 0, iy)] * dy;
        }
        if (d->ldry[p] == 1) {                           /* :119-124 */
            pfe = 0.0;
            qfcds[ix] = 0.0;
            c[p] = 0.0;
            continue;
        }
        const double hp = d->h[p], he = d->h[IX(d, ie, iy)], hn = d->h[IX(d, ix, in)];
        const double qxl = d->fe1[p], qyl = d->fn1[p];
        const double fw = c[IX(d, iw, iy)], fp = c[p], fe = c[IX(d, ie, iy)];
        const double fee = c[IX(d, iee, iy)], fs = c[IX(d, ix, is)];
        const double fn = c[IX(d, ix, in)], fnn = c[IX(d, ix, inn)];
        const double pfw = pfe;
#ifdef ORC_PLANT_ADE_CARRY
        const double qfs = 0.0;
#else
        const double qfs = qfcds[ix];
#endif
        double fem, pfde, fnm, qfdn;
        /* x direction (:149-194) */
        if (d->ldry[IX(d, ie, iy)] == 0) {
            double hnue = fmax(hp * nt * he * nt, .0001);
            double hne = sqrt(hnue) / sigc / dx * dy * D;
            pfde = -hne * (fe - fp);
            if (qxl > 0.0) {
#ifdef ORC_PLANT_QUICK
                if (d->ldry[IX(d, iw, iy)] == 0) fem = -.125 * fw + .875 * fp + .25 * fe;
#else
                if (d->ldry[IX(d, iw, iy)] == 0) fem = -.125 * fw + .75 * fp + .375 * fe;
#endif
                else fem = 0.5 * fp + 0.5 * fe;
            } else {
                if (d->ldry[IX(d, iee, iy)] == 0) fem = .375 * fp + .75 * fe - .125 * fee;
                else fem = 0.5 * fp + 0.5 * fe;
            }
        } else {
            fem = 0.; pfde = 0.;
        }
        fem = fmax(0., fem);
        if (ix == nr) fem = c[IX(d, nr + 1, iy)];
        const double pfce = qxl * fem * dy;
        pfe = pfce + pfde;
        if (pfe < 0) {
            if (d->ldry[IX(d, ie, iy)] == 0) {
                double cvolrate = -(fe * he) * areae / dtfl;
                pfe = fmax(pfe, cvolrate);
            } else {
                pfe = 0.;
            }
        }
        /* y direction (:196-241) */
        if (d->ldry[IX(d, ix, in)] == 0) {
            double hnue = fmax(.0001, hp * nt * hn * nt);
            double hnn = sqrt(hnue) / sigc / dy * dx * D;
            qfdn = -hnn * (fn - fp);
            if (qyl > 0.0) {
                if (d->ldry[IX(d, ix, is)] == 0) fnm = -.125 * fs + .75 * fp + .375 * fn;
                else fnm = 0.5 * fp + .5 * fn;
            } else {
                if (d->ldry[IX(d, ix, inn)] == 0) fnm = .375 * fp + .75 * fn - .125 * fnn;
                else fnm = .5 * fp + .5 * fn;
            }
        } else {
            fnm = 0.; qfdn = 0.;
        }
        fnm = fmax(0., fnm);
        if (iy == nc) fnm = c[IX(d, ix, nc + 1)];
        const double qfcn = qyl * fnm * dx;
        double qfn = qfcn + qfdn;
        if (qfn < 0) {
            if (d->ldry[IX(d, ix, in)] == 0) {
                double cvolrate = -(fn * hn) * arean / dtfl;
                qfn = fmax(qfn, cvolrate);
            } else {
                qfn = 0.;
            }
        }
        /* available material in the cell (:243-278) */
        const double cvolpot = (fp * hp) * area;
        const double cvolrat = cvolpot / dtfl + pfw + qfs;
        if (cvolrat > 0.0) {
            if (pfe > 0. && qfn > 0.) {
#ifdef ORC_PLANT_ADE_LIMIT
                if (0) {
#else
                if (pfe + qfn > cvolrat) {
#endif
                    double cf = qfn / (pfe + qfn);
                    pfe = (1. - cf) * cvolrat;
                    qfn = cf * cvolrat;
                }
            } else if (pfe > 0.) {
#ifndef ORC_PLANT_ADE_LIMIT1
                pfe = fmin(pfe, (cvolrat - qfn));
#endif
            } else if (qfn > 0.) {
                qfn = fmin(qfn, (cvolrat - pfe));
            }
        } else {
            if (pfe >= 0. && qfn < 0.) {
                double cbilan = cvolrat - qfn;
#ifdef ORC_PLANT_ADE_PASS
                if (cbilan > 0.) pfe = pfe;
#else
                if (cbilan > 0.) pfe = fmin(pfe, cbilan);
#endif
#ifndef ORC_PLANT_ADE_NEG
                else pfe = 0.;
#endif
            } else if (pfe < 0. && qfn >= 0.) {
                double cbilan = cvolrat - pfe;
                if (cbilan > 0.) qfn = fmin(qfn, cbilan);
                else qfn = 0.;
            } else if (pfe >= 0. && qfn >= 0.) {
#ifndef ORC_PLANT_ADE_BILAN
                pfe = 0.0; qfn = 0.0;
#endif
            }
        }
        /* new concentration (:280-288) */
        const double dc = (pfw - pfe + qfs - qfn) * dtfl / area;
        double con = c[p] + dc / hp;
#ifndef ORC_PLANT_ADE_CLAMP
        con = fmin(cmaxr[p], con);
        con = fmax(cminr[p], con);
#endif
        c[p] = con;
        qfcds[ix] = qfn;
    }
    free(own);
}

/* ------------------------------------------------------------------ */
/* read_functions.cpp:175-225: weir cells and bed from the parsed DEM   */
/* ------------------------------------------------------------------ */
void orc_read_geo(orc_domain *d, const float *dem, double ks_input)
{
    const int nr = d->nrows, nc = d->ncols;
    const size_t mm = (size_t)d->mrows * d->mcols;
    float *raw = (float *)calloc(mm, sizeof(float));   /* zb_raw, zero-filled */
    /* first file line is irow = NROWS (read_functions.cpp:145,163) */
    for (int line = 0; line < nr; line++) {
        const int irow = nr - line;
        for (int j = 0; j < nc; j++) raw[IX(d, irow, j + 1)] = dem[(size_t)line * nc + j];
    }
    const float nod = (float)d->nodata;
    for (int icol = 1; icol <= nc; icol++) {
        for (int irow = 1; irow <= nr; irow++) {
            const size_t p = IX(d, irow, icol);
            const float zbp = raw[p];
            float zbp_corr;
            if ((double)zbp == d->nodata || zbp == 0.0f) {
                const float zbn = raw[IX(d, irow + 1, icol)], zbs = raw[IX(d, irow - 1, icol)];
                const float zbe = raw[IX(d, irow, icol + 1)], zbw = raw[IX(d, irow, icol - 1)];
                int n = 0;
                zbp_corr = 0.0f;
                d->weir[p] = 1.0f;
                if ((double)zbn != d->nodata && zbn != 0.0f && irow < nr) { zbp_corr += zbn; n++; }
                if ((double)zbs != d->nodata && zbs != 0.0f && irow > 1) { zbp_corr += zbs; n++; }
                if ((double)zbe != d->nodata && zbe != 0.0f && icol < nc) { zbp_corr += zbe; n++; }
                if ((double)zbw != d->nodata && zbw != 0.0f && icol > 1) { zbp_corr += zbw; n++; }
                if (zbp_corr == 0.0f) zbp_corr = zbp;
                else zbp_corr /= n;
            } else {
                d->weir[p] = 0.0f;
                zbp_corr = fabsf(zbp);
            }
            d->zb[p] = (double)zbp_corr;
            d->ks[p] = ks_input;
        }
    }
    (void)nod;
    free(raw);
}

/* ------------------------------------------------------------------ */
/* initiate.cpp:27-231, cold start                                      */
/* ------------------------------------------------------------------ */
/* initiate.cpp:42-90: ghost copies of the raw bed, then the vertex to
 * cell-centre average over the non-(+-99999) corners */
static void init_bed(orc_domain *d)
{
    const int nr = d->nrows, nc = d->ncols;
    const int nr1 = nr + 1, nc1 = nc + 1;
    double *zbs1 = (double *)calloc((size_t)d->mrows, sizeof(double));
    double *zb = d->zb;
    /* boundary elevations (:43-55) */
    for (int irow = 0; irow <= nr1; irow++) {
        zbs1[irow] = zb[IX(d, irow, 1)];
        zb[IX(d, irow, 0)] = zb[IX(d, irow, 1)];
        zb[IX(d, irow, nc1)] = zb[IX(d, irow, nc)];
    }
    for (int icol = 0; icol <= nc1; icol++) {
        zb[IX(d, 0, icol)] = zb[IX(d, 1, icol)];
        zb[IX(d, nr1, icol)] = zb[IX(d, nr, icol)];
    }
    zb[IX(d, 0, nc1)] = zb[IX(d, 1, nc)];
    zb[IX(d, nr1, nc1)] = zb[IX(d, nr, nc)];
    /* vertices to cell centres (:57-90) */
    for (int icol = 1; icol <= nc; icol++) {
        double zbsw = zbs1[1];
        const int icol1 = icol + 1;
        double zbnw = zb[IX(d, 1, icol1)];
        zbs1[1] = zbnw;
        for (int irow = 1; irow <= nr; irow++) {
            const int irow1 = irow + 1;
            const double zbse = zbs1[irow1];
            const double zbne = zb[IX(d, irow1, icol1)];
            unsigned a = 0;
            double zbsum = 0;
            if (fabs(zbsw) != 99999) { zbsum = zbsum + zbsw; a = a + 1; }
            if (fabs(zbse) != 99999) { zbsum = zbsum + zbse; a = a + 1; }
            if (fabs(zbnw) != 99999) { zbsum = zbsum + zbnw; a = a + 1; }
            if (fabs(zbne) != 99999) { zbsum = zbsum + zbne; a = a + 1; }
            zb[IX(d, irow, icol)] = (zbsum) / a;
            zbs1[irow1] = zbne;
            zbsw = zbse;
            zbnw = zbne;
        }
    }
    free(zbs1);
}

/* initiate.cpp:93-107: depth from the (still zero) z, discharges from the
 * (zero) velocities, background soil mass, for cells with |zb| != 99999 */
static void init_prelude(orc_domain *d)
{
    const int nr = d->nrows, nc = d->ncols;
    const double *zb = d->zb;
    for (int icol = 1; icol <= nc; icol++) {
        for (int irow = 1; irow <= nr; irow++) {
            const size_t p = IX(d, irow, icol);
            if (fabs(zb[p]) != 99999) {
                d->h[p] = fmax(d->z[p] - zb[p], 0.0);
                d->z[p] = zb[p] + d->h[p];
                d->qx[p] = (d->ux ? d->ux[p] : 0.0) * d->h[p];
                d->qy[p] = (d->uy ? d->uy[p] : 0.0) * d->h[p];
                if (d->soil) d->soil[p] = d->soil_bckgrd;
            }
        }
    }
}

/* initiate.cpp:192-228: ghost rows then ghost columns by linear extrapolation */
static void init_boundary(orc_domain *d)
{
    const int nr = d->nrows, nc = d->ncols;
    const int nr1 = nr + 1, nc1 = nc + 1;
    double *zb = d->zb;
    for (int icol = 0; icol <= nc1; icol++) {
        zb[IX(d, 0, icol)] = 1.5 * zb[IX(d, 1, icol)] - .5 * zb[IX(d, 2, icol)];
        zb[IX(d, nr1, icol)] = 1.5 * zb[IX(d, nr, icol)] - .5 * zb[IX(d, nr - 1, icol)];
        d->z[IX(d, 0, icol)] = 1.5 * d->z[IX(d, 1, icol)] - .5 * d->z[IX(d, 2, icol)];
        d->z[IX(d, nr1, icol)] = 1.5 * d->z[IX(d, nr, icol)] - .5 * d->z[IX(d, nr - 1, icol)];
        d->h[IX(d, 0, icol)] = fmax(0.0, d->z[IX(d, 0, icol)] - zb[IX(d, 0, icol)]);
        d->h[IX(d, nr1, icol)] = fmax(0.0, d->z[IX(d, nr1, icol)] - zb[IX(d, nr1, icol)]);
        d->qx[IX(d, 0, icol)] = 0.0;
        d->qy[IX(d, nr1, icol)] = 0.0;
        d->qxf[IX(d, 0, icol)] = 0.0;
    }
    for (int irow = 0; irow <= nr1; irow++) {
        zb[IX(d, irow, 0)] = 1.5 * zb[IX(d, irow, 1)] - .5 * zb[IX(d, irow, 2)];
        zb[IX(d, irow, nc1)] = 1.5 * zb[IX(d, irow, nc)] - .5 * zb[IX(d, irow, nc - 1)];
        d->z[IX(d, irow, 0)] = 1.5 * d->z[IX(d, irow, 1)] - .5 * d->z[IX(d, irow, 2)];
        d->z[IX(d, irow, nc1)] = 1.5 * d->z[IX(d, irow, nc)] - .5 * d->z[IX(d, irow, nc - 1)];
        d->h[IX(d, irow, 0)] = fmax(0.0, d->z[IX(d, irow, 0)] - zb[IX(d, irow, 0)]);
        d->h[IX(d, irow, nc1)] = fmax(0.0, d->z[IX(d, irow, nc1)] - zb[IX(d, irow, nc1)]);
        d->qx[IX(d, irow, 0)] = 0.0;
        d->qy[IX(d, irow, nc1)] = 0.0;
        d->qyf[IX(d, irow, 0)] = 0.0;
    }
}

void orc_initiation(orc_domain *d)
{
    const int nr = d->nrows, nc = d->ncols;
    init_bed(d);
    init_prelude(d);
    /* cold-start branch (:166-187) */
    for (int icol = 1; icol <= nc; icol++) {
        for (int irow = 1; irow <= nr; irow++) {
            const size_t p = IX(d, irow, icol);
            d->h[p] = 0.0;
            d->z[p] = d->zb[p];
            if (d->ux) d->ux[p] = 0.0;
            if (d->uy) d->uy[p] = 0.0;
            d->qx[p] = 0.0; d->qy[p] = 0.0; d->us[p] = 0.0;
            if (d->soil) d->soil[p] = 0.0;
            d->twet[p] = 0.0;
            d->ldry[p] = 1.0f;
            d->conc[0][p] = d->ade_solver ? 0.0 : -1.0;
        }
    }
    init_boundary(d);
}

int orc_initiation_restart(orc_domain *d, const double *table, int64_t trows)
{
    init_bed(d);
    init_prelude(d);
    /* the solute grids are allocated after the prelude (:126-129): zero */
    const size_t mm = (size_t)d->mrows * d->mcols;
    for (int k = 0; k < d->nchem; k++) memset(d->conc[k], 0, mm * sizeof(double));
    /* overlay of the file's rows (:132-153); row a of the reference's
     * filedata is table row a-1 (its row 0 is the header line) */
    for (int64_t a = 0; a < trows; a++) {
        const double *r = table + a * 16;
        const unsigned irow = (unsigned)(int)r[0];
        const unsigned icol = (unsigned)(int)r[1];
        if (irow >= (unsigned)d->mrows || icol >= (unsigned)d->mcols) return -1;
        const size_t p = IX(d, irow, icol);
        d->z[p] = r[4];
        d->h[p] = r[5];
        if (d->ux) d->ux[p] = r[6];
        if (d->uy) d->uy[p] = r[7];
#ifdef ORC_PLANT_RESTART_QX
        d->qx[p] = r[8] / (double)d->dxy;
        d->qy[p] = r[9] / (double)d->dxy;
#else
        d->qx[p] = r[8];                    /* written as qx*dxy, read back as is */
        d->qy[p] = r[9];
#endif
        d->us[p] = r[10];
        d->conc[0][p] = d->ade_solver ? fmax(r[11], 0.0) : -1.0;
        if (d->soil) d->soil[p] = r[12];
        d->twet[p] = r[15];
        d->ldry[p] = 0.0f;
    }
    init_boundary(d);
    return 0;
}

/* ------------------------------------------------------------------ */
/* WINTRAsolver_calc.cpp:26-57                                          */
/* ------------------------------------------------------------------ */
void orc_wintra(orc_domain *d, int ichem)
{
    /* :33-36 form f = tanh(...) from SWE_MAX, SWE_STD and the melt volume;
     * the release below (:46-54) does not use it */
    const int nr = d->nrows, nc = d->ncols;
    double *c = d->conc[ichem];
    for (int icol = 1; icol <= nc; icol++) {
        for (int irow = 1; irow <= nr; irow++) {
            const size_t p = IX(d, irow, icol);
            const double hp = d->h[p];
            const double zbp = d->zb[p];
            if (hp > d->hdry && zbp != d->nodata) {
#ifdef ORC_PLANT_WINTRA_AREA
                const double deltam = d->soil[p] * d->soil_rate / 3600 * d->dtfl;
                d->soil[p] = d->soil[p] - deltam;
                c[p] = c[p] + deltam / hp;
#else
                const double deltam = d->soil[p] * d->soil_rate / 3600 * d->dtfl;
                d->soil[p] = d->soil[p] - deltam;
                c[p] = c[p] + deltam / (hp * (double)d->arbase);
#endif
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* one iteration of the time loop: main.cpp:271-406                     */
/* ------------------------------------------------------------------ */
static double orc_wtime(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* orc_step with the wall time of each phase added to phase_s[0..3] (CFL,
 * add_meteo + add_inflow, hydrodynamics_calc, adesolver_calc + WINTRA) when
 * phase_s is not NULL; the arithmetic is orc_step's */
int orc_step_timed(orc_domain *d, double print_next, int32_t *it, double *hpall_out, double *phase_s)
{
    int flags = 0;
    double t0 = phase_s ? orc_wtime() : 0.0, t1;
    double hpall = orc_courant(d, print_next);
    if (phase_s) { t1 = orc_wtime(); phase_s[0] += t1 - t0; t0 = t1; }
    orc_add_meteo(d);
    if (orc_add_inflow(d)) flags |= 2;
    if (phase_s) { t1 = orc_wtime(); phase_s[1] += t1 - t0; t0 = t1; }
    if (hpall != 0) {
        (*it)++;
        orc_hydrodynamics(d);
        if (phase_s) { t1 = orc_wtime(); phase_s[2] += t1 - t0; t0 = t1; }
        if (d->ade_solver) {
            double *scratch = (double *)calloc((size_t)d->mrows * d->mcols * 3, sizeof(double));
            for (int k = 0; k < d->nchem; k++) orc_adesolver(d, *it, k, scratch);
            free(scratch);
        }
        /* main.cpp:399-404; the module is switched off without the ADE
         * solver (main.cpp:207-213) */
        if (d->wintra && d->ade_solver && d->soil)
            for (int k = 0; k < d->nchem; k++) orc_wintra(d, k);
        if (phase_s) { t1 = orc_wtime(); phase_s[3] += t1 - t0; t0 = t1; }
        flags |= 1;
    }
    if (hpall_out) *hpall_out = hpall;
    return flags;
}

int orc_step(orc_domain *d, double print_next, int32_t *it, double *hpall_out)
{
    return orc_step_timed(d, print_next, it, hpall_out, NULL);
}
