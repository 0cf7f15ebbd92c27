// fx_ade32.hip -- fast advection-dispersion (FX_ADE_FAST) for the FP32-mixed
// path, one LDS-tiled kernel for all solutes of a step.
//
// The reference (ADEsolver_calc.cpp:27-291) sweeps the grid sequentially, so
// each face sees already-updated W/S concentrations.  This form evaluates the
// same face terms from the step's pass-1 concentrations everywhere (a Jacobi
// update), which makes every cell independent:
//   * concentration is carried as solute mass per unit area m = c*h, so the
//     melt mixing of add_meteo (common.cpp:226-233) is m += meteoi*cmelt where
//     the cell held water before the melt; pass 1's rescale (:64-72) is
//     reproduced from the depths before and after the step;
//   * east/north face flux of a wet cell: diffusive -sqrt(max(hp*he,1e-4))/
//     sigc/dx*dy*D*(ce-cp) plus advective fe_1*fem*dy with the reference's
//     upwind-biased face value (-1/8, 3/4, 3/8, or 1/2,1/2 next to dry cells),
//     clamped >= 0, and the inflow limit of the upstream cell (:149-241);
//   * outflow limiting against the cell's available solute (:243-278), with
//     the W/S inflow terms taken before their own cells' limiting;
//   * new concentration clamped to the min/max of the four neighbours (:284-285).
// Each face flux is computed once per face value and used by both cells, so
// the transport itself conserves solute mass.
#include "fx_params.h"

#pragma clang fp contract(fast)

#define ATX 64
#define ATY 8
#define ALR (ATX + 5)
#define ALC (ATY + 5)
#define APR (ATX + 2)
#define APC (ATY + 2)
#define AFR (ATX + 1)
#define AFC (ATY + 1)
#define ANT 256

namespace {

struct a32p {
    const double *d[2];
    const float *ff;
    const uint8_t *flags;
    float *mass[2][FX_MAX_CHEM];
    const float *melt_mult;
    int64_t mm;
};

__global__ __launch_bounds__(ANT) void k32_ade(fx_params P, a32p A, fx_ctl *ctl)
{
    if (ctl->halt) return;
    const int cur = ctl->cur;
    const double *d_old = A.d[cur], *d_new = A.d[cur ^ 1];
    const bool transport = ctl->gate && ctl->it > 1;
    const double meteoi = P.meteo ? ctl->meteoi : 0.0;
    const float dt = (float)ctl->dtfl;
    const int ld = P.ld, mcols = P.ncols + 2;
    const int r0 = 1 + blockIdx.x * ATX, c0 = 1 + blockIdx.y * ATY;
    const int tid = threadIdx.x + threadIdx.y * 64;
    const float hdry = (float)P.hdry;
    const float dx = (float)P.dxy, area = (float)P.arbase;
    const float area_dt = area / dt;

    __shared__ float s_c[ALR * ALC], s_cr[ALR * ALC], s_h[ALR * ALC], s_fe[ALR * ALC], s_fn[ALR * ALC];
    __shared__ float s_h0[ALR * ALC], s_melt[ALR * ALC];
    __shared__ uint8_t s_w[ALR * ALC], s_wp[ALR * ALC];
    __shared__ float s_pe[APR * APC], s_pn[APR * APC];
    __shared__ float s_e[AFR * AFC], s_n[AFR * AFC];

    // geometry shared by all solutes: depth after the update, wet now (loop-4
    // meaning), wet before (ldry_prev), the step's face discharges
    for (int i = tid; i < ALR * ALC; i += ANT) {
        const int lr = i % ALR, lc = i / ALR;
        const int r = r0 - 3 + lr, c = c0 - 3 + lc;
        float h = 0.f, fe = 0.f, fn = 0.f, h0 = 0.f, mlt = 0.f;
        uint8_t w = 0, wp = 0;
        if (r >= 0 && r < ld && c >= 0 && c < mcols) {
            const size_t p = (size_t)r + (size_t)c * ld;
            const uint8_t f = A.flags[p];
            if (f & FXF_INNER) {
                const double hn = fmax(0.0, d_new[p]);
                const double dold = d_old[p];
                const double melt = (f & FXF_NODATA) ? 0.0
                                    : (P.melt_mult ? meteoi * (double)A.melt_mult[p] : meteoi);
                h = (float)hn;
                w = (!(hn < P.hdry) && !(f & FXF_WEIR)) ? 1 : 0;
                wp = (!(fmax(0.0, dold) < P.hdry) && !(f & FXF_WEIR)) ? 1 : 0;
                h0 = (float)fmax(0.0, dold + melt);
                mlt = dold > 0.0 ? (float)melt : 0.f;   // solute arrives with melt only into water
                fe = A.ff[p];
                fn = A.ff[A.mm + p];
            }
        }
        s_h[i] = h; s_fe[i] = fe; s_fn[i] = fn; s_w[i] = w; s_wp[i] = wp;
        s_h0[i] = h0; s_melt[i] = mlt;
    }

    for (int k = 0; k < P.nchem; k++) {
        const float *m = A.mass[cur][k];
        float *m_out = A.mass[cur ^ 1][k];
        const double cm = ctl->cmelt[k];
        const float D = (float)P.D[k];
        const float kdiff = 2.f * D;          // 1/sigc/dx*dy with sigc = 0.5, dx = dy
        __syncthreads();
        // concentration after this step's melt mixing (craw, common.cpp:226-233)
        // and after pass 1 (c1, ADEsolver_calc.cpp:64-72): c*h0/h where wet now
        // and before, 0 where dry now, unchanged where wet now but dry before
        for (int i = tid; i < ALR * ALC; i += ANT) {
            const int lr = i % ALR, lc = i / ALR;
            const int r = r0 - 3 + lr, c = c0 - 3 + lc;
            float c1 = 0.f, cr = 0.f;
            if (s_h0[i] > 0.f || s_w[i]) {
                const size_t p = (size_t)r + (size_t)c * ld;
                const float mv = m[p] + s_melt[i] * (float)cm;
                const float h0 = s_h0[i];
                cr = h0 > 0.f ? mv / h0 : 0.f;
                if (s_w[i]) c1 = s_wp[i] ? mv / s_h[i] : cr;
            }
            s_c[i] = transport ? c1 : cr;
            s_cr[i] = cr;
        }
        __syncthreads();
        if (transport) {
            // face fluxes before the outflow limiter (E and N face of Q)
            for (int i = tid; i < APR * APC; i += ANT) {
                const int pr = i % APR, pc = i / APR;
                const int li = (pr + 1) + (pc + 1) * ALR;     // Q at rows r0-2.., cols c0-2..
                float pe = 0.f, pn = 0.f;
                if (s_w[li]) {
                    const float hp = s_h[li], cp = s_c[li];
                    // x direction
                    if (s_w[li + 1]) {
                        const float he = s_h[li + 1], ce = s_c[li + 1];
                        const float hne = sqrtf(fmaxf(hp * he, 1e-4f)) * kdiff;
                        const float pfde = -hne * (ce - cp);
                        const float q = s_fe[li];
                        float fem;
                        if (q > 0.f) fem = s_w[li - 1] ? -.125f * s_c[li - 1] + .75f * cp + .375f * ce : .5f * (cp + ce);
                        else fem = s_w[li + 2] ? .375f * cp + .75f * ce - .125f * s_c[li + 2] : .5f * (cp + ce);
                        fem = fmaxf(0.f, fem);
                        pe = q * fem * dx + pfde;
                        if (pe < 0.f) pe = fmaxf(pe, -(ce * he) * area_dt);
                    } else {
                        pe = 0.f;      // dry east neighbour: fem = 0, no diffusion
                    }
                    if (s_w[li + ALR]) {
                        const float hn = s_h[li + ALR], cn = s_c[li + ALR];
                        const float hnn = sqrtf(fmaxf(hp * hn, 1e-4f)) * kdiff;
                        const float qfdn = -hnn * (cn - cp);
                        const float q = s_fn[li];
                        float fnm;
                        if (q > 0.f) fnm = s_w[li - ALR] ? -.125f * s_c[li - ALR] + .75f * cp + .375f * cn : .5f * (cp + cn);
                        else fnm = s_w[li + 2 * ALR] ? .375f * cp + .75f * cn - .125f * s_c[li + 2 * ALR] : .5f * (cp + cn);
                        fnm = fmaxf(0.f, fnm);
                        pn = q * fnm * dx + qfdn;
                        if (pn < 0.f) pn = fmaxf(pn, -(cn * hn) * area_dt);
                    } else {
                        pn = 0.f;
                    }
                }
                s_pe[i] = pe;
                s_pn[i] = pn;
            }
            __syncthreads();
            // outflow limiter of Q against its available solute
            for (int i = tid; i < AFR * AFC; i += ANT) {
                const int fr = i % AFR, fc = i / AFR;
                const int li = (fr + 2) + (fc + 2) * ALR;
                const int pi = (fr + 1) + (fc + 1) * APR;
                float pfe = s_pe[pi], qfn = s_pn[pi];
                if (s_w[li]) {
                    const float pfw = s_pe[pi - 1], qfs = s_pn[pi - APR];
                    const float cvolrat = s_c[li] * s_h[li] * area_dt + pfw + qfs;
                    if (cvolrat > 0.f) {
                        if (pfe > 0.f && qfn > 0.f) {
                            if (pfe + qfn > cvolrat) {
                                const float cf = qfn / (pfe + qfn);
                                pfe = (1.f - cf) * cvolrat;
                                qfn = cf * cvolrat;
                            }
                        } else if (pfe > 0.f) {
                            pfe = fminf(pfe, cvolrat - qfn);
                        } else if (qfn > 0.f) {
                            qfn = fminf(qfn, cvolrat - pfe);
                        }
                    } else {
                        if (pfe >= 0.f && qfn < 0.f) {
                            const float cb = cvolrat - qfn;
                            pfe = cb > 0.f ? fminf(pfe, cb) : 0.f;
                        } else if (pfe < 0.f && qfn >= 0.f) {
                            const float cb = cvolrat - pfe;
                            qfn = cb > 0.f ? fminf(qfn, cb) : 0.f;
                        } else if (pfe >= 0.f && qfn >= 0.f) {
                            pfe = 0.f;
                            qfn = 0.f;
                        }
                    }
                } else {
                    pfe = 0.f;
                    qfn = 0.f;
                }
                s_e[i] = pfe;
                s_n[i] = qfn;
            }
            __syncthreads();
        }
        // update own cells
        for (int j = 0; j < ATY / 4; j++) {
            const int lr = threadIdx.x, lc = threadIdx.y + 4 * j;
            const int r = r0 + lr, c = c0 + lc;
            if (r > P.nrows || c > P.ncols) continue;
            const size_t p = (size_t)r + (size_t)c * ld;
            const int li = (lr + 3) + (lc + 3) * ALR;
            float out;
            if (!transport) {
                out = m[p] + s_melt[li] * (float)cm;
            } else if (!s_w[li]) {
                out = 0.f;                               // dry after the update (:69-71,119-124)
            } else {
                const int fi = (lr + 1) + (lc + 1) * AFR;
                const float hp = s_h[li];
                const float dc = (s_e[fi - 1] - s_e[fi] + s_n[fi - AFR] - s_n[fi]) * dt / area;
                float con = s_c[li] + dc / hp;
                // bounds of pass 1: W and S already rescaled, E and N not yet (:51-62)
                const float cw = s_c[li - 1], ce = s_cr[li + 1], cs = s_c[li - ALR], cn = s_cr[li + ALR];
                con = fminf(fmaxf(fmaxf(cw, ce), fmaxf(cs, cn)), con);
                con = fmaxf(fminf(fminf(cw, ce), fminf(cs, cn)), con);
                out = con * hp;
            }
            m_out[p] = out;
        }
    }
}

} // namespace

int fx32_launch_ade(fx_domain d)
{
    fx_params P = fx_make_params(d);
    a32p A;
    A.d[0] = d->s32.d[0];
    A.d[1] = d->s32.d[1];
    A.ff = d->s32.ff;
    A.flags = d->s32.flags;
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < FX_MAX_CHEM; k++) A.mass[b][k] = d->s32.mass[b][k];
    A.melt_mult = d->melt_mult;
    A.mm = d->ncell;
    dim3 g((P.nrows + ATX - 1) / ATX, (P.ncols + ATY - 1) / ATY);
    fx_prof_begin(d, FX_K_ADE32);
    hipLaunchKernelGGL(k32_ade, g, dim3(64, 4), 0, d->stream, P, A, d->ctl);
    fx_prof_end(d, FX_K_ADE32);
    return fx_hip_check(hipGetLastError(), "k32_ade");
}
