// fx_ade.hip -- adesolver_calc (ADEsolver_calc.cpp:27-291) in FP64, exact.
//
// The reference updates concentrations in place in one sequential sweep,
// irow fastest.  Cell (ix,iy) therefore sees the FINAL values of its W and S
// neighbours (and the east-face flux carried from W, the north-face flux
// carried from S through qfcds) and the pass-1 values of everything to its E
// and N.  That is a wavefront dependence, mapped here as a systolic array:
//
//  * one lane owns one column iy and walks ix = 1..NROWS (contiguous memory);
//  * lane l runs one row behind lane l-1, so at every iteration the S
//    neighbour's result (north-face flux and final concentration) is exactly
//    one lane down and moves up with one cross-lane shift, no LDS, no barrier;
//  * a 64-lane wave covers 64 columns; consecutive waves are chained through
//    global memory: the producer publishes its last lane's progress every few
//    rows and the consumer waits (bounded) before it needs a row.  Waves take
//    column blocks in ticket order (atomic counter), so a wave only ever
//    waits on a wave that started earlier.
//
// Pass 1 (the neighbour bounds and the rescale to the new depth) is a plain
// parallel kernel writing the rescaled copy B; the sweep reads B for cells
// not yet visited, the raw input A for the E/N bound terms, and writes final
// values back into A.
#include "fx_params.h"

#define ADE_W 64
#define ADE_PUBLISH 8
#define ADE_SPIN_LIMIT 20000000LL

namespace {

// pass 1 over the whole padded grid: ghosts copied, interior rescaled
// (ADEsolver_calc.cpp:64-72).  Also resets the wavefront's counters.
__global__ __launch_bounds__(256) void k_ade_pass1(fx_params P, fx_state64 s, int ichem,
                                                   fx_ctl *ctl, unsigned mode,
                                                   uint32_t *progress, int nblocks)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate || ctl->it <= 1)) return;
    const int64_t mm = (int64_t)P.ld * (P.ncols + 2);
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) ctl->ade_ticket = 0;
    if (blockIdx.x == 0)
        for (int b = threadIdx.x; b < nblocks; b += 256) progress[b] = 0;
    if (i >= mm) return;
    const double *A = s.conc[ichem];
    double *B = s.conc_tmp;
    const int ix = (int)(i % P.ld), iy = (int)(i / P.ld);
    double c = A[i];
    if (ix >= 1 && ix <= P.nrows && iy >= 1 && iy <= P.ncols) {
        const double hnew = s.h[i];
        if (s.ldry[i] == 0 && s.ldry_prev[i] == 0) c = c * s.h0[i] / hnew;
        else if (s.ldry[i] == 1) c = 0.0;
    }
    B[i] = c;
}

__device__ inline double shift_up(double v)
{
    return __shfl_up(v, 1, ADE_W);
}

__global__ __launch_bounds__(ADE_W) void k_ade_sweep(fx_params P, fx_state64 s, int ichem,
                                                     fx_ctl *ctl, unsigned mode,
                                                     uint32_t *progress, double *qedge)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate || ctl->it <= 1)) return;
    const int lane = threadIdx.x;
    uint32_t ticket = 0;
    if (lane == 0) ticket = atomicAdd(&ctl->ade_ticket, 1u);
    ticket = __shfl(ticket, 0, ADE_W);
    const int b = (int)ticket;
    const int nr = P.nrows, nc = P.ncols;
    const size_t ld = P.ld;
    const int iy = b * ADE_W + 1 + lane;
    const bool col_ok = iy <= nc;
    double *A = s.conc[ichem];
    const double *B = s.conc_tmp;
    const double *fe1 = s.fe1, *fn1 = s.fn1, *h = s.h;
    const uint8_t *ldry = s.ldry;
    const double dx = P.dxy, dy = P.dxy, area = P.arbase, areae = area, arean = area;
    const double dtfl = ctl->dtfl, D = P.D[ichem];
    const double nt = 1, sigc = 0.5;
    const int is = iy - 1, in = iy + 1, inn = (iy + 2 < nc + 1) ? iy + 2 : nc + 1;

    double pfe = 0.0;          // east-face flux carried along the column
    double fw_final = 0.0;     // final concentration of the W cell
    double q_out = 0.0, c_out = 0.0;   // this lane's result of the last iteration
    uint32_t seen = 0;         // producer progress already observed (lane 0)
    bool bail = false;

    const int T = nr + ADE_W - 1;
    for (int t = 0; t < T; t++) {
        // receive the S neighbour's (qfn, final c) for this lane's row
        double q_in = shift_up(q_out), c_in = shift_up(c_out);
        const int ix = t - lane + 1;
        const bool active = col_ok && ix >= 1 && ix <= nr;
        if (lane == 0 && active) {
            if (b == 0) {
                q_in = 0.0;                                  // qfcds starts at zero
                c_in = A[(size_t)ix];                        // ghost column 0
            } else {
                if ((uint32_t)ix > seen) {
                    long long spin = 0;
                    for (;;) {
                        uint32_t pr = __hip_atomic_load(&progress[b - 1], __ATOMIC_ACQUIRE,
                                                        __HIP_MEMORY_SCOPE_AGENT);
                        if (pr >= (uint32_t)ix) { seen = pr; break; }
                        if (++spin > ADE_SPIN_LIMIT) { bail = true; break; }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                q_in = __hip_atomic_load(&qedge[(size_t)(b - 1) * ld + ix], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
                c_in = __hip_atomic_load(&A[(size_t)ix + (size_t)(iy - 1) * ld], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (__shfl(bail ? 1 : 0, 0, ADE_W)) {
            if (lane == 0) atomicOr(&ctl->ade_error, 1u);
            return;
        }
        if (active) {
            const size_t p = (size_t)ix + (size_t)iy * ld;
            const bool last_cell = (ix == nr && iy == nc);   // a < NCOLS*NROWS
            if (last_cell) {
                q_out = 0.0;
                c_out = B[p];
                A[p] = c_out;
            } else {
                const int iw = ix - 1, ie = ix + 1, iee = (ix + 2 < nr + 1) ? ix + 2 : nr + 1;
                if (ix == 1) pfe = A[(size_t)iy * ld] * fe1[(size_t)iy * ld] * dy;   // :106-116
                if (ix == 1) fw_final = A[(size_t)iy * ld];
                if (ldry[p] == 1) {                                                  // :119-124
                    pfe = 0.0;
                    q_out = 0.0;
                    c_out = 0.0;
                } else {
                    const size_t pE = p + 1, pEE = (size_t)iee + (size_t)iy * ld;
                    const size_t pW = p - 1, pS = p - ld, pN = p + ld;
                    const size_t pNN = (size_t)ix + (size_t)inn * ld;
                    // neighbour bounds of pass 1: W,S rescaled, E,N raw (:51-62)
                    const double cmaxr = fmax(B[pW], fmax(A[pE], fmax(B[pS], A[pN])));
                    const double cminr = fmin(B[pW], fmin(A[pE], fmin(B[pS], A[pN])));
                    const double hp = h[p], he = h[pE], hn = h[pN];
                    const double qxl = fe1[p], qyl = fn1[p];
                    const double fw = fw_final, fp = B[p], fe = B[pE], fee = B[pEE];
                    const double fs = c_in, fn = B[pN], fnn = B[pNN];
                    const double pfw = pfe, qfs = q_in;
                    double fem, pfde, fnm, qfdn;
                    if (ldry[pE] == 0) {
                        const double hnue = fmax(hp * nt * he * nt, .0001);
                        const double hne = sqrt(hnue) / sigc / dx * dy * D;
                        pfde = -hne * (fe - fp);
                        if (qxl > 0.0) {
                            if (ldry[pW] == 0) fem = -.125 * fw + .75 * fp + .375 * fe;
                            else fem = 0.5 * fp + 0.5 * fe;
                        } else {
                            if (ldry[pEE] == 0) fem = .375 * fp + .75 * fe - .125 * fee;
                            else fem = 0.5 * fp + 0.5 * fe;
                        }
                    } else {
                        fem = 0.; pfde = 0.;
                    }
                    fem = fmax(0., fem);
                    if (ix == nr) fem = A[(size_t)(nr + 1) + (size_t)iy * ld];
                    const double pfce = qxl * fem * dy;
                    pfe = pfce + pfde;
                    if (pfe < 0) {
                        if (ldry[pE] == 0) {
                            const double cvolrate = -(fe * he) * areae / dtfl;
                            pfe = fmax(pfe, cvolrate);
                        } else {
                            pfe = 0.;
                        }
                    }
                    if (ldry[pN] == 0) {
                        const double hnue = fmax(.0001, hp * nt * hn * nt);
                        const double hnn = sqrt(hnue) / sigc / dy * dx * D;
                        qfdn = -hnn * (fn - fp);
                        if (qyl > 0.0) {
                            if (ldry[pS] == 0) fnm = -.125 * fs + .75 * fp + .375 * fn;
                            else fnm = 0.5 * fp + .5 * fn;
                        } else {
                            if (ldry[pNN] == 0) fnm = .375 * fp + .75 * fn - .125 * fnn;
                            else fnm = .5 * fp + .5 * fn;
                        }
                    } else {
                        fnm = 0.; qfdn = 0.;
                    }
                    fnm = fmax(0., fnm);
                    if (iy == nc) fnm = A[(size_t)ix + (size_t)(nc + 1) * ld];
                    const double qfcn = qyl * fnm * dx;
                    double qfn = qfcn + qfdn;
                    if (qfn < 0) {
                        if (ldry[pN] == 0) {
                            const double cvolrate = -(fn * hn) * arean / dtfl;
                            qfn = fmax(qfn, cvolrate);
                        } else {
                            qfn = 0.;
                        }
                    }
                    const double cvolpot = (fp * hp) * area;
                    const double cvolrat = cvolpot / dtfl + pfw + qfs;
                    if (cvolrat > 0.0) {
                        if (pfe > 0. && qfn > 0.) {
                            if (pfe + qfn > cvolrat) {
                                const double cf = qfn / (pfe + qfn);
                                pfe = (1. - cf) * cvolrat;
                                qfn = cf * cvolrat;
                            }
                        } else if (pfe > 0.) {
                            pfe = fmin(pfe, (cvolrat - qfn));
                        } else if (qfn > 0.) {
                            qfn = fmin(qfn, (cvolrat - pfe));
                        }
                    } else {
                        if (pfe >= 0. && qfn < 0.) {
                            const double cbilan = cvolrat - qfn;
                            if (cbilan > 0.) pfe = fmin(pfe, cbilan);
                            else pfe = 0.;
                        } else if (pfe < 0. && qfn >= 0.) {
                            const double cbilan = cvolrat - pfe;
                            if (cbilan > 0.) qfn = fmin(qfn, cbilan);
                            else qfn = 0.;
                        } else if (pfe >= 0. && qfn >= 0.) {
                            pfe = 0.0; qfn = 0.0;
                        }
                    }
                    const double dc = (pfw - pfe + qfs - qfn) * dtfl / area;
                    double con = fp + dc / hp;
                    con = fmin(cmaxr, con);
                    con = fmax(cminr, con);
                    q_out = qfn;
                    c_out = con;
                }
                A[p] = c_out;
                (void)iw;
            }
            fw_final = c_out;
            if (lane == ADE_W - 1) qedge[(size_t)b * ld + ix] = q_out;
        }
        // publish the last lane's progress every few rows and at the end
        if (lane == ADE_W - 1 && active && ((ix % ADE_PUBLISH) == 0 || ix == nr)) {
            __threadfence();
            __hip_atomic_store(&progress[b], (uint32_t)ix, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // a block whose last lane is past NCOLS never publishes from the loop
    if (lane == 0 && !(iy + ADE_W - 1 <= nc)) {
        __threadfence();
        __hip_atomic_store(&progress[b], (uint32_t)nr, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

} // namespace

int fx64_launch_ade_exact(fx_domain d, unsigned mode, int ichem, int it_override)
{
    (void)it_override;
    fx_params P = fx_make_params(d);
    const int64_t mm = (int64_t)d->mrows * d->mcols;
    const int nblocks = (P.ncols + ADE_W - 1) / ADE_W;
    if (nblocks > d->ade_blocks) return fx_fail(FX_ERR_STATE, "ADE block buffers too small");
    fx_prof_begin(d, FX_K_ADE_PASS1);
    hipLaunchKernelGGL(k_ade_pass1, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, d->stream,
                       P, d->s64, ichem, d->ctl, mode, d->ade_progress, nblocks);
    fx_prof_end(d, FX_K_ADE_PASS1);
    FX_HIP(hipGetLastError());
    double *qedge = (double *)(d->ade_progress + ((d->ade_blocks + 63) / 64) * 64);
    fx_prof_begin(d, FX_K_ADE_SWEEP);
    hipLaunchKernelGGL(k_ade_sweep, dim3(nblocks), dim3(ADE_W), 0, d->stream,
                       P, d->s64, ichem, d->ctl, mode, d->ade_progress, qedge);
    fx_prof_end(d, FX_K_ADE_SWEEP);
    return fx_hip_check(hipGetLastError(), "k_ade_sweep");
}
