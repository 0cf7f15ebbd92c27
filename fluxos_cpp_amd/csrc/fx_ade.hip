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
//    rows, and the consumer fetches the W boundary (the producer's north-face
//    flux and final concentration) ADE_HK rows at a time, one block ahead of
//    use, after a bounded wait for the rows it needs; so no step of the sweep
//    waits on a memory round trip.  Waves take column blocks in ticket order
//    (atomic counter), so a wave only ever waits on a wave that started
//    earlier.
//
// Pass 1 (the neighbour bounds and the rescale to the new depth) is a plain
// parallel kernel writing the rescaled copy B; the sweep reads B for cells
// not yet visited, the raw input A for the E/N bound terms, and writes final
// values back into A.
#include "fx_params.h"

#define ADE_W 64
#define ADE_PUBLISH 8
#define ADE_HK 8              // W-boundary rows fetched per batch (<= ADE_W, a multiple of ADE_PUBLISH)
#define ADE_SPIN_LIMIT 20000000LL

namespace {

// pass 1 over the whole padded grid: ghosts copied, interior rescaled
// (ADEsolver_calc.cpp:64-72).  Also resets the wavefront's counters.
__global__ __launch_bounds__(256) void k_ade_pass1(fx_params P, fx_state64 s, int ichem, int slot,
                                                   fx_ctl *ctl, unsigned mode,
                                                   uint32_t *progress, int nblocks)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate || ctl->it <= 1)) return;
    const int64_t mm = (int64_t)P.ld * (P.ncols + 2);
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) ctl->ade_ticket = 0;
    if (blockIdx.x == 0)
        for (int b = threadIdx.x; b < nblocks; b += 256) progress[(size_t)slot * nblocks + b] = 0;
    if (i >= mm) return;
    const double *A = s.conc[ichem];
    double *B = s.conc_tmp + (size_t)slot * mm;
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

// Inputs of one row of a lane's column that do not depend on the sweep's
// carried values, so they are loaded ADE_DEPTH rows ahead (row r, column iy):
struct arow {
    double Bee;     // B(min(r+2, NROWS+1), iy): pass-1 value two rows on
    double Ae;      // A(r+1, iy): raw value of the E cell (ghost row at r = NROWS)
    double he;      // h(r+1, iy)
    double qx, qy;  // fe_1(r, iy), fn_1(r, iy)
    double Bs;      // B(r, iy-1)
    double An, Bn;  // A(r, iy+1) raw (ghost column at iy = NCOLS), B(r, iy+1)
    double hn;      // h(r, iy+1)
    double Bnn;     // B(r, min(iy+2, NCOLS+1))
    uint32_t dry;   // ldry bits: 1 = EE, 2 = S, 4 = N, 8 = NN
};

struct actx64 {
    const double *A, *B, *h, *fe1, *fn1;
    const uint8_t *ldry;
    size_t ld;
    int nr, nc, iy, inn;
    bool col_ok;
};

__device__ inline void arow_load(const actx64 &X, int r, arow &a)
{
    if (!X.col_ok || r < 1 || r > X.nr) return;
    const size_t ld = X.ld, iy = (size_t)X.iy;
    const size_t p = (size_t)r + iy * ld;
    const size_t iee = (size_t)((r + 2 < X.nr + 1) ? r + 2 : X.nr + 1);
    a.Bee = X.B[iee + iy * ld];
    a.Ae = X.A[p + 1];
    a.he = X.h[p + 1];
    a.qx = X.fe1[p];
    a.qy = X.fn1[p];
    a.Bs = X.B[p - ld];
    a.An = X.A[p + ld];
    a.Bn = X.B[p + ld];
    a.hn = X.h[p + ld];
    const size_t pnn = (size_t)r + (size_t)X.inn * ld;
    a.Bnn = X.B[pnn];
    a.dry = (X.ldry[iee + iy * ld] == 1 ? 1u : 0u) | (X.ldry[p - ld] == 1 ? 2u : 0u) |
            (X.ldry[p + ld] == 1 ? 4u : 0u) | (X.ldry[pnn] == 1 ? 8u : 0u);
}

#ifndef ADE_DEPTH
#define ADE_DEPTH 4           // rows of loads in flight ahead of use
#endif

// nsol solutes (k0 .. k0+nsol-1) sweep concurrently, each its own chain of
// waves: ticket t is block t / nsol of solute slot t % nsol, so a wave only
// waits on a wave of its own solute with a smaller ticket
__global__ __launch_bounds__(ADE_W) void k_ade_sweep(fx_params P, fx_state64 s, int k0, int nsol,
                                                     fx_ctl *ctl, unsigned mode,
                                                     uint32_t *progress_all, double *qedge_all, int nblocks)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate || ctl->it <= 1)) return;
    const int lane = threadIdx.x;
    uint32_t ticket = 0;
    if (lane == 0) ticket = atomicAdd(&ctl->ade_ticket, 1u);
    ticket = __shfl(ticket, 0, ADE_W);
    const int slot = (int)ticket % nsol;
    const int b = (int)ticket / nsol;
    const int ichem = k0 + slot;
    const int nr = P.nrows, nc = P.ncols;
    const size_t ld = P.ld;
    const int64_t mm = (int64_t)ld * (nc + 2);
    uint32_t *progress = progress_all + (size_t)slot * nblocks;
    double *qedge = qedge_all + (size_t)slot * nblocks * ld;
    const int iy = b * ADE_W + 1 + lane;
    const bool col_ok = iy <= nc;
    double *A = s.conc[ichem];
    const uint8_t *ldry = s.ldry;
    const double dx = P.dxy, dy = P.dxy, area = P.arbase, areae = area, arean = area;
    const double dtfl = ctl->dtfl, D = P.D[ichem];
    const double nt = 1, sigc = 0.5;
    actx64 X;
    X.A = A; X.B = s.conc_tmp + (size_t)slot * mm; X.h = s.h; X.fe1 = s.fe1; X.fn1 = s.fn1; X.ldry = ldry;
    X.ld = ld; X.nr = nr; X.nc = nc; X.iy = iy; X.col_ok = col_ok;
    X.inn = (iy + 2 < nc + 1) ? iy + 2 : nc + 1;

    double pfe = 0.0;          // east-face flux carried along the column
    double fw_final = 0.0;     // final concentration of the W cell
    double q_out = 0.0, c_out = 0.0;   // this lane's result of the last iteration
    uint32_t seen = 0;         // producer progress already observed (lane 0)
    bool bail = false;
    // sliding window of the lane's own column: pass-1 values B, wet flags and
    // depth at rows ix-1, ix, ix+1 (the row ix+2 values arrive with each row)
    double Bw = 0.0, Bp = 0.0, Be = 0.0, hp_w = 0.0;
    uint32_t dw = 0, dp = 0, de = 0;
    double a0 = 0.0, f0 = 0.0;  // ghost row 0 of the column: raw value, fe_1
    if (col_ok) {
        const size_t c0 = (size_t)iy * ld;
        Bw = X.B[c0];
        Bp = X.B[c0 + 1];
        Be = X.B[c0 + (nr >= 2 ? 2 : 1)];
        hp_w = X.h[c0 + 1];
        dw = ldry[c0] == 1;
        dp = ldry[c0 + 1] == 1;
        de = ldry[c0 + (nr >= 2 ? 2 : 1)] == 1;
        a0 = A[c0];
        f0 = X.fe1[c0];
    }
    arow ring[ADE_DEPTH];
#pragma unroll
    for (int k = 0; k < ADE_DEPTH; k++) {      // rows of steps 0 .. ADE_DEPTH-1
        ring[k] = arow{};
        arow_load(X, k - lane + 1, ring[k]);
    }

    // W boundary of lane 0 (column b*ADE_W: the previous wave's last column,
    // or ghost column 0 with qfcds = 0): lane j of cur / nxt holds row
    // base + j of the block in use / of the next block
    const size_t wcol = (size_t)b * ADE_W * ld;
    double cur_q = 0.0, cur_c = 0.0, nxt_q = 0.0, nxt_c = 0.0;
    auto fetch = [&](int r0) -> bool {
        if (r0 > nr) return true;
        if (b > 0) {
            const uint32_t need = (uint32_t)min(r0 + ADE_HK - 1, nr);
            if (lane == 0 && seen < need) {
                long long spin = 0;
                for (;;) {
                    const uint32_t pr = __hip_atomic_load(&progress[b - 1], __ATOMIC_ACQUIRE,
                                                          __HIP_MEMORY_SCOPE_AGENT);
                    if (pr >= need) { seen = pr; break; }
                    if (++spin > ADE_SPIN_LIMIT) { bail = true; break; }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            if (__shfl(bail ? 1 : 0, 0, ADE_W)) return false;
        }
        const int r = r0 + lane;
        if (lane < ADE_HK && r <= nr) {
            nxt_q = b > 0 ? __hip_atomic_load(&qedge[(size_t)(b - 1) * ld + r], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT)
                          : 0.0;
            nxt_c = __hip_atomic_load(&A[(size_t)r + wcol], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return true;
    };
    if (!fetch(1)) bail = true;
    cur_q = nxt_q;
    cur_c = nxt_c;
    if (!bail && !fetch(1 + ADE_HK)) bail = true;
    if (bail) {
        if (lane == 0) atomicOr(&ctl->ade_error, 1u);
        return;
    }

    const int T = nr + ADE_W - 1;
    auto step = [&](int t, arow &R) {
        // lane 0 works on row t + 1: a new block of the W boundary every ADE_HK steps
        if (t > 0 && t % ADE_HK == 0) {
            cur_q = nxt_q;
            cur_c = nxt_c;
            if (!fetch(t + 1 + ADE_HK)) return false;
        }
        const double qb = __shfl(cur_q, t % ADE_HK, ADE_W), cb = __shfl(cur_c, t % ADE_HK, ADE_W);
        // receive the S neighbour's (qfn, final c) for this lane's row
        double q_in = shift_up(q_out), c_in = shift_up(c_out);
        const int ix = t - lane + 1;
        const bool active = col_ok && ix >= 1 && ix <= nr;
        if (lane == 0) { q_in = qb; c_in = cb; }
        if (active) {
            const size_t p = (size_t)ix + (size_t)iy * ld;
            const double Bee = R.Bee;
            const uint32_t dee = R.dry & 1u;
            const bool last_cell = (ix == nr && iy == nc);   // a < NCOLS*NROWS
            if (last_cell) {
                q_out = 0.0;
                c_out = Bp;
                A[p] = c_out;
            } else {
                if (ix == 1) pfe = a0 * f0 * dy;                                     // :106-116
                if (ix == 1) fw_final = a0;
                if (dp) {                                                            // :119-124
                    pfe = 0.0;
                    q_out = 0.0;
                    c_out = 0.0;
                } else {
                    // neighbour bounds of pass 1: W,S rescaled, E,N raw (:51-62)
                    const double cmaxr = fmax(Bw, fmax(R.Ae, fmax(R.Bs, R.An)));
                    const double cminr = fmin(Bw, fmin(R.Ae, fmin(R.Bs, R.An)));
                    const double hp = hp_w, he = R.he, hn = R.hn;
                    const double qxl = R.qx, qyl = R.qy;
                    const double fw = fw_final, fp = Bp, fe = Be, fee = Bee;
                    const double fs = c_in, fn = R.Bn, fnn = R.Bnn;
                    const double pfw = pfe, qfs = q_in;
                    double fem, pfde, fnm, qfdn;
                    if (!de) {
                        const double hnue = fmax(hp * nt * he * nt, .0001);
                        const double hne = sqrt(hnue) / sigc / dx * dy * D;
                        pfde = -hne * (fe - fp);
                        if (qxl > 0.0) {
                            if (!dw) fem = -.125 * fw + .75 * fp + .375 * fe;
                            else fem = 0.5 * fp + 0.5 * fe;
                        } else {
                            if (!dee) fem = .375 * fp + .75 * fe - .125 * fee;
                            else fem = 0.5 * fp + 0.5 * fe;
                        }
                    } else {
                        fem = 0.; pfde = 0.;
                    }
                    fem = fmax(0., fem);
                    if (ix == nr) fem = R.Ae;                       // A(NROWS+1, iy)
                    const double pfce = qxl * fem * dy;
                    pfe = pfce + pfde;
                    if (pfe < 0) {
                        if (!de) {
                            const double cvolrate = -(fe * he) * areae / dtfl;
                            pfe = fmax(pfe, cvolrate);
                        } else {
                            pfe = 0.;
                        }
                    }
                    if (!(R.dry & 4u)) {
                        const double hnue = fmax(.0001, hp * nt * hn * nt);
                        const double hnn = sqrt(hnue) / sigc / dy * dx * D;
                        qfdn = -hnn * (fn - fp);
                        if (qyl > 0.0) {
                            if (!(R.dry & 2u)) fnm = -.125 * fs + .75 * fp + .375 * fn;
                            else fnm = 0.5 * fp + .5 * fn;
                        } else {
                            if (!(R.dry & 8u)) fnm = .375 * fp + .75 * fn - .125 * fnn;
                            else fnm = .5 * fp + .5 * fn;
                        }
                    } else {
                        fnm = 0.; qfdn = 0.;
                    }
                    fnm = fmax(0., fnm);
                    if (iy == nc) fnm = R.An;                       // A(ix, NCOLS+1)
                    const double qfcn = qyl * fnm * dx;
                    double qfn = qfcn + qfdn;
                    if (qfn < 0) {
                        if (!(R.dry & 4u)) {
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
            }
            fw_final = c_out;
            if (lane == ADE_W - 1) qedge[(size_t)b * ld + ix] = q_out;
            // slide the column window down one row
            Bw = Bp; Bp = Be; Be = Bee;
            dw = dp; dp = de; de = dee;
            hp_w = R.he;
        }
        // publish the last lane's progress every few rows and at the end
        if (lane == ADE_W - 1 && active && ((ix % ADE_PUBLISH) == 0 || ix == nr)) {
            __threadfence();
            __hip_atomic_store(&progress[b], (uint32_t)ix, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        // this slot now fetches the row of step t + ADE_DEPTH
        arow_load(X, t + ADE_DEPTH - lane + 1, R);
        return true;
    };
    bool ok = true;
    int t = 0;
    for (; ok && t + ADE_DEPTH <= T; t += ADE_DEPTH) {
#pragma unroll
        for (int k = 0; k < ADE_DEPTH; k++)
            if (ok) ok = step(t + k, ring[k]);
    }
#pragma unroll
    for (int k = 0; k < ADE_DEPTH - 1; k++)
        if (ok && t + k < T) ok = step(t + k, ring[k]);
    if (!ok) {
        if (lane == 0) atomicOr(&ctl->ade_error, 1u);
        return;
    }
    // a block whose last lane is past NCOLS never publishes from the loop
    if (lane == 0 && !(iy + ADE_W - 1 <= nc)) {
        __threadfence();
        __hip_atomic_store(&progress[b], (uint32_t)nr, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

} // namespace

int fx64_launch_ade_exact(fx_domain d, unsigned mode, int k0, int nsol)
{
    fx_params P = fx_make_params(d);
    const int64_t mm = (int64_t)d->mrows * d->mcols;
    const int nblocks = (P.ncols + ADE_W - 1) / ADE_W;
    if (nblocks > d->ade_blocks || nsol < 1 || k0 < 0 || k0 + nsol > d->desc.nchem)
        return fx_fail(FX_ERR_STATE, "ADE block buffers too small");
    for (int k = 0; k < nsol; k++) {
        fx_prof_begin(d, FX_K_ADE_PASS1);
        hipLaunchKernelGGL(k_ade_pass1, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, d->stream,
                           P, d->s64, k0 + k, k, d->ctl, mode, d->ade_progress, nblocks);
        fx_prof_end(d, FX_K_ADE_PASS1);
        FX_HIP(hipGetLastError());
    }
    const int64_t nb = (int64_t)d->ade_blocks * d->desc.nchem;
    double *qedge = (double *)(d->ade_progress + ((nb + 63) / 64) * 64);
    fx_prof_begin(d, FX_K_ADE_SWEEP);
    hipLaunchKernelGGL(k_ade_sweep, dim3(nblocks * nsol), dim3(ADE_W), 0, d->stream,
                       P, d->s64, k0, nsol, d->ctl, mode, d->ade_progress, qedge, nblocks);
    fx_prof_end(d, FX_K_ADE_SWEEP);
    return fx_hip_check(hipGetLastError(), "k_ade_sweep");
}
