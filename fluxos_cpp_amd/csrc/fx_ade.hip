// fx_ade.hip -- adesolver_calc (ADEsolver_calc.cpp:27-291) in FP64, exact.
//
// The reference updates concentrations in place in one sequential sweep,
// irow fastest.  Cell (ix,iy) therefore sees the FINAL values of its W and S
// neighbours (and the east-face flux carried from W, the north-face flux
// carried from S through qfcds) and the pass-1 values of everything to its E
// and N.  That is a wavefront dependence, mapped here as a systolic array:
//
//  * one lane owns one column iy and walks ix = 1..NROWS;
//  * lane l runs one row behind lane l-1, so at every step the S neighbour's
//    result (north-face flux and final concentration) is exactly one lane
//    down and moves up with one cross-lane shift;
//  * the sweeping wave does only the dependent chain: every term of a cell's
//    update that does not read the carried values is computed beforehand by
//    three helper waves of the same block, with row-contiguous loads, and
//    handed over through an LDS ring of 8-step chunks;
//  * 64-column blocks are chained through global memory: the helpers publish
//    the block's last column (north-face flux, final value) chunk by chunk
//    and fetch the previous block's, after a bounded wait, one chunk ahead
//    of use.  Blocks take column ranges in ticket order (atomic counter), so
//    a block only ever waits on one that started earlier.
//
// Pass 1 (the rescale to the new depth) is a plain parallel kernel writing
// the rescaled copy B; the records read B for cells not yet visited and the
// raw input A for the E/N bound terms; the helpers write the final values
// back into A.
#include "fx_params.h"

#define ADE_W 64              // columns per block, one per lane of the sweeping wave
#define ADE_CHK 8             // steps per chunk of the record ring
#define ADE_NP 7              // double pairs per cell record
#define ADE_WAVES 8           // waves per block: the chain, the boundary wave, six record waves
#define ADE_SPIN_LIMIT 4000000LL   // polls of a bounded wait (each a load and a short sleep)

namespace {

// Edge words (the north-face flux and final value of a block's last column,
// row by row, for the next block) hold this bit pattern until written: a
// signalling NaN that no arithmetic produces.  Each word is written once per
// sweep with one 64-bit atomic store, so a reader that sees anything else
// sees the final value, with no fence on either side.
#define ADE_EMPTY 0x7ff4deadbeef0001ull

// pass 1 over the whole padded grid: ghosts copied, interior rescaled
// (ADEsolver_calc.cpp:64-72).  Also resets the block tickets and empties the
// edge words of this solute.
__global__ __launch_bounds__(256) void k_ade_pass1(fx_params P, fx_state64 s, int ichem, int slot,
                                                   fx_ctl *ctl, unsigned mode,
                                                   unsigned long long *edges, int nblocks)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate || ctl->it <= 1)) return;
    const int64_t mm = (int64_t)P.ld * (P.ncols + 2);
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) ctl->ade_ticket = 0;
    const int64_t ne = (int64_t)nblocks * P.ld;      // <= mm
    if (i < ne) {
        unsigned long long *e = edges + (size_t)slot * 2 * ne;
        e[i] = ADE_EMPTY;
        e[ne + i] = ADE_EMPTY;
    }
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

// Record of one cell for the sweep: everything of the cell's update that
// does not depend on the values the sweep carries (the final concentration
// and east-face flux of the W cell, the final concentration and north-face
// flux of the S cell), as seven double pairs.  The flag word says which face
// still needs the carried value and whether the cell is updated at all.
#define AR_ACTIVE 1u    // wet and not the sweep's unvisited last cell
#define AR_CX 2u        // east face: QUICK from the W cell's final value
#define AR_CY 4u        // north face: QUICK from the S cell's final value
#define AR_LAST 8u      // (NROWS, NCOLS): the sweep stops one cell short
// neither ACTIVE nor LAST: a dry cell, set to 0

// slot (step tt of a chunk, lane l) sits at lane position (l + 8 tt) mod 64 of
// its row: the sweep reads a whole row (conflict-free), the helpers write 8
// steps of 8 columns at once (conflict-free too)
__device__ inline int ade_pos(int l, int tt) { return (l + 8 * tt) & (ADE_W - 1); }

struct ade_lds {
    double2 rec[2][ADE_NP][ADE_CHK][ADE_W];   // record ring, two chunks
    double out[2][ADE_CHK][ADE_W];            // final concentrations of the chunk
    double2 bnd[2][ADE_CHK];                  // (qfcds, final c) of the column before the block
    int ticket, bail;
};

// The same arrays one element off, passed as separate kernel arguments:
// loads of neighbouring elements through them stay separate loads into the
// registers that carry them to the next chunk, instead of one wide load whose
// halves must be copied out (a copy waits for the load).
struct ade_off {
    const double *Bw, *Be;      // pass-1 copies, slot 0, one element before / after
    const double *he;           // h + 1
    const uint8_t *lw, *le;     // ldry - 1, ldry + 1
};

struct ade_blk {
    const double *B, *h, *fe1, *fn1;
    const double *Bw, *Be, *he;              // B - 1, B + 1, h + 1
    double *A;
    const uint8_t *ldry, *lw, *le;           // ldry, ldry - 1, ldry + 1
    size_t ld;
    int nr, nc, iy0;
    double D, dtfl, dx, dy, area;
};

// The inputs of one cell's record (row ix, column iy), gathered ahead of use.
struct ade_in {
    double fp, fw, fe, fee, fs, fn, fnn, Ae, An, hp, he, hn, qxl, qyl;
    uint8_t lp, lw, le, lee, ls, ln, lnn;    // ldry of P W E EE S N NN, compared on use
    bool ok;
};

__device__ inline void ade_gather(const ade_blk &K, int ix, int iy, bool ok, ade_in &g)
{
    g.ok = ok;
    if (!ok) return;
    const size_t ld = K.ld;
    const size_t p = (size_t)ix + (size_t)iy * ld;
    const size_t iee = (size_t)min(ix + 2, K.nr + 1) + (size_t)iy * ld;
    const size_t pnn = (size_t)ix + (size_t)min(iy + 2, K.nc + 1) * ld;
    g.fp = K.B[p]; g.fw = K.Bw[p]; g.fe = K.Be[p]; g.fee = K.B[iee];
    g.fs = K.B[p - ld]; g.fn = K.B[p + ld]; g.fnn = K.B[pnn];
    g.Ae = K.A[p + 1]; g.An = K.A[p + ld];
    g.hp = K.h[p]; g.he = K.he[p]; g.hn = K.h[p + ld];
    g.qxl = K.fe1[p]; g.qyl = K.fn1[p];
    g.lp = K.ldry[p]; g.lw = K.lw[p]; g.le = K.le[p]; g.lee = K.ldry[iee];
    g.ls = K.ldry[p - ld]; g.ln = K.ldry[p + ld]; g.lnn = K.ldry[pnn];
}

// The chain-independent part of ADEsolver_calc.cpp:92-289 for one cell, in
// the reference's operation order (bitwise): the neighbour bounds of pass 1
// (:51-62), the diffusive face fluxes (:150-154, :197-201), the face values
// of the cases that do not read the W or S cell's final value, with the
// >= 0 clamp, the boundary overrides and the inflow limits (:155-194,
// :202-241), and cvolpot/dtfl (:244-245).
__device__ inline void ade_record(const ade_blk &K, int ix, int iy, const ade_in &g, double2 (&r)[ADE_NP])
{
    const double nt = 1, sigc = 0.5;
    const double fp = g.fp;
    uint32_t fl = 0;
    double hp = 0, cvp = 0, cmx = 0, cmn = 0, ax = 0, qx = 0, px = 0, lx = 0, ay = 0, qy = 0, py = 0, ly = 0;
    if (ix == K.nr && iy == K.nc) {
        fl = AR_LAST;
    } else if (g.lp != 1) {
        fl = AR_ACTIVE;
        hp = g.hp;
        const double fw = g.fw, fe = g.fe, fee = g.fee, fs = g.fs, fn = g.fn, fnn = g.fnn;
        const double Ae = g.Ae, An = g.An, he = g.he, hn = g.hn, qxl = g.qxl, qyl = g.qyl;
        const bool dw = g.lw == 1, de = g.le == 1, dee = g.lee == 1;
        const bool ds = g.ls == 1, dn = g.ln == 1, dnn = g.lnn == 1;
        // inflow limits of the E and N faces (:189-190, :236-237), used by both
        // the chain and the precomputed cases
        const double lim_e = -(fe * he) * K.area / K.dtfl, lim_n = -(fn * hn) * K.area / K.dtfl;
        // pass-1 bounds: W and S rescaled, E and N raw
        cmx = fmax(fw, fmax(Ae, fmax(fs, An)));
        cmn = fmin(fw, fmin(Ae, fmin(fs, An)));
        {   // X direction
            double fem = 0., pfde = 0.;
            if (!de) {
                const double hnue = fmax(hp * nt * he * nt, .0001);
                const double hne = sqrt(hnue) / sigc / K.dx * K.dy * K.D;
                pfde = -hne * (fe - fp);
                if (qxl > 0.0) fem = dw ? 0.5 * fp + 0.5 * fe : 0.;
                else fem = !dee ? .375 * fp + .75 * fe - .125 * fee : 0.5 * fp + 0.5 * fe;
            }
            if (!de && qxl > 0.0 && !dw && ix != K.nr) {
                fl |= AR_CX;
                ax = .375 * fe;
                qx = qxl;
                px = pfde;
                lx = lim_e;
            } else {
                fem = fmax(0., fem);
                if (ix == K.nr) fem = Ae;                    // A(NROWS+1, iy)
                const double pfce = qxl * fem * K.dy;
                double pfe = pfce + pfde;
                if (pfe < 0) pfe = !de ? fmax(pfe, lim_e) : 0.;
                px = pfe;
            }
        }
        {   // Y direction
            double fnm = 0., qfdn = 0.;
            if (!dn) {
                const double hnue = fmax(.0001, hp * nt * hn * nt);
                const double hnn = sqrt(hnue) / sigc / K.dy * K.dx * K.D;
                qfdn = -hnn * (fn - fp);
                if (qyl > 0.0) fnm = ds ? 0.5 * fp + .5 * fn : 0.;
                else fnm = !dnn ? .375 * fp + .75 * fn - .125 * fnn : .5 * fp + .5 * fn;
            }
            if (!dn && qyl > 0.0 && !ds && iy != K.nc) {
                fl |= AR_CY;
                ay = .375 * fn;
                qy = qyl;
                py = qfdn;
                ly = lim_n;
            } else {
                fnm = fmax(0., fnm);
                if (iy == K.nc) fnm = An;                    // A(ix, NCOLS+1)
                const double qfcn = qyl * fnm * K.dx;
                double qfn = qfcn + qfdn;
                if (qfn < 0) qfn = !dn ? fmax(qfn, lim_n) : 0.;
                py = qfn;
            }
        }
        cvp = (fp * hp) * K.area / K.dtfl;
    }
    r[0] = make_double2(fp, hp);
    r[1] = make_double2(cvp, cmx);
    r[2] = make_double2(cmn, ax);
    r[3] = make_double2(qx, px);
    r[4] = make_double2(lx, ay);
    r[5] = make_double2(qy, py);
    r[6] = make_double2(ly, __longlong_as_double((long long)fl));
}

// slot m of record-thread h for chunk k: 384 threads cover the chunk's 512
// slots in two rounds; a wave covers 8 steps of 8 columns (8 contiguous rows
// of 8 columns)
#define ADE_FILLERS (64 * (ADE_WAVES - 2))
#define ADE_FROUNDS ((ADE_CHK * ADE_W + ADE_FILLERS - 1) / ADE_FILLERS)

__device__ inline void ade_slot(const ade_blk &K, int k, int j, int &tt, int &l, int &ix, int &iy, bool &ok)
{
    tt = j % ADE_CHK;
    l = j / ADE_CHK;
    ix = k * ADE_CHK + tt - l + 1;
    iy = K.iy0 + l;
    ok = ix >= 1 && ix <= K.nr && iy <= K.nc;
}

// lane i receives lane i-1's v, lane 0 receives below0 (a wave shift by one
// lane; the lane without a source keeps the old operand)
__device__ inline double ade_from_below(double v, double below0)
{
    const long long x = __double_as_longlong(v), o = __double_as_longlong(below0);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// The chain (wave 0): the dependent part of ADEsolver_calc.cpp:92-289, one
// cell per lane and step.
template <bool AREA_P2>
__device__ inline void ade_chain(const ade_blk &K, ade_lds &S, const double *fe1, int lane, int T, int NK,
                                 bool has_next, int b, unsigned long long *qedge, unsigned long long *cedge)
{
    const int nr = K.nr, nc = K.nc;
    const size_t ld = K.ld;
    const int iy = K.iy0 + lane;
    const bool col_ok = iy <= nc;
    const double dx = K.dx, dy = K.dy, area = K.area, dtfl = K.dtfl, inv_area = 1.0 / K.area;
    const bool edge_lane = lane == ADE_W - 1 && has_next;
    // W cell of row 1: ghost row 0, its value and east-face flux (:105-116)
    double fw = 0.0, pfw = 0.0;
    if (col_ok) {
        const size_t c0 = (size_t)iy * ld;
        fw = K.A[c0];
        pfw = K.A[c0] * fe1[c0] * dy;
    }
    double q_out = 0.0, c_out = 0.0;
    for (int k = 0; k <= NK; k++) {
        __syncthreads();
        if (k == NK || S.bail) continue;
        const int buf = k & 1;
        double2 bd_next = S.bnd[buf][0];
        for (int tt = 0; tt < ADE_CHK; tt++) {
            const int t = k * ADE_CHK + tt;
            if (t >= T) break;
            // the S cell's result: one lane down, or the previous block's
            // last column for lane 0 (read one step ahead)
            const double2 bd = bd_next;
            if (tt + 1 < ADE_CHK) bd_next = S.bnd[buf][tt + 1];
            const double qfs = ade_from_below(q_out, bd.x), fs = ade_from_below(c_out, bd.y);
            const int ix = t - lane + 1;
            if (!(col_ok && ix >= 1 && ix <= nr)) continue;
            const int pos = ade_pos(lane, tt);
            const double2 r0 = S.rec[buf][0][tt][pos], r1 = S.rec[buf][1][tt][pos];
            const double2 r2 = S.rec[buf][2][tt][pos], r3 = S.rec[buf][3][tt][pos];
            const double2 r4 = S.rec[buf][4][tt][pos], r5 = S.rec[buf][5][tt][pos];
            const double2 r6 = S.rec[buf][6][tt][pos];
            const uint32_t fl = (uint32_t)__double_as_longlong(r6.y);
            const double fp = r0.x, hp = r0.y, cvp = r1.x, cmx = r1.y, cmn = r2.x;
            const double A1 = .75 * fp;
            // east face (:149-194) from the W cell's final value
            double pfe;
            {
                const double fem = fmax(0., -.125 * fw + A1 + r2.y);
                double pc = r3.x * fem * dy + r3.y;
                if (pc < 0) pc = fmax(pc, r4.x);
                pfe = (fl & AR_CX) ? pc : r3.y;
            }
            // north face (:196-241) from the S cell's final value
            double qfn;
            {
                const double fnm = fmax(0., -.125 * fs + A1 + r4.y);
                double qc = r5.x * fnm * dx + r5.y;
                if (qc < 0) qc = fmax(qc, r6.x);
                qfn = (fl & AR_CY) ? qc : r5.y;
            }
            // available material (:244-279)
            const double cvolrat = cvp + pfw + qfs;
            const bool pos_ = cvolrat > 0.0f;
            const bool pg = pfe > 0., qg = qfn > 0., pge = pfe >= 0., qge = qfn >= 0.;
            const double cb = cvolrat - qfn, cb2 = cvolrat - pfe;
            const double m1 = fmin(pfe, cb), m2 = fmin(qfn, cb2);
            // the reference's branches as selects: B (:254-255), C (:256-257),
            // D (:260-266), E (:267-273), F (:274-277); A (:248-253) below
            const bool cB = pos_ && pg && !qg, cC = pos_ && !pg && qg;
            const bool cD = !pos_ && pge && qfn < 0., cE = !pos_ && pfe < 0. && qge, cF = !pos_ && pge && qge;
            double pfe2 = (cB || (cD && cb > 0.)) ? m1 : ((cD || cF) ? 0. : pfe);
            double qfn2 = (cC || (cE && cb2 > 0.)) ? m2 : ((cE || cF) ? 0. : qfn);
            if (pos_ && pg && qg && pfe + qfn > cvolrat) {
                const double cf = qfn / (pfe + qfn);
                pfe2 = (1. - cf) * cvolrat;
                qfn2 = cf * cvolrat;
            }
            // new concentration (:281-285)
            double dc = (pfw - pfe2 + qfs - qfn2) * dtfl;
            dc = AREA_P2 ? dc * inv_area : dc / area;
            double con = fp + dc / hp;
            con = fmin(cmx, con);
            con = fmax(cmn, con);
            if (fl & AR_ACTIVE) {
                q_out = qfn2; c_out = con; pfw = pfe2;
            } else if (fl & AR_LAST) {
                q_out = 0.0; c_out = fp;
            } else {                                                   // dry (:118-124)
                q_out = 0.0; c_out = 0.0; pfw = 0.0;
            }
            fw = c_out;
            S.out[buf][tt][pos] = c_out;
            if (edge_lane) {
                __hip_atomic_store(&qedge[(size_t)b * ld + ix], (unsigned long long)__double_as_longlong(q_out),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&cedge[(size_t)b * ld + ix], (unsigned long long)__double_as_longlong(c_out),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// The sweep proper: one block (workgroup) of eight waves per 64 columns and
// solute.  Wave 0 is the systolic array -- lane l owns column iy0 + l and runs
// one row behind lane l-1, so the S cell's result is one lane down -- and does
// only the dependent chain, reading each cell's record from LDS.  Its last
// lane stores the block's last column (north-face flux and final value) row
// by row into the edge words for the next block.  Wave 4 (on wave 0's SIMD,
// and light) fetches the previous block's edge words for lane 0 one chunk
// ahead and writes finished chunks back to A; the other six waves gather the
// inputs of the chunk after next and build the records of the next chunk,
// between two barriers while wave 0 sweeps.
//
// Blocks take column ranges in ticket order and solute slot t % nsol, so a
// block only waits on one that started earlier; every wait is bounded, and a
// block whose wait expires sets ade_error, fills its own edge words so the
// blocks after it do not wait, and runs its barriers to the end.
__global__ __launch_bounds__(64 * ADE_WAVES) void k_ade_sweep(fx_params P, fx_state64 s, int k0, int nsol,
                                                   ade_off O, fx_ctl *ctl, unsigned mode,
                                                   unsigned long long *edges, int nblocks)
{
    if ((mode & FXM_GATED) && (ctl->halt || !ctl->gate || ctl->it <= 1)) return;
    __shared__ ade_lds S;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    if (tid == 0) {
        S.ticket = (int)atomicAdd(&ctl->ade_ticket, 1u);
        S.bail = 0;
    }
    __syncthreads();
    const int ticket = S.ticket;
    const int slot = ticket % nsol, b = ticket / nsol, ichem = k0 + slot;
    const int nr = P.nrows, nc = P.ncols;
    const size_t ld = P.ld;
    const int64_t mm = (int64_t)ld * (nc + 2);
    const int64_t ne = (int64_t)nblocks * ld;
    unsigned long long *qedge = edges + (size_t)slot * 2 * ne;
    unsigned long long *cedge = qedge + ne;
    ade_blk K;
    K.A = s.conc[ichem];
    K.B = s.conc_tmp + (size_t)slot * mm;
    K.h = s.h; K.fe1 = s.fe1; K.fn1 = s.fn1; K.ldry = s.ldry;
    K.Bw = O.Bw + (size_t)slot * mm; K.Be = O.Be + (size_t)slot * mm; K.he = O.he;
    K.lw = O.lw; K.le = O.le;
    K.ld = ld; K.nr = nr; K.nc = nc; K.iy0 = b * ADE_W + 1;
    K.D = P.D[ichem]; K.dtfl = ctl->dtfl; K.dx = P.dxy; K.dy = P.dxy; K.area = P.arbase;
    const int T = nr + ADE_W - 1;               // steps: lane 63 ends on row NROWS
    const int NK = (T + ADE_CHK - 1) / ADE_CHK;
    // the block after this one reads its last column; the last block has none
    const bool has_next = K.iy0 + ADE_W - 1 < nc;

    if (wave == 0) {
        // division by the cell area: an exact product when the area is a
        // power of two (a 1 m or 2 m grid), else the IEEE division
        const uint64_t ab = (uint64_t)__double_as_longlong(K.area);
        const bool area_p2 = (ab & 0x000fffffffffffffull) == 0 && ((ab >> 52) & 0x7ff) != 0 && ((ab >> 52) & 0x7ff) != 0x7ff;
        if (area_p2) ade_chain<true>(K, S, s.fe1, lane, T, NK, has_next, b, qedge, cedge);
        else ade_chain<false>(K, S, s.fe1, lane, T, NK, has_next, b, qedge, cedge);
    } else if (wave == 4) {
        // ---------- boundary of lane 0 and write-back ----------
        // chunk k's boundary: lane 0's rows 8k+1 .. 8k+8, from the previous
        // block's edge words, or the ghost column 0 with qfcds = 0 (:140, :145)
        auto fetch_bnd = [&](int k) {
            const int r = k * ADE_CHK + 1 + lane;
            const bool need = lane < ADE_CHK && r <= nr;
            double q = 0.0, c = 0.0;
            if (b == 0) {
                if (need) c = K.A[(size_t)r];                      // A(r, 0)
            } else {
                bool have = !need;
                long long spin = 0;
                for (;;) {
                    if (!have) {
                        const unsigned long long qb = __hip_atomic_load(&qedge[(size_t)(b - 1) * ld + r], __ATOMIC_RELAXED,
                                                                        __HIP_MEMORY_SCOPE_AGENT);
                        const unsigned long long cb = __hip_atomic_load(&cedge[(size_t)(b - 1) * ld + r], __ATOMIC_RELAXED,
                                                                        __HIP_MEMORY_SCOPE_AGENT);
                        if (qb != ADE_EMPTY && cb != ADE_EMPTY) {
                            have = true;
                            q = __longlong_as_double((long long)qb);
                            c = __longlong_as_double((long long)cb);
                        }
                    }
                    if (__all(have)) break;
                    if (++spin > ADE_SPIN_LIMIT) { if (lane == 0) S.bail = 1; break; }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            if (need) S.bnd[k & 1][lane] = make_double2(q, c);
        };
        fetch_bnd(0);
        for (int k = 0; k <= NK; k++) {
            __syncthreads();
            if (k >= 1) {                                              // chunk k-1 is final
                const int buf = (k - 1) & 1;
#pragma unroll
                for (int m = 0; m < ADE_CHK * ADE_W / 64; m++) {
                    int tt, l, ix, iy;
                    bool ok;
                    ade_slot(K, k - 1, lane + 64 * m, tt, l, ix, iy, ok);
                    if (ok) K.A[(size_t)ix + (size_t)iy * ld] = S.out[buf][tt][ade_pos(l, tt)];
                }
            }
            if (k + 1 < NK && !S.bail) fetch_bnd(k + 1);
        }
        if (S.bail && has_next && lane == 0) {
            // the step is void; let the blocks after this one run to their end
            atomicOr(&ctl->ade_error, 1u);
            for (int r = 1; r <= nr; r++) {
                __hip_atomic_store(&qedge[(size_t)b * ld + r], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&cedge[(size_t)b * ld + r], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (S.bail && lane == 0) {
            atomicOr(&ctl->ade_error, 1u);
        }
    } else {
        // ---------- records: gather chunk k+2, build chunk k+1 ----------
        const int h = (wave < 4 ? wave - 1 : wave - 2) * 64 + lane;
        ade_in g[ADE_FROUNDS];
        auto gather = [&](int k) {
#pragma unroll
            for (int m = 0; m < ADE_FROUNDS; m++) {
                int tt, l, ix, iy;
                bool ok;
                ade_slot(K, k, h + ADE_FILLERS * m, tt, l, ix, iy, ok);
                ade_gather(K, ix, iy, ok && k < NK && h + ADE_FILLERS * m < ADE_CHK * ADE_W, g[m]);
            }
        };
        auto build = [&](int k) {
            const int buf = k & 1;
#pragma unroll
            for (int m = 0; m < ADE_FROUNDS; m++) {
                int tt, l, ix, iy;
                bool ok;
                ade_slot(K, k, h + ADE_FILLERS * m, tt, l, ix, iy, ok);
                if (g[m].ok) {
                    double2 r[ADE_NP];
                    ade_record(K, ix, iy, g[m], r);
                    const int pos = ade_pos(l, tt);
#pragma unroll
                    for (int q = 0; q < ADE_NP; q++) S.rec[buf][q][tt][pos] = r[q];
                }
            }
        };
        gather(0);
        build(0);
        gather(1);
        for (int k = 0; k <= NK; k++) {
            __syncthreads();
            if (k + 1 < NK) build(k + 1);
            if (k + 2 < NK) gather(k + 2);
        }
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
    // edge words of slot k: 2 * nblocks * mrows after the first k slots
    unsigned long long *edges = (unsigned long long *)d->ade_edges;
    for (int k = 0; k < nsol; k++) {
        fx_prof_begin(d, FX_K_ADE_PASS1);
        hipLaunchKernelGGL(k_ade_pass1, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, d->stream,
                           P, d->s64, k0 + k, k, d->ctl, mode, edges, nblocks);
        fx_prof_end(d, FX_K_ADE_PASS1);
        FX_HIP(hipGetLastError());
    }
    ade_off O;
    const uintptr_t e = sizeof(double);
    O.Bw = (const double *)((uintptr_t)d->s64.conc_tmp - e);
    O.Be = (const double *)((uintptr_t)d->s64.conc_tmp + e);
    O.he = (const double *)((uintptr_t)d->s64.h + e);
    O.lw = (const uint8_t *)((uintptr_t)d->s64.ldry - 1);
    O.le = (const uint8_t *)((uintptr_t)d->s64.ldry + 1);
    fx_prof_begin(d, FX_K_ADE_SWEEP);
    hipLaunchKernelGGL(k_ade_sweep, dim3(nblocks * nsol), dim3(64 * ADE_WAVES), 0, d->stream,
                       P, d->s64, k0, nsol, O, d->ctl, mode, edges, nblocks);
    fx_prof_end(d, FX_K_ADE_SWEEP);
    return fx_hip_check(hipGetLastError(), "k_ade_sweep");
}
