// fx_ade32.hip -- fast advection-dispersion (FX_ADE_FAST) for the FP32-mixed
// path: one column-sweep kernel for up to four solutes of a step.
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
//
// Layout of the sweep: one wave per strip of 58 rows.  Lane k holds row
// r0-3+k of every column the wave visits; a cell's stencil reaches rows
// r-3..r+2 (the limited flux of the W neighbour needs the W neighbour's own
// W face), so lanes 3..60 own rows r0..r0+57.  Cross-row values move by lane
// shuffles; cross-column values are a four-column register window (S, P, N,
// NN) that shifts by one column per step.  Nothing is staged in LDS and every
// cell is read from HBM once per strip (plus five halo columns per block).
// HBM per cell: d before and after the step (FP64), the step's two face
// discharges and the flag byte (written by the hydro sweep), and the solute
// masses read and written.
#include "fx_kern32.h"

// multiply-adds are fused within one source expression only: the column
// routines are inlined several times (block prologue, loop body, tail), and
// fusing across statements can differ between copies, which would make a
// cell's rounding depend on its place in the tile and results on the tiling
#pragma clang fp contract(on)

#define AOWN FX32_AOWN        // rows owned per strip
#define ANC 4                 // solutes per launch

namespace {

// one cell's HBM inputs in flight
template <int NC>
struct apref {
    double dold, dnew;
    float fe, fn, mult;
    float m[NC];
    uint8_t f;                // flag byte (kept a byte: a widening after the
                              // load would wait for the load there)
};

// per-cell geometry of one column, shared by the solutes
struct ageo {
    float h, sh, fe, fn;      // depth, its square root, the step's face discharges
    bool w;
};

struct actx {
    const fx_params *P;
    const a32p *A;
    const double *d_old, *d_new;
    const float *m_in[ANC];
    float *m_out[ANC];
    double meteoi, inflowi;
    int inflow_c;             // the inflow cell's column if it is in this lane's row, else -1
    float cm[ANC], cin[ANC], kdiff[ANC];
    float hdry, dx, area_dt, dt_area;
    int row, mcols;
    uint32_t ld;
    bool row_ok, transport;
};

template <int NC>
__device__ inline void a_load(const actx &X, int c, apref<NC> &a)
{
    a.f = 0;
    if (!X.row_ok || c < 0 || c >= X.mcols) return;      // column test is uniform
    const uint32_t p = (uint32_t)X.row + (uint32_t)c * X.ld;
    const uint32_t o4 = p * 4u;
    const a32p &A = *X.A;
    a.f = at_off(A.flags, p);
    a.dold = at_off(X.d_old, p * 8u);
    a.dnew = at_off(X.d_new, p * 8u);
    const float2 ff = at_off(A.ff, p * 8u);
    a.fe = ff.x;
    a.fn = ff.y;
    a.mult = X.P->melt_mult ? at_off(X.P->melt_mult, o4) : 1.0f;
#pragma unroll
    for (int k = 0; k < NC; k++) a.m[k] = at_off(X.m_in[k], o4);
}

// derived values of a freshly loaded cell of column col: geometry, and per
// solute the pass-1 concentration c1 (craw where the cell was dry before the
// step, 0 where dry now) and the raw concentration cr after the melt mixing.
// Without transport c1 carries the mixed mass itself, which is the update.
// The inflow cell is found by its column, not by a bit of the prefetched
// flags: a test on a loaded value right after the load would wait for it there.
template <int NC>
__device__ inline void a_derive(const actx &X, const apref<NC> &a, int col, ageo &g, float *c1, float *cr)
{
    const fx_params &P = *X.P;
    const bool inner = a.f & FXF_INNER;
    const bool weir = a.f & FXF_WEIR;
    const double melt = (!inner || (a.f & FXF_NODATA)) ? 0.0 : X.meteoi * (double)a.mult;
    // max(0, d) >= hdry > 0 is d >= hdry (a NaN depth is dry in both forms)
    const bool w = inner && a.dnew >= P.hdry && !weir;
    const bool wp = inner && a.dold >= P.hdry && !weir;
    double dm = a.dold + melt;
    float inf = 0.f;                                   // inflow depth carrying solute
    if (col == X.inflow_c) {                           // add_inflow (common.cpp:324-376)
        if (dm <= 0.0) dm = 0.0;
        if (dm > 0.0) inf = (float)X.inflowi;          // solute arrives only into water
        dm += X.inflowi;
    }
    const float h0 = inner ? (float)fmax(0.0, dm) : 0.f;
    const float mlt = (inner && a.dold > 0.0) ? (float)melt : 0.f;   // solute arrives with melt only into water
    // the depth and its square root are read only where the cell is wet
    // (faces towards a dry cell carry nothing, and a dry cell computes none)
    g.h = w ? (float)a.dnew : 0.f;
    g.sh = fx_sqrt(g.h);
    g.w = w;
    g.fe = inner ? a.fe : 0.f;
    g.fn = inner ? a.fn : 0.f;
    // one reciprocal per depth, shared by the solutes (h >= hdry > 0 where wet)
    const float ih0 = h0 > 0.f ? __builtin_amdgcn_rcpf(h0) : 0.f;
    const float ihn = (w && wp) ? __builtin_amdgcn_rcpf(g.h) : 0.f;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const float mv = inner ? a.m[k] + mlt * X.cm[k] + inf * X.cin[k] : 0.f;
        const float r = mv * ih0;
        const float c = w ? (wp ? mv * ihn : r) : 0.f;
        c1[k] = X.transport ? c : mv;
        cr[k] = r;
    }
}

// face flux of one solute before the outflow limiter, from the cell P (value
// cp) towards its neighbour Q (cq, wet flag wq); cu is the value one further
// upstream of P, cd one further downstream of Q, wu/wd their wet flags, q the
// face discharge; sq = sqrt(max(hp*hq, 1e-4)) and cap = -hq*area/dt are the
// solute-independent parts (ADEsolver_calc.cpp:149-241)
__device__ inline float a_face(float q, float cp, float cq, bool wq, float cu, bool wu, float cd, bool wd,
                               float sq, float cap, float kdiff, float dx)
{
    if (!wq) return 0.f;                          // dry neighbour: fem = 0, no diffusion
    const float pd = -(sq * kdiff) * (cq - cp);
    // upwind-biased face value: -1/8 of the far upstream value, 3/4 of the
    // upstream one and 3/8 of the downstream one, upstream being P for
    // q > 0 and Q otherwise; the plain mean where the far value is dry.
    // Written as the mean less 1/8 of the upwind curvature far - 2 up + down,
    // where up + down = cp + cq whichever way the flow goes.
    const bool pos = q > 0.f;
    const float far = pos ? cu : cd, up = pos ? cp : cq;
    const float sum = cp + cq;
    const float curv = (pos ? wu : wd) ? (far + sum) - 3.f * up : 0.f;
    const float fm = fx_max0(.5f * sum - .125f * curv);
    float pf = q * fm * dx + pd;
    if (pf < 0.f) pf = fmaxf(pf, cq * cap);
    return pf;
}

// outflow limiter of one cell against its available solute
__device__ inline void a_limit(float cv, float &pfe, float &qfn)
{
    if (cv > 0.f) {
        if (pfe > 0.f && qfn > 0.f) {
            if (pfe + qfn > cv) {
                const float cf = qfn * __builtin_amdgcn_rcpf(pfe + qfn);
                pfe = (1.f - cf) * cv;
                qfn = cf * cv;
            }
        } else if (pfe > 0.f) {
            pfe = fminf(pfe, cv - qfn);
        } else if (qfn > 0.f) {
            qfn = fminf(qfn, cv - pfe);
        }
    } else {
        if (pfe >= 0.f && qfn < 0.f) {
            const float cb = cv - qfn;
            pfe = cb > 0.f ? fminf(pfe, cb) : 0.f;
        } else if (pfe < 0.f && qfn >= 0.f) {
            const float cb = cv - pfe;
            qfn = cb > 0.f ? fminf(qfn, cb) : 0.f;
        } else if (pfe >= 0.f && qfn >= 0.f) {
            pfe = 0.f;
            qfn = 0.f;
        }
    }
}

__device__ inline float up1(float v) { return __shfl_up(v, 1, 64); }     // value of row r-1
__device__ inline float dn1(float v) { return __shfl_down(v, 1, 64); }   // value of row r+1
__device__ inline float dn2(float v) { return __shfl_down(v, 2, 64); }   // value of row r+2

// register window of the sweep: index 0 = column c-1 (S), 1 = c (P),
// 2 = c+1 (N), 3 = c+2 (NN)
template <int NC>
struct awin {
    ageo g[4];
    float c1[4][NC], cr[4][NC];
    float pn_s[NC], n_s[NC];          // N-face flux of column c-1: before / after limiting
};

template <int NC>
__device__ inline void a_shift(awin<NC> &W)
{
#pragma unroll
    for (int j = 0; j < 3; j++) {
        W.g[j] = W.g[j + 1];
#pragma unroll
        for (int k = 0; k < NC; k++) { W.c1[j][k] = W.c1[j + 1][k]; W.cr[j][k] = W.cr[j + 1][k]; }
    }
}

// fluxes of column c (window slot 1); with upd, also the update of the own
// rows.  Returns nothing: the limited N-face flux is carried to column c+1.
template <int NC>
__device__ inline void a_column(const actx &X, awin<NC> &W, int c, bool upd, bool own, bool &bad)
{
    const ageo &gP = W.g[1], &gN = W.g[2], &gNN = W.g[3], &gS = W.g[0];
    // the wet flags of rows r-1, r+1 and r+2 from the wave's mask of column c
    // (an edge lane sees a dry row beyond the wave; it owns no cell)
    const uint64_t mw = __builtin_amdgcn_ballot_w64(gP.w);
    const bool wW = lane_bit(mw << 1), wE = lane_bit(mw >> 1), wEE = lane_bit(mw >> 2);
    const float hE = dn1(gP.h);
    const uint32_t p = (uint32_t)X.row + (uint32_t)c * X.ld;
    const bool act = X.transport && gP.w;
    // per cell, shared by the solutes: the faces' eddy terms and inflow
    // floors, h*area/dt of the limiter and dt/(area*h) of the update
    float sE = 0.f, sN = 0.f, hA = 0.f, kup = 0.f;
    if (act) {
        // sqrt(max(hp hq, 1e-4)) as max(sqrt(hp) sqrt(hq), 0.01): one square
        // root per cell instead of one per face
        sE = fx_max_nn(gP.sh * dn1(gP.sh), .01f);
        sN = fx_max_nn(gP.sh * gN.sh, .01f);
        hA = gP.h * X.area_dt;
        kup = X.dt_area * __builtin_amdgcn_rcpf(gP.h);
    }
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const float cp = W.c1[1][k];
        float pe = 0.f, pn = 0.f;
        const float cW = up1(cp), cE = dn1(cp), cEE = dn2(cp);
        const float crE = dn1(W.cr[1][k]);
        if (act) {
            pe = a_face(gP.fe, cp, cE, wE, cW, wW, cEE, wEE, sE, -hE * X.area_dt, X.kdiff[k], X.dx);
            pn = a_face(gP.fn, cp, W.c1[2][k], gN.w, W.c1[0][k], gS.w, W.c1[3][k], gNN.w, sN,
                        -gN.h * X.area_dt, X.kdiff[k], X.dx);
        }
        // limiting of this cell: inflow terms of its W and S faces before their
        // own cells' limiting
        const float pe_w = up1(pe);
        float e = pe, n = pn;
        if (gP.w) a_limit(cp * hA + pe_w + W.pn_s[k], e, n);
        else { e = 0.f; n = 0.f; }
        const float e_w = up1(e);
        if (upd && own) {
            float out;
            if (!X.transport) {
                out = cp;                                // melt mixing only
            } else if (!gP.w) {
                out = 0.f;                               // dry after the update (:69-71,119-124)
            } else {
                float con = cp + (e_w - e + W.n_s[k] - n) * kup;
                // bounds of pass 1: W and S already rescaled, E and N not yet (:51-62)
                const float cs = W.c1[0][k], cn = W.cr[2][k];
                // clamped to [lo, hi] (lo <= hi) by one median of three
                const float hi = fmaxf(fmaxf(cW, crE), fmaxf(cs, cn));
                const float lo = fminf(fminf(cW, crE), fminf(cs, cn));
                con = __builtin_amdgcn_fmed3f(con, lo, hi);
                out = con * gP.h;
            }
            bad |= X.transport && gP.w && !fx_finitef(out);   // reported as ctl->status bit 2
            at_off(X.m_out[k], p * 4u) = out;
        }
        W.pn_s[k] = pn;
        W.n_s[k] = n;
    }
}

// Record b of the fold of the step sweep's CFL partials for the next clock:
// max wave speed, max depth, weir volume and melt area over the partials of
// chunk b, in a fixed order, by the block's one wave.  The clock reads the
// records after this launch has ended.
__device__ inline void fold_record(const a32fold &F, int b)
{
    const int chunk = (F.npart + F.nrec - 1) / F.nrec;
    const int lo = min(F.npart, b * chunk), hi = min(F.npart, lo + chunk);
    double a = 0.0, h = 0.0, w = 0.0, m = 0.0;
    for (int i = lo + (int)threadIdx.x; i < hi; i += 64) {
        const double2 *q = reinterpret_cast<const double2 *>(F.partials + FX32_NPV * (size_t)i);
        const double2 x = q[0], y = q[1];
        a = fmax(a, x.x);
        h = fmax(h, x.y);
        w += y.x;
        m += y.y;
    }
    a = wmax(a); h = wmax(h); w = wsum(w); m = wsum(m);
    if (threadIdx.x == 0) {
        double2 *r = reinterpret_cast<double2 *>(F.rec + FX32_NPV * (size_t)b);
        r[0] = make_double2(a, h);
        r[1] = make_double2(w, m);
    }
}

// one block of the sweep: strip bx (rows), column block by
template <int NC>
__device__ inline void ade_tile(const fx_params &P, const a32p &A, fx_ctl *ctl, int bx, int by,
                                const a32fold &F)
{
    if (ctl->halt) return;
    if (F.rec && (int)blockIdx.x < F.nrec) fold_record(F, (int)blockIdx.x);
    const int cur = ctl->cur;
    actx X;
    X.P = &P;
    X.A = &A;
    X.d_old = A.d[cur];
    X.d_new = A.d[cur ^ 1];
    X.transport = ctl->gate && ctl->it > 1;
    X.meteoi = P.meteo ? ctl->meteoi : 0.0;
    X.inflowi = ctl->inflowi;
    const float dt = (float)ctl->dtfl;
    X.hdry = (float)P.hdry;
    X.dx = (float)P.dxy;
    X.area_dt = (float)P.arbase / dt;
    X.dt_area = dt / (float)P.arbase;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        X.m_in[k] = A.mass[cur][A.k0 + k];
        X.m_out[k] = A.mass[cur ^ 1][A.k0 + k];
        X.cm[k] = (float)ctl->cmelt[A.k0 + k];
        X.cin[k] = (float)ctl->cin[A.k0 + k];
        X.kdiff[k] = 2.f * (float)P.D[A.k0 + k];     // 1/sigc/dx*dy with sigc = 0.5, dx = dy
    }
    const int lane = threadIdx.x;
    const int r0 = 1 + bx * AOWN;
    const int c0 = 1 + by * P.ach;
    const int c_last = min(c0 + P.ach - 1, P.ncols);
    X.row = r0 - 3 + lane;
    X.row_ok = X.row >= 0 && X.row < P.ld;
    X.mcols = P.ncols + 2;
    X.inflow_c = -1;                                  // the inflow cell is interior (host-checked)
    if (P.inflow && X.row == (int)((uint32_t)P.inflow_p % (uint32_t)P.ld))
        X.inflow_c = (int)((uint32_t)P.inflow_p / (uint32_t)P.ld);
    X.ld = (uint32_t)P.ld;
    const bool own = lane >= 3 && lane <= 3 + AOWN - 1 && X.row <= P.nrows;

    awin<NC> W;
    apref<NC> a, b;
    bool bad = false;
    // prologue: window = columns c0-3 .. c0; the N-face fluxes of column
    // c0-2 (before limiting) and of column c0-1 (both) feed column c0
    a_load<NC>(X, c0 - 3, a); a_derive<NC>(X, a, c0 - 3, W.g[0], W.c1[0], W.cr[0]);
    a_load<NC>(X, c0 - 2, a); a_derive<NC>(X, a, c0 - 2, W.g[1], W.c1[1], W.cr[1]);
    a_load<NC>(X, c0 - 1, a); a_derive<NC>(X, a, c0 - 1, W.g[2], W.c1[2], W.cr[2]);
    a_load<NC>(X, c0, a);     a_derive<NC>(X, a, c0, W.g[3], W.c1[3], W.cr[3]);
#pragma unroll
    for (int k = 0; k < NC; k++) { W.pn_s[k] = 0.f; W.n_s[k] = 0.f; }
    a_column<NC>(X, W, c0 - 2, false, own, bad);
    a_shift<NC>(W);
    a_load<NC>(X, c0 + 1, a); a_derive<NC>(X, a, c0 + 1, W.g[3], W.c1[3], W.cr[3]);
    a_column<NC>(X, W, c0 - 1, false, own, bad);
    if constexpr (NC >= 3) {
        // three or four solutes: one column of prefetch (a: c+2) keeps the
        // kernel at five waves per SIMD, which large grids of many rounds need
        a_load<NC>(X, c0 + 2, a);
        for (int c = c0; c <= c_last; c++) {
            a_shift<NC>(W);
            a_derive<NC>(X, a, c + 2, W.g[3], W.c1[3], W.cr[3]);
            if (c + 3 <= c_last + 2) a_load<NC>(X, c + 3, a);
            a_column<NC>(X, W, c, true, own, bad);
        }
        if (__any(bad) && lane == 0) atomicOr(&ctl->status, 4);
        return;
    }
    // steady state: prefetch two columns ahead (a: c+2, b: c+3)
    a_load<NC>(X, c0 + 2, a);
    a_load<NC>(X, c0 + 3, b);
    int c = c0;
    for (; c + 1 <= c_last; c += 2) {
        a_shift<NC>(W);
        a_derive<NC>(X, a, c + 2, W.g[3], W.c1[3], W.cr[3]);
        if (c + 4 <= c_last + 2) a_load<NC>(X, c + 4, a);
        a_column<NC>(X, W, c, true, own, bad);
        a_shift<NC>(W);
        a_derive<NC>(X, b, c + 3, W.g[3], W.c1[3], W.cr[3]);
        if (c + 5 <= c_last + 2) a_load<NC>(X, c + 5, b);
        a_column<NC>(X, W, c + 1, true, own, bad);
    }
    if (c <= c_last) {
        a_shift<NC>(W);
        a_derive<NC>(X, a, c + 2, W.g[3], W.c1[3], W.cr[3]);
        a_column<NC>(X, W, c, true, own, bad);
    }
    if (__any(bad) && lane == 0) atomicOr(&ctl->status, 4);
}

template <int NC>
__global__ __launch_bounds__(64) void k32_ade_sweep(fx_params P, a32p A, fx_ctl *ctl, int gx, int n, int chunk,
                                                    a32fold F)
{
    const int t = fx_xcd_tile((int)blockIdx.x, n, chunk);
    ade_tile<NC>(P, A, ctl, t % gx, t / gx, F);
}

// batched: block b runs block FX_BMAP_BLK(amap[b]) of member FX_BMAP_SLOT(amap[b]);
// every member of a batch carries the same number of solutes
template <int NC>
__global__ __launch_bounds__(64) void k32b_ade_sweep(const fx_bslot *__restrict__ slots, const uint32_t *__restrict__ amap, int k0)
{
    const uint32_t m = amap[blockIdx.x];
    const fx_bslot &B = slots[FX_BMAP_SLOT(m)];
    if (!B.ade_on) return;
    const int blk = (int)FX_BMAP_BLK(m);
    a32p A = B.A;
    A.k0 = k0;
    const a32fold F{};
    ade_tile<NC>(B.P, A, B.ctl, blk % B.agx, blk / B.agx, F);
}

template <int NC>
void launch_nc(fx_domain d, const fx_params &P, a32p A, dim3 g, const a32fold &F)
{
    const int n = (int)(g.x * g.y);
    hipLaunchKernelGGL(k32_ade_sweep<NC>, dim3(n), dim3(64), 0, d->stream, P, A, d->ctl, (int)g.x, n,
                       fx32_xcd_chunk((int)g.x, n), F);
}

inline dim3 ade_grid(const fx_params &P)
{
    return dim3((P.nrows + AOWN - 1) / AOWN, (P.ncols + P.ach - 1) / P.ach);
}

} // namespace

a32p make_a32p(fx_domain d)
{
    a32p A;
    A.d[0] = d->s32.d[0];
    A.d[1] = d->s32.d[1];
    A.ff = d->s32.ff;
    A.flags = d->s32.flags;
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < FX_MAX_CHEM; k++) A.mass[b][k] = d->s32.mass[b][k];
    A.melt_mult = d->melt_mult;
    A.mm = d->ncell;
    A.k0 = 0;
    return A;
}

int fx32_ade_gx(fx_domain d)
{
    return (int)ade_grid(fx_make_params(d)).x;
}

int fx32_ade_blocks(fx_domain d)
{
    const dim3 g = ade_grid(fx_make_params(d));
    return (int)(g.x * g.y);
}

int fx32b_launch_ade(hipStream_t st, const fx_bslot *slots, const uint32_t *amap, int nblocks, int nchem)
{
    for (int k0 = 0; k0 < nchem; k0 += ANC) {
        switch (std::min(ANC, nchem - k0)) {
        case 1: hipLaunchKernelGGL(k32b_ade_sweep<1>, dim3(nblocks), dim3(64), 0, st, slots, amap, k0); break;
        case 2: hipLaunchKernelGGL(k32b_ade_sweep<2>, dim3(nblocks), dim3(64), 0, st, slots, amap, k0); break;
        case 3: hipLaunchKernelGGL(k32b_ade_sweep<3>, dim3(nblocks), dim3(64), 0, st, slots, amap, k0); break;
        default: hipLaunchKernelGGL(k32b_ade_sweep<4>, dim3(nblocks), dim3(64), 0, st, slots, amap, k0); break;
        }
    }
    return fx_hip_check(hipGetLastError(), "k32b_ade_sweep");
}

// fold: the first launch also folds the step sweep's CFL partials into the
// records the next clock reads
int fx32_launch_ade(fx_domain d, bool fold)
{
    fx_params P = fx_make_params(d);
    a32p A = make_a32p(d);
    const dim3 g = ade_grid(P);
    fx_prof_begin(d, FX_K_ADE32);
    for (int k0 = 0; k0 < P.nchem; k0 += ANC) {
        A.k0 = k0;
        a32fold F{};
        if (fold && k0 == 0) {
            F.partials = d->s32.partials;
            F.rec = d->s32.rec;
            F.npart = (int)d->s32.npart;
            F.nrec = std::min<int>(FX32_REC, (int)(g.x * g.y));
        }
        switch (std::min(ANC, P.nchem - k0)) {
        case 1: launch_nc<1>(d, P, A, g, F); break;
        case 2: launch_nc<2>(d, P, A, g, F); break;
        case 3: launch_nc<3>(d, P, A, g, F); break;
        default: launch_nc<4>(d, P, A, g, F); break;
        }
    }
    fx_prof_end(d, FX_K_ADE32);
    return fx_hip_check(hipGetLastError(), "k32_ade_sweep");
}
