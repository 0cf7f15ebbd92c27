// fx_capi.hip -- the extern "C" entry points of libfluxos_hip.so
// (declared in include/fluxos_hip.h).  Host code only: domain lifetime,
// state transfer in the reference's arma::Mat layout, the granular seams and
// the device-resident loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fx_kern32.h"

static thread_local std::string g_err;

int fx_fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int fx_hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return FX_OK;
    if (e == hipErrorOutOfMemory) return fx_fail(FX_ERR_NOMEM, std::string(what) + ": out of device memory");
    return fx_fail(FX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void fx_prof_begin(fx_domain d, int id)
{
    if (!d->prof.on) return;
    fx_prof &p = d->prof;
    p.skipping = id >= 0 && id < FX_K_COUNT && (p.seq[id]++ % p.stride) != 0;
    if (p.skipping) return;
    if (p.pool.size() < 2) {
        for (int i = 0; i < 64; i++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            p.pool.push_back(e);
        }
    }
    fx_prof::rec r;
    r.id = id;
    r.a = p.pool.back(); p.pool.pop_back();
    r.b = p.pool.back(); p.pool.pop_back();
    (void)hipEventRecord(r.a, d->stream);
    p.pending.push_back(r);
}

void fx_prof_end(fx_domain d, int id)
{
    (void)id;
    if (!d->prof.on || d->prof.pending.empty() || d->prof.skipping) return;
    (void)hipEventRecord(d->prof.pending.back().b, d->stream);
    if (d->prof.pending.size() > 8192) (void)fx_prof_drain(d);
}

int fx_prof_drain(fx_domain d)
{
    fx_prof &p = d->prof;
    if (p.pending.empty()) return FX_OK;
    FX_HIP(hipStreamSynchronize(d->stream));
    for (auto &r : p.pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess && r.id >= 0 && r.id < FX_K_COUNT) {
            p.total_ms[r.id] += ms;
            p.count[r.id] += 1;
        }
        p.pool.push_back(r.a);
        p.pool.push_back(r.b);
    }
    p.pending.clear();
    return FX_OK;
}

fx_params fx_make_params(fx_domain d)
{
    fx_params P;
    memset(&P, 0, sizeof(P));
    const fx_domain_desc &D = d->desc;
    P.nrows = D.nrows;
    P.ncols = D.ncols;
    P.ld = d->mrows;
    P.nchem = D.nchem;
    P.ade = D.ade_solver ? 1 : 0;
    P.meteo = d->meteo;
    P.meteo_nrows = d->meteo_nrows;
    P.meteo_ncols = d->meteo_ncols;
    P.melt_mult = d->melt_mult;
    P.inflow = d->inflow;
    P.inflow_nrows = d->inflow_nrows;
    P.inflow_ncols = d->inflow_ncols;
    P.inflow_p = d->inflow ? d->inflow_irow + d->inflow_icol * (int64_t)d->mrows : -1;
    P.rates_nonneg = (d->meteo_nonneg && d->inflow_nonneg) ? 1 : 0;
    P.ch = d->ch;
    P.ach = d->ach;
    P.wintra = d->wintra ? 1 : 0;
    P.soil_rate = d->soil_rate;
    P.dxy = (double)D.dxy;
    P.arbase = (double)((int64_t)D.dxy * (int64_t)D.dxy);
    P.dxy_u = (uint32_t)D.dxy;
    P.hdry = D.hdry;
    P.ks0 = d->ks0;
    P.nodata = D.nodata;
    P.g = (double)D.gacc;
    P.cfl_dxy = D.cfl * (float)D.dxy;
    P.cvdef = (double)D.cvdef;
    P.nuem = (double)D.nuem;
    for (int k = 0; k < FX_MAX_CHEM; k++) P.D[k] = D.d_coef[k];
    if (D.precision == FX_FP32_MIXED) {
        P.d32[0] = d->s32.d[0];
        P.d32[1] = d->s32.d[1];
    }
    return P;
}

// FP32-mixed domains check the inflow cell's bed on the host (the FP64 kernels
// flag it as they go, common.cpp:324,377-380)
int fx_inflow_cell_ok(fx_domain d)
{
    if (d->desc.precision != FX_FP32_MIXED || !d->inflow || d->zb_host.empty()) return FX_OK;
    const size_t p = (size_t)d->inflow_irow + (size_t)d->inflow_icol * d->mrows;
    if (d->zb_host[p] == d->desc.nodata) return fx_fail(FX_ERR_ARG, "inflow ix,iy in NODATA_VALUE cell");
    return FX_OK;
}

namespace {

const int kBX = 64, kBY = 4;

int cfl_blocks(const fx_domain_desc &D)
{
    return ((D.nrows + kBX - 1) / kBX) * ((D.ncols + kBY - 1) / kBY);
}

int64_t fp64_bytes(const fx_domain_desc &D)
{
    const int64_t mm = (int64_t)(D.nrows + 2) * (D.ncols + 2);
    int64_t b = mm * 8 * 14 + mm * 3;
    b += mm * 8 * std::max(1, D.nchem);
    if (D.ade_solver) b += mm * 8 * std::max(1, D.nchem);    // pass-1 copies, one per solute
    return b;
}

int64_t fp32_bytes(const fx_domain_desc &D)
{
    const int64_t mm = (int64_t)(D.nrows + 2) * (D.ncols + 2);
    // d x2 (f64), (qx,qy) x2, zb (f64), f6 (6 per cell), ff (2 planes), tw x2, flags
    int64_t b = mm * 16 + mm * 4 * 4 + mm * 8 + mm * 4 * 6 + mm * 4 * 2 + mm * 4 * 2 + mm;
    if (D.ade_solver) b += mm * 8 * std::max(1, D.nchem);
    return b;
}

template <typename T>
int dalloc(fx_domain d, T **p, int64_t n)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)std::max<int64_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return fx_hip_check(e, "hipMalloc");
    e = hipMemsetAsync(q, 0, (size_t)std::max<int64_t>(n, 1) * sizeof(T), d->stream);
    if (e != hipSuccess) return fx_hip_check(e, "hipMemsetAsync");
    *p = (T *)q;
    return FX_OK;
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

int check_desc(const fx_domain_desc *D)
{
    if (!D) return fx_fail(FX_ERR_ARG, "desc is NULL");
    if (D->nrows < 3 || D->ncols < 3) return fx_fail(FX_ERR_ARG, "NROWS and NCOLS must be >= 3");
    if ((int64_t)(D->nrows + 2) * (D->ncols + 2) > ((int64_t)1 << 34))
        return fx_fail(FX_ERR_ARG, "domain too large");
    if (D->dxy <= 0) return fx_fail(FX_ERR_ARG, "CELLSIZE must be a positive integer");
    if (D->nchem < 0 || D->nchem > FX_MAX_CHEM) return fx_fail(FX_ERR_ARG, "nchem out of range");
    if (D->precision != FX_FP64_PARITY && D->precision != FX_FP32_MIXED)
        return fx_fail(FX_ERR_ARG, "unknown precision");
    if (D->precision == FX_FP32_MIXED && (int64_t)(D->nrows + 2) * (D->ncols + 2) >= FX32_MAX_CELLS)
        return fx_fail(FX_ERR_ARG, "FP32-mixed domains hold fewer than 2^29 cells (ghost rows included)");
    // the FP32 kernels take every depth above hdry as a normal FP32 number
    if (D->precision == FX_FP32_MIXED && !(D->hdry >= 1e-30))
        return fx_fail(FX_ERR_ARG, "FP32-mixed domains need ROUGNESS_HEIGHT (hdry) >= 1e-30 m");
    if (D->ade_mode < FX_ADE_OFF || D->ade_mode > FX_ADE_FAST) return fx_fail(FX_ERR_ARG, "unknown ade_mode");
    if (D->precision == FX_FP64_PARITY && D->ade_solver && D->ade_mode == FX_ADE_FAST)
        return fx_fail(FX_ERR_UNSUPPORTED, "FP64 parity domains run the exact ADE only");
    if (D->precision == FX_FP32_MIXED && D->ade_solver && D->ade_mode != FX_ADE_FAST)
        return fx_fail(FX_ERR_UNSUPPORTED, "FP32-mixed domains with the ADE solver run the fast ADE");
    return FX_OK;
}

int enter(fx_domain d)
{
    if (!d) return fx_fail(FX_ERR_ARG, "domain is NULL");
    FX_HIP(hipSetDevice(d->desc.device));
    return FX_OK;
}

// Entry points that change a domain's state, clock or tables.  A batch member
// belongs to its batch, whose slot holds the member's launch parameters and
// table addresses from admission on: it is refused until evicted.
int enter_alone(fx_domain d)
{
    int rc = enter(d);
    if (rc) return rc;
    if (d->in_batch) return fx_fail(FX_ERR_STATE, "domain is a batch member: evict it first");
    d->cfl_fresh = false;                   // the call may change the state
    return FX_OK;
}

// The captured loop iteration holds table addresses and launch parameters by
// value: drop it before anything it names is freed or changed.
void drop_graph(fx_domain d)
{
    if (d->graph_ok) (void)hipGraphExecDestroy(d->graph);
    d->graph_ok = false;
}

int pull_ctl(fx_domain d)
{
    FX_HIP(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(fx_ctl), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}

int push_ctl(fx_domain d)
{
    FX_HIP(hipMemcpyAsync(d->ctl, d->ctl_host, sizeof(fx_ctl), hipMemcpyHostToDevice, d->stream));
    return FX_OK;
}

bool ring_is_weir(const fx_domain_desc &D, int mrows, const float *weir)
{
    for (int icol = 1; icol <= D.ncols; icol++)
        if (weir[1 + (size_t)icol * mrows] != 1.0f || weir[D.nrows + (size_t)icol * mrows] != 1.0f) return false;
    for (int irow = 1; irow <= D.nrows; irow++)
        if (weir[irow + (size_t)1 * mrows] != 1.0f || weir[irow + (size_t)D.ncols * mrows] != 1.0f) return false;
    return true;
}

// the soil_mass grid, allocated (zeroed) on first need
int soil_alloc(fx_domain d)
{
    if (d->soil) return FX_OK;
    if (hipMalloc((void **)&d->soil, (size_t)d->ncell * 8) != hipSuccess) {
        d->soil = nullptr;
        return fx_fail(FX_ERR_NOMEM, "hipMalloc of the soil_mass grid failed");
    }
    d->bytes += d->ncell * 8;
    FX_HIP(hipMemsetAsync(d->soil, 0, (size_t)d->ncell * 8, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}

// double field pointer of an FP64 domain
double *f64_ptr(fx_domain d, int field)
{
    fx_state64 &s = d->s64;
    switch (field) {
    case FX_F_Z: return s.z;
    case FX_F_H: return s.h;
    case FX_F_QX: return s.qx;
    case FX_F_QY: return s.qy;
    case FX_F_FE1: return s.fe1;
    case FX_F_FE2: return s.fe2;
    case FX_F_FE3: return s.fe3;
    case FX_F_FN1: return s.fn1;
    case FX_F_FN2: return s.fn2;
    case FX_F_FN3: return s.fn3;
    case FX_F_TWET: return s.twet;
    case FX_F_H0: return s.h0;
    case FX_F_ZB: return s.zb;
    default:
        if (field >= FX_F_CONC0 && field < FX_F_CONC0 + d->desc.nchem) return s.conc[field - FX_F_CONC0];
        return nullptr;
    }
}

uint8_t *u8_ptr(fx_domain d, int field)
{
    if (field == FX_F_LDRY) return d->s64.ldry;
    if (field == FX_F_LDRY_PREV) return d->s64.ldry_prev;
    return nullptr;
}

int launch_step64(fx_domain d)
{
    int rc;
    const unsigned G = FXM_GATED;
    if ((rc = fx64_launch_cfl(d, G))) return rc;
    if ((rc = fx64_launch_melt(d, G | FXM_MELT | FXM_INFLOW | FXM_LOOP1))) return rc;
    if ((rc = fx64_launch_flux(d, G))) return rc;
    if ((rc = fx64_launch_update(d, G))) return rc;
    if (d->desc.ade_solver && d->desc.ade_mode != FX_ADE_OFF &&
        (rc = fx64_launch_ade_exact(d, G, 0, d->desc.nchem)))          // the solutes are independent
        return rc;
    if (d->wintra && (rc = fx64_launch_wintra(d, G, 0, d->desc.nchem))) return rc;
    return fx_launch_step_end(d);
}

// first: the call's first iteration (FP32: no fast ADE of this call has
// folded the CFL partials yet)
int launch_step(fx_domain d, bool first)
{
    if (d->desc.precision == FX_FP64_PARITY) return launch_step64(d);
    return fx32_launch_step(d, first);
}

} // namespace

extern "C" {

const char *fx_last_error(void) { return g_err.c_str(); }
const char *fx_version(void) { return "fluxos_hip 0.1 (gfx950)"; }

int64_t fx_domain_bytes(const fx_domain_desc *desc)
{
    if (check_desc(desc)) return -1;
    return desc->precision == FX_FP64_PARITY ? fp64_bytes(*desc) : fp32_bytes(*desc);
}

int fx_domain_create(const fx_domain_desc *desc, fx_domain *out)
{
    if (!out) return fx_fail(FX_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = check_desc(desc);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fx_fail(FX_ERR_NOGPU, "no HIP device available");
    if (desc->device < 0 || desc->device >= ndev) return fx_fail(FX_ERR_ARG, "device ordinal out of range");
    FX_HIP(hipSetDevice(desc->device));
    fx_domain d = new fx_domain_s();
    d->desc = *desc;
    if (d->desc.nchem < 1) d->desc.nchem = 1;   // conc_SW[0] always exists (initiate.cpp:127-130)
    d->mrows = desc->nrows + 2;
    d->mcols = desc->ncols + 2;
    d->ncell = (int64_t)d->mrows * d->mcols;
    d->meteo_nonneg = true;
    d->inflow_nonneg = true;
    d->graph_ok = false;
    d->ch = FX32_CH;
    d->ach = FX32_ACH;
    d->in_batch = false;
    d->cfl_fresh = false;
    auto bail = [&](int code) { fx_domain_destroy(d); return code; };
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fx_fail(FX_ERR_HIP, "hipStreamCreate failed"));
    if ((rc = dalloc(d, &d->ctl, 1))) return bail(rc);
    if (hipHostMalloc((void **)&d->ctl_host, sizeof(fx_ctl), hipHostMallocDefault) != hipSuccess)
        return bail(fx_fail(FX_ERR_HIP, "hipHostMalloc failed"));
    memset(d->ctl_host, 0, sizeof(fx_ctl));
    d->npartials = cfl_blocks(d->desc);
    if ((rc = dalloc(d, &d->partials, 2 * (int64_t)d->npartials))) return bail(rc);
    const int64_t mm = d->ncell;
    if (desc->precision == FX_FP64_PARITY) {
        fx_state64 &s = d->s64;
        double **dbl[] = {&s.z, &s.zb, &s.h, &s.qx, &s.qy, &s.ks, &s.fe1, &s.fe2, &s.fe3,
                          &s.fn1, &s.fn2, &s.fn3, &s.twet, &s.h0};
        for (double **p : dbl)
            if ((rc = dalloc(d, p, mm))) return bail(rc);
        if ((rc = dalloc(d, &s.ldry, mm))) return bail(rc);
        if ((rc = dalloc(d, &s.ldry_prev, mm))) return bail(rc);
        if ((rc = dalloc(d, &s.weir, mm))) return bail(rc);
        for (int k = 0; k < d->desc.nchem; k++)
            if ((rc = dalloc(d, &s.conc[k], mm))) return bail(rc);
        if (d->desc.ade_solver) {
            // pass-1 copies and wavefront buffers per solute: the solutes of a
            // step sweep concurrently
            if ((rc = dalloc(d, &s.conc_tmp, mm * d->desc.nchem))) return bail(rc);
            d->ade_blocks = (d->desc.ncols + 63) / 64;
            const int64_t nb = (int64_t)d->ade_blocks * d->desc.nchem;
            // edge words (north-face flux, final value) of every block's last column
            if ((rc = dalloc(d, &d->ade_edges, 2 * nb * d->mrows))) return bail(rc);
        }
        d->bytes = fp64_bytes(d->desc);
    } else {
        if ((rc = fx32_init_state(d))) return bail(rc);
        d->bytes = fp32_bytes(d->desc);
    }
    if (hipStreamSynchronize(d->stream) != hipSuccess) return bail(fx_fail(FX_ERR_HIP, "sync failed"));
    *out = d;
    return FX_OK;
}

int fx_domain_destroy(fx_domain d)
{
    if (d && d->in_batch) return fx_fail(FX_ERR_STATE, "domain is a batch member: evict it first");
    if (!d) return FX_OK;
    (void)hipSetDevice(d->desc.device);
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    if (d->graph_ok) (void)hipGraphExecDestroy(d->graph);
    fx_state64 &s = d->s64;
    void *ptrs[] = {s.z, s.zb, s.h, s.qx, s.qy, s.ks, s.fe1, s.fe2, s.fe3, s.fn1, s.fn2, s.fn3,
                    s.twet, s.h0, s.ldry, s.ldry_prev, s.weir, s.conc_tmp};
    for (void *p : ptrs) dfree(p);
    for (int k = 0; k < FX_MAX_CHEM; k++) dfree(s.conc[k]);
    fx_state32 &t = d->s32;
    void *ptrs32[] = {t.d[0], t.d[1], t.q[0], t.q[1], t.zb, t.f6, t.ff,
                      t.flags, t.tw, t.partials, t.part2, t.ticket, t.rec};
    for (void *p : ptrs32) dfree(p);
    for (int k = 0; k < FX_MAX_CHEM; k++) { dfree(t.mass[0][k]); dfree(t.mass[1][k]); }
    for (auto &r : d->prof.pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : d->prof.pool) (void)hipEventDestroy(e);
    dfree(d->soil);
    dfree(d->ctl);
    dfree(d->partials);
    dfree(d->ade_edges);
    dfree(d->meteo);
    dfree(d->melt_mult);
    dfree(d->inflow);
    if (d->ctl_host) (void)hipHostFree(d->ctl_host);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
    return FX_OK;
}

int fx_set_static(fx_domain d, const double *zb, const double *ks, const float *weir)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if (!zb || !ks || !weir) return fx_fail(FX_ERR_ARG, "zb, ks and weir are required");
    const int64_t mm = d->ncell;
    for (int64_t i = 0; i < mm; i++)
        if (weir[i] != 0.0f && weir[i] != 1.0f) return fx_fail(FX_ERR_ARG, "weir values must be 0 or 1");
    if (!ring_is_weir(d->desc, d->mrows, weir))
        return fx_fail(FX_ERR_UNSUPPORTED,
                       "the outer interior ring must be NODATA (weir) cells: the reference's weir "
                       "boundary reads an uninitialised volrat otherwise");
    std::vector<uint8_t> w(mm);
    for (int64_t i = 0; i < mm; i++) w[i] = weir[i] == 1.0f ? 1 : 0;
    if (d->desc.precision == FX_FP64_PARITY) {
        FX_HIP(hipMemcpyAsync(d->s64.zb, zb, mm * 8, hipMemcpyHostToDevice, d->stream));
        FX_HIP(hipMemcpyAsync(d->s64.ks, ks, mm * 8, hipMemcpyHostToDevice, d->stream));
        FX_HIP(hipMemcpyAsync(d->s64.weir, w.data(), mm, hipMemcpyHostToDevice, d->stream));
        FX_HIP(hipStreamSynchronize(d->stream));
    } else {
        // ks enters the FP32 path as the interior roughness: it must be uniform
        const double ks0 = ks[1 + (size_t)1 * d->mrows];
        d->ks0 = ks0;
        for (int icol = 1; icol <= d->desc.ncols; icol++)
            for (int irow = 1; irow <= d->desc.nrows; irow++)
                if (ks[irow + (size_t)icol * d->mrows] != ks0)
                    return fx_fail(FX_ERR_UNSUPPORTED, "FP32-mixed domains need a uniform ks");
        d->zb_host.assign(zb, zb + mm);
        if ((rc = fx32_set_field(d, FX_F_ZB, zb))) return rc;
        if ((rc = fx32_set_field(d, -1 /* weir */, w.data()))) return rc;
    }
    d->has_static = true;
    d->melt_area_ok = false;
    return FX_OK;
}

int fx_set_field(fx_domain d, int32_t field, const void *host)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if (!host) return fx_fail(FX_ERR_ARG, "host buffer is NULL");
    const int64_t mm = d->ncell;
    if (field == FX_F_US) {
        const double *u = (const double *)host;
        for (int64_t i = 0; i < mm; i++)
            if (u[i] != 0.0) return fx_fail(FX_ERR_UNSUPPORTED, "us is identically zero in the reference");
        return FX_OK;
    }
    if (field == FX_F_ZB) return fx_fail(FX_ERR_ARG, "zb is set with fx_set_static");
    if (field == FX_F_SOIL) {
        if ((rc = soil_alloc(d))) return rc;
        FX_HIP(hipMemcpyAsync(d->soil, host, mm * 8, hipMemcpyHostToDevice, d->stream));
        FX_HIP(hipStreamSynchronize(d->stream));
        return FX_OK;
    }
    if (d->desc.precision == FX_FP32_MIXED) return fx32_set_field(d, field, host);
    if (double *p = f64_ptr(d, field)) {
        FX_HIP(hipMemcpyAsync(p, host, mm * 8, hipMemcpyHostToDevice, d->stream));
        FX_HIP(hipStreamSynchronize(d->stream));
        return FX_OK;
    }
    if (uint8_t *p = u8_ptr(d, field)) {
        const float *f = (const float *)host;
        std::vector<uint8_t> w(mm);
        for (int64_t i = 0; i < mm; i++) {
            if (f[i] != 0.0f && f[i] != 1.0f) return fx_fail(FX_ERR_ARG, "flag values must be 0 or 1");
            w[i] = f[i] == 1.0f ? 1 : 0;
        }
        FX_HIP(hipMemcpyAsync(p, w.data(), mm, hipMemcpyHostToDevice, d->stream));
        FX_HIP(hipStreamSynchronize(d->stream));
        return FX_OK;
    }
    return fx_fail(FX_ERR_ARG, "unknown field");
}

int fx_get_field(fx_domain d, int32_t field, void *host)
{
    int rc = enter(d);
    if (rc) return rc;
    if (!host) return fx_fail(FX_ERR_ARG, "host buffer is NULL");
    const int64_t mm = d->ncell;
    if (field == FX_F_US) {
        memset(host, 0, mm * 8);
        return FX_OK;
    }
    if (field == FX_F_SOIL) {
        if (!d->soil) { memset(host, 0, mm * 8); return FX_OK; }
        FX_HIP(hipMemcpyAsync(host, d->soil, mm * 8, hipMemcpyDeviceToHost, d->stream));
        FX_HIP(hipStreamSynchronize(d->stream));
        return FX_OK;
    }
    if (d->desc.precision == FX_FP32_MIXED) return fx32_get_field(d, field, host);
    if (double *p = f64_ptr(d, field)) {
        FX_HIP(hipMemcpyAsync(host, p, mm * 8, hipMemcpyDeviceToHost, d->stream));
        FX_HIP(hipStreamSynchronize(d->stream));
        return FX_OK;
    }
    if (uint8_t *p = u8_ptr(d, field)) {
        std::vector<uint8_t> w(mm);
        FX_HIP(hipMemcpyAsync(w.data(), p, mm, hipMemcpyDeviceToHost, d->stream));
        FX_HIP(hipStreamSynchronize(d->stream));
        float *f = (float *)host;
        for (int64_t i = 0; i < mm; i++) f[i] = (float)w[i];
        return FX_OK;
    }
    return fx_fail(FX_ERR_ARG, "unknown field");
}

int fx_set_wintra(fx_domain d, int32_t status, double soil_release_rate)
{
    int rc = enter(d);
    if (rc) return rc;
    d->cfl_fresh = false;
    if (d->in_batch) return fx_fail(FX_ERR_STATE, "domain is a batch member");
    if (!std::isfinite(soil_release_rate)) return fx_fail(FX_ERR_ARG, "soil_release_rate must be finite");
    if (status) {
        if (!d->desc.ade_solver)
            return fx_fail(FX_ERR_STATE, "WINTRA needs the ADE solver (the reference turns it off without)");
        if ((rc = soil_alloc(d))) return rc;
    }
    d->wintra = status != 0;
    d->soil_rate = soil_release_rate;
    drop_graph(d);
    return FX_OK;
}

int fx_wintrasolver_calc(fx_domain d, int32_t ichem)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if (d->desc.precision != FX_FP64_PARITY)
        return fx_fail(FX_ERR_UNSUPPORTED, "granular calls need an FP64 parity domain");
    if (!d->wintra) return fx_fail(FX_ERR_STATE, "WINTRA is off for this domain (fx_set_wintra)");
    if (ichem < 0 || ichem >= d->desc.nchem) return fx_fail(FX_ERR_ARG, "ichem out of range");
    if ((rc = fx64_launch_wintra(d, 0, ichem, 1))) return rc;
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}

int fx_set_clock(fx_domain d, double tim, double dtfl, int32_t it)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if ((rc = pull_ctl(d))) return rc;
    d->ctl_host->tim = tim;
    d->ctl_host->dtfl = dtfl;
    d->ctl_host->it = it;
    if ((rc = push_ctl(d))) return rc;
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}

int fx_get_clock(fx_domain d, double *tim, double *dtfl, int32_t *it)
{
    int rc = enter(d);
    if (rc) return rc;
    if ((rc = pull_ctl(d))) return rc;
    if (tim) *tim = d->ctl_host->tim;
    if (dtfl) *dtfl = d->ctl_host->dtfl;
    if (it) *it = d->ctl_host->it;
    return FX_OK;
}

int fx_set_meteo(fx_domain d, int32_t nrows, int32_t ncols, const float *table, const float *melt_mult)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    const bool has_table = table && nrows > 0;
    if (has_table && ncols < 2) return fx_fail(FX_ERR_ARG, "meteo table needs at least 2 columns");
    if (has_table && nrows > 1 << 20) return fx_fail(FX_ERR_ARG, "meteo table too long");
    FX_HIP(hipStreamSynchronize(d->stream));
    drop_graph(d);
    dfree(d->meteo);
    d->meteo = nullptr;
    d->meteo_nrows = 0;
    dfree(d->melt_mult);
    d->melt_mult = nullptr;
    d->melt_area_ok = false;
    d->meteo_nonneg = true;
    if (has_table) {
        for (int32_t r = 0; r < nrows; r++)
            if (!(table[(size_t)r * ncols + 1] >= 0.0f)) d->meteo_nonneg = false;
        FX_HIP(hipMalloc((void **)&d->meteo, (size_t)nrows * ncols * sizeof(float)));
        FX_HIP(hipMemcpy(d->meteo, table, (size_t)nrows * ncols * sizeof(float), hipMemcpyHostToDevice));
        d->meteo_nrows = nrows;
        d->meteo_ncols = ncols;
    }
    if (melt_mult) {
        for (int64_t i = 0; i < d->ncell; i++)
            if (!(melt_mult[i] >= 0.0f)) { d->meteo_nonneg = false; break; }
        FX_HIP(hipMalloc((void **)&d->melt_mult, d->ncell * sizeof(float)));
        FX_HIP(hipMemcpy(d->melt_mult, melt_mult, d->ncell * sizeof(float), hipMemcpyHostToDevice));
    }
    return FX_OK;
}

int fx_set_inflow(fx_domain d, int32_t nrows, int32_t ncols, const float *table, int64_t irow, int64_t icol)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    const bool has_table = table && nrows > 0;
    if (has_table) {
        if (ncols < 2) return fx_fail(FX_ERR_ARG, "inflow table needs at least 2 columns");
        if (nrows > 1 << 20) return fx_fail(FX_ERR_ARG, "inflow table too long");
        if (irow < 0 || irow >= d->mrows || icol < 0 || icol >= d->mcols)
            return fx_fail(FX_ERR_ARG, "inflow cell outside the grid");
        if (irow < 1 || irow > d->desc.nrows || icol < 1 || icol > d->desc.ncols)
            return fx_fail(FX_ERR_UNSUPPORTED, "inflow cell on the ghost ring: no interior cell receives it");
    }
    FX_HIP(hipStreamSynchronize(d->stream));
    drop_graph(d);
    dfree(d->inflow);
    d->inflow = nullptr;
    d->inflow_nrows = 0;
    d->inflow_nonneg = true;
    if (has_table) {
        for (int32_t r = 0; r < nrows; r++)
            if (!(table[(size_t)r * ncols + 1] >= 0.0f)) d->inflow_nonneg = false;
        FX_HIP(hipMalloc((void **)&d->inflow, (size_t)nrows * ncols * sizeof(float)));
        FX_HIP(hipMemcpy(d->inflow, table, (size_t)nrows * ncols * sizeof(float), hipMemcpyHostToDevice));
        d->inflow_nrows = nrows;
        d->inflow_ncols = ncols;
        d->inflow_irow = irow;
        d->inflow_icol = icol;
    }
    return FX_OK;
}

int fx_courant_condition(fx_domain d, double print_next, double *dtfl, double *hpall)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if (d->desc.precision != FX_FP64_PARITY)
        return fx_fail(FX_ERR_UNSUPPORTED, "granular calls need an FP64 parity domain");
    if ((rc = pull_ctl(d))) return rc;
    d->ctl_host->print_next = print_next;
    if ((rc = push_ctl(d))) return rc;
    if ((rc = fx64_launch_cfl(d, 0))) return rc;
    if ((rc = pull_ctl(d))) return rc;
    if (dtfl) *dtfl = d->ctl_host->dtfl;
    if (hpall) *hpall = d->ctl_host->hpall;
    return FX_OK;
}

int fx_add_meteo(fx_domain d)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if (d->desc.precision != FX_FP64_PARITY)
        return fx_fail(FX_ERR_UNSUPPORTED, "granular calls need an FP64 parity domain");
    if ((rc = fx_launch_forcing(d))) return rc;
    if ((rc = fx64_launch_melt(d, FXM_MELT))) return rc;
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}

int fx_add_inflow(fx_domain d)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if (d->desc.precision != FX_FP64_PARITY)
        return fx_fail(FX_ERR_UNSUPPORTED, "granular calls need an FP64 parity domain");
    if ((rc = fx_launch_forcing(d))) return rc;
    if ((rc = fx64_launch_melt(d, FXM_INFLOW))) return rc;
    if ((rc = pull_ctl(d))) return rc;
    if (d->ctl_host->status & 8) {
        d->ctl_host->status &= ~8;
        if ((rc = push_ctl(d))) return rc;
        FX_HIP(hipStreamSynchronize(d->stream));
        return fx_fail(FX_ERR_ARG, "inflow ix,iy in NODATA_VALUE cell");
    }
    return FX_OK;
}

int fx_hydrodynamics_calc(fx_domain d)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if (d->desc.precision != FX_FP64_PARITY)
        return fx_fail(FX_ERR_UNSUPPORTED, "granular calls need an FP64 parity domain");
    if (!d->has_static) return fx_fail(FX_ERR_STATE, "fx_set_static has not been called");
    if ((rc = fx64_launch_melt(d, FXM_LOOP1))) return rc;
    if ((rc = fx64_launch_flux(d, 0))) return rc;
    if ((rc = fx64_launch_update(d, 0))) return rc;
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}

int fx_adesolver_calc(fx_domain d, int32_t it, int32_t ichem)
{
    int rc = enter_alone(d);
    if (rc) return rc;
    if (d->desc.precision != FX_FP64_PARITY)
        return fx_fail(FX_ERR_UNSUPPORTED, "granular calls need an FP64 parity domain");
    if (!d->desc.ade_solver) return fx_fail(FX_ERR_STATE, "the ADE solver is off for this domain");
    if (ichem < 0 || ichem >= d->desc.nchem) return fx_fail(FX_ERR_ARG, "ichem out of range");
    if (it <= 1) return FX_OK;                      // ADEsolver_calc.cpp:45,74-77
    if ((rc = fx64_launch_ade_exact(d, 0, ichem, 1))) return rc;
    if ((rc = pull_ctl(d))) return rc;
    if (d->ctl_host->ade_error) return fx_fail(FX_ERR_HIP, "ADE wavefront wait expired");
    return FX_OK;
}

int fx_advance(fx_domain d, double t_end, double print_next, int32_t max_steps, fx_step_info *info)
{
    // FP32: when nothing changed the state since the last call's last sweep,
    // that sweep's CFL partials are the current state's, and the first
    // iteration's clock folds them as any later iteration's does
    const bool fresh = d && d->cfl_fresh;
    int rc = enter_alone(d);
    if (rc) return rc;
    if (!d->has_static) return fx_fail(FX_ERR_STATE, "fx_set_static has not been called");
    if (max_steps <= 0) return fx_fail(FX_ERR_ARG, "max_steps must be positive");
    if ((rc = pull_ctl(d))) return rc;
    fx_ctl *c = d->ctl_host;
    fx_step_info loc{};
    if (!info) info = &loc;
    memset(info, 0, sizeof(*info));
    info->print_next = print_next;
    if (c->tim > t_end) {
        info->tim = c->tim; info->dtfl = c->dtfl; info->hpall = c->hpall; info->it = c->it;
        info->status = 2;
        return FX_OK;
    }
    if ((rc = fx_inflow_cell_ok(d))) return rc;
    if (d->desc.precision == FX_FP32_MIXED && !fresh) {
        if ((rc = fx32_launch_cfl_initial(d))) return rc;
        if ((rc = pull_ctl(d))) return rc;
    }
    c->print_next = print_next;
    c->t_end = t_end;
    c->halt = 0;
    c->pending_halt = 0;
    c->steps = 0;
    c->solver_steps = 0;
    c->status &= ~4;
    c->ade_error = 0;
    if ((rc = push_ctl(d))) return rc;
    const bool fp32 = d->desc.precision == FX_FP32_MIXED;
    int32_t done = 0;
    double dt_est = c->dtfl;
    double tim0 = c->tim;
    while (done < max_steps) {
        int n = 8;
        if (dt_est > 0.0 && std::isfinite(dt_est)) {
            double need = (print_next - tim0) / dt_est;
            n = (int)std::min(256.0, std::max(1.0, std::ceil(need) + 1.0));
        }
        n = std::min(n, max_steps - done);
        for (int i = 0; i < n; i++)
            if ((rc = launch_step(d, done == 0 && i == 0))) return rc;
        if ((rc = pull_ctl(d))) return rc;
        done = c->steps;
        // FP32: the last step's flip and halt are the next clock's, so a step
        // that reached its print time ends the loop by pending_halt
        if (c->halt || (fp32 && c->pending_halt) || (c->status & 12) || c->ade_error) break;
        dt_est = c->dtfl;
        tim0 = c->tim;
    }
    if (fp32) {                                    // the last step's flip and halt
        if ((rc = fx32_launch_step_close(d))) return rc;
        if ((rc = pull_ctl(d))) return rc;
    }
    info->tim = c->tim;
    info->dtfl = c->dtfl;
    info->hpall = c->hpall;
    info->it = c->it;
    info->steps = c->steps;
    info->solver_steps = c->solver_steps;
    info->status = (c->tim >= print_next ? 1 : 0) | (c->tim > t_end ? 2 : 0) | (c->status & 4);
    if (c->ade_error) return fx_fail(FX_ERR_HIP, "ADE wavefront wait expired");
    if (c->status & 8) {
        c->status &= ~8;
        if ((rc = push_ctl(d))) return rc;
        FX_HIP(hipStreamSynchronize(d->stream));
        return fx_fail(FX_ERR_ARG, "inflow ix,iy in NODATA_VALUE cell");
    }
    if (c->status & 4) return fx_fail(FX_ERR_NONFINITE, "non-finite value in the state");
    d->cfl_fresh = d->desc.precision == FX_FP32_MIXED;
    return FX_OK;
}

int fx_set_tiling(fx_domain d, int32_t ch, int32_t ach)
{
    int rc = enter(d);
    if (rc) return rc;
    d->cfl_fresh = false;                   // the partials follow the tiles
    if (d->desc.precision != FX_FP32_MIXED) return fx_fail(FX_ERR_UNSUPPORTED, "tiling applies to FP32-mixed domains");
    if (d->in_batch) return fx_fail(FX_ERR_STATE, "domain is a batch member: its tiling is fixed");
    if (ch < FX32_CH_MIN || ch > 4096 || ach < FX32_CH_MIN || ach > 4096)
        return fx_fail(FX_ERR_ARG, "column-block lengths must be 8..4096");
    return fx32_set_tiling(d, ch, ach);
}

int fx_get_tiling(fx_domain d, int32_t *ch, int32_t *ach)
{
    if (!d) return fx_fail(FX_ERR_ARG, "domain is NULL");
    if (ch) *ch = d->ch;
    if (ach) *ach = d->ach;
    return FX_OK;
}

int fx_mass_balance(fx_domain d, double out[3])
{
    int rc = enter(d);
    if (rc) return rc;
    if (!out) return fx_fail(FX_ERR_ARG, "out is NULL");
    double vol = 0.0;
    if (d->desc.precision == FX_FP64_PARITY) rc = fx64_launch_volume(d, &vol);
    else rc = fx32_launch_volume(d, &vol);
    if (rc) return rc;
    if ((rc = pull_ctl(d))) return rc;
    double pend = 0.0;
    if (d->desc.precision == FX_FP32_MIXED && (rc = fx32_pending_weir(d, &pend))) return rc;
    out[0] = vol;
    out[1] = d->ctl_host->led_melt;
    out[2] = d->ctl_host->led_weir + pend;
    return FX_OK;
}

int fx_solute_balance(fx_domain d, int32_t ichem, double out[2])
{
    int rc = enter(d);
    if (rc) return rc;
    if (!out) return fx_fail(FX_ERR_ARG, "out is NULL");
    if (!d->desc.ade_solver) return fx_fail(FX_ERR_STATE, "the ADE solver is off for this domain");
    if (ichem < 0 || ichem >= d->desc.nchem) return fx_fail(FX_ERR_ARG, "ichem out of range");
    double mass = 0.0;
    if (d->desc.precision == FX_FP64_PARITY) {
        if ((rc = fx64_launch_solute(d, ichem, &mass))) return rc;
        out[0] = mass;
        out[1] = 0.0;
        return FX_OK;
    }
    if ((rc = fx32_launch_solute(d, ichem, &mass))) return rc;
    double area = 0.0;
    if ((rc = fx32_pending(d, nullptr, &area))) return rc;
    if ((rc = pull_ctl(d))) return rc;
    const fx_ctl *c = d->ctl_host;
    double in = c->sled[ichem];
    if (d->meteo) in += c->meteoi * area * (double)d->desc.dxy * d->desc.dxy * c->cmelt[ichem];
    out[0] = mass;
    out[1] = in;
    return FX_OK;
}

int fx_profile(fx_domain d, int32_t on)
{
    int rc = enter(d);
    if (rc) return rc;
    if ((rc = fx_prof_drain(d))) return rc;
    for (int i = 0; i < FX_K_COUNT; i++) { d->prof.total_ms[i] = 0.0; d->prof.count[i] = 0; d->prof.seq[i] = 0; }
    d->prof.on = on != 0;
    d->prof.stride = on > 1 ? on : 1;
    d->prof.skipping = false;
    return FX_OK;
}

int fx_profile_read(fx_domain d, int32_t kernel, double *total_ms, int64_t *launches)
{
    int rc = enter(d);
    if (rc) return rc;
    if (kernel < 0 || kernel >= FX_K_COUNT) return fx_fail(FX_ERR_ARG, "kernel id out of range");
    if ((rc = fx_prof_drain(d))) return rc;
    if (total_ms) *total_ms = d->prof.total_ms[kernel];
    if (launches) *launches = d->prof.count[kernel];
    return FX_OK;
}

int fx_synchronize(fx_domain d)
{
    int rc = enter(d);
    if (rc) return rc;
    FX_HIP(hipStreamSynchronize(d->stream));
    return FX_OK;
}

} // extern "C"
