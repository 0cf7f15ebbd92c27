// fx_batch.hip -- batches of FP32-mixed domains advanced together (host code).
//
// A batch holds up to max_members domains ("members").  One batched step
// runs one iteration of the reference's loop (main.cpp:271-406) for every
// member, each with its own CFL step, clock, forcing and print time, in three
// launches that cover all members at once:
//   k32b_clock   one block per member: the previous step's buffer flip and
//                print/end halt, then fold its CFL partials, dt, forcing;
//   k32b_sweep   every member's step-sweep blocks, packed in one grid;
//   k32b_ade     every member's ADE blocks (if the batch carries solutes);
// and after a run's last step
//   k32b_end     that step's flip and halt and a compact status per member.
// The per-block member map (tmap/amap) is rebuilt when members join or
// leave; a run of several steps is captured once as a hipGraph and replayed
// (environment FLUXOS_BATCH_GRAPH=0 launches every step directly instead,
// for tools that do not follow graph launches).
// A member whose clock reaches its print time (or passes t_end) halts: its
// blocks return at once until the host resumes or evicts it, so members
// finish independently while the rest keep stepping.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fx_kern32.h"

namespace {

constexpr int GRAPH_STEPS = 8;     // steps per captured graph

} // namespace

struct fx_batch_s {
    int device = 0;
    int nchem = 0;
    int ade = 0;
    int maxm = 0;
    hipStream_t stream = nullptr;
    fx_bslot *slots_dev = nullptr;
    std::vector<fx_bslot> slots_host;
    std::vector<fx_domain> member;     // nullptr = free slot
    std::vector<char> reported;        // halt already reported by poll
    std::vector<double> t_end;
    std::vector<int> nab;              // ADE blocks of each member
    fx_bstat *stat_dev = nullptr;
    fx_bstat *stat_host = nullptr;     // pinned
    uint32_t *tmap = nullptr, *amap = nullptr;
    int64_t tmap_cap = 0, amap_cap = 0;
    int ntiles = 0, natiles = 0, nslot = 0;
    fx_ctl *idle_ctl = nullptr;        // halted control block of free slots
    bool dirty = true;
    bool use_graph = true;             // FLUXOS_BATCH_GRAPH=0: direct launches
    hipGraphExec_t graph = nullptr;
    bool graph_ok = false;
    int64_t steps_run = 0;
};

namespace {

int set_dev(fx_batch b)
{
    FX_HIP(hipSetDevice(b->device));
    return FX_OK;
}

void drop_graph(fx_batch b)
{
    if (b->graph_ok) (void)hipGraphExecDestroy(b->graph);
    b->graph_ok = false;
}

int grow(uint32_t **p, int64_t *cap, int64_t need)
{
    if (need <= *cap) return FX_OK;
    if (*p) FX_HIP(hipFree(*p));
    *p = nullptr;
    const int64_t n = std::max<int64_t>(need, 2 * *cap);
    FX_HIP(hipMalloc((void **)p, (size_t)n * sizeof(uint32_t)));
    *cap = n;
    return FX_OK;
}

// The fast ADE's tile width of each member for this membership (the ADE's
// results do not depend on it): the member's own while the batch's ADE blocks
// fill the GPU's wave slots, halved (down to 16 columns) while they are fewer
// than ADE_FILL.  That is the case when few members are left, whose waves are
// then too few for the GPU and walk long column runs one after the other.
constexpr int ADE_FILL = 4096;

void ade_widths(fx_batch b)
{
    std::vector<int> w(b->maxm, 0);
    auto blocks = [&](int s) {
        const fx_bslot &B = b->slots_host[s];
        return (int64_t)B.agx * ((B.P.ncols + w[s] - 1) / w[s]);
    };
    int64_t tot = 0;
    for (int s = 0; s < b->maxm; s++)
        if (b->member[s]) {
            w[s] = fx_make_params(b->member[s]).ach;
            tot += blocks(s);
        }
    for (bool narrower = true; tot < ADE_FILL && narrower;) {
        narrower = false;
        tot = 0;
        for (int s = 0; s < b->maxm; s++) {
            if (!b->member[s]) continue;
            if (w[s] > 16) {
                w[s] = std::max(16, w[s] / 2);
                narrower = true;
            }
            tot += blocks(s);
        }
    }
    for (int s = 0; s < b->maxm; s++)
        if (b->member[s]) {
            b->slots_host[s].P.ach = w[s];
            b->nab[s] = (int)blocks(s);
        }
}

// rebuild the slot table and the block -> member maps after a membership change
int rebuild(fx_batch b)
{
    std::vector<uint32_t> t, a;
    b->nslot = 0;
    ade_widths(b);
    for (int s = 0; s < b->maxm; s++) {
        fx_domain d = b->member[s];
        if (!d) continue;
        b->nslot = s + 1;
        const fx_bslot &B = b->slots_host[s];
        for (int k = 0; k < B.npart; k++) t.push_back(FX_BMAP(s, k));
        if (B.ade_on)
            for (int k = 0; k < b->nab[s]; k++) a.push_back(FX_BMAP(s, k));
    }
    // workgroup w runs entry fx_xcd_tile(w) of the member-ordered list (runs of
    // tiles dealt to the XCDs)
    if (const int chunk = fx32_xcd_chunk_batch()) {
        auto permute = [chunk](std::vector<uint32_t> &v) {
            std::vector<uint32_t> o(v.size());
            const int n = (int)v.size();
            for (int w = 0; w < n; w++) o[w] = v[fx_xcd_tile(w, n, chunk)];
            v.swap(o);
        };
        permute(t);
        permute(a);
    }
    int rc;
    if ((rc = grow(&b->tmap, &b->tmap_cap, std::max<int64_t>(1, (int64_t)t.size())))) return rc;
    if ((rc = grow(&b->amap, &b->amap_cap, std::max<int64_t>(1, (int64_t)a.size())))) return rc;
    if (!t.empty())
        FX_HIP(hipMemcpyAsync(b->tmap, t.data(), t.size() * 4, hipMemcpyHostToDevice, b->stream));
    if (!a.empty())
        FX_HIP(hipMemcpyAsync(b->amap, a.data(), a.size() * 4, hipMemcpyHostToDevice, b->stream));
    FX_HIP(hipMemcpyAsync(b->slots_dev, b->slots_host.data(), sizeof(fx_bslot) * b->maxm,
                          hipMemcpyHostToDevice, b->stream));
    FX_HIP(hipStreamSynchronize(b->stream));
    b->ntiles = (int)t.size();
    b->natiles = (int)a.size();
    drop_graph(b);
    b->dirty = false;
    return FX_OK;
}

int launch_one(fx_batch b)
{
    int rc;
    if ((rc = fx32b_launch_clock(b->stream, b->slots_dev, b->nslot))) return rc;
    if ((rc = fx32b_launch_sweep(b->stream, b->slots_dev, b->tmap, b->ntiles))) return rc;
    if (b->ade && b->natiles > 0 &&
        (rc = fx32b_launch_ade(b->stream, b->slots_dev, b->amap, b->natiles, b->nchem)))
        return rc;
    return FX_OK;
}

int capture(fx_batch b)
{
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return FX_ERR_HIP;
    int rc = FX_OK;
    for (int i = 0; i < GRAPH_STEPS && rc == FX_OK; i++) rc = launch_one(b);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != FX_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return FX_ERR_HIP;
    }
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return FX_ERR_HIP;
    }
    b->graph_ok = true;
    return FX_OK;
}

} // namespace

extern "C" {

int fx_batch_create(int32_t device, int32_t nchem, int32_t ade_solver, int32_t max_members, fx_batch *out)
{
    if (!out) return fx_fail(FX_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (max_members < 1 || max_members > 4095) return fx_fail(FX_ERR_ARG, "max_members must be 1..4095");
    if (nchem < 0 || nchem > FX_MAX_CHEM) return fx_fail(FX_ERR_ARG, "nchem out of range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fx_fail(FX_ERR_NOGPU, "no HIP device available");
    if (device < 0 || device >= ndev) return fx_fail(FX_ERR_ARG, "device ordinal out of range");
    FX_HIP(hipSetDevice(device));
    fx_batch b = new fx_batch_s();
    b->device = device;
    b->nchem = std::max(1, (int)nchem);
    b->ade = ade_solver ? 1 : 0;
    b->maxm = max_members;
    {
        const char *g = std::getenv("FLUXOS_BATCH_GRAPH");
        b->use_graph = !(g && g[0] == '0');
    }
    b->slots_host.assign(max_members, fx_bslot{});
    b->member.assign(max_members, nullptr);
    b->reported.assign(max_members, 0);
    b->t_end.assign(max_members, 0.0);
    b->nab.assign(max_members, 0);
    auto bail = [&](hipError_t e, const char *what) {
        int rc = fx_hip_check(e, what);
        fx_batch_destroy(b);
        return rc;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "stream");
    if ((e = hipMalloc((void **)&b->slots_dev, sizeof(fx_bslot) * max_members)) != hipSuccess) return bail(e, "slots");
    if ((e = hipMalloc((void **)&b->stat_dev, sizeof(fx_bstat) * max_members)) != hipSuccess) return bail(e, "stat");
    if ((e = hipHostMalloc((void **)&b->stat_host, sizeof(fx_bstat) * max_members, hipHostMallocDefault)) != hipSuccess)
        return bail(e, "stat host");
    if ((e = hipMalloc((void **)&b->idle_ctl, sizeof(fx_ctl))) != hipSuccess) return bail(e, "idle ctl");
    fx_ctl idle{};
    idle.halt = 1;
    if ((e = hipMemcpy(b->idle_ctl, &idle, sizeof(fx_ctl), hipMemcpyHostToDevice)) != hipSuccess)
        return bail(e, "idle ctl");
    for (auto &s : b->slots_host) s.ctl = b->idle_ctl;
    memset(b->stat_host, 0, sizeof(fx_bstat) * max_members);
    *out = b;
    return FX_OK;
}

int fx_batch_destroy(fx_batch b)
{
    if (!b) return FX_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (fx_domain m : b->member)          // members still in the batch leave it with their state
        if (m) m->in_batch = false;
    drop_graph(b);
    if (b->slots_dev) (void)hipFree(b->slots_dev);
    if (b->stat_dev) (void)hipFree(b->stat_dev);
    if (b->stat_host) (void)hipHostFree(b->stat_host);
    if (b->tmap) (void)hipFree(b->tmap);
    if (b->amap) (void)hipFree(b->amap);
    if (b->idle_ctl) (void)hipFree(b->idle_ctl);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return FX_OK;
}

int fx_batch_admit(fx_batch b, fx_domain d, double t_end, double print_next, int32_t *slot)
{
    if (!b || !d) return fx_fail(FX_ERR_ARG, "batch or domain is NULL");
    int rc;
    if ((rc = set_dev(b))) return rc;
    if (d->desc.device != b->device) return fx_fail(FX_ERR_ARG, "domain lives on another device");
    if (d->desc.precision != FX_FP32_MIXED) return fx_fail(FX_ERR_UNSUPPORTED, "batches hold FP32-mixed domains");
    if (!d->has_static) return fx_fail(FX_ERR_STATE, "fx_set_static has not been called");
    if (d->wintra) return fx_fail(FX_ERR_UNSUPPORTED, "WINTRA domains run alone (fx_advance)");
    if ((rc = fx_inflow_cell_ok(d))) return rc;
    const int dade = (d->desc.ade_solver && d->desc.ade_mode == FX_ADE_FAST) ? 1 : 0;
    if (dade != b->ade || (dade && d->desc.nchem != b->nchem))
        return fx_fail(FX_ERR_ARG, "the domain's solutes differ from the batch's");
    for (int s = 0; s < b->maxm; s++)
        if (b->member[s] == d) return fx_fail(FX_ERR_STATE, "domain is already a member");
    // a domain steps in one batch only: two batches would advance the same
    // state and control block on different streams
    if (d->in_batch) return fx_fail(FX_ERR_STATE, "domain is a member of another batch: evict it first");
    int s = 0;
    while (s < b->maxm && b->member[s]) s++;
    if (s == b->maxm) return fx_fail(FX_ERR_STATE, "batch is full");
    // the batch's queued steps must not overlap the member's set-up
    FX_HIP(hipStreamSynchronize(b->stream));
    // CFL partials of the current state, then the clock for a run to t_end
    if ((rc = fx32_launch_cfl_initial(d))) return rc;
    FX_HIP(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(fx_ctl), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    fx_ctl *c = d->ctl_host;
    c->print_next = print_next;
    c->t_end = t_end;
    c->halt = c->tim > t_end ? 1 : 0;
    c->pending_halt = 0;
    c->flip_pending = 0;
    c->steps = 0;
    c->solver_steps = 0;
    c->status &= ~4;
    c->ade_error = 0;
    FX_HIP(hipMemcpyAsync(d->ctl, c, sizeof(fx_ctl), hipMemcpyHostToDevice, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    fx_bslot &B = b->slots_host[s];
    B.P = fx_make_params(d);
    B.S = make_s32p(d);
    B.A = make_a32p(d);
    B.ctl = d->ctl;
    B.npart = fx32_npart(d);
    B.gx = (B.P.nrows + FX32_TH - 1) / FX32_TH;
    B.agx = fx32_ade_gx(d);
    B.ade_on = dade;
    b->nab[s] = fx32_ade_blocks(d);
    b->member[s] = d;
    d->in_batch = true;
    d->cfl_fresh = false;               // the batch steps it
    b->reported[s] = 0;
    b->t_end[s] = t_end;
    b->stat_host[s] = fx_bstat{c->tim, c->dtfl, c->it, 0, 0, c->halt, c->status, 0};
    b->dirty = true;
    if (slot) *slot = s;
    return FX_OK;
}

int fx_batch_resume(fx_batch b, int32_t slot, double print_next)
{
    if (!b || slot < 0 || slot >= b->maxm || !b->member[slot]) return fx_fail(FX_ERR_ARG, "no such member");
    int rc;
    if ((rc = set_dev(b))) return rc;
    fx_domain d = b->member[slot];
    FX_HIP(hipStreamSynchronize(b->stream));
    FX_HIP(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(fx_ctl), hipMemcpyDeviceToHost, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    fx_ctl *c = d->ctl_host;
    if (c->tim > c->t_end) return fx_fail(FX_ERR_STATE, "member has passed its end time");
    c->print_next = print_next;
    c->halt = 0;
    c->pending_halt = 0;
    FX_HIP(hipMemcpyAsync(d->ctl, c, sizeof(fx_ctl), hipMemcpyHostToDevice, d->stream));
    FX_HIP(hipStreamSynchronize(d->stream));
    b->reported[slot] = 0;
    b->stat_host[slot].halt = 0;
    return FX_OK;
}

int fx_batch_evict(fx_batch b, int32_t slot)
{
    if (!b || slot < 0 || slot >= b->maxm || !b->member[slot]) return fx_fail(FX_ERR_ARG, "no such member");
    int rc;
    if ((rc = set_dev(b))) return rc;
    FX_HIP(hipStreamSynchronize(b->stream));
    b->member[slot]->in_batch = false;
    b->member[slot] = nullptr;
    b->slots_host[slot] = fx_bslot{};
    b->slots_host[slot].ctl = b->idle_ctl;
    b->reported[slot] = 0;
    b->dirty = true;
    return FX_OK;
}

int fx_batch_run(fx_batch b, int32_t nsteps)
{
    if (!b) return fx_fail(FX_ERR_ARG, "batch is NULL");
    if (nsteps <= 0) return fx_fail(FX_ERR_ARG, "nsteps must be positive");
    int rc;
    if ((rc = set_dev(b))) return rc;
    if (b->dirty && (rc = rebuild(b))) return rc;
    if (b->nslot == 0) return FX_OK;
    if (b->use_graph && !b->graph_ok && nsteps >= GRAPH_STEPS) (void)capture(b);
    int left = nsteps;
    while (b->graph_ok && left >= GRAPH_STEPS) {
        FX_HIP(hipGraphLaunch(b->graph, b->stream));
        left -= GRAPH_STEPS;
    }
    for (; left > 0; left--)
        if ((rc = launch_one(b))) return rc;
    // each step's buffer flip and halt are done by the next step's clock; the
    // last step's, and the status the host polls, here
    if ((rc = fx32b_launch_end(b->stream, b->slots_dev, b->nslot, b->stat_dev))) return rc;
    b->steps_run += nsteps;
    return FX_OK;
}

int fx_batch_poll(fx_batch b, fx_batch_event *ev, int32_t max_events, int32_t *n_events)
{
    if (!b || !n_events || (max_events > 0 && !ev)) return fx_fail(FX_ERR_ARG, "bad arguments");
    int rc;
    if ((rc = set_dev(b))) return rc;
    *n_events = 0;
    if (b->nslot > 0 && b->steps_run > 0)
        FX_HIP(hipMemcpyAsync(b->stat_host, b->stat_dev, sizeof(fx_bstat) * b->nslot, hipMemcpyDeviceToHost,
                              b->stream));
    FX_HIP(hipStreamSynchronize(b->stream));
    for (int s = 0; s < b->nslot && *n_events < max_events; s++) {
        if (!b->member[s] || b->reported[s]) continue;
        const fx_bstat &st = b->stat_host[s];
        if (!st.halt && !(st.status & 4)) continue;
        fx_batch_event &e = ev[(*n_events)++];
        e.slot = s;
        e.kind = (st.status & 4) ? FX_EV_ERROR : (st.tim > b->t_end[s] ? FX_EV_DONE : FX_EV_PRINT);
        e.tim = st.tim;
        e.it = st.it;
        e.steps = st.steps;
        e.solver_steps = st.solver_steps;
        e.pad = 0;
        b->reported[s] = 1;
    }
    return FX_OK;
}

int fx_batch_info(fx_batch b, int32_t *n_members, int64_t *cells, int64_t *cell_updates)
{
    if (!b) return fx_fail(FX_ERR_ARG, "batch is NULL");
    int n = 0;
    int64_t c = 0, u = 0;
    for (int s = 0; s < b->maxm; s++) {
        fx_domain d = b->member[s];
        if (!d) continue;
        n++;
        const int64_t cells_s = (int64_t)d->desc.nrows * d->desc.ncols;
        c += cells_s;
        if (s < b->nslot) u += cells_s * b->stat_host[s].solver_steps;
    }
    if (n_members) *n_members = n;
    if (cells) *cells = c;
    if (cell_updates) *cell_updates = u;
    return FX_OK;
}

} // extern "C"
