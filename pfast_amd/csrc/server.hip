// server.hip -- the request server: traveltime-table and gradient requests on
// resident velocity models, admitted by device-memory footprint and batched with
// the other requests of the same grid shape into shared sweep launches.
//
// Scheduling, on one host thread driven by pfast_server_progress:
//   * The 3D models of one grid shape form a group with its own stream and chunk
//     table.  A group runs sweep cycles; one cycle is the 8 orderings over every
//     chunk (C sources) still sweeping, of every admitted request of that shape,
//     forward-phase and adjoint-phase chunks in the same cycle.  Requests join and
//     leave at cycle boundaries (continuous batching): every chunk sees exactly the
//     sequence of full cycles pfast_shots3d_run gives it alone, so its traveltimes,
//     lambda, gradient and misfit are bit-identical to a standalone run.
//     A cycle is issued in slices of kSlice launch steps with at most two
//     slices queued, so that the streams of other groups and of 2D requests, which
//     may share a hardware queue with a large group, wait for milliseconds of its
//     work rather than for a whole cycle.
//   * 2D requests run one launch per stage (forward, adjoints) on one of the 2D
//     streams; different models use different streams and overlap.
//   * Admission: first fit in arrival order while the device memory of the
//     resident models and admitted requests stays within the budget.  Freed
//     memory returns to a caching pool (no hipFree, which synchronises the device).
//   * Results stay resident until pfast_server_release; fetching them uses a copy
//     stream of the server and never waits for other requests' sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pfast.h"
#include "batch3d.h"
#include "context.h"
#include "pfast_internal.h"
#include "shots2d.h"
#include "shots3d.h"

namespace {

constexpr int kStreams2D = 2;
constexpr int kSlice = 8;     // launch steps (tile anti-diagonals) per queued slice of a cycle
constexpr int kQueued = 2;    // slices of a group cycle queued at once
constexpr int kMaxSrc2D = 1 << 16;  // per-request source limit of the 2D status staging

enum Phase { PH_NONE = 0, PH_FWD, PH_ADJ, PH_FINISH };

struct SModel {
    int dim = 3, group = -1, slot = 0;
    pfast_model3d *m3 = nullptr;
    pfast_model2d *m2 = nullptr;
    size_t bytes = 0;
    int refs = 0;  // requests that reference it and are not released
};

struct SRequest {
    long id = 0;
    int model = 0, kind = PFAST_REQ_TABLE, MAXSwp = 100, state = PFAST_REQ_QUEUED;
    int nsrc = 0;
    long nrec = 0;
    std::vector<int> src, srnum, rec;
    std::vector<double> nrm, tobs, trobs;
    size_t bytes = 0;
    double t_submit = 0, t_admit = -1, t_done = -1;
    std::string err;
    bool release = false;  // destroy as soon as its stream is idle
    int phase = PH_NONE, cycle = 0;
    hipEvent_t ev = nullptr;  // completion of the last issued stage
    bool ev_pending = false;
    // 3D
    pfast_shots3d *s3 = nullptr;
    std::vector<int> gid;    // group chunk id of each local chunk
    std::vector<char> live;  // local chunk still sweeping in the current phase
    std::vector<int> cyc_f, cyc_a;
    // 2D
    pfast_shots2d *s2 = nullptr;
};

struct Group {
    pfast::Grid3 g{};
    pfast_ctx *ctx = nullptr;
    std::vector<ChunkRef> table;  // by chunk id
    std::vector<long> owner;      // request id per chunk id, 0 = free
    std::vector<int> lchunk;      // local chunk of the owner
    std::vector<int> freeids;
    int cap = 0;                  // capacity of the device/pinned arrays below (chunk ids)
    ChunkRef *dtable = nullptr, *htable = nullptr;
    int *dact = nullptr, *hact = nullptr;
    unsigned *dchanged = nullptr, *hchanged = nullptr;
    bool dirty = true;
    hipEvent_t ev[kQueued] = {};  // end of the queued slices, by slice number mod kQueued
    bool inflight = false;        // a cycle is in progress
    int o = 0, D = 0;             // next slice: launch group (first ordering), first step
    int issued = 0, done = 0, nfwd = 0, nadj = 0;
    // exact tile skipping: records by chunk id (cap x ntile), zeroed when a chunk
    // starts a phase; epochs from the group's cycle counter
    int *tmod = nullptr, *tclean = nullptr;
    int ntile = 0;
    long cycle_no = 0;
    // dataflow launches (pfast_ctx_set_dataflow): state by chunk id, cycle epochs
    unsigned *flow = nullptr;
    size_t nflow = 0;
    int ftile = 0, fepoch = 0;
    bool dataflow_cycle = false;      // the cycle in flight runs as dataflow launches
    unsigned *habort = nullptr;       // pinned: the state's abort word after the cycle
    std::vector<int> reset;  // chunk ids whose records are zeroed before the next cycle
    std::vector<int> cyc_ids;  // chunk ids of the cycle in flight
    std::set<long> cyc_reqs;   // requests of the cycle in flight
};

struct Slot2D {
    pfast_ctx *ctx = nullptr;
    long running = 0;  // request with stages in flight on this stream
    int *hstatus = nullptr;
};

double elapsed(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct pfast_server {
    int device = 0;
    size_t budget = 0, in_use = 0;
    DevicePool pool;
    std::vector<std::unique_ptr<Group>> groups;
    Slot2D slots[kStreams2D];
    int next_slot = 0;
    hipStream_t copy = nullptr;
    std::map<int, SModel> models;
    int next_model = 1;
    std::map<long, SRequest> reqs;
    std::deque<long> queue;
    long next_id = 1;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    long cycles = 0;         // group sweep cycles issued
    double node_sweeps = 0;  // source node-sweeps issued by group cycles
};

namespace {

size_t round2m(size_t b) { return pfast::pool_round(b); }

size_t footprint3d(const pfast::Grid3 &g, int nsrc, long nrec, bool adjoint)
{
    const size_t N = (size_t)g.nrows * g.ncols * g.nz;
    const size_t nchunk = (nsrc + PFAST_CHUNK - 1) / PFAST_CHUNK;
    const size_t field = round2m(nchunk * N * PFAST_CHUNK * 4);
    // the adjoint field holds (lambda, T) pairs
    const size_t afield = adjoint ? round2m(nchunk * N * PFAST_CHUNK * 8) : 0;
    size_t b = field + afield + 2 * round2m(N * 8) +
               nchunk * (PFAST_CHUNK * 4 + 8 + sizeof(ChunkRef)) + (size_t)nsrc * 8 +
               (size_t)nrec * (8 + 4 + 12 + 8);
    if (adjoint) b += (size_t)nrec * (8 + 4 + 8 + 8 + 4 + 8 + 24 + 24 + 8) + (size_t)nsrc * 8;
    return b + 64 * 1024;  // size-class slack of the small tables
}

int grow_group(Group &G, int need)
{
    if (need <= G.cap) return PFAST_OK;
    int cap = std::max(64, G.cap);
    while (cap < need) cap *= 2;
    ChunkRef *dt = nullptr, *ht = nullptr;
    int *da = nullptr, *ha = nullptr;
    unsigned *dc = nullptr, *hc = nullptr;
    if (hipMalloc(&dt, cap * sizeof(ChunkRef)) != hipSuccess || hipMalloc(&da, cap * sizeof(int)) != hipSuccess ||
        hipMalloc(&dc, cap * sizeof(unsigned)) != hipSuccess ||
        hipHostMalloc((void **)&ht, cap * sizeof(ChunkRef)) != hipSuccess ||
        hipHostMalloc((void **)&ha, cap * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&hc, cap * sizeof(unsigned)) != hipSuccess) {
        (void)hipGetLastError();
        if (dt) (void)hipFree(dt);
        if (da) (void)hipFree(da);
        if (dc) (void)hipFree(dc);
        if (ht) (void)hipHostFree(ht);
        if (ha) (void)hipHostFree(ha);
        if (hc) (void)hipHostFree(hc);
        pfast::set_error("allocation of a group's chunk tables failed");
        return PFAST_ERR_NOMEM;
    }
    int *tm = nullptr, *tc = nullptr;
    if (G.ntile &&
        (hipMalloc(&tm, (size_t)cap * G.ntile * 5 * sizeof(int)) != hipSuccess ||
         hipMalloc(&tc, (size_t)cap * G.ntile * kTileClean * sizeof(int)) != hipSuccess ||
         hipMemsetAsync(tm, 0, (size_t)cap * G.ntile * 5 * sizeof(int), G.ctx->stream) != hipSuccess ||
         hipMemsetAsync(tc, 0, (size_t)cap * G.ntile * kTileClean * sizeof(int), G.ctx->stream) != hipSuccess)) {
        (void)hipGetLastError();
        for (void *q : {(void *)dt, (void *)da, (void *)dc, (void *)tm, (void *)tc})
            if (q) (void)hipFree(q);
        (void)hipHostFree(ht);
        (void)hipHostFree(ha);
        (void)hipHostFree(hc);
        pfast::set_error("allocation of a group's tile records failed");
        return PFAST_ERR_NOMEM;
    }
    if (G.dtable) {
        (void)hipFree(G.dtable);
        (void)hipFree(G.dact);
        (void)hipFree(G.dchanged);
        (void)hipHostFree(G.htable);
        (void)hipHostFree(G.hact);
        (void)hipHostFree(G.hchanged);
    }
    unsigned *fl = nullptr;
    const size_t fw = pfast::flow_words(cap, G.ftile);
    if (hipMalloc(&fl, fw * sizeof(unsigned)) != hipSuccess ||
        pfast::launch_flow_reset(G.ctx->stream, fl, fw) != hipSuccess) {
        (void)hipGetLastError();
        for (void *q : {(void *)dt, (void *)da, (void *)dc, (void *)tm, (void *)tc, (void *)fl})
            if (q) (void)hipFree(q);
        (void)hipHostFree(ht);
        (void)hipHostFree(ha);
        (void)hipHostFree(hc);
        pfast::set_error("allocation of a group's dataflow state failed");
        return PFAST_ERR_NOMEM;
    }
    if (G.flow) (void)hipFree(G.flow);
    G.flow = fl;
    G.nflow = fw;
    G.fepoch = 0;
    if (G.tmod) (void)hipFree(G.tmod);
    if (G.tclean) (void)hipFree(G.tclean);
    G.tmod = tm;
    G.tclean = tc;
    G.reset.clear();  // fresh records are zero
    G.dtable = dt;
    G.htable = ht;
    G.dact = da;
    G.hact = ha;
    G.dchanged = dc;
    G.hchanged = hc;
    G.cap = cap;
    G.dirty = true;
    return PFAST_OK;
}

int group_of(pfast_server *S, const pfast::Grid3 &g, int *out)
{
    for (size_t i = 0; i < S->groups.size(); i++) {
        const pfast::Grid3 &h = S->groups[i]->g;
        if (h.nrows == g.nrows && h.ncols == g.ncols && h.nz == g.nz && h.dx == g.dx &&
            h.dy == g.dy && h.dz == g.dz) {
            *out = (int)i;
            return PFAST_OK;
        }
    }
    auto G = std::make_unique<Group>();
    G->g = g;
    int rc = pfast_ctx_create(S->device, &G->ctx);
    if (rc) return rc;
    G->ctx->pool = &S->pool;
    {
        const SweepVariant &sv = pfast::sweep_variant(G->ctx->sweep);
        G->ftile = ((g.nrows + sv.ti - 1) / sv.ti) * ((g.ncols + sv.tj - 1) / sv.tj);
        if (G->ctx->tile_skip && sv.kind == 0) G->ntile = G->ftile;
    }
    for (auto &e : G->ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            for (auto &f : G->ev)
                if (f) (void)hipEventDestroy(f);
            pfast_ctx_destroy(G->ctx);
            pfast::set_error("hipEventCreate failed");
            return PFAST_ERR_HIP;
        }
    if (hipHostMalloc((void **)&G->habort, sizeof(unsigned)) != hipSuccess) {
        for (auto &f : G->ev) (void)hipEventDestroy(f);
        pfast_ctx_destroy(G->ctx);
        pfast::set_error("hipHostMalloc failed");
        return PFAST_ERR_NOMEM;
    }
    *G->habort = 0;
    S->groups.push_back(std::move(G));
    *out = (int)S->groups.size() - 1;
    return PFAST_OK;
}

void fail_request(pfast_server *S, SRequest &R, const std::string &msg)
{
    R.state = PFAST_REQ_FAILED;
    R.err = msg;
    R.t_done = elapsed(S->t0);
    R.phase = PH_NONE;
}

void free_chunks(Group &G, SRequest &R)
{
    for (int id : R.gid) {
        G.owner[id] = 0;
        G.freeids.push_back(id);
    }
    R.gid.clear();
    G.dirty = true;
}

// Admit a queued 3D request into its (idle) group: shot set, forward initialisation,
// chunk ids in the group table.
int admit3d(pfast_server *S, SRequest &R, SModel &M)
{
    Group &G = *S->groups[M.group];
    const bool grad = R.kind == PFAST_REQ_GRADIENT;
    int rc = pfast_shots3d_create(M.m3, R.nsrc, R.src.data(), R.srnum.data(),
                                  R.rec.empty() ? nullptr : R.rec.data(),
                                  grad ? R.nrm.data() : nullptr, grad ? R.tobs.data() : nullptr,
                                  &R.s3);
    if (rc) return rc;
    if ((rc = pfast::fwd_begin3d(R.s3))) return rc;
    const int nch = R.s3->nchunk;
    R.gid.resize(nch);
    for (int c = 0; c < nch; c++) {
        int id;
        if (!G.freeids.empty()) {
            id = G.freeids.back();
            G.freeids.pop_back();
        } else {
            id = (int)G.table.size();
            G.table.emplace_back();
            G.owner.push_back(0);
            G.lchunk.push_back(0);
        }
        G.table[id] = R.s3->hchunks[c];
        G.owner[id] = R.id;
        G.lchunk[id] = c;
        R.gid[c] = id;
        G.reset.push_back(id);
    }
    G.dirty = true;
    R.live.assign(nch, 1);
    R.cyc_f.assign(R.nsrc, 0);
    R.phase = PH_FWD;
    R.cycle = 0;
    return PFAST_OK;
}

int admit2d(pfast_server *S, SRequest &R, SModel &M)
{
    Slot2D &sl = S->slots[M.slot];
    const bool grad = R.kind == PFAST_REQ_GRADIENT;
    int rc = pfast_shots2d_create(M.m2, R.nsrc, R.src.data(), R.srnum.data(),
                                  grad ? R.rec.data() : nullptr, grad ? R.nrm.data() : nullptr,
                                  grad ? R.tobs.data() : nullptr,
                                  grad && !R.trobs.empty() ? R.trobs.data() : nullptr, &R.s2);
    if (rc) return rc;
    if ((rc = pfast::issue_fwd2d(R.s2, R.MAXSwp, sl.hstatus))) return rc;
    HIPCHK(hipEventRecord(R.ev, sl.ctx->stream));
    R.ev_pending = true;
    R.phase = PH_FWD;
    sl.running = R.id;
    return PFAST_OK;
}

bool stream_idle(hipStream_t st) { return hipStreamQuery(st) == hipSuccess; }

void try_admit(pfast_server *S)
{
    for (auto it = S->queue.begin(); it != S->queue.end();) {
        SRequest &R = S->reqs[*it];
        SModel &M = S->models[R.model];
        if (S->in_use + R.bytes > S->budget) {
            ++it;
            continue;
        }
        bool ready = M.dim == 3 ? !S->groups[M.group]->inflight : S->slots[M.slot].running == 0;
        if (!ready) {
            ++it;
            continue;
        }
        if (S->in_use + R.bytes + S->pool.cached > S->budget)
            pfast::pool_trim(&S->pool, S->budget - S->in_use - R.bytes);
        R.t_admit = elapsed(S->t0);
        R.state = PFAST_REQ_RUNNING;
        S->in_use += R.bytes;
        const int rc = M.dim == 3 ? admit3d(S, R, M) : admit2d(S, R, M);
        if (rc) {
            if (M.dim == 3 && !R.gid.empty()) free_chunks(*S->groups[M.group], R);
            fail_request(S, R, pfast_last_error());
        }
        it = S->queue.erase(it);
    }
}

// Keep at most kQueued slices of the cycle in flight queued on the group's stream;
// the last one is followed by the copy of the changed flags.
int pump_cycle(Group &G)
{
    hipStream_t st = G.ctx->stream;
    const int ns = pfast::group_steps3d(G.ctx, G.g), go = pfast::group_orderings3d(G.ctx);
    while (G.o < 8 && G.issued - G.done < kQueued) {
        const int D1 = std::min(G.D + kSlice, ns);
        const pfast::TileSkip skip{G.tmod, G.tclean, (int)G.cycle_no};
        int rc = pfast::issue_steps3d(G.ctx, G.g, G.dtable, G.dact, G.nfwd, G.dact + G.nfwd,
                                      G.nadj, G.dchanged, G.o, G.D, D1, nullptr,
                                      G.tmod ? &skip : nullptr);
        if (rc) return rc;
        G.D = D1;
        if (G.D == ns) {
            G.o += go;
            G.D = 0;
        }
        if (G.o == 8)
            HIPCHK(hipMemcpyAsync(G.hchanged, G.dchanged, G.table.size() * sizeof(unsigned),
                                  hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(G.ev[G.issued % kQueued], st));
        G.issued++;
    }
    return PFAST_OK;
}

// Start a sweep cycle over the live chunks of group G.
int issue_cycle(pfast_server *S, Group &G)
{
    const int n = (int)G.table.size();
    int rc = grow_group(G, n);
    if (rc) return rc;
    hipStream_t st = G.ctx->stream;
    if (G.dirty) {
        std::copy(G.table.begin(), G.table.end(), G.htable);
        HIPCHK(hipMemcpyAsync(G.dtable, G.htable, n * sizeof(ChunkRef), hipMemcpyHostToDevice, st));
        G.dirty = false;
    }
    std::vector<int> fwd, adj;
    G.cyc_reqs.clear();
    long real = 0;
    for (int id = 0; id < n; id++) {
        if (!G.owner[id]) continue;
        SRequest &R = S->reqs[G.owner[id]];
        const int c = G.lchunk[id];
        if (!R.live[c]) continue;
        (R.phase == PH_ADJ ? adj : fwd).push_back(id);
        G.cyc_reqs.insert(R.id);
        real += std::min(PFAST_CHUNK, R.nsrc - c * PFAST_CHUNK);
    }
    if (fwd.empty() && adj.empty()) return PFAST_OK;
    for (long id : G.cyc_reqs) S->reqs[id].cycle++;
    if (G.tmod) {
        // a launch epoch is 1 + ((cycle * 8 + o) * nd + D): keep it below INT_MAX by
        // forgetting every record before the next cycle would pass it
        const long nd = pfast::diagonals3d(G.ctx, G.g);
        if ((G.cycle_no + 2) * 8 * nd >= 0x7fffffffL) {
            HIPCHK(hipMemsetAsync(G.tmod, 0, (size_t)G.cap * G.ntile * 5 * sizeof(int), st));
            HIPCHK(hipMemsetAsync(G.tclean, 0, (size_t)G.cap * G.ntile * kTileClean * sizeof(int), st));
            G.cycle_no = 0;
            G.reset.clear();
        }
        for (int id : G.reset) {
            HIPCHK(hipMemsetAsync(G.tmod + (size_t)id * G.ntile * 5, 0, G.ntile * 5 * sizeof(int), st));
            HIPCHK(hipMemsetAsync(G.tclean + (size_t)id * G.ntile * kTileClean, 0, G.ntile * kTileClean * sizeof(int), st));
        }
        G.reset.clear();
    }
    std::copy(fwd.begin(), fwd.end(), G.hact);
    std::copy(adj.begin(), adj.end(), G.hact + fwd.size());
    const int na = (int)(fwd.size() + adj.size());
    HIPCHK(hipMemcpyAsync(G.dact, G.hact, na * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(G.dchanged, 0, n * sizeof(unsigned), st));
    G.nfwd = (int)fwd.size();
    G.nadj = (int)adj.size();
    G.o = G.D = G.issued = G.done = 0;
    G.cycle_no++;
    G.cyc_ids = fwd;
    G.cyc_ids.insert(G.cyc_ids.end(), adj.begin(), adj.end());
    G.inflight = true;
    S->cycles++;
    S->node_sweeps += 8.0 * G.g.nrows * G.g.ncols * G.g.nz * (double)real;
    if (pfast::use_dataflow(G.ctx, na, G.ftile, G.nflow)) {
        // the whole cycle as dataflow launches (forward chunks, then adjoint chunks);
        // they share the GPU with other groups' work block by block
        if (++G.fepoch > 0x7FFF) {
            HIPCHK(pfast::launch_flow_reset(st, G.flow, G.nflow));
            G.fepoch = 1;
        }
        const pfast::TileSkip skip{G.tmod, G.tclean, (int)G.cycle_no};
        const pfast::TileSkip *sp = G.tmod ? &skip : nullptr;
        int rc2 = pfast::issue_cycle3d_flow(G.ctx, G.g, G.dtable, G.dact, G.nfwd, false,
                                            G.dchanged, G.flow, G.cap, G.fepoch, st, sp);
        if (!rc2)
            rc2 = pfast::issue_cycle3d_flow(G.ctx, G.g, G.dtable, G.dact + G.nfwd, G.nadj, true,
                                            G.dchanged, G.flow, G.cap, G.fepoch, st, sp);
        if (rc2) return rc2;
        HIPCHK(hipMemcpyAsync(G.hchanged, G.dchanged, n * sizeof(unsigned), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(G.habort, G.flow + G.nflow - 2, sizeof(unsigned), hipMemcpyDeviceToHost, st));
        HIPCHK(hipEventRecord(G.ev[0], st));
        G.issued = 1;
        G.o = 8;
        G.dataflow_cycle = true;
        return PFAST_OK;
    }
    G.dataflow_cycle = false;
    return pump_cycle(G);
}

// A finished cycle: convergence per chunk, phase ends per request.
void finish_cycle(pfast_server *S, Group &G, int *ndone)
{
    for (int id : G.cyc_ids) {
        SRequest &R = S->reqs[G.owner[id]];
        const int c = G.lchunk[id];
        const unsigned bits = G.hchanged[id];
        std::vector<int> &cyc = R.phase == PH_ADJ ? R.cyc_a : R.cyc_f;
        for (int k = 0; k < PFAST_CHUNK; k++) {
            const int s = c * PFAST_CHUNK + k;
            if (s < R.nsrc && !((bits >> k) & 1u) && cyc[s] == 0) cyc[s] = R.cycle;
        }
        if (!bits) R.live[c] = 0;
    }
    for (long rid : G.cyc_reqs) {
        SRequest &R = S->reqs[rid];
        bool any = false;
        for (char l : R.live) any |= l != 0;
        if (any && R.cycle < R.MAXSwp) continue;
        std::vector<int> &cyc = R.phase == PH_ADJ ? R.cyc_a : R.cyc_f;
        for (int &v : cyc)
            if (v == 0) v = R.cycle;
        if (R.phase == PH_FWD && R.kind == PFAST_REQ_TABLE) {
            free_chunks(G, R);
            R.phase = PH_NONE;
            R.state = PFAST_REQ_DONE;
            R.t_done = elapsed(S->t0);
            (*ndone)++;
        } else if (R.phase == PH_FWD) {
            int rc = pfast::adj_begin3d(R.s3);
            if (rc) {
                free_chunks(G, R);
                fail_request(S, R, pfast_last_error());
                (*ndone)++;
                continue;
            }
            R.phase = PH_ADJ;
            R.cycle = 0;
            R.cyc_a.assign(R.nsrc, 0);
            R.live.assign(R.live.size(), 1);
            G.reset.insert(G.reset.end(), R.gid.begin(), R.gid.end());
        } else {
            free_chunks(G, R);
            int rc = pfast::adj_end3d(R.s3);
            if (!rc && hipEventRecord(R.ev, G.ctx->stream) != hipSuccess) rc = PFAST_ERR_HIP;
            if (rc) {
                fail_request(S, R, pfast_last_error());
                (*ndone)++;
                continue;
            }
            R.ev_pending = true;
            R.phase = PH_FINISH;
        }
    }
}

// A finished stage event of request R (3D gradient reduction, 2D stages).
void finish_stage(pfast_server *S, SRequest &R, int *ndone)
{
    SModel &M = S->models[R.model];
    if (M.dim == 3) {  // PH_FINISH
        R.phase = PH_NONE;
        R.state = PFAST_REQ_DONE;
        R.t_done = elapsed(S->t0);
        (*ndone)++;
        return;
    }
    Slot2D &sl = S->slots[M.slot];
    if (R.phase == PH_FWD) {
        int rc = pfast::check_fwd2d(R.s2, sl.hstatus);
        if (!rc && R.kind == PFAST_REQ_GRADIENT) {
            rc = pfast::issue_adj2d(R.s2, R.MAXSwp);
            if (!rc) rc = pfast::issue_grad2d(R.s2, nullptr);
            if (!rc && hipEventRecord(R.ev, sl.ctx->stream) != hipSuccess) rc = PFAST_ERR_HIP;
            if (!rc) {
                R.ev_pending = true;
                R.phase = PH_ADJ;
                return;
            }
        }
        sl.running = 0;
        if (rc) {
            fail_request(S, R, pfast_last_error());
        } else {
            R.phase = PH_NONE;
            R.state = PFAST_REQ_DONE;
            R.t_done = elapsed(S->t0);
        }
        (*ndone)++;
        return;
    }
    // PH_ADJ of 2D
    R.s2->have_adj = true;
    sl.running = 0;
    R.phase = PH_NONE;
    R.state = PFAST_REQ_DONE;
    R.t_done = elapsed(S->t0);
    (*ndone)++;
}

hipStream_t stream_of(pfast_server *S, const SRequest &R)
{
    const SModel &M = S->models.at(R.model);
    return M.dim == 3 ? S->groups[M.group]->ctx->stream : S->slots[M.slot].ctx->stream;
}

// Destroy a finished request's objects.  A done request has no work left on the
// device (its last cycle or stage completed and its chunk ids left every active
// list), so its memory goes back to the pool at once; a failed one may have
// stages in flight and waits for its stream to drain.
bool try_destroy(pfast_server *S, SRequest &R)
{
    if (R.state == PFAST_REQ_RUNNING) return false;
    if (R.state == PFAST_REQ_FAILED && (R.s3 || R.s2) && !stream_idle(stream_of(S, R)))
        return false;
    if (R.s3) pfast::shots3d_free(R.s3);
    if (R.s2) pfast::shots2d_free(R.s2);
    R.s3 = nullptr;
    R.s2 = nullptr;
    if (R.ev) (void)hipEventDestroy(R.ev);
    R.ev = nullptr;
    if (R.t_admit >= 0) S->in_use -= R.bytes;
    S->models[R.model].refs--;
    return true;
}

int progress_once(pfast_server *S, int *ndone)
{
    for (auto &G : S->groups) {
        if (!G->inflight) continue;
        hipError_t q = hipSuccess;
        while (G->done < G->issued && (q = hipEventQuery(G->ev[G->done % kQueued])) == hipSuccess)
            G->done++;
        if (q == hipSuccess && (G->o < 8 || G->done < G->issued)) {
            int rc = pump_cycle(*G);
            if (rc) return rc;
            continue;
        }
        if (q == hipErrorNotReady) continue;
        G->inflight = false;
        const bool aborted = q == hipSuccess && G->dataflow_cycle && *G->habort;
        if (q != hipSuccess || aborted) {
            const std::string why = aborted ? std::string("a dataflow dependency wait timed out")
                                            : std::string(hipGetErrorString(q));
            for (long rid : G->cyc_reqs) {
                SRequest &R = S->reqs[rid];
                free_chunks(*G, R);
                fail_request(S, R, "sweep cycle failed: " + why);
                (*ndone)++;
            }
            if (aborted) {  // the state is reset before the group sweeps again
                *G->habort = 0;
                if (pfast::launch_flow_reset(G->ctx->stream, G->flow, G->nflow) != hipSuccess)
                    (void)hipGetLastError();
                G->fepoch = 0;
            }
            continue;
        }
        finish_cycle(S, *G, ndone);
    }
    for (auto &kv : S->reqs) {
        SRequest &R = kv.second;
        if (!R.ev_pending) continue;
        const hipError_t q = hipEventQuery(R.ev);
        if (q == hipErrorNotReady) continue;
        R.ev_pending = false;
        if (q != hipSuccess) {
            SModel &M = S->models[R.model];
            if (M.dim == 2) S->slots[M.slot].running = 0;
            fail_request(S, R, std::string("request stage failed: ") + hipGetErrorString(q));
            (*ndone)++;
            continue;
        }
        finish_stage(S, R, ndone);
    }
    for (auto it = S->reqs.begin(); it != S->reqs.end();) {
        if (it->second.release && try_destroy(S, it->second))
            it = S->reqs.erase(it);
        else
            ++it;
    }
    try_admit(S);
    for (auto &G : S->groups)
        if (!G->inflight) {
            int rc = issue_cycle(S, *G);
            if (rc) return rc;
        }
    return PFAST_OK;
}

SRequest *find(pfast_server *S, long ticket)
{
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) {
        pfast::set_error("invalid argument: unknown ticket");
        return nullptr;
    }
    return &it->second;
}

int enqueue(pfast_server *S, SRequest &&R, long *ticket)
{
    SModel &M = S->models[R.model];
    size_t resident = 0;
    for (auto &kv : S->models) resident += kv.second.bytes;
    if (resident + R.bytes > S->budget) {
        pfast::set_error("request needs " + std::to_string(R.bytes) +
                         " bytes of device memory, more than the server's budget");
        return PFAST_ERR_NOMEM;
    }
    R.id = S->next_id++;
    R.t_submit = elapsed(S->t0);
    if (hipEventCreateWithFlags(&R.ev, hipEventDisableTiming) != hipSuccess) {
        pfast::set_error("hipEventCreate failed");
        return PFAST_ERR_HIP;
    }
    M.refs++;
    *ticket = R.id;
    S->queue.push_back(R.id);
    S->reqs.emplace(R.id, std::move(R));
    return PFAST_OK;
}

}  // namespace

extern "C" {

int pfast_server_create(int device, size_t budget_bytes, pfast_server **out)
{
    ARGCHK(out && budget_bytes > 0, "bad argument");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    ARGCHK(device >= 0 && device < ndev, "no such device");
    HIPCHK(hipSetDevice(device));
    auto *S = new pfast_server();
    S->device = device;
    S->budget = budget_bytes;
    int rc = PFAST_OK;
    if (hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking) != hipSuccess) {
        pfast::set_error("hipStreamCreate failed");
        rc = PFAST_ERR_HIP;
    }
    for (int i = 0; !rc && i < kStreams2D; i++) {
        rc = pfast_ctx_create(device, &S->slots[i].ctx);
        if (!rc) {
            S->slots[i].ctx->pool = &S->pool;
            if (hipHostMalloc((void **)&S->slots[i].hstatus, kMaxSrc2D * sizeof(int)) != hipSuccess) {
                pfast::set_error("hipHostMalloc failed");
                rc = PFAST_ERR_NOMEM;
            }
        }
    }
    if (rc) {
        pfast_server_destroy(S);
        return rc;
    }
    *out = S;
    return PFAST_OK;
}

int pfast_server_destroy(pfast_server *S)
{
    if (!S) return PFAST_OK;
    (void)hipSetDevice(S->device);
    (void)hipDeviceSynchronize();
    for (auto &kv : S->reqs) {
        SRequest &R = kv.second;
        if (R.s3) pfast_shots3d_destroy(R.s3);
        if (R.s2) pfast_shots2d_destroy(R.s2);
        if (R.ev) (void)hipEventDestroy(R.ev);
    }
    S->reqs.clear();
    for (auto &kv : S->models) {
        if (kv.second.m3) pfast_model3d_destroy(kv.second.m3);
        if (kv.second.m2) pfast_model2d_destroy(kv.second.m2);
    }
    S->models.clear();
    for (auto &G : S->groups) {
        if (G->dtable) {
            (void)hipFree(G->dtable);
            (void)hipFree(G->dact);
            (void)hipFree(G->dchanged);
            (void)hipHostFree(G->htable);
            (void)hipHostFree(G->hact);
            (void)hipHostFree(G->hchanged);
        }
        if (G->tmod) (void)hipFree(G->tmod);
        if (G->tclean) (void)hipFree(G->tclean);
        if (G->flow) (void)hipFree(G->flow);
        if (G->habort) (void)hipHostFree(G->habort);
        for (auto &e : G->ev)
            if (e) (void)hipEventDestroy(e);
        pfast_ctx_destroy(G->ctx);
    }
    S->groups.clear();
    for (auto &sl : S->slots) {
        if (sl.hstatus) (void)hipHostFree(sl.hstatus);
        if (sl.ctx) pfast_ctx_destroy(sl.ctx);
    }
    if (S->copy) (void)hipStreamDestroy(S->copy);
    pfast::pool_trim(&S->pool, 0);
    delete S;
    return PFAST_OK;
}

int pfast_server_add_model3d(pfast_server *S, const double *vp, int nrows, int ncols, int nz,
                             double dx, double dy, double dz, int *model_id)
{
    ARGCHK(S && vp && model_id, "NULL argument");
    HIPCHK(hipSetDevice(S->device));
    const size_t N = (size_t)nrows * ncols * nz;
    const size_t bytes = round2m(N * 8) + round2m(N * 4) +
                         round2m((size_t)pfast::tile8_size(nrows, ncols, nz) * 4);
    ARGCHK(S->in_use + bytes <= S->budget, "resident models exceed the server's memory budget");
    SModel M;
    M.dim = 3;
    pfast::Grid3 g{nrows, ncols, nz, dx, dy, dz, dx == dy && dy == dz};
    int rc = group_of(S, g, &M.group);
    if (rc) return rc;
    rc = pfast_model3d_create(S->groups[M.group]->ctx, vp, nrows, ncols, nz, dx, dy, dz, &M.m3);
    if (rc) return rc;
    M.bytes = bytes;
    S->in_use += bytes;
    *model_id = S->next_model++;
    S->models[*model_id] = M;
    return PFAST_OK;
}

int pfast_server_add_model2d(pfast_server *S, const double *vp, int nrows, int ncols, double dx,
                             double dz, const double *Z, int NumofR, int *model_id)
{
    ARGCHK(S && vp && model_id, "NULL argument");
    HIPCHK(hipSetDevice(S->device));
    const size_t bytes = round2m((size_t)nrows * ncols * 8) + round2m((size_t)NumofR * ncols * 24 + 8) +
                         round2m((size_t)nrows * ncols * 4);
    ARGCHK(S->in_use + bytes <= S->budget, "resident models exceed the server's memory budget");
    SModel M;
    M.dim = 2;
    M.slot = S->next_slot;
    S->next_slot = (S->next_slot + 1) % kStreams2D;
    int rc = pfast_model2d_create(S->slots[M.slot].ctx, vp, nrows, ncols, dx, dz, Z, NumofR, &M.m2);
    if (rc) return rc;
    M.bytes = bytes;
    S->in_use += bytes;
    *model_id = S->next_model++;
    S->models[*model_id] = M;
    return PFAST_OK;
}

int pfast_server_remove_model(pfast_server *S, int model_id)
{
    ARGCHK(S, "NULL server");
    auto it = S->models.find(model_id);
    ARGCHK(it != S->models.end(), "unknown model");
    ARGCHK(it->second.refs == 0, "model still has unreleased requests");
    if (it->second.m3) pfast_model3d_destroy(it->second.m3);
    if (it->second.m2) pfast_model2d_destroy(it->second.m2);
    S->in_use -= it->second.bytes;
    S->models.erase(it);
    return PFAST_OK;
}

int pfast_server_submit3d(pfast_server *S, int model_id, int kind, int nsrc, const int *src,
                          const int *srnum, const int *rec, const double *nxyz,
                          const double *T_obs, int MAXSwp, long *ticket)
{
    ARGCHK(S && src && srnum && ticket && nsrc > 0 && MAXSwp > 0, "bad argument");
    ARGCHK(kind == PFAST_REQ_TABLE || kind == PFAST_REQ_GRADIENT, "unknown request kind");
    auto it = S->models.find(model_id);
    ARGCHK(it != S->models.end() && it->second.dim == 3, "unknown 3D model");
    long nrec = 0;
    for (int i = 0; i < nsrc; i++) {
        ARGCHK(srnum[i] >= 0, "negative srnum");
        nrec += srnum[i];
    }
    ARGCHK(kind == PFAST_REQ_TABLE || (rec && nxyz && T_obs), "a gradient request needs receivers, normals and picks");
    SRequest R;
    R.model = model_id;
    R.kind = kind;
    R.MAXSwp = MAXSwp;
    R.nsrc = nsrc;
    R.nrec = nrec;
    R.src.assign(src, src + 3 * (size_t)nsrc);
    R.srnum.assign(srnum, srnum + nsrc);
    if (rec) R.rec.assign(rec, rec + 3 * nrec);
    if (kind == PFAST_REQ_GRADIENT) {
        R.nrm.assign(nxyz, nxyz + 3 * nrec);
        R.tobs.assign(T_obs, T_obs + nrec);
    }
    const pfast_model3d *m = it->second.m3;
    // the sweep kernels address one chunk's field with 32-bit byte offsets, as for a shot set
    const long plane = (long)m->nrows * m->ncols, N = plane * m->nz;
    ARGCHK((N + 64L * plane) * PFAST_CHUNK * (kind == PFAST_REQ_GRADIENT ? 8 : 4) <= (1L << 32),
           "grid too large for the batched solver (one source chunk's field plus 64 levels "
           "must stay within 4 GiB)");
    R.bytes = footprint3d(pfast::grid_of(m), nsrc, nrec, kind == PFAST_REQ_GRADIENT);
    return enqueue(S, std::move(R), ticket);
}

int pfast_server_submit2d(pfast_server *S, int model_id, int kind, int nsrc, const int *src,
                          const int *srnum, const int *rec, const double *nxy,
                          const double *TT_obs, const double *TR_obs, int MAXSwp, long *ticket)
{
    ARGCHK(S && src && srnum && ticket && nsrc > 0 && MAXSwp > 0, "bad argument");
    ARGCHK(nsrc <= kMaxSrc2D, "too many sources in one 2D request");
    ARGCHK(kind == PFAST_REQ_TABLE || kind == PFAST_REQ_GRADIENT, "unknown request kind");
    auto it = S->models.find(model_id);
    ARGCHK(it != S->models.end() && it->second.dim == 2, "unknown 2D model");
    const pfast_model2d *m = it->second.m2;
    long nrec = 0;
    for (int i = 0; i < nsrc; i++) {
        ARGCHK(srnum[i] >= 0, "negative srnum");
        nrec += srnum[i];
    }
    ARGCHK(kind == PFAST_REQ_TABLE || (rec && nxy && TT_obs && (m->NumofR == 0 || TR_obs)),
           "a gradient request needs receivers, normals and picks");
    SRequest R;
    R.model = model_id;
    R.kind = kind;
    R.MAXSwp = MAXSwp;
    R.nsrc = nsrc;
    R.nrec = nrec;
    R.src.assign(src, src + 2 * (size_t)nsrc);
    R.srnum.assign(srnum, srnum + nsrc);
    if (kind == PFAST_REQ_GRADIENT) {
        R.rec.assign(rec, rec + 2 * nrec);
        R.nrm.assign(nxy, nxy + 2 * nrec);
        R.tobs.assign(TT_obs, TT_obs + nrec);
        if (m->NumofR) R.trobs.assign(TR_obs, TR_obs + nrec * m->NumofR);
    }
    R.bytes = pfast::footprint2d(m->nrows, m->ncols, m->NumofR, nsrc, nrec, kind == PFAST_REQ_GRADIENT);
    return enqueue(S, std::move(R), ticket);
}

int pfast_server_progress(pfast_server *S, int wait_us, int *ncompleted)
{
    ARGCHK(S, "NULL server");
    HIPCHK(hipSetDevice(S->device));
    const auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    for (;;) {
        int rc = progress_once(S, &done);
        if (rc) return rc;
        if (done || elapsed(t0) * 1e6 >= wait_us) break;
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (ncompleted) *ncompleted = done;
    return PFAST_OK;
}

int pfast_server_status(pfast_server *S, long ticket, int *state, double *times)
{
    ARGCHK(S, "NULL server");
    SRequest *R = find(S, ticket);
    if (!R) return PFAST_ERR_ARG;
    if (state) *state = R->state;
    if (times) {
        times[0] = R->t_submit;
        times[1] = R->t_admit;
        times[2] = R->t_done;
    }
    if (R->state == PFAST_REQ_FAILED) pfast::set_error(R->err);
    return PFAST_OK;
}

int pfast_server_cycles(pfast_server *S, long ticket, int *cycles)
{
    ARGCHK(S && cycles, "NULL argument");
    SRequest *R = find(S, ticket);
    if (!R) return PFAST_ERR_ARG;
    ARGCHK(R->state == PFAST_REQ_DONE, "request not done");
    if (R->s3) {
        for (int u = 0; u < R->nsrc; u++) {  // the shot set's internal order -> the caller's
            const int i = R->s3->iperm[u];
            cycles[2 * u] = R->cyc_f[i];
            cycles[2 * u + 1] = R->cyc_a.empty() ? 0 : R->cyc_a[i];
        }
        return PFAST_OK;
    }
    return pfast_shots2d_info(R->s2, cycles, nullptr);
}

int pfast_server_gradient(pfast_server *S, long ticket, double *grad, double *grad_r,
                          double *misfit)
{
    ARGCHK(S, "NULL server");
    SRequest *R = find(S, ticket);
    if (!R) return PFAST_ERR_ARG;
    if (R->state == PFAST_REQ_FAILED) {
        pfast::set_error(R->err);
        return PFAST_ERR_FATAL;
    }
    ARGCHK(R->state == PFAST_REQ_DONE && R->kind == PFAST_REQ_GRADIENT, "no gradient result");
    HIPCHK(hipSetDevice(S->device));
    if (R->s3) {
        const pfast_model3d *m = R->s3->m;
        const long N = (long)m->nrows * m->ncols * m->nz;
        std::vector<double> ee(R->nsrc);
        if (grad) HIPCHK(hipMemcpyAsync(grad, R->s3->grad, N * sizeof(double), hipMemcpyDeviceToHost, S->copy));
        HIPCHK(hipMemcpyAsync(ee.data(), R->s3->EE, R->nsrc * sizeof(double), hipMemcpyDeviceToHost, S->copy));
        HIPCHK(hipStreamSynchronize(S->copy));
        if (misfit) {
            double sum = 0, comp = 0;  // Kahan over sources, as pfast_shots3d_gradient
            for (int u = 0; u < R->nsrc; u++) {
                const double e = ee[R->s3->iperm[u]];
                const double y = e - comp, t = sum + y;
                comp = (t - sum) - y;
                sum = t;
            }
            misfit[0] = sum;
        }
        return PFAST_OK;
    }
    double E3[3];
    int rc = pfast_shots2d_gradient(R->s2, nullptr, grad, grad_r, nullptr, E3);
    if (!rc && misfit) {
        misfit[0] = E3[0];
        misfit[1] = E3[1];
    }
    return rc;
}

int pfast_server_field(pfast_server *S, long ticket, int src, double *T)
{
    ARGCHK(S && T, "NULL argument");
    SRequest *R = find(S, ticket);
    if (!R) return PFAST_ERR_ARG;
    ARGCHK(R->state == PFAST_REQ_DONE, "request not done");
    ARGCHK(src >= 0 && src < R->nsrc, "source index out of range");
    HIPCHK(hipSetDevice(S->device));
    if (R->s3) {
        const pfast_model3d *m = R->s3->m;
        const long N = (long)m->nrows * m->ncols * m->nz;
        HIPCHK(pfast::launch_gather_field(S->copy, R->s3->T, R->s3->iperm[src], N, 1, 1,
                                          R->s3->fieldbuf));
        HIPCHK(hipMemcpyAsync(T, R->s3->fieldbuf, N * sizeof(double), hipMemcpyDeviceToHost, S->copy));
        HIPCHK(hipStreamSynchronize(S->copy));
        return PFAST_OK;
    }
    return pfast_shots2d_field(R->s2, 0, src, 0, T);
}

int pfast_server_release(pfast_server *S, long ticket)
{
    ARGCHK(S, "NULL server");
    SRequest *R = find(S, ticket);
    if (!R) return PFAST_ERR_ARG;
    if (R->state == PFAST_REQ_QUEUED) {
        S->queue.erase(std::find(S->queue.begin(), S->queue.end(), ticket));
        S->models[R->model].refs--;
        if (R->ev) (void)hipEventDestroy(R->ev);
        S->reqs.erase(ticket);
        return PFAST_OK;
    }
    R->release = true;
    if (try_destroy(S, *R)) S->reqs.erase(ticket);
    return PFAST_OK;
}

int pfast_server_footprint(pfast_server *S, long ticket, size_t *bytes)
{
    ARGCHK(S && bytes, "NULL argument");
    SRequest *R = find(S, ticket);
    if (!R) return PFAST_ERR_ARG;
    *bytes = R->bytes;
    return PFAST_OK;
}

int pfast_server_info(pfast_server *S, size_t *in_use, int *queued, int *running, long *cycles,
                      double *node_sweeps)
{
    ARGCHK(S, "NULL server");
    if (in_use) *in_use = S->in_use;
    if (queued) *queued = (int)S->queue.size();
    if (running) {
        int n = 0;
        for (auto &kv : S->reqs) n += kv.second.state == PFAST_REQ_RUNNING;
        *running = n;
    }
    if (cycles) *cycles = S->cycles;
    if (node_sweeps) *node_sweeps = S->node_sweeps;
    return PFAST_OK;
}

}  // extern "C"
