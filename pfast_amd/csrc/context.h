// context.h -- the library context (device stream, profiling state, scratch) and
// the device-allocation helpers shared by the host translation units of libpfast.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pfast.h"
#include "pfast_internal.h"

// Caching device allocator of a request server: blocks freed by models and shot
// sets return here instead of to hipFree (which synchronises the device) and are
// handed out again for requests of the same size class.
struct DevicePool {
    std::map<size_t, std::vector<void *>> freel;  // size class -> cached blocks
    std::map<void *, size_t> sizes;               // every block of the pool
    size_t cached = 0;                            // bytes in freel
};

struct pfast_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool timing = false;
    // per timed kind (pfast::TimedKind): device ms, launches, node-sweeps of real sources
    // issued (every node of every active chunk) and swept (tiles that skipping left out
    // excluded; counted on the device while timing is on)
    double ms[5] = {0, 0, 0, 0, 0};
    long launches[5] = {0, 0, 0, 0, 0};
    double units[5] = {0, 0, 0, 0, 0};
    unsigned long long *dswept = nullptr;  // device counters, one per kind
    int sweep = 0;                // batched sweep kernel variant
    int sweep2d = 0;              // batched 2D sweep form (0 row window, 1 plane loop)
    int nmodels = 0;              // live models bound to this context
    std::vector<hipEvent_t> evpool;
    std::vector<std::pair<int, int>> pending;  // (kind, event index of start)
    size_t evnext = 0;
    // grow-only device scratch for the single-source forms
    std::map<std::string, std::pair<void *, size_t>> cache;
    DevicePool *pool = nullptr;  // set for the contexts of a request server
    // batched 3D cycles: the active chunks split over `nstreams` streams (stream and
    // aux[0..nstreams-2]) so that one stream's launch tails fill with another's work
    int nstreams = 1;
    // exact tile skipping in the column-window sweeps: 0 off, 1 whole tiles, 2 whole
    // tiles and start levels (SweepArgs::starts)
    int tile_skip = 2;
    int dataflow = 2;   // batched 3D cycles as dataflow launches: 0 never, 1 always, 2 auto
    int paired = 1;     // per-diagonal launches pair orderings 2p and 2p+1 (plan.h)
    // batched 3D phases: sweep cycles issued between two convergence checks on the host
    // (a converged chunk retires on the device at the end of its cycle either way), and
    // whether those cycles are replayed from a captured hipGraph
    int check_every = 1;
    int graphs = 0;
    // internal source order of new batched 3D shot sets (pfast_source_order): 0 the
    // caller's, 1 Morton order of the source nodes, 2 balanced clusters (chunks of
    // neighbours)
    int source_order = 2;
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> fork_join;  // [0] fork on stream, [1..] joins on aux
};

namespace pfast {

// Timed kinds: batched 3D forward / adjoint sweeps launched per tile anti-diagonal,
// 2D shot-set solves, batched 3D forward / adjoint sweeps as dataflow launches.
enum TimedKind { kFwdDiag = 0, kAdjDiag = 1, k2D = 2, kFwdFlow = 3, kAdjFlow = 4 };
constexpr int kTimedKinds = 5;

// An event pair around one launch when timing is enabled; folded into the
// totals by time_collect after a stream synchronisation.
void time_begin(pfast_ctx *ctx, int kind, hipStream_t st = nullptr);  // nullptr: ctx->stream
void time_end(pfast_ctx *ctx, hipStream_t st = nullptr);
void time_collect(pfast_ctx *ctx);

template <class T>
int dalloc(T **p, size_t count)
{
    if (hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
        set_error("hipMalloc failed (" + std::to_string(count * sizeof(T)) + " bytes)");
        return PFAST_ERR_NOMEM;
    }
    return PFAST_OK;
}

template <class T>
int dupload(T **p, const T *h, size_t count, hipStream_t st)
{
    int rc = dalloc(p, count);
    if (rc) return rc;
    if (count && h) HIPCHK(hipMemcpyAsync(*p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
    return PFAST_OK;
}

// Pool size class: powers of two up to 1 MiB, then multiples of 2 MiB.
size_t pool_round(size_t bytes);
void *pool_get(DevicePool *p, size_t bytes);  // nullptr when the device is out of memory
void pool_put(DevicePool *p, void *q);
void pool_trim(DevicePool *p, size_t keep);   // hipFree cached blocks down to `keep` bytes

// Allocation of context-owned objects (models, shot sets): the context's pool
// when it has one, else hipMalloc / hipFree.
template <class T>
int dalloc(pfast_ctx *ctx, T **p, size_t count)
{
    if (!ctx || !ctx->pool) return dalloc(p, count);
    *p = (T *)pool_get(ctx->pool, std::max<size_t>(count, 1) * sizeof(T));
    if (!*p) {
        set_error("device allocation failed (" + std::to_string(count * sizeof(T)) + " bytes)");
        return PFAST_ERR_NOMEM;
    }
    return PFAST_OK;
}

template <class T>
int dupload(pfast_ctx *ctx, T **p, const T *h, size_t count)
{
    int rc = dalloc(ctx, p, count);
    if (rc) return rc;
    if (count && h)
        HIPCHK(hipMemcpyAsync(*p, h, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return PFAST_OK;
}

inline void dfree(pfast_ctx *ctx, void *p)
{
    if (!p) return;
    if (ctx && ctx->pool)
        pool_put(ctx->pool, p);
    else
        (void)hipFree(p);
}

}  // namespace pfast
