// capi.hip -- extern "C" boundary of libpfast (see include/pfast.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pfast.h"
#include "batch3d.h"
#include "flow.h"
#include "context.h"
#include "pfast_internal.h"
#include "plain_kernels.h"
#include "plan.h"
#include "shots3d.h"

namespace pfast {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace pfast

// ------------------------------------------------------------------ context

namespace {

int ensure_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        pfast::set_error("no HIP device visible");
        return PFAST_ERR_HIP;
    }
    return PFAST_OK;
}

int scratch(pfast_ctx *ctx, const char *name, size_t bytes, void **out)
{
    auto &e = ctx->cache[name];
    if (e.second < bytes) {
        if (e.first) (void)hipFree(e.first);
        e.first = nullptr;
        e.second = 0;
        if (hipMalloc(&e.first, bytes) != hipSuccess) {
            pfast::set_error(std::string("hipMalloc failed for ") + name);
            return PFAST_ERR_NOMEM;
        }
        e.second = bytes;
    }
    *out = e.first;
    return PFAST_OK;
}

template <class T>
int scratch_t(pfast_ctx *ctx, const char *name, size_t count, T **out)
{
    void *p = nullptr;
    int rc = scratch(ctx, name, std::max<size_t>(count, 1) * sizeof(T), &p);
    *out = (T *)p;
    return rc;
}

std::mutex g_default_mu;
pfast_ctx *g_default = nullptr;

int default_ctx(pfast_ctx **out)
{
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        int rc = pfast_ctx_create(dev, &g_default);
        if (rc) return rc;
    }
    *out = g_default;
    return PFAST_OK;
}

}  // namespace

namespace pfast {
void time_begin(pfast_ctx *ctx, int kind, hipStream_t st)
{
    if (!ctx->timing) return;
    if (ctx->evnext + 2 > ctx->evpool.size()) {
        size_t add = std::max<size_t>(256, ctx->evpool.size());
        for (size_t i = 0; i < add; i++) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            ctx->evpool.push_back(e);
        }
    }
    ctx->pending.push_back({kind, (int)ctx->evnext});
    (void)hipEventRecord(ctx->evpool[ctx->evnext], st ? st : ctx->stream);
    ctx->evnext += 2;
}

void time_end(pfast_ctx *ctx, hipStream_t st)
{
    if (!ctx->timing) return;
    (void)hipEventRecord(ctx->evpool[ctx->pending.back().second + 1], st ? st : ctx->stream);
}

// After a stream sync: fold the recorded pairs into the totals.
void time_collect(pfast_ctx *ctx)
{
    if (!ctx->timing) return;
    for (auto &pr : ctx->pending) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ctx->evpool[pr.second], ctx->evpool[pr.second + 1]);
        ctx->ms[pr.first] += ms;
        ctx->launches[pr.first] += 1;
    }
    ctx->pending.clear();
    ctx->evnext = 0;
}

}  // namespace pfast

extern "C" {

const char *pfast_last_error(void) { return pfast::g_err.c_str(); }
int pfast_version(void) { return 10000; }

int pfast_device_count(int *n)
{
    ARGCHK(n, "n is NULL");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return PFAST_OK;
}

int pfast_ctx_create(int device, pfast_ctx **out)
{
    ARGCHK(out, "out is NULL");
    int rc = ensure_device();
    if (rc) return rc;
    HIPCHK(hipSetDevice(device));
    auto *c = new pfast_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        pfast::set_error("hipStreamCreate failed");
        return PFAST_ERR_HIP;
    }
    *out = c;
    return PFAST_OK;
}

int pfast_ctx_destroy(pfast_ctx *ctx)
{
    if (!ctx) return PFAST_OK;
    ARGCHK(ctx->nmodels == 0, "context still has live models: destroy them first");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->cache)
        if (e.second.first) (void)hipFree(e.second.first);
    if (ctx->dswept) (void)hipFree(ctx->dswept);
    for (auto e : ctx->evpool) (void)hipEventDestroy(e);
    for (auto e : ctx->fork_join) (void)hipEventDestroy(e);
    for (auto a : ctx->aux) {
        (void)hipStreamSynchronize(a);
        (void)hipStreamDestroy(a);
    }
    (void)hipStreamDestroy(ctx->stream);
    if (ctx == g_default) g_default = nullptr;
    delete ctx;
    return PFAST_OK;
}

int pfast_ctx_sync(pfast_ctx *ctx)
{
    ARGCHK(ctx, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    pfast::time_collect(ctx);
    return PFAST_OK;
}

int pfast_ctx_set_sweep(pfast_ctx *ctx, int variant)
{
    ARGCHK(ctx, "ctx is NULL");
    ARGCHK(variant >= 0 && variant < pfast::sweep_variant_count(), "unknown sweep variant");
    ctx->sweep = variant;
    return PFAST_OK;
}

int pfast_ctx_set_sweep2d(pfast_ctx *ctx, int variant)
{
    ARGCHK(ctx, "ctx is NULL");
    ARGCHK(variant == 0 || variant == 1, "unknown 2D sweep variant");
    ctx->sweep2d = variant;
    return PFAST_OK;
}

const char *pfast_sweep_name(int variant)
{
    if (variant < 0 || variant >= pfast::sweep_variant_count()) return nullptr;
    return pfast::sweep_variant(variant).name;
}

int pfast_ctx_set_streams(pfast_ctx *ctx, int nstreams)
{
    ARGCHK(ctx && nstreams >= 1 && nstreams <= 4, "nstreams must be 1..4");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    while ((int)ctx->aux.size() < nstreams - 1) {
        hipStream_t a;
        HIPCHK(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
        ctx->aux.push_back(a);
    }
    while ((int)ctx->fork_join.size() < nstreams) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->fork_join.push_back(e);
    }
    ctx->nstreams = nstreams;
    return PFAST_OK;
}

int pfast_ctx_set_tile_skip(pfast_ctx *ctx, int mode)
{
    ARGCHK(ctx, "ctx is NULL");
    ARGCHK(mode >= 0 && mode <= 2, "tile skip mode must be 0, 1 or 2");
    ctx->tile_skip = mode;
    return PFAST_OK;
}

int pfast_ctx_set_paired(pfast_ctx *ctx, int enable)
{
    ARGCHK(ctx, "ctx is NULL");
    ctx->paired = enable ? 1 : 0;
    return PFAST_OK;
}

int pfast_ctx_set_dataflow(pfast_ctx *ctx, int mode)
{
    ARGCHK(ctx, "ctx is NULL");
    ARGCHK(mode >= 0 && mode <= 2, "dataflow mode must be 0, 1 or 2");
    ctx->dataflow = mode;
    return PFAST_OK;
}

int pfast_ctx_set_check_interval(pfast_ctx *ctx, int cycles)
{
    ARGCHK(ctx, "ctx is NULL");
    ARGCHK(cycles >= 1 && cycles <= PFAST_CHECK_INTERVAL_MAX,
           "check interval outside 1.." + std::to_string(PFAST_CHECK_INTERVAL_MAX));
    ctx->check_every = cycles;
    return PFAST_OK;
}

int pfast_ctx_set_source_order(pfast_ctx *ctx, int mode)
{
    ARGCHK(ctx, "ctx is NULL");
    ARGCHK(mode >= 0 && mode <= 2, "source order must be 0, 1 or 2");
    ctx->source_order = mode;
    return PFAST_OK;
}

int pfast_ctx_set_graphs(pfast_ctx *ctx, int enable)
{
    ARGCHK(ctx, "ctx is NULL");
    ctx->graphs = enable ? 1 : 0;
    return PFAST_OK;
}

int pfast_ctx_timing_enable(pfast_ctx *ctx, int enable)
{
    ARGCHK(ctx, "ctx is NULL");
    if (enable && !ctx->dswept) {
        HIPCHK(hipSetDevice(ctx->device));
        int rc = pfast::dalloc(&ctx->dswept, pfast::kTimedKinds);
        if (rc) return rc;
        HIPCHK(hipMemsetAsync(ctx->dswept, 0, pfast::kTimedKinds * sizeof(unsigned long long),
                              ctx->stream));
    }
    ctx->timing = enable != 0;
    return PFAST_OK;
}

int pfast_ctx_timing_reset(pfast_ctx *ctx)
{
    ARGCHK(ctx, "ctx is NULL");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->pending.clear();
    ctx->evnext = 0;
    for (int k = 0; k < pfast::kTimedKinds; k++) {
        ctx->ms[k] = 0;
        ctx->launches[k] = 0;
        ctx->units[k] = 0;
    }
    if (ctx->dswept) {
        HIPCHK(hipMemsetAsync(ctx->dswept, 0, pfast::kTimedKinds * sizeof(unsigned long long),
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return PFAST_OK;
}

int pfast_ctx_work_swept(pfast_ctx *ctx, int kind, double *node_sweeps)
{
    ARGCHK(ctx && kind >= 0 && kind < pfast::kTimedKinds && node_sweeps, "bad argument");
    *node_sweeps = 0;
    if (!ctx->dswept) return PFAST_OK;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (auto a : ctx->aux) HIPCHK(hipStreamSynchronize(a));
    unsigned long long v = 0;
    HIPCHK(hipMemcpy(&v, ctx->dswept + kind, sizeof(v), hipMemcpyDeviceToHost));
    *node_sweeps = (double)v;
    return PFAST_OK;
}

int pfast_ctx_work(pfast_ctx *ctx, int kind, double *node_sweeps)
{
    ARGCHK(ctx && kind >= 0 && kind < pfast::kTimedKinds && node_sweeps, "bad argument");
    *node_sweeps = ctx->units[kind];
    return PFAST_OK;
}

int pfast_ctx_timing(pfast_ctx *ctx, int kind, double *ms, long *launches)
{
    ARGCHK(ctx && kind >= 0 && kind < pfast::kTimedKinds, "bad ctx/kind");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    pfast::time_collect(ctx);
    if (ms) *ms = ctx->ms[kind];
    if (launches) *launches = ctx->launches[kind];
    return PFAST_OK;
}

}  // extern "C"

// ======================================================= single-source FP64 (3D)
namespace {

// Forward FSM of one source on device buffers (vp, U, p: N doubles; flag: 1 int).
// Same cycle/convergence rule as src/T_pfsm3d.h:152.
int fsm3d_dev(pfast_ctx *ctx, const double *vp, double *U, double *p, int *flag, int si, int sj,
              int sk, int MAXSwp, double dx, double dy, double dz, int nrows, int ncols, int nz,
              int *cycles_out)
{
    hipStream_t st = ctx->stream;
    const long N = (long)nrows * ncols * nz;
    HIPCHK(pfast::launch_init3d_f64(st, U, N, ((long)sk * nrows + si) * ncols + sj));
    int count = 0;
    for (;;) {
        if (!(count < MAXSwp)) break;
        count++;
        HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), st));
        for (int o = 0; o < 8; o++) {
            const int s3 = o < 4 ? -1 : 1, s1 = (o & 2) ? 1 : -1, s2 = (o & 1) ? 1 : -1;
            pfast::time_begin(ctx, 0);
            HIPCHK(pfast::launch_fsm3d_ordering_f64(st, vp, U, p, flag, si, sj, sk, s1, s2, s3, dx,
                                                    dy, dz, nrows, ncols, nz, o == 0, o == 7));
            pfast::time_end(ctx);
        }
        int h = 0;
        HIPCHK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        pfast::time_collect(ctx);
        if (!h) break;
    }
    if (cycles_out) *cycles_out = count;
    return PFAST_OK;
}

// Adjoint of one source on device buffers, src/T_padjoint3d.h:73-172.
// T_obs/rec/nxyz point at this source's slice of the receiver table.
int adj3d_dev(pfast_ctx *ctx, double *U, const double *T, unsigned char *tag, double *p,
              int *flag, double *dEE, const double *T_obs, const int *rec, const double *nxyz,
              int nrec, int si, int sj, int sk, int MAXSwp, double dx, double dy, double dz,
              int nrows, int ncols, int nz, double *EE_out, int *cycles_out)
{
    hipStream_t st = ctx->stream;
    const long N = (long)nrows * ncols * nz;
    HIPCHK(hipMemsetAsync(U, 0, N * sizeof(double), st));
    HIPCHK(hipMemsetAsync(tag, 0, N, st));
    HIPCHK(pfast::launch_adj3d_bc_f64(st, U, tag, T, T_obs, rec, nxyz, 0, nrec, si, sj, sk, dx,
                                      dy, dz, nrows, ncols, nz, dEE, flag));
    int notconv = 0;
    double EE = 0;
    HIPCHK(hipMemcpyAsync(&notconv, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&EE, dEE, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int count = 0, swept = 0;
    for (;;) {
        if (!(count < MAXSwp)) { count++; break; }
        count++;
        if (!notconv) break;
        swept++;
        HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), st));
        for (int o = 0; o < 8; o++) {
            const int s3 = o < 4 ? -1 : 1, s1 = (o & 2) ? 1 : -1, s2 = (o & 1) ? 1 : -1;
            pfast::time_begin(ctx, 1);
            HIPCHK(pfast::launch_adj3d_ordering_f64(st, U, T, tag, p, flag, s1, s2, s3, dx, dy, dz,
                                                    nrows, ncols, nz, o == 0, o == 7));
            pfast::time_end(ctx);
        }
        HIPCHK(hipMemcpyAsync(&notconv, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        pfast::time_collect(ctx);
    }
    if (EE_out) *EE_out = EE;
    if (cycles_out) *cycles_out = swept;
    return PFAST_OK;
}

}  // namespace

extern "C" {

int pfast_fsm3d(const double *vp, double *U, const int *src, int srcid, int MAXSwp, double dx,
                double dy, double dz, int nrows, int ncols, int nz, int pe, int *ncycles)
{
    (void)pe;
    ARGCHK(vp && U && src, "NULL array");
    ARGCHK(nrows >= 2 && ncols >= 2 && nz >= 2, "grid dimensions must be >= 2");
    ARGCHK(srcid >= 0, "srcid < 0");
    const int si = src[srcid * 3], sj = src[srcid * 3 + 1], sk = src[srcid * 3 + 2];
    ARGCHK(si >= 0 && si < nrows && sj >= 0 && sj < ncols && sk >= 0 && sk < nz,
           "source node outside the grid");
    pfast_ctx *ctx;
    int rc = default_ctx(&ctx);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const long N = (long)nrows * ncols * nz;
    double *dvp, *dU, *dp;
    int *flag;
    if ((rc = scratch_t(ctx, "p3.vp", N, &dvp)) || (rc = scratch_t(ctx, "p3.U", N, &dU)) ||
        (rc = scratch_t(ctx, "p3.p", N, &dp)) || (rc = scratch_t(ctx, "p3.flag", 4, &flag)))
        return rc;
    HIPCHK(hipMemcpyAsync(dvp, vp, N * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    rc = fsm3d_dev(ctx, dvp, dU, dp, flag, si, sj, sk, MAXSwp, dx, dy, dz, nrows, ncols, nz,
                   ncycles);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(U, dU, N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return PFAST_OK;
}

}  // extern "C"

// ----------------------------------------------- allocation-free forward (3D)
// Device scratch of pfast_fsm3d_noalloc: the FP64 velocities, traveltimes and the
// convergence copy of one source plus the cycle flag, in a file-scope array that each
// GPU the library runs on holds until the process ends.  In host code `scratch` is
// only a stand-in for the symbol: its device address comes from hipGetSymbolAddress,
// once per device, and is the only form host code uses.  It has external linkage: the
// runtime resolves hipGetSymbolAddress by the name in the code object's symbol table,
// where a `static` array is a local symbol it cannot find (without -fgpu-rdc).
__device__ __align__(256) unsigned char scratch[PFAST_NOALLOC_BYTES];

namespace {

std::mutex g_noalloc_mu;                 // calls on one device share its scratch
unsigned char *g_noalloc_ptr[64] = {};   // device address of `scratch`, per device

int noalloc_scratch(int dev, unsigned char **out)
{
    if (dev < 0 || dev >= 64) {
        pfast::set_error("device index out of range for the allocation-free scratch");
        return PFAST_ERR_ARG;
    }
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        // `::scratch`: inside this namespace the unqualified name is the grow-only
        // scratch allocator above
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
        if (e != hipSuccess || !p) {
            pfast::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return PFAST_ERR_HIP;
        }
        g_noalloc_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_noalloc_ptr[dev];
    return PFAST_OK;
}

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

int pfast_fsm3d_noalloc(const double *vp, double *U, const int *src, int srcid, int MAXSwp,
                        double dx, double dy, double dz, int nrows, int ncols, int nz, int pe,
                        int *ncycles)
{
    (void)pe;
    ARGCHK(vp && U && src, "NULL array");
    ARGCHK(nrows >= 2 && ncols >= 2 && nz >= 2, "grid dimensions must be >= 2");
    ARGCHK(srcid >= 0, "srcid < 0");
    const long N = (long)nrows * ncols * nz;
    if (N > PFAST_NOALLOC_MAX_NODES) {
        pfast::set_error("grid too large for pfast_fsm3d_noalloc: " + std::to_string(N) +
                         " nodes, at most " + std::to_string((long)PFAST_NOALLOC_MAX_NODES));
        return PFAST_ERR_ARG;
    }
    const int si = src[srcid * 3], sj = src[srcid * 3 + 1], sk = src[srcid * 3 + 2];
    ARGCHK(si >= 0 && si < nrows && sj >= 0 && sj < ncols && sk >= 0 && sk < nz,
           "source node outside the grid");
    int rc = ensure_device();
    if (rc) return rc;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    unsigned char *base = nullptr;
    if ((rc = noalloc_scratch(dev, &base))) return rc;
    const size_t nb = align256((size_t)N * sizeof(double));
    double *dvp = reinterpret_cast<double *>(base);
    double *dU = reinterpret_cast<double *>(base + nb);
    double *dp = reinterpret_cast<double *>(base + 2 * nb);
    int *flag = reinterpret_cast<int *>(base + 3 * nb);
    // the null stream and no timing: this call creates no context, stream or event
    pfast_ctx local;
    local.device = dev;
    local.stream = nullptr;
    HIPCHK(hipMemcpyAsync(dvp, vp, N * sizeof(double), hipMemcpyHostToDevice, nullptr));
    rc = fsm3d_dev(&local, dvp, dU, dp, flag, si, sj, sk, MAXSwp, dx, dy, dz, nrows, ncols, nz,
                   ncycles);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(U, dU, N * sizeof(double), hipMemcpyDeviceToHost, nullptr));
    HIPCHK(hipStreamSynchronize(nullptr));
    return PFAST_OK;
}

int pfast_adjoint_fsm3d(double *U, const double *T, const double *T_obs, const int *rec,
                        const double *nxyz, const int *src, const int *srnum, int srcid,
                        int MAXSwp, double dx, double dy, double dz, int nrows, int ncols, int nz,
                        int pe, double *EE, int *ncycles)
{
    (void)pe;
    ARGCHK(U && T && src && srnum, "NULL array");
    ARGCHK(nrows >= 3 && ncols >= 3 && nz >= 3, "grid dimensions must be >= 3");
    pfast_ctx *ctx;
    int rc = default_ctx(&ctx);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const long N = (long)nrows * ncols * nz;
    long ch0 = 0;
    for (int s = 0; s < srcid; s++) ch0 += srnum[s];
    const int nrec = srnum[srcid];
    ARGCHK(nrec == 0 || (T_obs && rec && nxyz), "NULL receiver table");
    for (int r = 0; r < nrec; r++) {
        const int *q = rec + (ch0 + r) * 3;
        ARGCHK(q[0] >= 0 && q[0] < nrows && q[1] >= 0 && q[1] < ncols && q[2] >= 0 && q[2] < nz,
               "receiver node outside the grid");
    }
    double *dU, *dT, *dp, *dobs, *dn, *dEE;
    int *drec, *flag;
    unsigned char *tag;
    if ((rc = scratch_t(ctx, "a3.U", N, &dU)) || (rc = scratch_t(ctx, "a3.T", N, &dT)) ||
        (rc = scratch_t(ctx, "a3.p", N, &dp)) || (rc = scratch_t(ctx, "a3.tag", N, &tag)) ||
        (rc = scratch_t(ctx, "a3.obs", nrec, &dobs)) ||
        (rc = scratch_t(ctx, "a3.n", 3 * (size_t)nrec, &dn)) ||
        (rc = scratch_t(ctx, "a3.rec", 3 * (size_t)nrec, &drec)) ||
        (rc = scratch_t(ctx, "a3.EE", 1, &dEE)) || (rc = scratch_t(ctx, "a3.flag", 1, &flag)))
        return rc;
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync(dT, T, N * sizeof(double), hipMemcpyHostToDevice, st));
    if (nrec) {
        HIPCHK(hipMemcpyAsync(dobs, T_obs + ch0, nrec * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dn, nxyz + ch0 * 3, 3 * nrec * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(drec, rec + ch0 * 3, 3 * nrec * sizeof(int), hipMemcpyHostToDevice, st));
    }
    const int si = src[srcid * 3], sj = src[srcid * 3 + 1], sk = src[srcid * 3 + 2];
    rc = adj3d_dev(ctx, dU, dT, tag, dp, flag, dEE, dobs, drec, dn, nrec, si, sj, sk, MAXSwp, dx,
                   dy, dz, nrows, ncols, nz, EE, ncycles);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(U, dU, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PFAST_OK;
}

// ======================================================= single-source FP64 (2D)
int pfast_rt_fsm2d(const double *vp, double *U, double *TTd, double *TTu, const double *Z,
                   const int *src, const int *rec, const int *srnum, int srcid, int NumofR,
                   int MAXSwp, double dx, double dz, int nrows, int ncols, int pe, int *ncycles,
                   int *k2vec)
{
    (void)rec;
    (void)srnum;
    ARGCHK(vp && U && src, "NULL array");
    ARGCHK(NumofR >= 0 && (NumofR == 0 || (TTd && TTu && Z)), "reflector arrays");
    ARGCHK(nrows >= 2 && ncols >= 2, "grid dimensions must be >= 2");
    const int si = src[srcid * 2], sj = src[srcid * 2 + 1];
    ARGCHK(si >= 0 && si < nrows && sj >= 0 && sj < ncols, "source node outside the grid");
    pfast_ctx *ctx;
    int rc = default_ctx(&ctx);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const long N = (long)nrows * ncols;
    const int R = std::max(NumofR, 1);
    double *dvp, *dU, *dTd, *dTu, *dp, *dZ;
    int *dsrc, *dcyc, *dst, *dk2;
    if ((rc = scratch_t(ctx, "r2.vp", N, &dvp)) || (rc = scratch_t(ctx, "r2.U", N, &dU)) ||
        (rc = scratch_t(ctx, "r2.Td", N * R, &dTd)) || (rc = scratch_t(ctx, "r2.Tu", N * R, &dTu)) ||
        (rc = scratch_t(ctx, "r2.p", N, &dp)) || (rc = scratch_t(ctx, "r2.Z", (size_t)R * ncols * 3, &dZ)) ||
        (rc = scratch_t(ctx, "r2.src", 2, &dsrc)) || (rc = scratch_t(ctx, "r2.cyc", 1 + 2 * R, &dcyc)) ||
        (rc = scratch_t(ctx, "r2.st", 1, &dst)) || (rc = scratch_t(ctx, "r2.k2", ncols, &dk2)))
        return rc;
    int hs[2] = {si, sj};
    HIPCHK(hipMemcpyAsync(dvp, vp, N * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dsrc, hs, sizeof(hs), hipMemcpyHostToDevice, st));
    if (NumofR)
        HIPCHK(hipMemcpyAsync(dZ, Z, (size_t)NumofR * ncols * 3 * sizeof(double), hipMemcpyHostToDevice, st));
    Grid2 g{nrows, ncols, dx, dz, 1.0 / (dx * dx), 1.0 / (dz * dz)};
    HIPCHK(pfast::launch_rt_fsm2d_f64(st, dvp, dU, dTd, dTu, dp, dZ, dsrc, 1, NumofR, MAXSwp, g,
                                      dcyc, dst, dk2));
    int status = 0;
    std::vector<int> cyc(1 + 2 * R);
    HIPCHK(hipMemcpyAsync(&status, dst, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cyc.data(), dcyc, cyc.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(U, dU, N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (NumofR) {
        HIPCHK(hipMemcpyAsync(TTd, dTd, N * NumofR * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(TTu, dTu, N * NumofR * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    if (k2vec) HIPCHK(hipMemcpyAsync(k2vec, dk2, ncols * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (ncycles) std::copy(cyc.begin(), cyc.begin() + 1 + 2 * NumofR, ncycles);
    if (status == 1) {
        pfast::set_error("no k2 above reflector inside model (src/RT_pfsm2d.h:232)");
        return PFAST_ERR_FATAL;
    }
    if (status == 2) {
        pfast::set_error("negative reflected traveltime (src/RT_pfsm2d.h:225)");
        return PFAST_ERR_FATAL;
    }
    return PFAST_OK;
}

static int adj2d_common(bool refl, double *Uout, const double *T, const double *obs, double Cr,
                        const int *rec, const double *nxyz, const int *src, const int *srnum,
                        int srcid, int NumofR, int MAXSwp, double dx, double dz, int nrows,
                        int ncols, double *EE, int *ncycles)
{
    ARGCHK(Uout && T && src && srnum, "NULL array");
    ARGCHK(nrows >= 3 && ncols >= 3, "grid dimensions must be >= 3");
    pfast_ctx *ctx;
    int rc = default_ctx(&ctx);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const long N = (long)nrows * ncols;
    long ch0 = 0;
    for (int s = 0; s < srcid; s++) ch0 += srnum[s];
    const int nrec = srnum[srcid];
    const int R = refl ? std::max(NumofR, 1) : 1;
    const int ostride = refl ? NumofR : 1;
    ARGCHK(nrec == 0 || (obs && rec && nxyz), "NULL receiver table");
    for (int r = 0; r < nrec; r++) {
        const int *q = rec + (ch0 + r) * 2;
        ARGCHK(q[0] >= 0 && q[0] < nrows && q[1] >= 0 && q[1] < ncols, "receiver node outside the grid");
    }
    double *dU, *dU2, *dT, *dp, *dobs, *dn, *dEE;
    int *drec, *dsrc, *dsr, *dcyc;
    long *dch0;
    unsigned char *tag;
    if ((rc = scratch_t(ctx, "j2.U", N, &dU)) || (rc = scratch_t(ctx, "j2.U2", N, &dU2)) ||
        (rc = scratch_t(ctx, "j2.T", N * R, &dT)) || (rc = scratch_t(ctx, "j2.p", N, &dp)) ||
        (rc = scratch_t(ctx, "j2.tag", N, &tag)) ||
        (rc = scratch_t(ctx, "j2.obs", (size_t)nrec * ostride, &dobs)) ||
        (rc = scratch_t(ctx, "j2.n", 2 * (size_t)nrec, &dn)) ||
        (rc = scratch_t(ctx, "j2.rec", 2 * (size_t)nrec, &drec)) ||
        (rc = scratch_t(ctx, "j2.src", 2, &dsrc)) || (rc = scratch_t(ctx, "j2.sr", 1, &dsr)) ||
        (rc = scratch_t(ctx, "j2.ch0", 1, &dch0)) || (rc = scratch_t(ctx, "j2.EE", 1, &dEE)) ||
        (rc = scratch_t(ctx, "j2.cyc", R, &dcyc)))
        return rc;
    const int hs[2] = {src[srcid * 2], src[srcid * 2 + 1]};
    const long zero = 0;
    HIPCHK(hipMemcpyAsync(dT, T, N * (refl ? NumofR : 1) * sizeof(double), hipMemcpyHostToDevice, st));
    if (nrec) {
        HIPCHK(hipMemcpyAsync(dobs, obs + ch0 * ostride, (size_t)nrec * ostride * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dn, nxyz + ch0 * 2, 2 * nrec * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(drec, rec + ch0 * 2, 2 * nrec * sizeof(int), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(dsrc, hs, sizeof(hs), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dsr, &nrec, sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dch0, &zero, sizeof(long), hipMemcpyHostToDevice, st));
    Grid2 g{nrows, ncols, dx, dz, 1.0 / (dx * dx), 1.0 / (dz * dz)};
    if (refl)
        HIPCHK(pfast::launch_radj2d_f64(st, dU, dU2, dp, tag, dT, dobs, Cr, drec, dn, dch0, dsr,
                                        dsrc, 1, NumofR, MAXSwp, g, dEE, dcyc));
    else
        HIPCHK(pfast::launch_tadj2d_f64(st, dU, dp, tag, dT, dobs, Cr, drec, dn, dch0, dsr, dsrc,
                                        1, MAXSwp, g, dEE, dcyc));
    std::vector<int> cyc(R);
    double hEE = 0;
    HIPCHK(hipMemcpyAsync(Uout, dU, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&hEE, dEE, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cyc.data(), dcyc, R * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (EE) *EE = hEE;
    if (ncycles) std::copy(cyc.begin(), cyc.begin() + (refl ? NumofR : 1), ncycles);
    return PFAST_OK;
}

int pfast_tadjoint_fsm2d(double *U, const double *T, const double *T_obs, double Cr,
                         const int *rec, const double *nxyz, const int *src, const int *srnum,
                         int srcid, int MAXSwp, double dx, double dz, int nrows, int ncols, int pe,
                         double *EE, int *ncycles)
{
    (void)pe;
    return adj2d_common(false, U, T, T_obs, Cr, rec, nxyz, src, srnum, srcid, 0, MAXSwp, dx, dz,
                        nrows, ncols, EE, ncycles);
}

int pfast_radjoint_fsm2d(double *UR, const double *Td, const double *Tu, const double *TR_obs,
                         double Cr, const int *rec, const double *nxyz, const int *src,
                         const int *srnum, int srcid, int NumofR, int MAXSwp, double dx,
                         double dz, int nrows, int ncols, int pe, double *EE, int *ncycles)
{
    (void)pe;
    (void)Td;
    if (NumofR <= 0) {
        ARGCHK(UR, "NULL array");
        std::fill(UR, UR + (long)nrows * ncols, 0.0);
        if (EE) *EE = 0.0;
        return PFAST_OK;
    }
    return adj2d_common(true, UR, Tu, TR_obs, Cr, rec, nxyz, src, srnum, srcid, NumofR, MAXSwp,
                        dx, dz, nrows, ncols, EE, ncycles);
}

}  // extern "C"

// ======================================================= batched 3D
namespace {

// One batch of sweep cycles of a phase on the model's stream: per cycle, the cycle's
// start (cycle number, change bits), the sweeps of every chunk in the device list(s)
// (dataflow launches, or per tile anti-diagonal split over the context's streams as
// `off` says), and the cycle's end (converged sources, retired chunks), all reading the
// cycle number and the live chunks from the device phase state.
int issue_batch3d(pfast_shots3d *s, bool adjoint, bool flow, int ncycles, int nlist,
                  const std::vector<int> &off, const pfast::TileSkip *sp)
{
    pfast_model3d *m = s->m;
    pfast_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    const pfast::Grid3 g = pfast::grid_of(m);
    const int K = (int)off.size() - 1;
    for (int c = 0; c < ncycles; c++) {
        HIPCHK(pfast::launch_cycle_begin(st, s->phase, s->changed, s->nchunk));
        if (flow) {
            int rc = pfast::issue_cycle3d_flow(ctx, g, s->chunks, s->active, nlist, adjoint,
                                               s->changed, s->flow, s->nchunk, 0, st, sp, s->phase);
            if (rc) return rc;
        } else {
            if (K > 1) {
                HIPCHK(hipEventRecord(ctx->fork_join[0], st));
                for (int k = 1; k < K; k++) HIPCHK(hipStreamWaitEvent(ctx->aux[k - 1], ctx->fork_join[0], 0));
            }
            for (int k = 0; k < K; k++) {
                hipStream_t sk = k ? ctx->aux[k - 1] : st;
                const int *lk = s->active + off[k];
                const int nk = off[k + 1] - off[k];
                int rc = adjoint ? pfast::issue_cycle3d(ctx, g, s->chunks, nullptr, 0, lk, nk, s->changed, sk, sp, s->phase)
                                 : pfast::issue_cycle3d(ctx, g, s->chunks, lk, nk, nullptr, 0, s->changed, sk, sp, s->phase);
                if (rc) return rc;
            }
            for (int k = 1; k < K; k++) {
                HIPCHK(hipEventRecord(ctx->fork_join[k], ctx->aux[k - 1]));
                HIPCHK(hipStreamWaitEvent(st, ctx->fork_join[k], 0));
            }
        }
        HIPCHK(pfast::launch_cycle_end(st, s->phase, s->changed, s->nchunk, s->nsrc));
    }
    return PFAST_OK;
}

constexpr size_t kMaxGraphs = 16;  // captured batches kept per shot set

void drop_graphs(pfast_shots3d *s)
{
    for (auto &gr : s->graphs) (void)hipGraphExecDestroy(gr.exec);
    s->graphs.clear();
}

// The batch as a hipGraph: captured on first use of its key (everything its launch
// arguments depend on), replayed afterwards.
int launch_batch3d_graph(pfast_shots3d *s, bool adjoint, bool flow, int ncycles, int nlist,
                         const std::vector<int> &off, const std::vector<int> &lists,
                         const pfast::TileSkip *sp)
{
    pfast_ctx *ctx = s->m->ctx;
    hipStream_t st = ctx->stream;
    std::vector<long> key = {adjoint, flow, ncycles, nlist, ctx->sweep, ctx->tile_skip,
                             sp != nullptr, (long)(size_t)s->flow, (long)(size_t)s->tmod};
    key.insert(key.end(), off.begin(), off.end());
    key.insert(key.end(), lists.begin(), lists.end());
    for (auto &gr : s->graphs)
        if (gr.key == key) {
            HIPCHK(hipGraphLaunch(gr.exec, st));
            return PFAST_OK;
        }
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = issue_batch3d(s, adjoint, flow, ncycles, nlist, off, sp);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    HIPCHK(e);
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(ei);
    if (s->graphs.size() >= kMaxGraphs) {
        (void)hipGraphExecDestroy(s->graphs.front().exec);
        s->graphs.erase(s->graphs.begin());
    }
    s->graphs.push_back({std::move(key), exec});
    HIPCHK(hipGraphLaunch(exec, st));
    return PFAST_OK;
}

// Sweep cycles of the batched FP32 solver until every source is stationary.  The host
// checks convergence every ctx->check_every cycles; in between, a chunk that has
// converged retires on the device at the end of its cycle (its later launches return
// at once), so the fields, cycle counts and results do not depend on the interval.
int run_sweeps3d(pfast_shots3d *s, bool adjoint, int MAXSwp, std::vector<int> &cyc)
{
    pfast_model3d *m = s->m;
    pfast_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    if (!s->phase) {
        int rc = pfast::dalloc(ctx, &s->phase, (size_t)2 + s->nchunk + s->nsrc);
        if (rc) return rc;
    }
    if (!s->hphase && hipHostMalloc((void **)&s->hphase, ((size_t)2 + s->nchunk + s->nsrc) * sizeof(int)) != hipSuccess) {
        s->hphase = nullptr;
        pfast::set_error("hipHostMalloc failed");
        return PFAST_ERR_NOMEM;
    }
    std::vector<int> act(s->nchunk);
    for (int i = 0; i < s->nchunk; i++) act[i] = i;
    cyc.assign(s->nsrc, 0);
    int cycle = 0;
    // tile skip records of this phase (column-window variants)
    pfast::TileSkip skip{nullptr, nullptr, 0};
    const SweepVariant &sv = pfast::sweep_variant(ctx->sweep);
    if (ctx->tile_skip && sv.kind == 0) {
        const int ntile = ((m->nrows + sv.ti - 1) / sv.ti) * ((m->ncols + sv.tj - 1) / sv.tj);
        if (ntile != s->ntile) {
            drop_graphs(s);
            pfast::dfree(ctx, s->tmod);
            pfast::dfree(ctx, s->tclean);
            s->tmod = s->tclean = nullptr;
            s->ntile = 0;
            int rc;
            if ((rc = pfast::dalloc(ctx, &s->tmod, (size_t)s->nchunk * ntile * 5)) ||
                (rc = pfast::dalloc(ctx, &s->tclean, (size_t)s->nchunk * ntile * kTileClean)))
                return rc;
            s->ntile = ntile;
        }
        HIPCHK(hipMemsetAsync(s->tmod, 0, (size_t)s->nchunk * ntile * 5 * sizeof(int), st));
        HIPCHK(hipMemsetAsync(s->tclean, 0, (size_t)s->nchunk * ntile * kTileClean * sizeof(int), st));
        skip.tmod = s->tmod;
        skip.tclean = s->tclean;
    }
    const pfast::TileSkip *sp = skip.tmod ? &skip : nullptr;
    HIPCHK(pfast::launch_phase_init(st, s->phase, s->nchunk, s->nsrc));
    const int ftile = ((m->nrows + sv.ti - 1) / sv.ti) * ((m->ncols + sv.tj - 1) / sv.tj);
    const size_t fwords = pfast::flow_words(s->nchunk, ftile);
    bool flow_ready = false;  // dataflow state allocated and reset for this phase
    std::vector<int> lists, off;
    while (!act.empty() && cycle < MAXSwp) {
        const int ncyc = std::max(1, std::min(ctx->check_every, MAXSwp - cycle));
        long real = 0;
        for (int ch : act) real += std::min(PFAST_CHUNK, s->nsrc - ch * PFAST_CHUNK);
        // both schedules update every node in the same order: the choice is per batch;
        // dataflow launch epochs come from the cycle number (16 bits, two per cycle)
        const bool flow = pfast::use_dataflow(ctx, (int)act.size(), ftile, fwords) &&
                          cycle + ncyc <= 0x7FFE;
        const int kind = flow ? (adjoint ? pfast::kAdjFlow : pfast::kFwdFlow)
                              : (adjoint ? pfast::kAdjDiag : pfast::kFwdDiag);
        ctx->units[kind] += 8.0 * ncyc * (double)((long)m->nrows * m->ncols * m->nz) * real;
        if (flow && !flow_ready) {
            if (fwords != s->nflow) {
                drop_graphs(s);
                pfast::dfree(ctx, s->flow);
                s->flow = nullptr;
                s->nflow = 0;
                int rc = pfast::dalloc(ctx, &s->flow, fwords);
                if (rc) return rc;
                s->nflow = fwords;
            }
            HIPCHK(pfast::launch_flow_reset(st, s->flow, fwords));
            flow_ready = true;
        }
        // the live chunks: one list for the dataflow launches, else round-robin over the
        // context's streams (one list each)
        const int K = flow ? 1 : std::max(1, std::min(ctx->nstreams, (int)act.size()));
        lists.clear();
        off.assign(K + 1, 0);
        for (int k = 0; k < K; k++) {
            for (size_t i = k; i < act.size(); i += K) lists.push_back(act[i]);
            off[k + 1] = (int)lists.size();
        }
        HIPCHK(hipMemcpyAsync(s->active, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice, st));
        int rc = ctx->graphs && !ctx->timing
                     ? launch_batch3d_graph(s, adjoint, flow, ncyc, (int)act.size(), off, lists, sp)
                     : issue_batch3d(s, adjoint, flow, ncyc, (int)act.size(), off, sp);
        if (rc) return rc;
        cycle += ncyc;
        HIPCHK(hipMemcpyAsync(s->hphase, s->phase, (2 + (size_t)s->nchunk) * sizeof(int), hipMemcpyDeviceToHost, st));
        int aborted = 0;
        if (flow)
            HIPCHK(hipMemcpyAsync(&aborted, s->flow + fwords - 2, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        pfast::time_collect(ctx);
        if (aborted) {
            pfast::set_error("dataflow sweep: a dependency wait timed out (results void)");
            return PFAST_ERR_HIP;
        }
        std::vector<int> next;
        for (int ch : act)
            if (s->hphase[2 + ch]) next.push_back(ch);
        act.swap(next);
    }
    HIPCHK(hipMemcpyAsync(s->hphase, s->phase, ((size_t)2 + s->nchunk + s->nsrc) * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int i = 0; i < s->nsrc; i++) {
        const int c = s->hphase[2 + s->nchunk + i];
        cyc[i] = c ? c : cycle;
    }
    return PFAST_OK;
}

}  // namespace

namespace pfast {

int diagonals3d(const pfast_ctx *ctx, const Grid3 &g)
{
    const SweepVariant &sv = sweep_variant(ctx->sweep);
    return (g.nrows + sv.ti - 1) / sv.ti + (g.ncols + sv.tj - 1) / sv.tj - 1;
}

namespace {
// Launch arguments common to the sweep schedules of one grid under ctx's variant.
SweepArgs sweep_args(const pfast_ctx *ctx, const Grid3 &g, const ChunkRef *chunks,
                     unsigned *changed, const TileSkip *skip)
{
    const SweepVariant &sv = sweep_variant(ctx->sweep);
    SweepArgs a{};
    a.chunks = chunks;
    a.changed = changed;
    a.nrows = g.nrows;
    a.ncols = g.ncols;
    a.nz = g.nz;
    a.nti = (g.nrows + sv.ti - 1) / sv.ti;
    a.ntj = (g.ncols + sv.tj - 1) / sv.tj;
    a.hx = (float)g.dx;
    a.hy = (float)g.dy;
    a.hz = (float)g.dz;
    a.wx = (float)(1.0 / (g.dx * g.dx));
    a.wy = (float)(1.0 / (g.dy * g.dy));
    a.wz = (float)(1.0 / (g.dz * g.dz));
    // the largest t with fl(t * wx) <= 1e-9 (the product is monotone in t)
    float thr = 1e-9f / a.wx;
    while (thr > 0.0f && thr * a.wx > 1e-9f) thr = std::nextafter(thr, 0.0f);
    while (std::nextafter(thr, INFINITY) * a.wx <= 1e-9f) thr = std::nextafter(thr, INFINITY);
    a.adj_thr = thr;
    const bool skipping = skip && sv.kind == 0;
    a.tmod = skipping ? skip->tmod : nullptr;
    a.tclean = skipping ? skip->tclean : nullptr;
    a.starts = skipping && ctx->tile_skip == 2 ? 1 : 0;
    a.ntA = 0x7fffffff;
    return a;
}

// Wall-clock ticks a dataflow dependency wait may take before the launch gives up
// (30 s: far beyond any wait of a correct schedule).
unsigned long long flow_timeout_ticks(int device)
{
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
        khz <= 0)
        khz = 100000;
    return (unsigned long long)khz * 1000ull * 30ull;
}
}  // namespace

bool use_dataflow(const pfast_ctx *ctx, int nact, int ntile, size_t words)
{
    if (ctx->dataflow == 0 || sweep_variant(ctx->sweep).kind != 0) return false;
    // the dataflow kernels address a state word by a 26-bit index
    if (words >= (size_t(1) << 26)) return false;
    if (ctx->dataflow == 1) return true;
    return (long)nact * ntile <= kAutoDataflowTasks;
}

int issue_cycle3d_flow(pfast_ctx *ctx, const Grid3 &g, const ChunkRef *chunks, const int *act,
                       int nact, bool adjoint, unsigned *changed, unsigned *flow, int nids,
                       int epoch, hipStream_t st, const TileSkip *skip, const int *dphase)
{
    if (!st) st = ctx->stream;
    if (nact <= 0) return PFAST_OK;
    SweepArgs a = sweep_args(ctx, g, chunks, changed, skip);
    a.swept = ctx->timing && ctx->dswept ? ctx->dswept + (adjoint ? kAdjFlow : kFwdFlow) : nullptr;
    a.active = act;
    a.nact = nact;
    a.nact_p8 = nact;
    const int ntile = a.nti * a.ntj;
    const size_t nprog = flow_words(nids, ntile) - 4;
    FlowArgs f{};
    f.prog = flow;
    f.nprog = (unsigned)nprog;
    f.ticket = flow + nprog;
    f.abort = reinterpret_cast<int *>(flow + nprog + 2);
    f.ntasks = 4u * (unsigned)ntile * (unsigned)nact;
    f.cycle = skip && !dphase ? skip->cycle : 0;
    f.nd = a.nti + a.ntj - 1;
    f.timeout = flow_timeout_ticks(ctx->device);
    if (dphase) {
        f.phase = dphase;
        a.live = dphase + 2;
    }
    for (int half = 0; half < 2; half++) {  // orderings 0-3, then 4-7
        f.o0 = 4 * half;
        // with dphase the kernel adds (2 * cycle) << 16 (unsigned arithmetic wraps)
        f.base = (unsigned)((dphase ? 0 : 2 * epoch) - 1 + half) << pflow::kEpochShift;
        time_begin(ctx, adjoint ? kAdjFlow : kFwdFlow, st);
        HIPCHK(launch_sweep3d_flow(st, a, f, ctx->sweep, adjoint ? 1 : 0, g.uniform));
        time_end(ctx, st);
    }
    // the dataflow kernels do not keep the start levels: sweep these chunks' tiles from
    // their first level in later per-diagonal launches
    if (a.starts) HIPCHK(launch_starts_reset(st, a.tclean, act, nact, ntile));
    return PFAST_OK;
}

int group_orderings3d(const pfast_ctx *ctx) { return plan_group_orderings(ctx->paired != 0); }

int group_steps3d(const pfast_ctx *ctx, const Grid3 &g)
{
    const SweepVariant &sv = sweep_variant(ctx->sweep);
    return plan_group_steps(ctx->paired != 0, (g.nrows + sv.ti - 1) / sv.ti,
                            (g.ncols + sv.tj - 1) / sv.tj);
}

int issue_steps3d(pfast_ctx *ctx, const Grid3 &g, const ChunkRef *chunks, const int *fwd,
                  int nfwd, const int *adj, int nadj, unsigned *changed, int o, int k0, int k1,
                  hipStream_t st, const TileSkip *skip, const int *dphase)
{
    if (!st) st = ctx->stream;
    SweepArgs a = sweep_args(ctx, g, chunks, changed, skip);
    const int nti = a.nti, ntj = a.ntj;
    const bool skipping = a.tmod != nullptr;
    const bool paired = ctx->paired != 0;
    if (dphase) {
        a.phase = dphase;
        a.live = dphase + 2;
        a.ep_cycle = 8 * (nti + ntj - 1);
    }
    // with dphase the kernels add (cycle - 1) * ep_cycle to the epoch
    const int ecyc = dphase || !skipping ? 0 : skip->cycle;
    SweepArgs af = a, aa = a;
    if (ctx->timing && ctx->dswept) {
        af.swept = ctx->dswept + kFwdDiag;
        aa.swept = ctx->dswept + kAdjDiag;
    }
    af.active = fwd;
    af.nact = nfwd;
    af.nact_p8 = (nfwd + 7) / 8 * 8;
    aa.active = adj;
    aa.nact = nadj;
    aa.nact_p8 = (nadj + 7) / 8 * 8;
    const int nd = nti + ntj - 1, lag = pair_lag(nti, ntj);
    auto epoch = [&](int oo, int D) { return skipping ? 1 + ((ecyc * 8 + oo) * nd + D) : 0; };
    for (int k = k0; k < k1; k++) {
        const PlanStep p = plan_step(paired, nd, lag, o, k);
        if (p.nparts == 0) continue;
        const int o0 = p.o[0], D = p.D[0];
        const int ntA = plan_diag_tiles(nti, ntj, D);
        int nt = ntA;
        af.o = aa.o = o0;
        af.epoch = aa.epoch = epoch(o0, D);
        af.ntA = aa.ntA = 0x7fffffff;
        if (p.nparts == 2) {
            af.ntA = aa.ntA = ntA;
            af.DB = aa.DB = p.D[1];
            af.oB = aa.oB = p.o[1];
            af.epochB = aa.epochB = epoch(p.o[1], p.D[1]);
            nt += plan_diag_tiles(nti, ntj, p.D[1]);
        }
        const int s1 = ord_s1(o0), s2 = ord_s2(o0), s3 = ord_s3(o0);
        if (nfwd) {
            time_begin(ctx, kFwdDiag, st);
            HIPCHK(launch_sweep3d(st, af, ctx->sweep, 0, g.uniform, D, s1, s2, s3, nt));
            time_end(ctx, st);
        }
        if (nadj) {
            time_begin(ctx, kAdjDiag, st);
            HIPCHK(launch_sweep3d(st, aa, ctx->sweep, 1, g.uniform, D, s1, s2, s3, nt));
            time_end(ctx, st);
        }
    }
    return PFAST_OK;
}

int issue_cycle3d(pfast_ctx *ctx, const Grid3 &g, const ChunkRef *chunks, const int *fwd,
                  int nfwd, const int *adj, int nadj, unsigned *changed, hipStream_t st,
                  const TileSkip *skip, const int *dphase)
{
    const int ns = group_steps3d(ctx, g), go = group_orderings3d(ctx);
    for (int o = 0; o < 8; o += go) {
        int rc = issue_steps3d(ctx, g, chunks, fwd, nfwd, adj, nadj, changed, o, 0, ns, st, skip,
                               dphase);
        if (rc) return rc;
    }
    return PFAST_OK;
}

int fwd_begin3d(pfast_shots3d *s)
{
    const long N = (long)s->m->nrows * s->m->ncols * s->m->nz;
    pfast_model3d *m = s->m;
    HIPCHK(launch_init_fwd(m->ctx->stream, s->T, s->srcnode, N, s->nchunk));
    HIPCHK(launch_init_src_box(m->ctx->stream, s->T, m->vp, s->srcnode, N, s->nchunk, s->srcinit,
                               m->nrows, m->ncols, m->nz, m->dx, m->dy, m->dz));
    return PFAST_OK;
}

int adj_begin3d(pfast_shots3d *s)
{
    pfast_model3d *m = s->m;
    hipStream_t st = m->ctx->stream;
    const long N = (long)m->nrows * m->ncols * m->nz;
    HIPCHK(launch_adj_misfit(st, s->T, s->pnode, s->pobs, s->poff, s->nsrc, N, s->EE));
    // marks first, so that the traveltime copies of the (lambda, T) field carry them;
    // the boundary values read |T|
    HIPCHK(launch_adj_mark(st, s->T, s->pnode, s->psrc, s->npick, N, 0));
    HIPCHK(launch_adj_pack(st, s->L, s->T, N, s->nchunk));
    HIPCHK(launch_adj_bc(st, s->L, s->T, s->bnode, s->bsrc, s->bobs, s->bnrm, s->nb, m->nrows,
                         m->ncols, m->nz, m->dx, m->dy, m->dz));
    return PFAST_OK;
}

int adj_end3d(pfast_shots3d *s)
{
    pfast_model3d *m = s->m;
    hipStream_t st = m->ctx->stream;
    const long N = (long)m->nrows * m->ncols * m->nz;
    HIPCHK(launch_adj_mark(st, s->T, s->pnode, s->psrc, s->npick, N, 1));
    HIPCHK(launch_grad_reduce(st, s->L, m->vp, s->grad, N, s->nchunk, s->nsrc));
    return PFAST_OK;
}

size_t pool_round(size_t bytes)
{
    if (bytes <= (size_t(1) << 20)) {
        size_t r = 256;
        while (r < bytes) r <<= 1;
        return r;
    }
    const size_t g = size_t(2) << 20;
    return (bytes + g - 1) / g * g;
}

void *pool_get(DevicePool *p, size_t bytes)
{
    const size_t b = pool_round(bytes);
    auto it = p->freel.find(b);
    if (it != p->freel.end() && !it->second.empty()) {
        void *q = it->second.back();
        it->second.pop_back();
        p->cached -= b;
        return q;
    }
    void *q = nullptr;
    if (hipMalloc(&q, b) != hipSuccess) {
        (void)hipGetLastError();
        pool_trim(p, 0);
        if (hipMalloc(&q, b) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    p->sizes[q] = b;
    return q;
}

void pool_put(DevicePool *p, void *q)
{
    auto it = p->sizes.find(q);
    if (it == p->sizes.end()) {
        (void)hipFree(q);
        return;
    }
    p->freel[it->second].push_back(q);
    p->cached += it->second;
}

void pool_trim(DevicePool *p, size_t keep)
{
    for (auto it = p->freel.rbegin(); it != p->freel.rend() && p->cached > keep; ++it)
        while (!it->second.empty() && p->cached > keep) {
            void *q = it->second.back();
            it->second.pop_back();
            (void)hipFree(q);
            p->sizes.erase(q);
            p->cached -= it->first;
        }
}

}  // namespace pfast

extern "C" {

int pfast_model3d_create(pfast_ctx *ctx, const double *vp, int nrows, int ncols, int nz,
                         double dx, double dy, double dz, pfast_model3d **out)
{
    ARGCHK(ctx && vp && out, "NULL argument");
    ARGCHK(nrows >= 3 && ncols >= 3 && nz >= 3, "grid dimensions must be >= 3");
    ARGCHK(dx > 0 && dy > 0 && dz > 0, "spacings must be positive");
    HIPCHK(hipSetDevice(ctx->device));
    auto *m = new pfast_model3d();
    m->ctx = ctx;
    ctx->nmodels++;  // undone by pfast_model3d_destroy, also on the failure paths below
    m->nrows = nrows;
    m->ncols = ncols;
    m->nz = nz;
    m->dx = dx;
    m->dy = dy;
    m->dz = dz;
    m->uniform = dx == dy && dy == dz;
    const long N = (long)nrows * ncols * nz;
    int rc;
    if ((rc = pfast::dalloc(ctx, &m->vp, N)) || (rc = pfast::dalloc(ctx, &m->f, N)) ||
        (rc = pfast::dalloc(ctx, &m->ft, pfast::tile8_size(nrows, ncols, nz)))) {
        pfast_model3d_destroy(m);
        return rc;
    }
    rc = pfast_model3d_set_vp(m, vp);
    if (rc) {
        pfast_model3d_destroy(m);
        return rc;
    }
    *out = m;
    return PFAST_OK;
}

int pfast_model3d_set_vp(pfast_model3d *m, const double *vp)
{
    ARGCHK(m && vp, "NULL argument");
    const long N = (long)m->nrows * m->ncols * m->nz;
    hipStream_t st = m->ctx->stream;
    HIPCHK(hipMemcpyAsync(m->vp, vp, N * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(pfast::launch_cell_time(st, m->vp, m->f, N, m->dx, m->uniform));
    HIPCHK(pfast::launch_tile8_copy(st, m->f, m->ft, m->nrows, m->ncols, m->nz));
    HIPCHK(hipStreamSynchronize(st));
    return PFAST_OK;
}

int pfast_model3d_destroy(pfast_model3d *m)
{
    if (!m) return PFAST_OK;
    ARGCHK(m->nshots == 0, "model still has live shot sets: destroy them first");
    (void)hipStreamSynchronize(m->ctx->stream);
    m->ctx->nmodels--;
    pfast::dfree(m->ctx, m->vp);
    pfast::dfree(m->ctx, m->f);
    pfast::dfree(m->ctx, m->ft);
    delete m;
    return PFAST_OK;
}

int pfast_shots3d_destroy(pfast_shots3d *s)
{
    if (!s) return PFAST_OK;
    (void)hipStreamSynchronize(s->m->ctx->stream);
    pfast::shots3d_free(s);
    return PFAST_OK;
}

}  // extern "C"

namespace pfast {

void shots3d_free(pfast_shots3d *s)
{
    s->m->nshots--;
    void *ptrs[] = {s->srcnode, s->T, s->L, s->rnode, s->rsrc, s->rec3, s->nxyz, s->Tobs,
                    s->pnode, s->psrc, s->pobs, s->poff, s->bnode, s->bsrc, s->bobs, s->bnrm,
                    s->EE, s->grad, s->picks, s->fieldbuf, s->changed, s->active, s->chunks,
                    s->tmod, s->tclean, s->flow, s->phase};
    drop_graphs(s);
    for (void *p : ptrs) pfast::dfree(s->m->ctx, p);
    if (s->hphase) (void)hipHostFree(s->hphase);
    delete s;
}

}  // namespace pfast

extern "C" int pfast_source_order(int nsrc, const int *src, double dx, double dy, double dz,
                                  int mode, int *perm)
{
    ARGCHK(nsrc >= 0 && (nsrc == 0 || (src && perm)), "bad argument");
    ARGCHK(mode >= 0 && mode <= 2, "source order must be 0, 1 or 2");
    for (int i = 0; i < nsrc; i++) perm[i] = i;
    if (mode == 0 || nsrc <= 1) return PFAST_OK;
    // Morton key of the source node (row, column, depth), 21 bits each
    auto spread = [](unsigned v) {  // bit b of v -> bit 3b
        unsigned long long x = v & 0x1FFFFFu;
        x = (x | x << 32) & 0x1F00000000FFFFull;
        x = (x | x << 16) & 0x1F0000FF0000FFull;
        x = (x | x << 8) & 0x100F00F00F00F00Full;
        x = (x | x << 4) & 0x10C30C30C30C30C3ull;
        x = (x | x << 2) & 0x1249249249249249ull;
        return x;
    };
    std::vector<unsigned long long> key(nsrc);
    for (int i = 0; i < nsrc; i++)
        key[i] = spread((unsigned)std::max(src[3 * i], 0)) << 2 |
                 spread((unsigned)std::max(src[3 * i + 1], 0)) << 1 |
                 spread((unsigned)std::max(src[3 * i + 2], 0));
    std::stable_sort(perm, perm + nsrc, [&](int a, int b) { return key[a] < key[b]; });
    const int C = PFAST_CHUNK, k = (nsrc + C - 1) / C;
    if (mode == 1 || k < 2 || (double)nsrc * k > 1e8) return PFAST_OK;
    // mode 2: balanced clusters of C sources (the last one holds the remainder), by
    // Lloyd iterations from the Morton chunks with a greedy capacity-bounded assignment
    // (nearest pairs first); deterministic
    std::vector<double> P(3 * (size_t)nsrc), cen(3 * (size_t)k);
    for (int i = 0; i < nsrc; i++) {
        P[3 * i] = src[3 * i] * dy;
        P[3 * i + 1] = src[3 * i + 1] * dx;
        P[3 * i + 2] = src[3 * i + 2] * dz;
    }
    std::vector<int> lab(nsrc);
    for (int p = 0; p < nsrc; p++) lab[perm[p]] = p / C;
    auto centroids = [&]() {
        std::vector<int> n(k, 0);
        std::fill(cen.begin(), cen.end(), 0.0);
        for (int i = 0; i < nsrc; i++) {
            n[lab[i]]++;
            for (int d = 0; d < 3; d++) cen[3 * lab[i] + d] += P[3 * i + d];
        }
        for (int c = 0; c < k; c++)
            for (int d = 0; d < 3; d++) cen[3 * c + d] /= std::max(n[c], 1);
    };
    // each source competes for its M nearest centroids; one left without a place
    // (rare) takes the nearest cluster with room
    const int M = std::min(k, 16);
    std::vector<std::pair<double, long>> pairs, near((size_t)k);
    for (int it = 0; it < 12; it++) {
        centroids();
        pairs.clear();
        for (int i = 0; i < nsrc; i++) {
            for (int c = 0; c < k; c++) {
                double d2 = 0;
                for (int d = 0; d < 3; d++) {
                    const double t = P[3 * i + d] - cen[3 * c + d];
                    d2 += t * t;
                }
                near[c] = {d2, (long)i * k + c};
            }
            std::nth_element(near.begin(), near.begin() + (M - 1), near.end());
            pairs.insert(pairs.end(), near.begin(), near.begin() + M);
        }
        std::sort(pairs.begin(), pairs.end());
        std::vector<int> cap(k, C), nl(nsrc, -1);
        cap[k - 1] = nsrc - (k - 1) * C;
        for (const auto &q : pairs) {
            const int i = (int)(q.second / k), c = (int)(q.second % k);
            if (nl[i] < 0 && cap[c] > 0) {
                nl[i] = c;
                cap[c]--;
            }
        }
        for (int i = 0; i < nsrc; i++) {
            if (nl[i] >= 0) continue;
            double best = 0;
            for (int c = 0; c < k; c++) {
                if (!cap[c]) continue;
                double d2 = 0;
                for (int d = 0; d < 3; d++) {
                    const double t = P[3 * i + d] - cen[3 * c + d];
                    d2 += t * t;
                }
                if (nl[i] < 0 || d2 < best) {
                    best = d2;
                    nl[i] = c;
                }
            }
            cap[nl[i]]--;
        }
        const bool same = nl == lab;
        lab.swap(nl);
        if (same) break;
    }
    // clusters in Morton order of their first member, members in Morton order
    std::vector<int> first(k, nsrc);
    std::vector<int> rank(nsrc);
    for (int p = 0; p < nsrc; p++) rank[perm[p]] = p;
    for (int i = 0; i < nsrc; i++) first[lab[i]] = std::min(first[lab[i]], rank[i]);
    std::stable_sort(perm, perm + nsrc, [&](int a, int b) {
        const int fa = first[lab[a]], fb = first[lab[b]];
        return fa != fb ? fa < fb : rank[a] < rank[b];
    });
    return PFAST_OK;
}

namespace {
int shots3d_create_ordered(pfast_model3d *m, int nsrc, const int *src, const int *srnum,
                           const int *rec, const double *nxyz, const double *T_obs,
                           std::vector<int> perm, std::vector<int> iperm, std::vector<long> rmap,
                           pfast_shots3d **out);
}  // namespace

extern "C" {

int pfast_shots3d_create(pfast_model3d *m, int nsrc, const int *src_in, const int *srnum_in,
                         const int *rec_in, const double *nxyz_in, const double *T_obs_in,
                         pfast_shots3d **out)
{
    ARGCHK(m && src_in && srnum_in && out && nsrc > 0, "bad argument");
    for (int i = 0; i < nsrc; i++) ARGCHK(srnum_in[i] >= 0, "negative srnum");
    // internal source order: chunks of spatially close sources (the receiver rows move
    // with their source)
    std::vector<int> perm(nsrc);
    pfast_source_order(nsrc, src_in, m->dx, m->dy, m->dz, m->ctx->source_order, perm.data());
    std::vector<long> off_in(nsrc + 1, 0);
    for (int i = 0; i < nsrc; i++) off_in[i + 1] = off_in[i] + srnum_in[i];
    std::vector<int> src(3 * (size_t)nsrc), srnum(nsrc), rec, iperm(nsrc);
    std::vector<double> nxyz, T_obs;
    std::vector<long> rmap;
    rmap.reserve(off_in[nsrc]);
    for (int p = 0; p < nsrc; p++) {
        const int u = perm[p];
        iperm[u] = p;
        for (int d = 0; d < 3; d++) src[3 * p + d] = src_in[3 * u + d];
        srnum[p] = srnum_in[u];
        for (long r = off_in[u]; r < off_in[u + 1]; r++) rmap.push_back(r);
    }
    if (rec_in) {
        rec.resize(3 * rmap.size());
        for (size_t r = 0; r < rmap.size(); r++)
            for (int d = 0; d < 3; d++) rec[3 * r + d] = rec_in[3 * rmap[r] + d];
    }
    if (nxyz_in) {
        nxyz.resize(3 * rmap.size());
        for (size_t r = 0; r < rmap.size(); r++)
            for (int d = 0; d < 3; d++) nxyz[3 * r + d] = nxyz_in[3 * rmap[r] + d];
    }
    if (T_obs_in) {
        T_obs.resize(rmap.size());
        for (size_t r = 0; r < rmap.size(); r++) T_obs[r] = T_obs_in[rmap[r]];
    }
    return shots3d_create_ordered(m, nsrc, src.data(), srnum.data(), rec_in ? rec.data() : nullptr,
                                  nxyz_in ? nxyz.data() : nullptr, T_obs_in ? T_obs.data() : nullptr,
                                  std::move(perm), std::move(iperm), std::move(rmap), out);
}

}  // extern "C"

namespace {

// pfast_shots3d_create on inputs already in the internal source order.
int shots3d_create_ordered(pfast_model3d *m, int nsrc, const int *src, const int *srnum,
                           const int *rec, const double *nxyz, const double *T_obs,
                           std::vector<int> perm, std::vector<int> iperm, std::vector<long> rmap,
                           pfast_shots3d **out)
{
    const int nrows = m->nrows, ncols = m->ncols, nz = m->nz;
    const long N = (long)nrows * ncols * nz;
    // the sweep kernels address one chunk's field (C sources) with 32-bit byte offsets;
    // a level outside the column (up to 64 past either end) must still fall outside it
    ARGCHK((N + 64L * nrows * ncols) * PFAST_CHUNK * 4 <= (1L << 32),
           "grid too large for the batched solver (one source chunk's field plus 64 levels "
           "must stay within 4 GiB)");
    // the adjoint field holds (lambda, T) pairs: 8 bytes per node and source
    ARGCHK(!(rec && T_obs && nxyz) || (N + 64L * nrows * ncols) * PFAST_CHUNK * 8 <= (1L << 32),
           "grid too large for the batched adjoint (one source chunk's (lambda, T) field plus "
           "64 levels must stay within 4 GiB)");
    auto *s = new pfast_shots3d();
    s->m = m;
    m->nshots++;  // undone by pfast_shots3d_destroy, also on the failure paths below
    s->perm = std::move(perm);
    s->iperm = std::move(iperm);
    s->rmap = std::move(rmap);
    s->nsrc = nsrc;
    s->nchunk = (nsrc + PFAST_CHUNK - 1) / PFAST_CHUNK;
    s->src.assign(src, src + 3 * nsrc);
    s->srnum.assign(srnum, srnum + nsrc);
    s->ch0.assign(nsrc + 1, 0);
    for (int i = 0; i < nsrc; i++) s->ch0[i + 1] = s->ch0[i] + srnum[i];
    s->nrec = s->ch0[nsrc];
    for (int i = 0; i < nsrc; i++) {
        const int *q = src + 3 * i;
        if (!(q[0] >= 0 && q[0] < nrows && q[1] >= 0 && q[1] < ncols && q[2] >= 0 && q[2] < nz)) {
            pfast_shots3d_destroy(s);
            pfast::set_error("invalid argument: source node outside the grid");
            return PFAST_ERR_ARG;
        }
    }
    if (rec)
        for (long r = 0; r < s->nrec; r++) {
            const int *q = rec + 3 * r;
            if (!(q[0] >= 0 && q[0] < nrows && q[1] >= 0 && q[1] < ncols && q[2] >= 0 && q[2] < nz)) {
                pfast_shots3d_destroy(s);
                pfast::set_error("invalid argument: receiver node outside the grid");
                return PFAST_ERR_ARG;
            }
        }
    hipStream_t st = m->ctx->stream;
    std::vector<int> sn(s->nchunk * PFAST_CHUNK, -1);
    for (int i = 0; i < nsrc; i++)
        sn[i] = (int)(((long)src[3 * i + 2] * nrows + src[3 * i]) * ncols + src[3 * i + 1]);
    int rc = 0;
    const size_t fsz = (size_t)s->nchunk * N * PFAST_CHUNK;
    if ((rc = pfast::dupload(m->ctx, &s->srcnode, sn.data(), sn.size())) || (rc = pfast::dalloc(m->ctx, &s->T, fsz)) ||
        (rc = pfast::dalloc(m->ctx, &s->changed, s->nchunk)) || (rc = pfast::dalloc(m->ctx, &s->active, s->nchunk)) ||
        (rc = pfast::dalloc(m->ctx, &s->EE, nsrc)) || (rc = pfast::dalloc(m->ctx, &s->grad, N)) ||
        (rc = pfast::dalloc(m->ctx, &s->fieldbuf, N))) {
        pfast_shots3d_destroy(s);
        return rc;
    }
    // receiver table: node index and owning source of each entry
    std::vector<long> rnode(s->nrec);
    std::vector<int> rsrc(s->nrec);
    if (rec) {
        for (int i = 0; i < nsrc; i++)
            for (long r = s->ch0[i]; r < s->ch0[i + 1]; r++) {
                const int *q = rec + 3 * r;
                rnode[r] = ((long)q[2] * nrows + q[0]) * ncols + q[1];
                rsrc[r] = i;
            }
        if ((rc = pfast::dupload(m->ctx, &s->rnode, rnode.data(), s->nrec)) ||
            (rc = pfast::dupload(m->ctx, &s->rsrc, rsrc.data(), s->nrec)) ||
            (rc = pfast::dupload(m->ctx, &s->rec3, rec, 3 * s->nrec)) || (rc = pfast::dalloc(m->ctx, &s->picks, s->nrec))) {
            pfast_shots3d_destroy(s);
            return rc;
        }
    }
    if (rec && T_obs && nxyz) {
        s->have_adj_data = true;
        std::vector<long> pnode, bnode;
        std::vector<int> psrc, bsrc;
        std::vector<double> pobs, bobs, bnrm;
        std::vector<long> poff(nsrc + 1, 0);
        for (int i = 0; i < nsrc; i++) {
            const long sn0 = sn[i];
            std::map<long, long> last;  // node -> last picked receiver (reference overwrite order)
            for (long r = s->ch0[i]; r < s->ch0[i + 1]; r++) {
                if (!(T_obs[r] < 100)) continue;
                pnode.push_back(rnode[r]);
                psrc.push_back(i);
                pobs.push_back(T_obs[r]);
                last[rnode[r]] = r;
            }
            poff[i + 1] = (long)pnode.size();
            for (auto &kv : last) {
                if (kv.first == sn0) continue;  // receiver at the source: lambda stays 0
                const long r = kv.second;
                bnode.push_back(kv.first);
                bsrc.push_back(i);
                bobs.push_back(T_obs[r]);
                bnrm.insert(bnrm.end(), nxyz + 3 * r, nxyz + 3 * r + 3);
            }
        }
        s->npick = (int)pnode.size();
        s->nb = (int)bnode.size();
        if ((rc = pfast::dalloc(m->ctx, &s->L, 2 * fsz)) || (rc = pfast::dupload(m->ctx, &s->pnode, pnode.data(), pnode.size())) ||
            (rc = pfast::dupload(m->ctx, &s->psrc, psrc.data(), psrc.size())) ||
            (rc = pfast::dupload(m->ctx, &s->pobs, pobs.data(), pobs.size())) ||
            (rc = pfast::dupload(m->ctx, &s->poff, poff.data(), poff.size())) ||
            (rc = pfast::dupload(m->ctx, &s->bnode, bnode.data(), bnode.size())) ||
            (rc = pfast::dupload(m->ctx, &s->bsrc, bsrc.data(), bsrc.size())) ||
            (rc = pfast::dupload(m->ctx, &s->bobs, bobs.data(), bobs.size())) ||
            (rc = pfast::dupload(m->ctx, &s->bnrm, bnrm.data(), bnrm.size())) ||
            (rc = pfast::dupload(m->ctx, &s->nxyz, nxyz, 3 * s->nrec)) ||
            (rc = pfast::dupload(m->ctx, &s->Tobs, T_obs, s->nrec))) {
            pfast_shots3d_destroy(s);
            return rc;
        }
    }
    s->hchunks.resize(s->nchunk);
    for (int i = 0; i < s->nchunk; i++) {
        ChunkRef &c = s->hchunks[i];
        c.field[0] = s->T + (size_t)i * N * PFAST_CHUNK;
        c.field[1] = s->L ? s->L + (size_t)i * N * PFAST_CHUNK * 2 : nullptr;
        c.slow = m->f;
        c.slow8 = m->ft;
        for (int k = 0; k < PFAST_CHUNK; k++) c.srcnode[k] = sn[i * PFAST_CHUNK + k];
    }
    if ((rc = pfast::dupload(m->ctx, &s->chunks, s->hchunks.data(), s->hchunks.size()))) {
        pfast_shots3d_destroy(s);
        return rc;
    }
    HIPCHK(hipStreamSynchronize(st));
    *out = s;
    return PFAST_OK;
}

}  // namespace

extern "C" {

int pfast_shots3d_set_source_init(pfast_shots3d *s, int halfwidth)
{
    ARGCHK(s, "NULL shots");
    ARGCHK(halfwidth >= 0 && halfwidth <= PFAST_SRCINIT_MAX,
           "source box half-width outside 0.." + std::to_string(PFAST_SRCINIT_MAX));
    hipStream_t st = s->m->ctx->stream;
    HIPCHK(hipSetDevice(s->m->ctx->device));
    s->srcinit = halfwidth;
    for (ChunkRef &c : s->hchunks) c.srcinit = halfwidth;
    HIPCHK(hipMemcpyAsync(s->chunks, s->hchunks.data(), s->hchunks.size() * sizeof(ChunkRef),
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return PFAST_OK;
}

int pfast_shots3d_run(pfast_shots3d *s, int mode, int MAXSwp)
{
    ARGCHK(s, "NULL shots");
    ARGCHK(mode == PFAST_FWD_ONLY || mode == PFAST_FWD_ADJ, "mode");
    ARGCHK(mode == PFAST_FWD_ONLY || s->have_adj_data, "shot set has no picks for the adjoint");
    ARGCHK(MAXSwp >= 0, "MAXSwp < 0");
    pfast_model3d *m = s->m;
    hipStream_t st = m->ctx->stream;
    HIPCHK(hipSetDevice(m->ctx->device));
    int rc = pfast::fwd_begin3d(s);
    if (!rc) rc = run_sweeps3d(s, false, MAXSwp, s->fcyc);
    if (rc || mode == PFAST_FWD_ONLY) return rc;
    if ((rc = pfast::adj_begin3d(s)) || (rc = run_sweeps3d(s, true, MAXSwp, s->acyc)) ||
        (rc = pfast::adj_end3d(s)))
        return rc;
    HIPCHK(hipStreamSynchronize(st));
    return PFAST_OK;
}

// A tuner replaces the setting in use only by a candidate at least 2% faster.
constexpr double kTuneMargin = 1.02;

int pfast_shots3d_autotune(pfast_shots3d *s, const int *variants, int nvariants, int *chosen,
                           double *ms)
{
    ARGCHK(s && chosen, "NULL argument");
    pfast_ctx *ctx = s->m->ctx;
    const int nall = pfast::sweep_variant_count();
    std::vector<int> cand;
    if (variants) {
        ARGCHK(nvariants > 0, "no candidate variants");
        for (int i = 0; i < nvariants; i++)
            ARGCHK(variants[i] >= 0 && variants[i] < nall, "unknown sweep variant");
        cand.assign(variants, variants + nvariants);
    } else {
        for (int v = 0; v < nall; v++) cand.push_back(v);
    }
    // every variant computes the same results, so the choice is by time alone: one warm-up
    // cycle, then two timed forward cycles from the initial field and, with picks, two
    // timed adjoint cycles on that field, per candidate (the variants set the forward and
    // the adjoint batch depth apart)
    const int keep = ctx->sweep;
    int best = keep;
    double tbest = 0.0, tkeep = 0.0;
    for (size_t i = 0; i < cand.size(); i++) {
        ctx->sweep = cand[i];
        int rc = pfast_shots3d_run(s, PFAST_FWD_ONLY, 1);
        const auto t0 = std::chrono::steady_clock::now();
        if (!rc) rc = pfast_shots3d_run(s, PFAST_FWD_ONLY, 2);
        if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = PFAST_ERR_HIP;
        if (rc) {
            ctx->sweep = keep;
            return rc;
        }
        if (s->have_adj_data) {
            rc = pfast::adj_begin3d(s);
            if (!rc) rc = run_sweeps3d(s, true, 2, s->acyc);
            if (!rc) rc = pfast::adj_end3d(s);
            if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = PFAST_ERR_HIP;
            if (rc) {
                ctx->sweep = keep;
                return rc;
            }
        }
        const double dt =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms) ms[i] = dt;
        if (cand[i] == keep) tkeep = dt;
        if (i == 0 || dt < tbest) {
            tbest = dt;
            best = cand[i];
        }
    }
    // the variant in use stays unless another is clearly faster (timing noise is ~1%)
    if (tkeep > 0.0 && tkeep <= kTuneMargin * tbest) best = keep;
    ctx->sweep = best;
    *chosen = best;
    return PFAST_OK;
}

int pfast_shots3d_autotune_schedule(pfast_shots3d *s, int *schedule, int *interval, double *ms)
{
    ARGCHK(s && schedule && interval, "NULL argument");
    pfast_ctx *ctx = s->m->ctx;
    // the schedule decides which launches a cycle takes, the interval how many cycles
    // pass between host checks: both act over a whole phase, so each candidate is timed
    // on a full forward solve (and a full adjoint one with picks); results never change
    const int keep_df = ctx->dataflow, keep_iv = ctx->check_every;
    const int cand[][2] = {{2, 1}, {0, 1}, {1, 1}, {2, 2}, {2, 4}};
    const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
    int best = 0, cur = -1;
    double tbest = 0.0;
    std::vector<double> t(ncand);
    for (int i = 0; i < ncand; i++) {
        if (cand[i][0] == keep_df && cand[i][1] == keep_iv) cur = i;
        ctx->dataflow = cand[i][0];
        ctx->check_every = cand[i][1];
        int rc = pfast_shots3d_run(s, PFAST_FWD_ONLY, 1);  // warm-up (first use of the schedule)
        const auto t0 = std::chrono::steady_clock::now();
        if (!rc) rc = pfast_shots3d_run(s, s->have_adj_data ? PFAST_FWD_ADJ : PFAST_FWD_ONLY, 100);
        if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = PFAST_ERR_HIP;
        if (rc) {
            ctx->dataflow = keep_df;
            ctx->check_every = keep_iv;
            return rc;
        }
        const double dt =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms) ms[i] = dt;
        t[i] = dt;
        if (i == 0 || dt < tbest) {
            tbest = dt;
            best = i;
        }
    }
    // the setting in use stays unless another is clearly faster
    if (cur >= 0 && t[cur] <= kTuneMargin * tbest) best = cur;
    ctx->dataflow = *schedule = cand[best][0];
    ctx->check_every = *interval = cand[best][1];
    return PFAST_OK;
}

int pfast_shots3d_gradient(pfast_shots3d *s, double *grad, double *misfit)
{
    ARGCHK(s && s->have_adj_data, "no adjoint results");
    pfast_model3d *m = s->m;
    hipStream_t st = m->ctx->stream;
    const long N = (long)m->nrows * m->ncols * m->nz;
    if (grad) HIPCHK(hipMemcpyAsync(grad, s->grad, N * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> ee(s->nsrc);
    HIPCHK(hipMemcpyAsync(ee.data(), s->EE, s->nsrc * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (misfit) {
        double sum = 0, comp = 0;  // Kahan over sources, in the caller's order
        for (int u = 0; u < s->nsrc; u++) {
            const double e = ee[s->iperm[u]];
            const double y = e - comp, t = sum + y;
            comp = (t - sum) - y;
            sum = t;
        }
        *misfit = sum;
    }
    return PFAST_OK;
}

int pfast_shots3d_picks(pfast_shots3d *s, double *picks)
{
    ARGCHK(s && picks && s->rnode, "no receiver table");
    pfast_model3d *m = s->m;
    hipStream_t st = m->ctx->stream;
    const long N = (long)m->nrows * m->ncols * m->nz;
    HIPCHK(pfast::launch_gather_picks(st, s->T, s->rnode, s->rsrc, (int)s->nrec, N, s->picks));
    std::vector<double> p(s->nrec);
    HIPCHK(hipMemcpyAsync(p.data(), s->picks, s->nrec * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (long r = 0; r < s->nrec; r++) picks[s->rmap[r]] = p[r];
    return PFAST_OK;
}

int pfast_shots3d_field(pfast_shots3d *s, int which, int src, double *field)
{
    ARGCHK(s && field && src >= 0 && src < s->nsrc && (which == 0 || which == 1), "bad argument");
    ARGCHK(which == 0 || s->L, "no adjoint field");
    pfast_model3d *m = s->m;
    hipStream_t st = m->ctx->stream;
    const long N = (long)m->nrows * m->ncols * m->nz;
    HIPCHK(pfast::launch_gather_field(st, which ? s->L : s->T, s->iperm[src], N, which ? 2 : 1,
                                      which == 0, s->fieldbuf));
    HIPCHK(hipMemcpyAsync(field, s->fieldbuf, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PFAST_OK;
}

int pfast_shots3d_cycles(pfast_shots3d *s, int *cycles)
{
    ARGCHK(s && cycles, "bad argument");
    for (int u = 0; u < s->nsrc; u++) {
        const int i = s->iperm[u];
        cycles[2 * u] = i < (int)s->fcyc.size() ? s->fcyc[i] : 0;
        cycles[2 * u + 1] = i < (int)s->acyc.size() ? s->acyc[i] : 0;
    }
    return PFAST_OK;
}

// Single-source FP64 forward+adjoint of source `srcid` of a resident shot set:
// the reference's per-source loop body (src/T_PFAST3D.cpp:388-393) on the GPU,
// on the same resident model.  grad (device, N doubles) accumulates -lambda/v^3.
int pfast_shots3d_plain_source(pfast_shots3d *s, int srcid, int MAXSwp, double *EE,
                               int *cycles2, double *T_out, double *lam_out)
{
    ARGCHK(s && s->have_adj_data && srcid >= 0 && srcid < s->nsrc, "bad argument");
    srcid = s->iperm[srcid];
    pfast_model3d *m = s->m;
    pfast_ctx *ctx = m->ctx;
    const long N = (long)m->nrows * m->ncols * m->nz;
    double *T, *U, *p, *dEE;
    int *flag;
    unsigned char *tag;
    int rc;
    if ((rc = scratch_t(ctx, "ps.T", N, &T)) || (rc = scratch_t(ctx, "ps.U", N, &U)) ||
        (rc = scratch_t(ctx, "ps.p", N, &p)) || (rc = scratch_t(ctx, "ps.tag", N, &tag)) ||
        (rc = scratch_t(ctx, "ps.EE", 1, &dEE)) || (rc = scratch_t(ctx, "ps.flag", 1, &flag)))
        return rc;
    const int si = s->src[3 * srcid], sj = s->src[3 * srcid + 1], sk = s->src[3 * srcid + 2];
    int fc = 0, ac = 0;
    rc = fsm3d_dev(ctx, m->vp, T, p, flag, si, sj, sk, MAXSwp, m->dx, m->dy, m->dz, m->nrows,
                   m->ncols, m->nz, &fc);
    if (rc) return rc;
    const long c0 = s->ch0[srcid];
    rc = adj3d_dev(ctx, U, T, tag, p, flag, dEE, s->Tobs + c0, s->rec3 + 3 * c0, s->nxyz + 3 * c0,
                   s->srnum[srcid], si, sj, sk, MAXSwp, m->dx, m->dy, m->dz, m->nrows, m->ncols,
                   m->nz, EE, &ac);
    if (rc) return rc;
    if (cycles2) {
        cycles2[0] = fc;
        cycles2[1] = ac;
    }
    if (T_out) HIPCHK(hipMemcpyAsync(T_out, T, N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (lam_out) HIPCHK(hipMemcpyAsync(lam_out, U, N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return PFAST_OK;
}

}  // extern "C"
