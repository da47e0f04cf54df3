// shots2d.hip -- batched 2D shot sets: every source of a 2D survey on one GPU.
//
// One work-group per source runs the whole RT_fsm2d (transmission, then per
// reflector the down-going solve, the reflection seeding and the up-going solve),
// then the transmission and reflection adjoints, with FP32 sweeps (plain.hip,
// policy Fast32).  The per-source fields stay resident; the gradients of the 2D
// driver's per-source loop (src/RT_PFAST2D.cpp:422-435)
//     gT -= lambdaT_s / v^3,  gR -= lambdaR_s / v^3,  gRT = gT (1 - WR) + gR WR
// and the misfits Et, Er are reduced across sources in FP64 (Kahan) on the device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/pfast.h"
#include "context.h"
#include "pfast_internal.h"
#include "plain_kernels.h"
#include "shots2d.h"

namespace {

// g = -(sum_s lambda_s) / v^3 for the transmission and reflection adjoints, and
// gRT = gT (1 - WR) + gR WR (WR = NULL: WR = 0).  Kahan over sources.
__global__ void grad2d_reduce(const float *__restrict__ LT, const float *__restrict__ LR,
                              const double *__restrict__ vp, const double *__restrict__ WR,
                              double *__restrict__ gT, double *__restrict__ gR,
                              double *__restrict__ gRT, long N, int nsrc, int reflect)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double sT = 0, cT = 0, sR = 0, cR = 0;
    for (int s = 0; s < nsrc; s++) {
        const double yT = (double)LT[(size_t)s * N + n] - cT, tT = sT + yT;
        cT = (tT - sT) - yT;
        sT = tT;
        if (reflect) {
            const double yR = (double)LR[(size_t)s * N + n] - cR, tR = sR + yR;
            cR = (tR - sR) - yR;
            sR = tR;
        }
    }
    const double v = vp[n], v3 = v * v * v;
    const double a = -sT / v3, b = -sR / v3;
    gT[n] = a;
    gR[n] = b;
    const double w = WR ? WR[n] : 0.0;
    gRT[n] = a * (1. - w) + b * w;
}

// The FP32 sweeps' slowness, 1 / vp rounded to FP32 once per model (the value the
// plane loop forms per node as (float)(1.0 / vp[n])).
__global__ void slow2d(const double *__restrict__ vp, float *__restrict__ slow, long N)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) slow[n] = (float)(1.0 / vp[n]);
}

__global__ void gather2d(const float *__restrict__ F, long N, double *__restrict__ out)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) out[n] = (double)F[n];
}

double kahan_sum(const std::vector<double> &v)
{
    double sum = 0, comp = 0;
    for (double e : v) {
        const double y = e - comp, t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    return sum;
}

Grid2 grid_of(const pfast_model2d *m)
{
    Grid2 g;
    g.nrows = m->nrows;
    g.ncols = m->ncols;
    g.dx = m->dx;
    g.dz = m->dz;
    g.wx = 1.0 / (m->dx * m->dx);
    g.wz = 1.0 / (m->dz * m->dz);
    g.window = m->ctx->sweep2d == 0;
    g.slow = m->slow;
    return g;
}

}  // namespace

namespace pfast {

int issue_fwd2d(pfast_shots2d *s, int MAXSwp, int *status_host)
{
    pfast_model2d *m = s->m;
    pfast_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    const Grid2 g = grid_of(m);
    s->have_fwd = s->have_adj = false;
    time_begin(ctx, 2);
    HIPCHK(launch_rt_fsm2d_f32(st, m->vp, s->TT, s->TTd, s->TTu, m->Z, s->src, s->nsrc, m->NumofR,
                               MAXSwp, g, s->fcyc, s->status, s->k2));
    time_end(ctx);
    HIPCHK(hipMemcpyAsync(status_host, s->status, s->nsrc * sizeof(int), hipMemcpyDeviceToHost, st));
    return PFAST_OK;
}

int check_fwd2d(pfast_shots2d *s, const int *status)
{
    for (int i = 0; i < s->nsrc; i++)
        if (status[i]) {
            set_error(std::string("source ") + std::to_string(i) +
                      (status[i] == 1 ? ": no k2 above the reflector in a column (the "
                                        "reference exits, src/RT_pfsm2d.h:232,257,281)"
                                      : ": negative reflected traveltime seed (the "
                                        "reference exits, src/RT_pfsm2d.h:227,253,274)"));
            return PFAST_ERR_FATAL;
        }
    s->have_fwd = true;
    return PFAST_OK;
}

int issue_adj2d(pfast_shots2d *s, int MAXSwp)
{
    pfast_model2d *m = s->m;
    pfast_ctx *ctx = m->ctx;
    hipStream_t st = ctx->stream;
    const Grid2 g = grid_of(m);
    const int R = m->NumofR;
    time_begin(ctx, 2);
    if (R > 0 && g.window) {
        // the two adjoints of each source run side by side (tag scratch [2][nsrc][N])
        const size_t F = (size_t)s->nsrc * m->nrows * m->ncols;
        HIPCHK(launch_adj2d_pair_f32(st, s->LT, s->tag, s->TT, s->TT_obs, s->LR, s->Uw,
                                     s->tag + F, s->TTu, s->TR_obs, 1.0, s->rec, s->nxy, s->dch0,
                                     s->srnum, s->src, s->nsrc, R, MAXSwp, g, s->EEt, s->tcyc,
                                     s->EEr, s->rcyc));
        time_end(ctx);
    } else {
        HIPCHK(launch_tadj2d_f32(st, s->LT, s->tag, s->TT, s->TT_obs, 1.0, s->rec, s->nxy, s->dch0,
                                 s->srnum, s->src, s->nsrc, MAXSwp, g, s->EEt, s->tcyc));
        if (R > 0)
            HIPCHK(launch_radj2d_f32(st, s->LR, s->Uw, s->tag, s->TTu, s->TR_obs, 1.0, s->rec,
                                     s->nxy, s->dch0, s->srnum, s->src, s->nsrc, R, MAXSwp, g,
                                     s->EEr, s->rcyc));
        time_end(ctx);
        if (R == 0) HIPCHK(hipMemsetAsync(s->EEr, 0, s->nsrc * sizeof(double), st));
    }
    return PFAST_OK;
}

int issue_grad2d(pfast_shots2d *s, const double *WR_dev)
{
    pfast_model2d *m = s->m;
    const long N = (long)m->nrows * m->ncols;
    hipLaunchKernelGGL(grad2d_reduce, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                       m->ctx->stream, s->LT, s->LR, m->vp, WR_dev, s->gT, s->gR, s->gRT, N,
                       s->nsrc, m->NumofR > 0 ? 1 : 0);
    HIPCHK(hipGetLastError());
    return PFAST_OK;
}

size_t footprint2d(int nrows, int ncols, int NumofR, int nsrc, long nrec, bool adjoint)
{
    const size_t N = (size_t)nrows * ncols, F = (size_t)nsrc * N, R1 = NumofR ? NumofR : 1;
    size_t b = F * 4 * (1 + 2 * R1) + N * 8 + (size_t)nsrc * (4 * (4 + 2 * NumofR + ncols) + 8);
    if (adjoint)
        b += F * (4 * 3 + 2) + N * 8 * 4 + (size_t)nrec * (8 + 16 + 8 + 8 * NumofR) +
             (size_t)nsrc * (8 * 2 + 4 * (1 + R1));
    return b;
}

}  // namespace pfast

extern "C" {

int pfast_model2d_create(pfast_ctx *ctx, const double *vp, int nrows, int ncols, double dx,
                         double dz, const double *Z, int NumofR, pfast_model2d **out)
{
    ARGCHK(ctx && vp && out, "NULL argument");
    ARGCHK(nrows >= 3 && ncols >= 3, "grid dimensions must be >= 3");
    // the FP32 row-window sweeps address a source's fields and the FP64 velocities with
    // 32-bit byte offsets
    ARGCHK((long)nrows * ncols <= (1L << 28), "2D grid too large (at most 2^28 nodes)");
    ARGCHK(dx > 0 && dz > 0, "spacings must be positive");
    ARGCHK(NumofR >= 0 && (NumofR == 0 || Z), "reflectors need Z");
    HIPCHK(hipSetDevice(ctx->device));
    auto *m = new pfast_model2d();
    m->ctx = ctx;
    ctx->nmodels++;  // undone by pfast_model2d_destroy, also on the failure paths below
    m->nrows = nrows;
    m->ncols = ncols;
    m->NumofR = NumofR;
    m->dx = dx;
    m->dz = dz;
    const long N = (long)nrows * ncols;
    int rc = pfast::dupload(ctx, &m->vp, vp, N);
    if (!rc) rc = pfast::dupload(ctx, &m->Z, Z, (size_t)NumofR * ncols * 3);
    if (!rc) rc = pfast::dalloc(ctx, &m->slow, N);
    if (!rc)
        hipLaunchKernelGGL(slow2d, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream,
                           m->vp, m->slow, N);
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) {
        pfast::set_error("upload of the 2D model failed");
        rc = PFAST_ERR_HIP;
    }
    if (rc) {
        pfast_model2d_destroy(m);
        return rc;
    }
    *out = m;
    return PFAST_OK;
}

int pfast_model2d_set_vp(pfast_model2d *m, const double *vp)
{
    ARGCHK(m && vp, "NULL argument");
    const long N = (long)m->nrows * m->ncols;
    HIPCHK(hipMemcpyAsync(m->vp, vp, N * sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
    hipLaunchKernelGGL(slow2d, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, m->ctx->stream,
                       m->vp, m->slow, N);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->ctx->stream));
    return PFAST_OK;
}

int pfast_model2d_destroy(pfast_model2d *m)
{
    if (!m) return PFAST_OK;
    ARGCHK(m->nshots == 0, "model still has live shot sets: destroy them first");
    (void)hipStreamSynchronize(m->ctx->stream);
    m->ctx->nmodels--;
    pfast::dfree(m->ctx, m->vp);
    pfast::dfree(m->ctx, m->Z);
    pfast::dfree(m->ctx, m->slow);
    delete m;
    return PFAST_OK;
}

int pfast_shots2d_destroy(pfast_shots2d *s)
{
    if (!s) return PFAST_OK;
    (void)hipStreamSynchronize(s->m->ctx->stream);
    pfast::shots2d_free(s);
    return PFAST_OK;
}

}  // extern "C"

namespace pfast {

void shots2d_free(pfast_shots2d *s)
{
    s->m->nshots--;
    void *ptrs[] = {s->src, s->rec, s->srnum, s->dch0, s->nxy, s->TT_obs, s->TR_obs, s->TT,
                    s->TTd, s->TTu, s->LT, s->LR, s->Uw, s->tag, s->EEt, s->EEr, s->fcyc,
                    s->tcyc, s->rcyc, s->status, s->k2, s->gT, s->gR, s->gRT, s->WR,
                    s->fieldbuf};
    for (void *p : ptrs) pfast::dfree(s->m->ctx, p);
    delete s;
}

}  // namespace pfast

extern "C" {

int pfast_shots2d_create(pfast_model2d *m, int nsrc, const int *src, const int *srnum,
                         const int *rec, const double *nxy, const double *TT_obs,
                         const double *TR_obs, pfast_shots2d **out)
{
    ARGCHK(m && src && srnum && out && nsrc > 0, "bad argument");
    for (int i = 0; i < nsrc; i++) {
        ARGCHK(srnum[i] >= 0, "negative srnum");
        ARGCHK(src[2 * i] >= 0 && src[2 * i] < m->nrows && src[2 * i + 1] >= 0 &&
                   src[2 * i + 1] < m->ncols,
               "source node outside the grid");
    }
    const int nrows = m->nrows, ncols = m->ncols, R = m->NumofR;
    const long N = (long)nrows * ncols;
    auto *s = new pfast_shots2d();
    s->m = m;
    m->nshots++;  // undone by pfast_shots2d_destroy, also on the failure paths below
    s->nsrc = nsrc;
    s->ch0.assign(nsrc + 1, 0);
    for (int i = 0; i < nsrc; i++) s->ch0[i + 1] = s->ch0[i] + srnum[i];
    s->nrec = s->ch0[nsrc];
    const bool adj = rec && nxy && TT_obs && (R == 0 || TR_obs);
    if (adj) {
        for (long r = 0; r < s->nrec; r++)
            if (!(rec[2 * r] >= 0 && rec[2 * r] < nrows && rec[2 * r + 1] >= 0 &&
                  rec[2 * r + 1] < ncols)) {
                pfast_shots2d_destroy(s);
                pfast::set_error("invalid argument: receiver node outside the grid");
                return PFAST_ERR_ARG;
            }
    }
    hipStream_t st = m->ctx->stream;
    const size_t F = (size_t)nsrc * N;
    int rc = 0;
    auto ok = [&](int r) { return (rc = r) == 0; };
    if (!(ok(pfast::dupload(m->ctx, &s->src, src, 2 * (size_t)nsrc)) &&
          ok(pfast::dupload(m->ctx, &s->srnum, srnum, nsrc)) &&
          ok(pfast::dupload(m->ctx, &s->dch0, s->ch0.data(), nsrc + 1)) &&
          ok(pfast::dalloc(m->ctx, &s->TT, F)) && ok(pfast::dalloc(m->ctx, &s->TTd, F * (R ? R : 1))) &&
          ok(pfast::dalloc(m->ctx, &s->TTu, F * (R ? R : 1))) && ok(pfast::dalloc(m->ctx, &s->fcyc, (size_t)nsrc * (1 + 2 * R))) &&
          ok(pfast::dalloc(m->ctx, &s->status, nsrc)) && ok(pfast::dalloc(m->ctx, &s->k2, (size_t)nsrc * ncols)) &&
          ok(pfast::dalloc(m->ctx, &s->fieldbuf, N)))) {
        pfast_shots2d_destroy(s);
        return rc;
    }
    if (adj) {
        if (!(ok(pfast::dupload(m->ctx, &s->rec, rec, 2 * (size_t)s->nrec)) &&
              ok(pfast::dupload(m->ctx, &s->nxy, nxy, 2 * (size_t)s->nrec)) &&
              ok(pfast::dupload(m->ctx, &s->TT_obs, TT_obs, s->nrec)) &&
              ok(pfast::dupload(m->ctx, &s->TR_obs, TR_obs, (size_t)s->nrec * R)) &&
              ok(pfast::dalloc(m->ctx, &s->LT, F)) && ok(pfast::dalloc(m->ctx, &s->LR, F)) &&
              ok(pfast::dalloc(m->ctx, &s->Uw, F)) && ok(pfast::dalloc(m->ctx, &s->tag, 2 * F)) &&
              ok(pfast::dalloc(m->ctx, &s->EEt, nsrc)) && ok(pfast::dalloc(m->ctx, &s->EEr, nsrc)) &&
              ok(pfast::dalloc(m->ctx, &s->tcyc, nsrc)) && ok(pfast::dalloc(m->ctx, &s->rcyc, (size_t)nsrc * (R ? R : 1))) &&
              ok(pfast::dalloc(m->ctx, &s->gT, N)) && ok(pfast::dalloc(m->ctx, &s->gR, N)) &&
              ok(pfast::dalloc(m->ctx, &s->gRT, N)) && ok(pfast::dalloc(m->ctx, &s->WR, N)))) {
            pfast_shots2d_destroy(s);
            return rc;
        }
        // reflection adjoint of a source with no reflectors: zero
        if (R == 0) (void)hipMemsetAsync(s->LR, 0, F * sizeof(float), st);
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        pfast_shots2d_destroy(s);
        pfast::set_error("upload of the 2D shot set failed");
        return PFAST_ERR_HIP;
    }
    *out = s;
    return PFAST_OK;
}

int pfast_shots2d_run(pfast_shots2d *s, int mode, int MAXSwp)
{
    ARGCHK(s && (mode == PFAST_FWD_ONLY || mode == PFAST_FWD_ADJ) && MAXSwp > 0, "bad argument");
    ARGCHK(mode == PFAST_FWD_ONLY || s->LT, "shot set created without receiver data");
    pfast_ctx *ctx = s->m->ctx;
    std::vector<int> status(s->nsrc);
    int rc = pfast::issue_fwd2d(s, MAXSwp, status.data());
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    pfast::time_collect(ctx);
    if ((rc = pfast::check_fwd2d(s, status.data())) || mode == PFAST_FWD_ONLY) return rc;
    if ((rc = pfast::issue_adj2d(s, MAXSwp))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    pfast::time_collect(ctx);
    s->have_adj = true;
    return PFAST_OK;
}

int pfast_shots2d_gradient(pfast_shots2d *s, const double *WR, double *gT, double *gR,
                           double *gRT, double *E3)
{
    ARGCHK(s && s->have_adj, "no adjoint results");
    pfast_model2d *m = s->m;
    hipStream_t st = m->ctx->stream;
    const long N = (long)m->nrows * m->ncols;
    if (WR) HIPCHK(hipMemcpyAsync(s->WR, WR, N * sizeof(double), hipMemcpyHostToDevice, st));
    const int rc = pfast::issue_grad2d(s, WR ? s->WR : nullptr);
    if (rc) return rc;
    if (gT) HIPCHK(hipMemcpyAsync(gT, s->gT, N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (gR) HIPCHK(hipMemcpyAsync(gR, s->gR, N * sizeof(double), hipMemcpyDeviceToHost, st));
    if (gRT) HIPCHK(hipMemcpyAsync(gRT, s->gRT, N * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> et(s->nsrc), er(s->nsrc);
    HIPCHK(hipMemcpyAsync(et.data(), s->EEt, s->nsrc * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(er.data(), s->EEr, s->nsrc * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (E3) {
        E3[0] = kahan_sum(et);
        E3[1] = kahan_sum(er);
        E3[2] = E3[0] + E3[1];
    }
    return PFAST_OK;
}

// which: 0 transmission T, 1 down-going TTd of reflector ir, 2 up-going TTu of
// reflector ir, 3 transmission adjoint, 4 reflection adjoint (sum over reflectors).
int pfast_shots2d_field(pfast_shots2d *s, int which, int src, int ir, double *field)
{
    ARGCHK(s && field && src >= 0 && src < s->nsrc && which >= 0 && which <= 4, "bad argument");
    pfast_model2d *m = s->m;
    const int R = m->NumofR;
    ARGCHK(which == 0 || which >= 3 || (ir >= 0 && ir < R), "bad reflector index");
    ARGCHK(which < 3 ? s->have_fwd : s->have_adj, "field not computed");
    hipStream_t st = m->ctx->stream;
    const long N = (long)m->nrows * m->ncols;
    const float *F = which == 0 ? s->TT + (size_t)src * N
                     : which == 1 ? s->TTd + ((size_t)src * R + ir) * N
                     : which == 2 ? s->TTu + ((size_t)src * R + ir) * N
                     : which == 3 ? s->LT + (size_t)src * N
                                  : s->LR + (size_t)src * N;
    hipLaunchKernelGGL(gather2d, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, F, N,
                       s->fieldbuf);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(field, s->fieldbuf, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PFAST_OK;
}

// cycles: per source 1 + 2R forward counts (transmission, then down / up per
// reflector), 1 transmission-adjoint count, R reflection-adjoint counts.
// k2vec: per source the last reflector's reflection rows (ncols ints).
int pfast_shots2d_info(pfast_shots2d *s, int *cycles, int *k2vec)
{
    ARGCHK(s && s->have_fwd, "no forward results");
    pfast_model2d *m = s->m;
    hipStream_t st = m->ctx->stream;
    const int R = m->NumofR, per = 1 + 2 * R + 1 + R;
    if (cycles) {
        std::vector<int> f((size_t)s->nsrc * (1 + 2 * R)), t(s->nsrc, 0), r((size_t)s->nsrc * (R ? R : 1), 0);
        HIPCHK(hipMemcpyAsync(f.data(), s->fcyc, f.size() * sizeof(int), hipMemcpyDeviceToHost, st));
        if (s->have_adj) {
            HIPCHK(hipMemcpyAsync(t.data(), s->tcyc, t.size() * sizeof(int), hipMemcpyDeviceToHost, st));
            if (R) HIPCHK(hipMemcpyAsync(r.data(), s->rcyc, (size_t)s->nsrc * R * sizeof(int), hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < s->nsrc; i++) {
            int *c = cycles + (size_t)i * per;
            for (int k = 0; k < 1 + 2 * R; k++) c[k] = f[(size_t)i * (1 + 2 * R) + k];
            c[1 + 2 * R] = t[i];
            for (int k = 0; k < R; k++) c[2 + 2 * R + k] = r[(size_t)i * R + k];
        }
    }
    if (k2vec) {
        HIPCHK(hipMemcpyAsync(k2vec, s->k2, (size_t)s->nsrc * m->ncols * sizeof(int),
                              hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return PFAST_OK;
}

}  // extern "C"
