// optim.hip -- gradient regularisation (GPU, hipFFT), the device-resident L-BFGS
// memory with its two-loop recursion and FP64 compensated dot products, and the
// optimiser helpers the reference's drivers call (host arithmetic in the reference's
// order).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/pfast.h"
#include "context.h"
#include "pfast_internal.h"

namespace {

// Wavenumbers exactly as src/T_padjoint3d.h:218-232 / src/RT_padjoint2d.h:520-530
// build them: k[j] = -pi/h + 2*pi*j/(n*h), then the two halves of length n/2
// swapped.  For even n this is the usual FFT order; for odd n it is a shifted,
// permuted vector (the DC bin does not get k = 0), reproduced as is.
std::vector<double> ref_wavenumbers(int n, double h)
{
    const double PI_ = 3.141592653589793;
    std::vector<double> k(n);
    const double kmax = PI_ / h, ks = 2 * PI_ / n / h;
    for (int j = 0; j < n; j++) k[j] = -kmax + ks * j;
    for (int j = 0; j < n / 2; j++) std::swap(k[j], k[n / 2 + j]);
    return k;
}

__global__ void pack_real(const double *__restrict__ in, hipfftDoubleComplex *__restrict__ out,
                          long N)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) out[n] = make_hipDoubleComplex(in[n], 0.0);
}

// F /= 1 + nux^2 kx^2 + nuy^2 ky^2 + nuz^2 kz^2 (same expression order as the
// reference, src/T_padjoint3d.h:237-238).  2D: ky = kz = NULL.
__global__ void smooth_filter(hipfftDoubleComplex *__restrict__ F, const double *__restrict__ kx,
                              const double *__restrict__ ky, const double *__restrict__ kz,
                              double nux, double nuy, double nuz, int nrows, int ncols, long N)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int j = (int)(n % ncols);
    const int i = (int)((n / ncols) % nrows);
    const long k = n / ((long)nrows * ncols);
    double d = 1 + nux * nux * kx[j] * kx[j] + nuy * nuy * ky[i] * ky[i];
    if (kz) d = d + nuz * nuz * kz[k] * kz[k];
    F[n].x = F[n].x / d;
    F[n].y = F[n].y / d;
}

// Real part of the inverse transform scaled by 1/N; air nodes (vp = 330) zeroed.
__global__ void unpack_mask(const hipfftDoubleComplex *__restrict__ F, const double *__restrict__ vp,
                            double *__restrict__ out, long N)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    out[n] = fabs(vp[n] - 330) < 1e-9 ? 0.0 : F[n].x / (double)N;
}

int regu_common(double *Dj, const double *vp, double nux, double nuy, double nuz, double dx,
                double dy, double dz, int nrows, int ncols, int nz, bool three)
{
    const long N = (long)nrows * ncols * (three ? nz : 1);
    double *dD = nullptr, *dvp = nullptr, *dk = nullptr;
    hipfftDoubleComplex *F = nullptr;
    hipfftHandle plan = 0;
    int rc = PFAST_OK;
    auto fail = [&](const std::string &m, int code) {
        pfast::set_error(m);
        rc = code;
    };
    std::vector<double> kx = ref_wavenumbers(ncols, dx);
    std::vector<double> ky = ref_wavenumbers(nrows, three ? dy : dz);
    std::vector<double> kz = three ? ref_wavenumbers(nz, dz) : std::vector<double>();
    const size_t nk = kx.size() + ky.size() + kz.size();
    if (hipMalloc(&dD, N * sizeof(double)) != hipSuccess ||
        hipMalloc(&dvp, N * sizeof(double)) != hipSuccess ||
        hipMalloc(&F, N * sizeof(hipfftDoubleComplex)) != hipSuccess ||
        hipMalloc(&dk, nk * sizeof(double)) != hipSuccess) {
        fail("hipMalloc failed in regularisation", PFAST_ERR_NOMEM);
    }
    if (!rc) {
        const bool ok =
            hipMemcpy(dD, Dj, N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dvp, vp, N * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dk, kx.data(), kx.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dk + kx.size(), ky.data(), ky.size() * sizeof(double), hipMemcpyHostToDevice) ==
                hipSuccess &&
            (!three || hipMemcpy(dk + kx.size() + ky.size(), kz.data(), kz.size() * sizeof(double),
                                 hipMemcpyHostToDevice) == hipSuccess);
        if (!ok) fail("hipMemcpy failed in regularisation", PFAST_ERR_HIP);
    }
    if (!rc) {
        hipfftResult fr = three ? hipfftPlan3d(&plan, nz, nrows, ncols, HIPFFT_Z2Z)
                                : hipfftPlan2d(&plan, nrows, ncols, HIPFFT_Z2Z);
        if (fr != HIPFFT_SUCCESS) fail("hipfftPlan failed", PFAST_ERR_HIP);
    }
    if (!rc) {
        const unsigned g = (unsigned)((N + 255) / 256);
        hipLaunchKernelGGL(pack_real, dim3(g), dim3(256), 0, 0, dD, F, N);
        if (hipGetLastError() != hipSuccess) fail("pack kernel launch failed", PFAST_ERR_HIP);
    }
    if (!rc && hipfftExecZ2Z(plan, F, F, HIPFFT_FORWARD) != HIPFFT_SUCCESS) fail("hipfftExec", PFAST_ERR_HIP);
    if (!rc) {
        const unsigned g = (unsigned)((N + 255) / 256);
        hipLaunchKernelGGL(smooth_filter, dim3(g), dim3(256), 0, 0, F, dk, dk + kx.size(),
                           three ? dk + kx.size() + ky.size() : nullptr, nux, three ? nuy : nuz,
                           nuz, nrows, ncols, N);
        if (hipGetLastError() != hipSuccess) fail("filter kernel launch failed", PFAST_ERR_HIP);
    }
    if (!rc && hipfftExecZ2Z(plan, F, F, HIPFFT_BACKWARD) != HIPFFT_SUCCESS) fail("hipfftExec", PFAST_ERR_HIP);
    if (!rc) {
        const unsigned g = (unsigned)((N + 255) / 256);
        hipLaunchKernelGGL(unpack_mask, dim3(g), dim3(256), 0, 0, F, dvp, dD, N);
        if (hipGetLastError() != hipSuccess) fail("unpack kernel launch failed", PFAST_ERR_HIP);
    }
    if (!rc && hipMemcpy(Dj, dD, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        fail("hipMemcpy failed in regularisation", PFAST_ERR_HIP);
    if (plan) (void)hipfftDestroy(plan);
    (void)hipFree(dD);
    (void)hipFree(dvp);
    (void)hipFree(F);
    (void)hipFree(dk);
    return rc;
}

}  // namespace

extern "C" {

int pfast_regu3d(double *Dj, const double *vp, double nux, double nuy, double nuz, double dx,
                 double dy, double dz, int nrows, int ncols, int nz)
{
    ARGCHK(Dj && vp && nrows > 0 && ncols > 0 && nz > 0, "bad argument");
    return regu_common(Dj, vp, nux, nuy, nuz, dx, dy, dz, nrows, ncols, nz, true);
}

int pfast_regu2d(double *Dj, const double *vp, double nux, double nuz, double dx, double dz,
                 int nrows, int ncols)
{
    ARGCHK(Dj && vp && nrows > 0 && ncols > 0, "bad argument");
    return regu_common(Dj, vp, nux, 0.0, nuz, dx, 0.0, dz, nrows, ncols, 1, false);
}

// ----------------------------------------------------------- host helpers
// Hestenes-Stiefel beta, src/T_padjoint3d.h:264-274 (2D :557-569, same order).
double pfast_beta_hs(const double *gk1, const double *gk0, const double *dk, long N)
{
    double a = 0.0, b = 0.0;
    for (long i = 0; i < N; i++) {
        a += gk1[i] * (gk1[i] - gk0[i]);
        b += -dk[i] * (gk1[i] - gk0[i]);
    }
    return a / b;
}

// CG_DESCENT beta (Hager-Zhang), src/T_padjoint3d.h:276-287.
double pfast_beta_cgdescent(const double *gk1, const double *gk0, const double *dk, long N)
{
    double dTy = 0.0, ny = 0.0, beta = 0.0;
    for (long i = 0; i < N; i++) {
        dTy += dk[i] * (-gk1[i] + gk0[i]);
        ny += (-gk1[i] + gk0[i]) * (-gk1[i] + gk0[i]);
    }
    for (long i = 0; i < N; i++) beta += ((-gk1[i] + gk0[i]) - 2 * dk[i] * ny / dTy) * (-gk1[i]) / dTy;
    return beta;
}

// L-BFGS pair (s, y, rho) into slot (Iterk-2) % lm; returns gamma = s.y / y.y,
// src/T_padjoint3d.h:289-304.
double pfast_gd2sk(const double *gk1, const double *gk0, const double *dk, double alpha1,
                   double *rhok, double *sk, double *yk, int Iterk, int lm, long N)
{
    double dTy = 0.0, ny = 0.0;
    for (long i = 0; i < N; i++) {
        dTy += alpha1 * dk[i] * (-gk1[i] + gk0[i]);
        ny += (-gk1[i] + gk0[i]) * (-gk1[i] + gk0[i]);
    }
    const double gamma = dTy / ny;
    const int slot = (Iterk - 2) % lm;
    rhok[slot] = 1.0 / dTy;
    for (long i = 0; i < N; i++) {
        sk[(long)slot * N + i] = alpha1 * dk[i];
        yk[(long)slot * N + i] = -gk1[i] + gk0[i];
    }
    return gamma;
}

// src/T_padjoint3d.h:306-315.
double pfast_zoom(double alpha0, double alpha1, double phi0, double phi1, double dphi0,
                  double dphi1)
{
    const double a = (2 * (phi0 - phi1) + (alpha1 - alpha0) * (dphi0 - dphi1)) /
                     ((alpha1 - alpha0) * (alpha1 - alpha0) * (alpha1 - alpha0));
    const double b = (3 * (phi1 - phi0) - (alpha1 - alpha0) * (dphi1 + 2 * dphi0)) /
                     ((alpha1 - alpha0) * (alpha1 - alpha0));
    if ((b * b - 3 * a * dphi0) >= 0) return alpha0 + (-b + sqrt(b * b - 3 * a * dphi0)) / 3.0 / a;
    return alpha1 - (alpha1 - alpha0) * dphi1 / (dphi1 - dphi0);
}

// Minimiser of the cubic through (alpha0, phi0, dphi0), (alpha1, phi1, dphi1),
// src/T_padjoint3d.h:316-326 and src/RT_padjoint2d.h:619-629.  The reference
// returns an undefined value when neither of its two cases applies; here that
// falls back to the secant step on the derivatives.
double pfast_cubic_interp(double alpha0, double alpha1, double phi0, double phi1, double dphi0,
                          double dphi1)
{
    const double a = (-2. * (phi1 - phi0) + (dphi0 + dphi1) * (alpha1 - alpha0)) /
                     ((alpha1 - alpha0) * (alpha1 - alpha0) * (alpha1 - alpha0));
    const double b = (3. * (phi1 - phi0) - (2. * dphi0 + dphi1) * (alpha1 - alpha0)) /
                     ((alpha1 - alpha0) * (alpha1 - alpha0));
    const double c = dphi0;
    if (fabs(a) > 0 && (b * b - 3 * a * c) > 0) return alpha0 + (-b + sqrt(b * b - 3 * a * c)) / (3. * a);
    if (a == 0 && fabs(b) > 0 && (b * b - 3 * a * c) < 0) return -c / (2. * b);
    return alpha1 - (alpha1 - alpha0) * dphi1 / (dphi1 - dphi0);
}

// 1 / max|g|, src/T_padjoint3d.h:328-334.
double pfast_magnitude_of(const double *g, long N)
{
    double m = fabs(g[0]);
    for (long k = 1; k < N; k++)
        if (fabs(g[k]) > m) m = fabs(g[k]);
    return 1. / m;
}

// 1 / |mean(g)|, src/T_padjoint3d.h:335-342.
double pfast_mean_of(const double *g, long N)
{
    double m = 0;
    for (long k = 0; k < N; k++) m = m + g[k];
    m = m / N;
    return 1. / fabs(m);
}

}  // extern "C"

// ------------------------------------------------- L-BFGS on the device (FP64)
namespace {

constexpr int kDotThreads = 256;
constexpr int kDotBlocks = 1024;

// (s, c): a sum and its running compensation (Neumaier); s + c is the value.
struct Comp {
    double s, c;
};

__device__ __forceinline__ Comp comp_add(Comp a, double x)
{
    const double t = a.s + x;
    a.c += fabs(a.s) >= fabs(x) ? (a.s - t) + x : (x - t) + a.s;
    a.s = t;
    return a;
}

__device__ __forceinline__ Comp comp_merge(Comp a, Comp b)
{
    Comp r = comp_add(a, b.s);
    r.c += b.c;
    return r;
}

// Per-block compensated partial sums of a[i] * b[i].
__global__ __launch_bounds__(kDotThreads) void dot_partial(const double *__restrict__ a,
                                                           const double *__restrict__ b, long N,
                                                           double2 *__restrict__ part)
{
    __shared__ double2 sh[kDotThreads];
    Comp acc{0.0, 0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
        acc = comp_add(acc, a[i] * b[i]);
    sh[threadIdx.x] = make_double2(acc.s, acc.c);
    __syncthreads();
    for (int w = kDotThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const Comp r = comp_merge(Comp{sh[threadIdx.x].x, sh[threadIdx.x].y},
                                      Comp{sh[threadIdx.x + w].x, sh[threadIdx.x + w].y});
            sh[threadIdx.x] = make_double2(r.s, r.c);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Operations that finish a dot product on the device: the compensated total t of the
// partials, then (op) sc[dst] = t; = scale * t with scale = sc[k]; the pair's
// rho = 1 / t and its s.y; gamma = sc[k] / t.
enum DotOp { kStore = 0, kScaled = 1, kRho = 2, kRatio = 3 };

__global__ __launch_bounds__(kDotThreads) void dot_final(const double2 *__restrict__ part, int nb,
                                                         double *__restrict__ sc, int op, int dst,
                                                         int k)
{
    __shared__ double2 sh[kDotThreads];
    Comp acc{0.0, 0.0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc = comp_merge(acc, Comp{part[i].x, part[i].y});
    sh[threadIdx.x] = make_double2(acc.s, acc.c);
    __syncthreads();
    for (int w = kDotThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const Comp r = comp_merge(Comp{sh[threadIdx.x].x, sh[threadIdx.x].y},
                                      Comp{sh[threadIdx.x + w].x, sh[threadIdx.x + w].y});
            sh[threadIdx.x] = make_double2(r.s, r.c);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double t = sh[0].x + sh[0].y;
    switch (op) {
    case kStore: sc[dst] = t; break;
    case kScaled: sc[dst] = sc[k] * t; break;
    case kRho: sc[dst] = 1.0 / t; sc[k] = t; break;
    default: sc[dst] = sc[k] / t; break;
    }
}

// s = alpha1 dk, y = -gk1 + gk0 (the pair of src/T_padjoint3d.h:289-304)
__global__ void pair_store(double *__restrict__ s, double *__restrict__ y,
                           const double *__restrict__ dk, const double *__restrict__ g1,
                           const double *__restrict__ g0, double alpha1, long N)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    s[i] = alpha1 * dk[i];
    y[i] = -g1[i] + g0[i];
}

// x += (sc[ka] - sc[kb]) * v   (kb < 0: x -= sc[ka] * v)
__global__ void axpy_sc(double *__restrict__ x, const double *__restrict__ v,
                        const double *__restrict__ sc, int ka, int kb, long N)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double c = kb < 0 ? -sc[ka] : sc[ka] - sc[kb];
    x[i] += c * v[i];
}

// x = sc[k] * v
__global__ void scale_sc(double *__restrict__ x, const double *__restrict__ v,
                         const double *__restrict__ sc, int k, long N)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    x[i] = sc[k] * v[i];
}

}  // namespace

struct pfast_lbfgs {
    pfast_ctx *ctx = nullptr;
    long N = 0;
    int LM = 0;
    double *sk = nullptr, *yk = nullptr;   // LM pairs, slot-major
    double *q = nullptr, *dk = nullptr, *g1 = nullptr, *g0 = nullptr;
    double2 *part = nullptr;               // per-block partial sums
    double *sc = nullptr;                  // scalars: see SC_* below
};

namespace {

// scalars: [0] s.y of the new pair, [1] y.y, [2] gamma, [3] the second loop's b,
// [4] a dot product result, then rho[LM], then a[LM]
constexpr int SC_SY = 0, SC_YY = 1, SC_GAMMA = 2, SC_B = 3, SC_DOT = 4, SC_RHO = 5;

inline dim3 grid_of(long N) { return dim3((unsigned)((N + 255) / 256)); }

// The compensated dot product a.b into the scalars (op, dst, k as dot_final).
hipError_t dot_dev(pfast_lbfgs *h, const double *a, const double *b, int op, int dst, int k)
{
    hipStream_t st = h->ctx->stream;
    const int nb = (int)std::min<long>(kDotBlocks, (h->N + kDotThreads - 1) / kDotThreads);
    hipLaunchKernelGGL(dot_partial, dim3(nb), dim3(kDotThreads), 0, st, a, b, h->N, h->part);
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(kDotThreads), 0, st, h->part, nb, h->sc, op, dst, k);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int pfast_lbfgs_create(pfast_ctx *ctx, long N, int LM, pfast_lbfgs **out)
{
    ARGCHK(ctx && out, "NULL argument");
    ARGCHK(N > 0 && LM >= 1, "N and LM must be positive");
    HIPCHK(hipSetDevice(ctx->device));
    auto *h = new pfast_lbfgs;
    h->ctx = ctx;
    h->N = N;
    h->LM = LM;
    int rc;
    if ((rc = pfast::dalloc(&h->sk, (size_t)N * LM)) || (rc = pfast::dalloc(&h->yk, (size_t)N * LM)) ||
        (rc = pfast::dalloc(&h->q, N)) || (rc = pfast::dalloc(&h->dk, N)) ||
        (rc = pfast::dalloc(&h->g1, N)) || (rc = pfast::dalloc(&h->g0, N)) ||
        (rc = pfast::dalloc(&h->part, kDotBlocks)) || (rc = pfast::dalloc(&h->sc, SC_RHO + 2 * LM))) {
        pfast_lbfgs_destroy(h);
        return rc;
    }
    *out = h;
    return PFAST_OK;
}

int pfast_lbfgs_destroy(pfast_lbfgs *h)
{
    if (!h) return PFAST_OK;
    for (void *p : {(void *)h->sk, (void *)h->yk, (void *)h->q, (void *)h->dk, (void *)h->g1,
                    (void *)h->g0, (void *)h->part, (void *)h->sc})
        if (p) (void)hipFree(p);
    delete h;
    return PFAST_OK;
}

int pfast_lbfgs_direction(pfast_lbfgs *h, int Iterk, const double *gk1, const double *gk0,
                          double *dk, double alpha1, double *gamma)
{
    ARGCHK(h && gk1 && gk0 && dk, "NULL argument");
    ARGCHK(Iterk >= 2, "the L-BFGS direction needs Iterk >= 2");
    HIPCHK(hipSetDevice(h->ctx->device));
    hipStream_t st = h->ctx->stream;
    const long N = h->N;
    const int LM = h->LM, SC_A = SC_RHO + LM;
    const size_t bytes = (size_t)N * sizeof(double);
    HIPCHK(hipMemcpyAsync(h->g1, gk1, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(h->g0, gk0, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(h->dk, dk, bytes, hipMemcpyHostToDevice, st));
    // the new pair in slot (Iterk - 2) % LM, rho = 1 / s.y, gamma = s.y / y.y
    const int slot = (Iterk - 2) % LM;
    double *s = h->sk + (size_t)slot * N, *y = h->yk + (size_t)slot * N;
    hipLaunchKernelGGL(pair_store, grid_of(N), dim3(256), 0, st, s, y, h->dk, h->g1, h->g0, alpha1, N);
    HIPCHK(hipGetLastError());
    HIPCHK(dot_dev(h, s, y, kRho, SC_RHO + slot, SC_SY));
    HIPCHK(dot_dev(h, y, y, kRatio, SC_GAMMA, SC_SY));
    // two-loop recursion over the last min(LM, Iterk - 1) pairs, newest first
    // (src/T_PFAST3D.cpp:462-481; ring slot i % LM)
    const int lo = Iterk - LM - 1 >= 0 ? Iterk - LM - 1 : 0;
    HIPCHK(hipMemcpyAsync(h->q, h->g1, bytes, hipMemcpyDeviceToDevice, st));
    for (int i = Iterk - 2; i >= lo; i--) {
        const int m = i % LM;
        HIPCHK(dot_dev(h, h->sk + (size_t)m * N, h->q, kScaled, SC_A + m, SC_RHO + m));
        hipLaunchKernelGGL(axpy_sc, grid_of(N), dim3(256), 0, st, h->q, h->yk + (size_t)m * N, h->sc,
                           SC_A + m, -1, N);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(scale_sc, grid_of(N), dim3(256), 0, st, h->dk, h->q, h->sc, SC_GAMMA, N);
    HIPCHK(hipGetLastError());
    for (int i = lo; i <= Iterk - 2; i++) {
        const int m = i % LM;
        HIPCHK(dot_dev(h, h->yk + (size_t)m * N, h->dk, kScaled, SC_B, SC_RHO + m));
        hipLaunchKernelGGL(axpy_sc, grid_of(N), dim3(256), 0, st, h->dk, h->sk + (size_t)m * N, h->sc,
                           SC_A + m, SC_B, N);
        HIPCHK(hipGetLastError());
    }
    double g = 0.0;
    HIPCHK(hipMemcpyAsync(dk, h->dk, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&g, h->sc + SC_GAMMA, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (gamma) *gamma = g;
    return PFAST_OK;
}

int pfast_lbfgs_dot(pfast_lbfgs *h, const double *a, const double *b, double *out)
{
    ARGCHK(h && a && b && out, "NULL argument");
    HIPCHK(hipSetDevice(h->ctx->device));
    hipStream_t st = h->ctx->stream;
    const size_t bytes = (size_t)h->N * sizeof(double);
    HIPCHK(hipMemcpyAsync(h->g1, a, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(h->g0, b, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(dot_dev(h, h->g1, h->g0, kStore, SC_DOT, 0));
    HIPCHK(hipMemcpyAsync(out, h->sc + SC_DOT, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PFAST_OK;
}

}  // extern "C"
