// optim.hip -- gradient regularisation (GPU, hipFFT) and the optimiser helpers
// the reference's drivers call (host arithmetic in the reference's order).
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/pfast.h"
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
