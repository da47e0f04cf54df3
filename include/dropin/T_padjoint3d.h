// Drop-in replacement of the reference's src/T_padjoint3d.h (3D adjoint state and
// the optimiser helpers the 3D driver calls).
//
// adjoint_fsm3d and regu3d run on the GPU through libpfast (pfast_adjoint_fsm3d,
// FP64 and bit-identical; pfast_regu3d, hipFFT).  The remaining helpers are host
// code with the reference's semantics; deviations are noted beside them.
#ifndef PFAST_DROPIN_T_PADJOINT3D_H
#define PFAST_DROPIN_T_PADJOINT3D_H

#include <cmath>

#include "T_pfsm3d.h"

#ifndef POS
#define POS(x) (((x) + fabs(x)) / 2.0)
#endif
#ifndef NEG
#define NEG(x) (((x)-fabs(x)) / 2.0)
#endif
#ifndef MAXV
#define MAXV 8000
#endif
#ifndef MINV
#define MINV 1400
#endif

inline void swap(double &a, double &b)
{
    double t = a;
    a = b;
    b = t;
}

// Adjoint local solve (host helper), src/T_padjoint3d.h:42-71.
inline double adjoint_slv3d(double *U, double *T, int *OneNode, double dx, double dy, double dz,
                            int nrows, int ncols, int nz, int *src, int srcid)
{
    (void)src;
    (void)srcid;
    const long plane = (long)nrows * ncols;
    const long n = OneNode[2] * plane + (long)OneNode[0] * ncols + OneNode[1];
    const double tc = T[n], txp = T[n + 1], txm = T[n - 1], typ = T[n + ncols], tym = T[n - ncols],
                 tzp = T[n + plane], tzm = T[n - plane];
    if (tc >= INF || txp >= INF || txm >= INF || typ >= INF || tym >= INF || tzp >= INF || tzm >= INF)
        return 0.0;
    double a1 = POS(-(txp - tc) / dx), a2 = NEG(-(tc - txm) / dx);
    double b1 = POS(-(typ - tc) / dy), b2 = NEG(-(tc - tym) / dy);
    double c1 = POS(-(tzp - tc) / dz), c2 = NEG(-(tc - tzm) / dz);
    double La1 = POS(-(tc - txm) / dx) * U[n - 1], La2 = NEG(-(txp - tc) / dx) * U[n + 1];
    double Lb1 = POS(-(tc - tym) / dy) * U[n - ncols], Lb2 = NEG(-(typ - tc) / dy) * U[n + ncols];
    double Lc1 = POS(-(tc - tzm) / dz) * U[n - plane], Lc2 = NEG(-(tzp - tc) / dz) * U[n + plane];
    double den = (a1 - a2) / dx + (b1 - b2) / dy + (c1 - c2) / dz;
    if (fabs(den) > 1e-9) return ((La1 - La2) / dx + (Lb1 - Lb2) / dy + (Lc1 - Lc2) / dz) / den;
    return 0.0;
}

// src/T_padjoint3d.h:73-172, on the GPU.  Returns this source's misfit EE.
inline double adjoint_fsm3d(double *U, double *T, double *T_obs, int *rec, double *nxyz, int *src,
                            int *srnum, int srcid, int MAXSwp, double dx, double dy, double dz,
                            int nrows, int ncols, int nz, int pe)
{
    double EE = 0.0;
    int cycles = 0;
    int rc = pfast_adjoint_fsm3d(U, T, T_obs, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dy, dz,
                                 nrows, ncols, nz, pe, &EE, &cycles);
    if (rc != PFAST_OK) {
        printf("PE_%d: adjoint_fsm3d failed: %s\n", pe, pfast_last_error());
        exit(1);
    }
    if (cycles >= MAXSwp)
        printf("|\tPE_%d: ....Warning! Adjoint Equation NOT Converged after %d sweeps...\n", pe,
               MAXSwp);
    return EE;
}

// src/T_padjoint3d.h:174-193.
inline bool check_conv(double *dvp, double stepsize, double *vp, double eps, double *dEE, int Iterk,
                       int nrows, int ncols, int nz)
{
    (void)eps;
    const long N = (long)nrows * ncols * nz;
    double chgE = Iterk >= 2 ? (dEE[Iterk - 2] - dEE[Iterk - 1]) / dEE[Iterk - 2] : 1;
    double mynorm = 0;
    for (long k = 0; k < N; k++) {
        double v = stepsize * stepsize * dvp[k] * dvp[k] / vp[k] / vp[k];
        if (v >= mynorm) mynorm = v;
    }
    mynorm = sqrt(mynorm);
    if (chgE <= 1e-2) return true;
    if (mynorm / nrows / ncols / nz >= 1) {
        printf("Over corrected..abort...\n");
        exit(1);
    }
    return false;
}

// Clamped update of non-air nodes, src/T_padjoint3d.h:195-203.
inline void updatevp(double *vp, double *dm, double stepsize, int nrows, int ncols, int nz)
{
    const long N = (long)nrows * ncols * nz;
    for (long k = 0; k < N; k++) {
        if (fabs(vp[k] - 330) > 1e-9) {
            vp[k] = vp[k] + stepsize * dm[k];
            if (vp[k] > MAXV) vp[k] = MAXV;
            if (vp[k] < MINV) vp[k] = MINV;
        }
    }
}

// FFT smoothing of the gradient, src/T_padjoint3d.h:205-262, on the GPU.
inline void regu3d(double *Dj, double *vp, double nux, double nuy, double nuz, double dx, double dy,
                   double dz, int nrows, int ncols, int nz)
{
    if (pfast_regu3d(Dj, vp, nux, nuy, nuz, dx, dy, dz, nrows, ncols, nz) != PFAST_OK) {
        printf("regu3d failed: %s\n", pfast_last_error());
        exit(1);
    }
}

inline double beta_HS(double *gk1, double *gk0, double *dk, int nrows, int ncols, int nz)
{
    return pfast_beta_hs(gk1, gk0, dk, (long)nrows * ncols * nz);
}

inline double beta_CGDESCENT(double *gk1, double *gk0, double *dk, int nrows, int ncols, int nz)
{
    return pfast_beta_cgdescent(gk1, gk0, dk, (long)nrows * ncols * nz);
}

// L-BFGS pair storage, src/T_padjoint3d.h:289-304.
inline double gd2sk(double *gk1, double *gk0, double *dk, double alpha1, double *rhok, double *sk,
                    double *yk, int Iterk, int lm, int nrows, int ncols, int nz)
{
    return pfast_gd2sk(gk1, gk0, dk, alpha1, rhok, sk, yk, Iterk, lm, (long)nrows * ncols * nz);
}

inline double zoom(double alpha0, double alpha1, double phi0, double phi1, double dphi0,
                   double dphi1)
{
    return pfast_zoom(alpha0, alpha1, phi0, phi1, dphi0, dphi1);
}

// src/T_padjoint3d.h:316-326.  Where the reference falls off the end of the
// function (undefined result) this returns the secant step.
inline double cubicInterp(double alpha0, double alpha1, double phi0, double phi1, double dphi0,
                          double dphi1)
{
    return pfast_cubic_interp(alpha0, alpha1, phi0, phi1, dphi0, dphi1);
}

inline double magnitudeOf(double *g, int nrows, int ncols, int nz)
{
    return pfast_magnitude_of(g, (long)nrows * ncols * nz);
}

inline double meanOf(double *g, int nrows, int ncols, int nz)
{
    return pfast_mean_of(g, (long)nrows * ncols * nz);
}

#endif
