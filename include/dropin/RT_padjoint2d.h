// Drop-in replacement of the reference's src/RT_padjoint2d.h (2D transmission and
// reflection adjoints, both overloads each, and the optimiser helpers).
#ifndef PFAST_DROPIN_RT_PADJOINT2D_H
#define PFAST_DROPIN_RT_PADJOINT2D_H

#include <cmath>

#include "RT_pfsm2d.h"

#ifndef POS
#define POS(x) (((x) + fabs(x)) / 2.0)
#endif
#ifndef NEG
#define NEG(x) (((x)-fabs(x)) / 2.0)
#endif
#ifndef MAXV
#define MAXV 4500
#endif
#ifndef MINV
#define MINV 1600
#endif

inline void swap(double &a, double &b)
{
    double t = a;
    a = b;
    b = t;
}

// Adjoint local solve (host helper), src/RT_padjoint2d.h:42-65.
inline double adjoint_slv2d(double *U, double *T, int *OneNode, double dx, double dz, int nrows,
                            int ncols)
{
    (void)nrows;
    const long n = (long)OneNode[0] * ncols + OneNode[1];
    const double tc = T[n], txp = T[n + 1], txm = T[n - 1], tzp = T[n + ncols], tzm = T[n - ncols];
    if (txp >= INF || tc >= INF || txm >= INF || tzp >= INF || tzm >= INF) return 0.0;
    double a1 = POS(-(txp - tc) / dx), a2 = NEG(-(tc - txm) / dx);
    double b1 = POS(-(tzp - tc) / dz), b2 = NEG(-(tc - tzm) / dz);
    double La1 = POS(-(tc - txm) / dx) * U[n - 1], La2 = NEG(-(txp - tc) / dx) * U[n + 1];
    double Lb1 = POS(-(tc - tzm) / dz) * U[n - ncols], Lb2 = NEG(-(tzp - tc) / dz) * U[n + ncols];
    double den = (a1 - a2) / dx + (b1 - b2) / dz;
    if (fabs(den) > 1e-9) return ((La1 - La2) / dx + (Lb1 - Lb2) / dz) / den;
    return 0.0;
}

inline double pfast_dropin_tadj(double *U, double *T, double *T_obs, double Cr, int *rec, double *nxyz,
                                int *src, int *srnum, int srcid, int MAXSwp, double dx, double dz,
                                int nrows, int ncols, int pe)
{
    double EE = 0.0;
    int cyc = 0;
    if (pfast_tadjoint_fsm2d(U, T, T_obs, Cr, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dz, nrows,
                             ncols, pe, &EE, &cyc) != PFAST_OK) {
        printf("PE_%d: Tadjoint_fsm2d failed: %s\n", pe, pfast_last_error());
        exit(1);
    }
    if (cyc >= MAXSwp)
        printf("|\tPE_%d: ....Warning!  Adjoint Equation of Transmission NOT Converged after %d "
               "sweeps...\n", pe, MAXSwp);
    return EE;
}

inline double pfast_dropin_radj(double *UR, double *Td, double *Tu, double *TR_obs, double Cr,
                                int *rec, double *nxyz, int *src, int *srnum, int srcid, int NumofR,
                                int MAXSwp, double dx, double dz, int nrows, int ncols, int pe)
{
    double EE = 0.0;
    if (pfast_radjoint_fsm2d(UR, Td, Tu, TR_obs, Cr, rec, nxyz, src, srnum, srcid, NumofR, MAXSwp,
                             dx, dz, nrows, ncols, pe, &EE, NULL) != PFAST_OK) {
        printf("PE_%d: Radjoint_fsm2d failed: %s\n", pe, pfast_last_error());
        exit(1);
    }
    return EE;
}

// Weighted overloads (Cr = 1-WR / WR), src/RT_padjoint2d.h:67 and :159.
inline double Tadjoint_fsm2d(double *U, double *T, double *T_obs, double WR, int *rec, double *nxyz,
                             int *src, int *srnum, int srcid, int MAXSwp, double dx, double dz,
                             int nrows, int ncols, int pe)
{
    return pfast_dropin_tadj(U, T, T_obs, 1 - WR, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dz,
                             nrows, ncols, pe);
}

inline double Radjoint_fsm2d(double *UR, double *Td, double *Tu, double *TR_obs, double WR, int *rec,
                             double *nxyz, int *src, int *srnum, int srcid, int NumofR, int MAXSwp,
                             double dx, double dz, int nrows, int ncols, int pe)
{
    return pfast_dropin_radj(UR, Td, Tu, TR_obs, WR, rec, nxyz, src, srnum, srcid, NumofR, MAXSwp,
                             dx, dz, nrows, ncols, pe);
}

// Plain overloads (Cr = 1), src/RT_padjoint2d.h:288 and :382 (the ones the driver calls).
inline double Tadjoint_fsm2d(double *U, double *T, double *T_obs, int *rec, double *nxyz, int *src,
                             int *srnum, int srcid, int MAXSwp, double dx, double dz, int nrows,
                             int ncols, int pe)
{
    return pfast_dropin_tadj(U, T, T_obs, 1.0, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dz, nrows,
                             ncols, pe);
}

inline double Radjoint_fsm2d(double *UR, double *Td, double *Tu, double *TR_obs, int *rec,
                             double *nxyz, int *src, int *srnum, int srcid, int NumofR, int MAXSwp,
                             double dx, double dz, int nrows, int ncols, int pe)
{
    return pfast_dropin_radj(UR, Td, Tu, TR_obs, 1.0, rec, nxyz, src, srnum, srcid, NumofR, MAXSwp,
                             dx, dz, nrows, ncols, pe);
}

// src/RT_padjoint2d.h:475-494 (2D stops at a relative misfit decrease <= 1e-3).
inline bool check_conv(double *dvp, double stepsize, double *vp, double *dEE, int Iterk, double dt,
                       int nrows, int ncols)
{
    (void)dt;
    const long N = (long)nrows * ncols;
    double chgE = Iterk >= 2 ? (dEE[Iterk - 2] - dEE[Iterk - 1]) / dEE[Iterk - 2] : 1;
    double mynorm = 0;
    for (long k = 0; k < N; k++) {
        double v = stepsize * stepsize * dvp[k] * dvp[k] / vp[k] / vp[k];
        if (v >= mynorm) mynorm = v;
    }
    mynorm = sqrt(mynorm);
    if (chgE <= 0.001) return true;
    if (mynorm / nrows / ncols >= 1) {
        printf("Over corrected..abort...\n");
        exit(1);
    }
    return false;
}

inline void updatevp(double *vp, double *dm, double stepsize, int nrows, int ncols)
{
    const long N = (long)nrows * ncols;
    for (long k = 0; k < N; k++) {
        if (fabs(vp[k] - 330) > 1e-9) {
            vp[k] = vp[k] + stepsize * dm[k];
            if (vp[k] > MAXV) vp[k] = MAXV;
            if (vp[k] < MINV) vp[k] = MINV;
        }
    }
}

inline void regu2d(double *Dj, double *vp, double nux, double nuz, double dx, double dz, int nrows,
                   int ncols)
{
    if (pfast_regu2d(Dj, vp, nux, nuz, dx, dz, nrows, ncols) != PFAST_OK) {
        printf("regu2d failed: %s\n", pfast_last_error());
        exit(1);
    }
}

inline double beta_HS(double *gk1, double *gk0, double *dk, int nrows, int ncols)
{
    return pfast_beta_hs(gk1, gk0, dk, (long)nrows * ncols);
}

inline double beta_CGDESCENT(double *gk1, double *gk0, double *dk, int nrows, int ncols)
{
    return pfast_beta_cgdescent(gk1, gk0, dk, (long)nrows * ncols);
}

inline double gd2sk(double *gk1, double *gk0, double *dk, double alpha1, double *rhok, double *sk,
                    double *yk, int Iterk, int lm, int nrows, int ncols)
{
    return pfast_gd2sk(gk1, gk0, dk, alpha1, rhok, sk, yk, Iterk, lm, (long)nrows * ncols);
}

inline double cubicInterp(double alpha0, double alpha1, double phi0, double phi1, double dphi0,
                          double dphi1)
{
    return pfast_cubic_interp(alpha0, alpha1, phi0, phi1, dphi0, dphi1);
}

inline double magnitudeOf(double *g, int nrows, int ncols) { return pfast_magnitude_of(g, (long)nrows * ncols); }

inline double meanOf(double *g, int nrows, int ncols) { return pfast_mean_of(g, (long)nrows * ncols); }

inline void RTinfo(int RTtag)
{
    const char *m = RTtag == 3 ? "Joint Inversion of Reflection and Transmission"
                               : (RTtag == 1 ? "Inversion of Transmission ONLY"
                                             : "Inversion of Reflection ONLY");
    printf("|| 2D %s ||\n", m);
}

#endif
