// ref_glue3d.cpp -- exposes the reference's own 3D routines (compiled from
// /root/reference/src headers, never copied) through extern "C" names so the
// tests can pin the oracle against them.  Built only into oracle/_ref/ by
// oracle/Makefile; test infrastructure, never part of the product.
#include "T_pfsm3d.h"
#include "T_padjoint3d.h"

extern "C" {
void ref_fsm3d(double *vp, double *U, int *src, int srcid, int MAXSwp, double dx, double dy,
               double dz, int nrows, int ncols, int nz)
{ fsm3d(vp, U, src, srcid, MAXSwp, dx, dy, dz, nrows, ncols, nz, 0); }
double ref_eikslv3d(double *U, double *vp, int *node, double dx, double dy, double dz,
                    int nrows, int ncols, int nz)
{ return eikslv3d(U, vp, node, dx, dy, dz, nrows, ncols, nz); }
double ref_adjoint_fsm3d(double *U, double *T, double *T_obs, int *rec, double *nxyz, int *src,
                         int *srnum, int srcid, int MAXSwp, double dx, double dy, double dz,
                         int nrows, int ncols, int nz)
{ return adjoint_fsm3d(U, T, T_obs, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dy, dz, nrows, ncols, nz, 0); }
void ref_regu3d(double *Dj, double *vp, double nux, double nuy, double nuz, double dx, double dy,
                double dz, int nrows, int ncols, int nz)
{ regu3d(Dj, vp, nux, nuy, nuz, dx, dy, dz, nrows, ncols, nz); }
void ref_updatevp3d(double *vp, double *dm, double step, int nrows, int ncols, int nz)
{ updatevp(vp, dm, step, nrows, ncols, nz); }
double ref_gd2sk3d(double *gk1, double *gk0, double *dk, double alpha1, double *rhok, double *sk,
                   double *yk, int Iterk, int lm, int nrows, int ncols, int nz)
{ return gd2sk(gk1, gk0, dk, alpha1, rhok, sk, yk, Iterk, lm, nrows, ncols, nz); }
double ref_magnitudeOf3d(double *g, int nrows, int ncols, int nz) { return magnitudeOf(g, nrows, ncols, nz); }
}
