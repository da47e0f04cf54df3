// ref_glue2d.cpp -- exposes the reference's own 2D routines (compiled from
// /root/reference/src headers, never copied) through extern "C" names so the
// tests can pin the oracle against them.  Built only into oracle/_ref/ by
// oracle/Makefile; test infrastructure, never part of the product.
#include "RT_pfsm2d.h"
#include "RT_padjoint2d.h"

extern "C" {
void ref_rt_fsm2d(double *vp, double *U, double *TTd, double *TTu, double *Z, int *src, int *rec,
                  int *srnum, int srcid, int NumofR, int MAXSwp, double dx, double dz, int nrows,
                  int ncols)
{ RT_fsm2d(vp, U, TTd, TTu, Z, src, rec, srnum, srcid, NumofR, MAXSwp, dx, dz, nrows, ncols, 0); }
double ref_eikslv2d(double *U, double *vp, int *node, double dx, double dz, int nrows, int ncols)
{ return eikslv2d(U, vp, node, dx, dz, nrows, ncols); }
double ref_tadjoint_fsm2d(double *U, double *T, double *T_obs, int *rec, double *nxyz, int *src,
                          int *srnum, int srcid, int MAXSwp, double dx, double dz, int nrows, int ncols)
{ return Tadjoint_fsm2d(U, T, T_obs, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dz, nrows, ncols, 0); }
double ref_tadjoint_fsm2d_wr(double *U, double *T, double *T_obs, double WR, int *rec, double *nxyz,
                             int *src, int *srnum, int srcid, int MAXSwp, double dx, double dz,
                             int nrows, int ncols)
{ return Tadjoint_fsm2d(U, T, T_obs, WR, rec, nxyz, src, srnum, srcid, MAXSwp, dx, dz, nrows, ncols, 0); }
double ref_radjoint_fsm2d(double *UR, double *Td, double *Tu, double *TR_obs, int *rec, double *nxyz,
                          int *src, int *srnum, int srcid, int NumofR, int MAXSwp, double dx,
                          double dz, int nrows, int ncols)
{ return Radjoint_fsm2d(UR, Td, Tu, TR_obs, rec, nxyz, src, srnum, srcid, NumofR, MAXSwp, dx, dz, nrows, ncols, 0); }
double ref_radjoint_fsm2d_wr(double *UR, double *Td, double *Tu, double *TR_obs, double WR, int *rec,
                             double *nxyz, int *src, int *srnum, int srcid, int NumofR, int MAXSwp,
                             double dx, double dz, int nrows, int ncols)
{ return Radjoint_fsm2d(UR, Td, Tu, TR_obs, WR, rec, nxyz, src, srnum, srcid, NumofR, MAXSwp, dx, dz, nrows, ncols, 0); }
void ref_regu2d(double *Dj, double *vp, double nux, double nuz, double dx, double dz, int nrows, int ncols)
{ regu2d(Dj, vp, nux, nuz, dx, dz, nrows, ncols); }
void ref_updatevp2d(double *vp, double *dm, double step, int nrows, int ncols) { updatevp(vp, dm, step, nrows, ncols); }
double ref_gd2sk2d(double *gk1, double *gk0, double *dk, double alpha1, double *rhok, double *sk,
                   double *yk, int Iterk, int lm, int nrows, int ncols)
{ return gd2sk(gk1, gk0, dk, alpha1, rhok, sk, yk, Iterk, lm, nrows, ncols); }
double ref_beta_HS2d(double *gk1, double *gk0, double *dk, int nrows, int ncols) { return beta_HS(gk1, gk0, dk, nrows, ncols); }
double ref_beta_CGDESCENT2d(double *gk1, double *gk0, double *dk, int nrows, int ncols) { return beta_CGDESCENT(gk1, gk0, dk, nrows, ncols); }
double ref_cubicInterp(double a0, double a1, double p0, double p1, double d0, double d1) { return cubicInterp(a0, a1, p0, p1, d0, d1); }
double ref_magnitudeOf2d(double *g, int nrows, int ncols) { return magnitudeOf(g, nrows, ncols); }
}
