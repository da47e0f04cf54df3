/*
 * pfast_oracle.h -- FP64 CPU restatement of the PFAST traveltime-and-gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (pfast_amd/, libpfast.so)
 * links, loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / CPU baseline.
 *
 * Every function restates one reference routine (file:line cited per function in
 * pfast_oracle.c) with the same loop order, the same floating-point operation order
 * and the same convergence rules, compiled with -ffp-contract=off so that the
 * results are bit-identical to the reference's FP64 code on the same inputs.
 *
 * Layouts are the reference's: 3D U[k*nrows*ncols + i*ncols + j], 2D U[i*ncols + j],
 * src/rec are DIM-strided int index triples (3D: row i, col j, depth k; 2D: row i
 * (depth), col j).
 */
#ifndef PFAST_ORACLE_H
#define PFAST_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_INF 9999999.0

/* ---- 3D ---------------------------------------------------------------- */
double orc_eikslv3d(const double *U, const double *vp, int i, int j, int k,
                    double dx, double dy, double dz, int nrows, int ncols, int nz);
/* returns the number of sweep cycles executed (reference `count` capped at MAXSwp) */
int orc_fsm3d(const double *vp, double *U, const int *src, int srcid, int MAXSwp,
              double dx, double dy, double dz, int nrows, int ncols, int nz);
double orc_adjoint_slv3d(const double *U, const double *T, int i, int j, int k,
                         double dx, double dy, double dz, int nrows, int ncols, int nz);
/* returns EE (misfit of this source); *ncycles receives the sweep cycles executed */
double orc_adjoint_fsm3d(double *U, const double *T, const double *T_obs, const int *rec,
                         const double *nxyz, const int *src, const int *srnum, int srcid,
                         int MAXSwp, double dx, double dy, double dz,
                         int nrows, int ncols, int nz, int *ncycles);

/* ---- 2D ---------------------------------------------------------------- */
double orc_eikslv2d(const double *U, const double *vp, int i, int j,
                    double dx, double dz, int nrows, int ncols);
/* status 0 ok, 1..3 = the reference's fatal exits (no k2 / negative reflected time).
 * ncycles[0] = transmission cycles, ncycles[1+2*ir] down-going, ncycles[2+2*ir] up-going.
 * k2vec_out (may be NULL) receives the last reflector's k2vec (ncols ints). */
int orc_rt_fsm2d(const double *vp, double *U, double *TTd, double *TTu, const double *Z,
                 const int *src, int srcid, int NumofR, int MAXSwp,
                 double dx, double dz, int nrows, int ncols, int *ncycles, int *k2vec_out);
double orc_adjoint_slv2d(const double *U, const double *T, int i, int j,
                         double dx, double dz, int nrows, int ncols);
double orc_tadjoint_fsm2d(double *U, const double *T, const double *T_obs, double Cr,
                          const int *rec, const double *nxyz, const int *src,
                          const int *srnum, int srcid, int MAXSwp, double dx, double dz,
                          int nrows, int ncols, int *ncycles);
double orc_radjoint_fsm2d(double *UR, const double *Tu, const double *TR_obs, double Cr,
                          const int *rec, const double *nxyz, const int *srnum, int srcid,
                          int NumofR, int MAXSwp, double dx, double dz,
                          int nrows, int ncols, int *ncycles);

/* ---- geometry and accumulation ------------------------------------------ */
/* idx = (int)ceil(coord/h) - 1 for n rows of a stride-`stride` double table, column `col` */
void orc_map_index(const double *table, int n, int stride, int col, double h,
                   int *out, int out_stride);
/* g[n] = g[n] - lam[n]/(vp[n]*vp[n]*vp[n]) */
void orc_grad_accum(double *g, const double *lam, const double *vp, long N);

/* ---- CPU baseline: rank-per-contiguous-source-block, P threads ----------- */
/* Forward + adjoint + gradient for sources [s0, s1) split over P threads like the
 * reference's MPI decomposition.  grad (N doubles) and *misfit are the sums.
 * Returns 0 on success. */
int orc_fwdadj3d_parallel(int P, const double *vp, const int *src, const int *srnum,
                          const int *rec, const double *nxyz, const double *T_obs,
                          int s0, int s1, int MAXSwp, double dx, double dy, double dz,
                          int nrows, int ncols, int nz, double *grad, double *misfit,
                          int *fcycles, int *acycles);

#ifdef __cplusplus
}
#endif
#endif
