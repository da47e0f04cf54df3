// plain_kernels.h -- launchers of the single-source FP64 and batched 2D kernels.
#pragma once
#include <hip/hip_runtime.h>

struct Grid2 {
    int nrows, ncols;
    double dx, dz, wx, wz;  // spacings (x = columns, z = rows) and 1/h^2
    // FP32 batched form: 1 = row-window sweeps and independent solves of a source in
    // blocks of their own (default), 0 = plane loop, one block per source (same results)
    int window = 1;
    const float *slow = nullptr;  // (float)(1 / vp) per node (the model's), for the row window
};

namespace pfast {
hipError_t launch_rt_fsm2d_f64(hipStream_t st, const double *vp, double *U, double *TTd,
                               double *TTu, double *p, const double *Z, const int *src, int nb,
                               int NumofR, int MAXSwp, Grid2 g, int *cycles, int *status,
                               int *k2out);
hipError_t launch_rt_fsm2d_f32(hipStream_t st, const double *vp, float *U, float *TTd,
                               float *TTu, const double *Z, const int *src, int nb, int NumofR,
                               int MAXSwp, Grid2 g, int *cycles, int *status, int *k2out);
hipError_t launch_tadj2d_f64(hipStream_t st, double *U, double *p, unsigned char *tag,
                             const double *T, const double *T_obs, double Cr, const int *rec,
                             const double *nxy, const long *ch0, const int *srnum,
                             const int *src, int nb, int MAXSwp, Grid2 g, double *EE,
                             int *cycles);
hipError_t launch_tadj2d_f32(hipStream_t st, float *U, unsigned char *tag, const float *T,
                             const double *T_obs, double Cr, const int *rec, const double *nxy,
                             const long *ch0, const int *srnum, const int *src, int nb,
                             int MAXSwp, Grid2 g, double *EE, int *cycles);
hipError_t launch_radj2d_f64(hipStream_t st, double *UR, double *U, double *p,
                             unsigned char *tag, const double *Tu, const double *TR_obs,
                             double Cr, const int *rec, const double *nxy, const long *ch0,
                             const int *srnum, const int *src, int nb, int NumofR, int MAXSwp,
                             Grid2 g, double *EE, int *cycles);
hipError_t launch_radj2d_f32(hipStream_t st, float *UR, float *U, unsigned char *tag,
                             const float *Tu, const double *TR_obs, double Cr, const int *rec,
                             const double *nxy, const long *ch0, const int *srnum,
                             const int *src, int nb, int NumofR, int MAXSwp, Grid2 g,
                             double *EE, int *cycles);
// Both FP32 adjoints of a batch in one launch (separate tag scratch for each).
hipError_t launch_adj2d_pair_f32(hipStream_t st, float *LT, unsigned char *tagT, const float *T,
                                 const double *T_obs, float *LR, float *Uw, unsigned char *tagR,
                                 const float *Tu, const double *TR_obs, double Cr, const int *rec,
                                 const double *nxy, const long *ch0, const int *srnum,
                                 const int *src, int nb, int NumofR, int MAXSwp, Grid2 g,
                                 double *EEt, int *tcyc, double *EEr, int *rcyc);
hipError_t launch_init3d_f64(hipStream_t st, double *U, long N, long srcnode);
hipError_t launch_fsm3d_ordering_f64(hipStream_t st, const double *vp, double *U, double *p,
                                     int *flag, int si, int sj, int sk, int s1, int s2, int s3,
                                     double dx, double dy, double dz, int nrows, int ncols,
                                     int nz, int first, int last);
hipError_t launch_adj3d_bc_f64(hipStream_t st, double *U, unsigned char *tag, const double *T,
                               const double *T_obs, const int *rec, const double *nxyz, long ch0,
                               int nrec, int si, int sj, int sk, double dx, double dy, double dz,
                               int nrows, int ncols, int nz, double *EE, int *flag);
hipError_t launch_adj3d_ordering_f64(hipStream_t st, double *U, const double *T,
                                     const unsigned char *tag, double *p, int *flag, int s1,
                                     int s2, int s3, double dx, double dy, double dz, int nrows,
                                     int ncols, int nz, int first, int last);
}  // namespace pfast
