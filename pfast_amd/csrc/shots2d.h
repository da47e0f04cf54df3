// shots2d.h -- the batched 2D model and shot-set objects and the stages of one
// run, shared by the C-ABI entry points and the request server.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "context.h"

struct pfast_model2d {
    pfast_ctx *ctx;
    int nrows, ncols, NumofR;
    double dx, dz;
    double *vp = nullptr;  // N doubles, reference layout vp[i * ncols + j]
    float *slow = nullptr; // N floats, (float)(1 / vp): the FP32 sweeps' slowness
    double *Z = nullptr;   // NumofR * ncols * 3 (depth, n_z, n_x), reference layout
    int nshots = 0;
};

struct pfast_shots2d {
    pfast_model2d *m;
    int nsrc;
    long nrec;
    std::vector<long> ch0;
    int *src = nullptr, *rec = nullptr, *srnum = nullptr;
    long *dch0 = nullptr;
    double *nxy = nullptr, *TT_obs = nullptr, *TR_obs = nullptr;
    float *TT = nullptr, *TTd = nullptr, *TTu = nullptr;  // [s][N], [s][R][N]
    float *LT = nullptr, *LR = nullptr, *Uw = nullptr;    // adjoints, work field
    unsigned char *tag = nullptr;
    double *EEt = nullptr, *EEr = nullptr;
    int *fcyc = nullptr, *tcyc = nullptr, *rcyc = nullptr, *status = nullptr, *k2 = nullptr;
    double *gT = nullptr, *gR = nullptr, *gRT = nullptr, *WR = nullptr, *fieldbuf = nullptr;
    bool have_fwd = false, have_adj = false;
};

namespace pfast {

// Stages of pfast_shots2d_run, asynchronous on the model's stream.
int issue_fwd2d(pfast_shots2d *s, int MAXSwp, int *status_host);  // RT_fsm2d; per-source status
                                                                   // copied to status_host
int check_fwd2d(pfast_shots2d *s, const int *status);               // fatal status -> error
int issue_adj2d(pfast_shots2d *s, int MAXSwp);                      // both adjoints
int issue_grad2d(pfast_shots2d *s, const double *WR_dev);           // gT, gR, gRT on the device
void shots2d_free(pfast_shots2d *s);  // no synchronisation: no device work may still use it
size_t footprint2d(int nrows, int ncols, int NumofR, int nsrc, long nrec, bool adjoint);

}  // namespace pfast
