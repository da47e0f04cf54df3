// Drop-in replacement of the reference's src/RT_pfsm2d.h (2D transmission +
// reflection fast sweeping).  Same names and signatures; RT_fsm2d runs on the GPU
// through libpfast (pfast_rt_fsm2d, FP64, bit-identical to the reference).
#ifndef PFAST_DROPIN_RT_PFSM2D_H
#define PFAST_DROPIN_RT_PFSM2D_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <iostream>

#include "../pfast.h"

#ifndef MAX
#define MAX(a, b) (((a) > (b)) ? (a) : (b))
#endif
#ifndef MIN
#define MIN(a, b) (((a) < (b)) ? (a) : (b))
#endif
#ifndef INF
#define INF 9999999
#endif
#ifndef DIM
#define DIM 2
#endif
#ifndef PI
#define PI (3.141592653589793)
#endif

using namespace std;

// src/RT_pfsm2d.h:39-57 (see the 3D twin for the unopened-file behaviour).
inline void readmod(double *mod, char *file, int nrows, int ncols)
{
    FILE *fp = fopen(file, "rb");
    if (!fp) {
        printf(" The file %s cannot be opened\n", file);
        exit(1);
    }
    printf(" \n...reading file:\n\t %s\n", file);
    size_t n = (size_t)nrows * ncols;
    if (fread(mod, sizeof(double), n, fp) != n) printf(" short read from %s\n", file);
    fclose(fp);
}

inline void writemod(double *mod, char *file, int nrows, int ncols)
{
    FILE *fp = fopen(file, "wb");
    if (!fp) {
        printf(" The file %s cannot be opened\n", file);
        return;
    }
    printf(" \n...writing file:\n\t %s\n", file);
    fwrite(mod, sizeof(double), (size_t)nrows * ncols, fp);
    fclose(fp);
}

// Local 2D Godunov update (host helper), src/RT_pfsm2d.h:71-108.
inline double eikslv2d(double *U, double *vp, int *node, double dx, double dz, int nrows, int ncols)
{
    const int i = node[0], j = node[1];
    const long n = (long)i * ncols + j;
    double s = 1.0 / vp[n], ha = dx, hb = dz, a, b, t, q, ta, tb, tc, disc;
    b = i - 1 < 0 ? U[n + ncols] : (i + 1 >= nrows ? U[n - ncols] : MIN(U[n - ncols], U[n + ncols]));
    a = j - 1 < 0 ? U[n + 1] : (j + 1 >= ncols ? U[n - 1] : MIN(U[n + 1], U[n - 1]));
    if (b > a) { t = a; a = b; b = t; t = ha; ha = hb; hb = t; }
    q = b + hb * s;
    if (q <= a) return q;
    ta = 1.0 / ha / ha + 1.0 / hb / hb;
    tb = -2. * (a / ha / ha + b / hb / hb);
    tc = a * a / ha / ha + b * b / hb / hb - s * s;
    disc = tb * tb - 4. * ta * tc;
    return disc >= 0 ? (-tb + sqrt(disc)) / 2. / ta : INF;
}

inline bool isconv(double *p, double *U, int nrows, int ncols)
{
    const long N = (long)nrows * ncols;
    for (long n = 0; n < N; n++)
        if (fabs(p[n] - U[n]) > 1e-9) return false;
    return true;
}

// src/RT_pfsm2d.h:116-380, on the GPU.  The reference's fatal conditions
// (no reflection point above a reflector, negative reflected time) end the
// program as there.
inline void RT_fsm2d(double *vp, double *U, double *TTd, double *TTu, double *Z, int *src, int *rec,
                     int *srnum, int srcid, int NumofR, int MAXSwp, double dx, double dz, int nrows,
                     int ncols, int pe)
{
    int cyc[1 + 2 * 64];
    int *cp = NumofR <= 64 ? cyc : (int *)malloc(sizeof(int) * (1 + 2 * NumofR));
    int rc = pfast_rt_fsm2d(vp, U, TTd, TTu, Z, src, rec, srnum, srcid, NumofR, MAXSwp, dx, dz,
                            nrows, ncols, pe, cp, NULL);
    if (rc != PFAST_OK) {
        printf("PE_%d: %s...abort...\n", pe, pfast_last_error());
        exit(1);
    }
    for (int ir = 0; ir < NumofR; ir++) {
        if (cp[1 + 2 * ir] >= MAXSwp)
            printf("|\tPE_%d: ....Warning! downgoing FSM equation NOT Converged after %d sweeps "
                   "for reflector %d...\n", pe, MAXSwp, ir + 1);
        if (cp[2 + 2 * ir] >= MAXSwp)
            printf("|\tPE_%d: ....Warning! upgoing FSM Equation NOT Converged after %d sweeps for "
                   "reflector %d...\n", pe, MAXSwp, ir + 1);
    }
    if (cp[0] >= MAXSwp)
        printf("|\tPE_%d: ....Warning! Eikonal Equation Solver NOT Converged after %d sweeps...\n",
               pe, MAXSwp);
    if (cp != cyc) free(cp);
}

inline void usage()
{
    printf(" usage: mpirun -np <ranks> RT_PFAST2D input_file [L-BFGS memory=5] [RTtag=3]\n");
}

inline void info() { printf(" * 2-D traveltime tomography (FSM + adjoint state), GPU engine libpfast\n"); }

#endif
