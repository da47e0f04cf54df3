// Drop-in replacement of the reference's src/T_pfsm3d.h (3D fast sweeping).
//
// Same function names and signatures, so a driver that includes "T_pfsm3d.h"
// compiles unchanged with -I<repo>/include/dropin and links -lpfast.  fsm3d runs on
// the GPU through libpfast (pfast_fsm3d, FP64, bit-identical to the reference's
// CPU result); the small helpers are host code with the reference's semantics.
// Like the reference headers, the functions are defined here (inline).
#ifndef PFAST_DROPIN_T_PFSM3D_H
#define PFAST_DROPIN_T_PFSM3D_H

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
#define DIM 3
#endif
#ifndef PI
#define PI (3.141592653589793)
#endif

using namespace std;

// readmod/writemod: raw little-endian FP64 arrays, src/T_pfsm3d.h:39-57.  A file
// that cannot be opened is reported and ends the program (the reference goes on
// to fread from a NULL stream).
inline void readmod(double *mod, char *file, int nrows, int ncols, int nz)
{
    FILE *fp = fopen(file, "rb");
    if (!fp) {
        printf(" The file %s cannot be opened\n", file);
        exit(1);
    }
    printf(" \n...reading file:\n\t %s\n", file);
    size_t n = (size_t)nrows * ncols * nz;
    if (fread(mod, sizeof(double), n, fp) != n) printf(" short read from %s\n", file);
    fclose(fp);
}

inline void writemod(double *mod, char *file, int nrows, int ncols, int nz)
{
    FILE *fp = fopen(file, "wb");
    if (!fp) {
        printf(" The file %s cannot be opened\n", file);
        return;
    }
    printf(" \n...writing file:\n\t %s\n", file);
    fwrite(mod, sizeof(double), (size_t)nrows * ncols * nz, fp);
    fclose(fp);
}

// Local Godunov update at one node (host helper), src/T_pfsm3d.h:73-120.
inline double eikslv3d(double *U, double *vp, int *node, double dx, double dy, double dz,
                       int nrows, int ncols, int nz)
{
    const long plane = (long)nrows * ncols;
    const int i = node[0], j = node[1], k = node[2];
    const long n = k * plane + (long)i * ncols + j;
    double s = 1.0 / vp[n], ha = dx, hb = dy, hc = dz, a, b, c, t, q, ta, tb, tc, disc;
    a = j - 1 < 0 ? U[n + 1] : (j + 1 >= ncols ? U[n - 1] : MIN(U[n + 1], U[n - 1]));
    b = i - 1 < 0 ? U[n + ncols] : (i + 1 >= nrows ? U[n - ncols] : MIN(U[n - ncols], U[n + ncols]));
    c = k - 1 < 0 ? U[n + plane] : (k + 1 >= nz ? U[n - plane] : MIN(U[n - plane], U[n + plane]));
    if (b > a) { t = a; a = b; b = t; t = ha; ha = hb; hb = t; }
    if (c > a) { t = a; a = c; c = t; t = ha; ha = hc; hc = t; }
    if (c > b) { t = b; b = c; c = t; t = hb; hb = hc; hc = t; }
    q = hc * s + c;
    if (q <= b) return q;
    ta = hc * hc + hb * hb;
    tb = -2. * (b * hc * hc + c * hb * hb);
    tc = hc * hc * b * b + hb * hb * c * c - hb * hb * hc * hc * s * s;
    disc = tb * tb - 4. * ta * tc;
    q = disc >= 0 ? (-tb + sqrt(disc)) / 2. / ta : INF;
    if (q <= a) return q;
    ta = 1.0 / ha / ha + 1.0 / hb / hb + 1.0 / hc / hc;
    tb = -2.0 * (a / ha / ha + b / hb / hb + c / hc / hc);
    tc = a * a / ha / ha + b * b / hb / hb + c * c / hc / hc - s * s;
    disc = tb * tb - 4. * ta * tc;
    return disc >= 0 ? (-tb + sqrt(disc)) / 2. / ta : INF;
}

// max|p-U| <= 1e-9, src/T_pfsm3d.h:122-129.
inline bool isconv(double *p, double *U, int nrows, int ncols, int nz)
{
    const long N = (long)nrows * ncols * nz;
    for (long n = 0; n < N; n++)
        if (fabs(p[n] - U[n]) > 1e-9) return false;
    return true;
}

// src/T_pfsm3d.h:131-185, on the GPU.
inline void fsm3d(double *vp, double *U, int *src, int srcid, int MAXSwp, double dx, double dy,
                  double dz, int nrows, int ncols, int nz, int pe)
{
    int cycles = 0;
    clock_t tic = clock();
    int rc = pfast_fsm3d(vp, U, src, srcid, MAXSwp, dx, dy, dz, nrows, ncols, nz, pe, &cycles);
    if (rc != PFAST_OK) {
        printf("PE_%d: fsm3d failed: %s\n", pe, pfast_last_error());
        exit(1);
    }
    if (cycles >= MAXSwp)
        printf("|\tPE_%d: ....Warning! Eikonal Equation NOT Converged after %d sweeps, %8.5f "
               "seconds...\n",
               pe, MAXSwp, (double)(clock() - tic) / CLOCKS_PER_SEC);
}

inline void usage()
{
    printf(" usage: mpirun -np <ranks> T_PFAST3D input_file [L-BFGS memory, default 5]\n");
}

inline void info()
{
    printf(" * 3-D traveltime tomography (FSM + adjoint state), GPU engine libpfast\n");
}

#endif
