/*
 * pfast_oracle.c -- FP64 CPU restatement of the PFAST hot path (test infrastructure).
 *
 * This file is the checker, never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it.  Build: see oracle/Makefile
 * (gcc -O2 -ffp-contract=off; plain IEEE double on x86-64 SSE2, no contraction).
 *
 * Each routine restates the reference routine cited beside it with identical
 * operation order, so that outputs are bit-identical to the reference's.
 */
#include "pfast_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define IDX3(i, j, k) ((long)(k) * nrows * ncols + (long)(i) * ncols + (j))
#define IDX2(i, j) ((long)(i) * ncols + (j))

static inline double dmin(double a, double b) { return a < b ? a : b; }
/* POS/NEG exactly as the reference macros ((x)+fabs(x))/2.0, ((x)-fabs(x))/2.0
 * (src/T_padjoint3d.h:8-14): keeps the sign-of-zero behaviour identical. */
static inline double posp(double x) { return (x + fabs(x)) / 2.0; }
static inline double negp(double x) { return (x - fabs(x)) / 2.0; }

/* Convergence test, src/T_pfsm3d.h:122-129 and src/RT_pfsm2d.h:110-114:
 * converged iff every |p-U| <= 1e-9. */
static int converged(const double *p, const double *U, long N)
{
    for (long n = 0; n < N; n++)
        if (fabs(p[n] - U[n]) > 1e-9) return 0;
    return 1;
}

/* ======================================================================= */
/* 3D local Godunov solve: src/T_pfsm3d.h:73-120                            */
/* ======================================================================= */
double orc_eikslv3d(const double *U, const double *vp, int i, int j, int k,
                    double dx, double dy, double dz, int nrows, int ncols, int nz)
{
    double s = 1.0 / vp[IDX3(i, j, k)];
    double ha = dx, hb = dy, hc = dz, ta, tb, tc, q, t;
    double a, b, c;
    /* x (column) neighbours, one-sided at the edges (:78-83) */
    if (j - 1 < 0) a = U[IDX3(i, j + 1, k)];
    else if (j + 1 >= ncols) a = U[IDX3(i, j - 1, k)];
    else a = dmin(U[IDX3(i, j + 1, k)], U[IDX3(i, j - 1, k)]);
    /* y (row) neighbours (:85-90) */
    if (i - 1 < 0) b = U[IDX3(i + 1, j, k)];
    else if (i + 1 >= nrows) b = U[IDX3(i - 1, j, k)];
    else b = dmin(U[IDX3(i - 1, j, k)], U[IDX3(i + 1, j, k)]);
    /* z (depth) neighbours (:92-97) */
    if (k - 1 < 0) c = U[IDX3(i, j, k + 1)];
    else if (k + 1 >= nz) c = U[IDX3(i, j, k - 1)];
    else c = dmin(U[IDX3(i, j, k - 1)], U[IDX3(i, j, k + 1)]);
    /* three compare-swaps giving c <= b <= a, spacings follow (:100-102) */
    if (b > a) { t = a; a = b; b = t; t = ha; ha = hb; hb = t; }
    if (c > a) { t = a; a = c; c = t; t = ha; ha = hc; hc = t; }
    if (c > b) { t = b; b = c; c = t; t = hb; hb = hc; hc = t; }
    if (!(c <= b && b <= a)) return 0; /* (:104) unreachable for non-NaN input */
    /* one-neighbour update (:106-107) */
    q = hc * s + c;
    if (q <= b) return q;
    /* two-neighbour quadratic (:109-113) */
    ta = hc * hc + hb * hb;
    tb = -2. * (b * hc * hc + c * hb * hb);
    tc = hc * hc * b * b + hb * hb * c * c - hb * hb * hc * hc * s * s;
    if ((tb * tb - 4. * ta * tc) >= 0) q = (-tb + sqrt(tb * tb - 4. * ta * tc)) / 2. / ta;
    else q = ORC_INF;
    if (q <= a) return q;
    /* three-neighbour quadratic (:115-119) */
    ta = 1.0 / ha / ha + 1.0 / hb / hb + 1.0 / hc / hc;
    tb = -2.0 * (a / ha / ha + b / hb / hb + c / hc / hc);
    tc = a * a / ha / ha + b * b / hb / hb + c * c / hc / hc - s * s;
    if ((tb * tb - 4. * ta * tc) >= 0) q = (-tb + sqrt(tb * tb - 4. * ta * tc)) / 2. / ta;
    else q = ORC_INF;
    return q;
}

/* ======================================================================= */
/* 3D fast sweeping: src/T_pfsm3d.h:131-185                                 */
/* ======================================================================= */
int orc_fsm3d(const double *vp, double *U, const int *src, int srcid, int MAXSwp,
              double dx, double dy, double dz, int nrows, int ncols, int nz)
{
    const long N = (long)nrows * ncols * nz;
    const int si = src[srcid * 3 + 0], sj = src[srcid * 3 + 1], sk = src[srcid * 3 + 2];
    double *p = (double *)malloc(sizeof(double) * N);
    int count = 0;
    /* point-source initialisation (:139-148) */
    for (long n = 0; n < N; n++) { U[n] = ORC_INF; p[n] = 0.0; }
    U[IDX3(si, sj, sk)] = 0.0;
    /* `while (!isconv && count++ < MAXSwp)` (:152) */
    for (;;) {
        if (converged(p, U, N)) break;
        if (!(count < MAXSwp)) { count++; break; }
        count++;
        memcpy(p, U, sizeof(double) * N);
        /* 8 orderings: s3 (depth) outermost, then s1 (rows), s2 (cols) (:158-165) */
        for (int s3 = -1; s3 <= 1; s3 += 2)
            for (int s1 = -1; s1 <= 1; s1 += 2)
                for (int s2 = -1; s2 <= 1; s2 += 2)
                    for (int kk = 0; kk < nz; kk++) {
                        int k = s3 < 0 ? nz - 1 - kk : kk;
                        for (int ii = 0; ii < nrows; ii++) {
                            int i = s1 < 0 ? nrows - 1 - ii : ii;
                            for (int jj = 0; jj < ncols; jj++) {
                                int j = s2 < 0 ? ncols - 1 - jj : jj;
                                if (i == si && j == sj && k == sk) continue; /* (:168) */
                                double q = orc_eikslv3d(U, vp, i, j, k, dx, dy, dz, nrows, ncols, nz);
                                U[IDX3(i, j, k)] = dmin(U[IDX3(i, j, k)], q); /* (:170) */
                            }
                        }
                    }
    }
    free(p);
    return count > MAXSwp ? MAXSwp : count;
}

/* ======================================================================= */
/* 3D adjoint local solve: src/T_padjoint3d.h:42-71                         */
/* ======================================================================= */
double orc_adjoint_slv3d(const double *U, const double *T, int i, int j, int k,
                         double dx, double dy, double dz, int nrows, int ncols, int nz)
{
    const double tc = T[IDX3(i, j, k)];
    const double txp = T[IDX3(i, j + 1, k)], txm = T[IDX3(i, j - 1, k)];
    const double typ = T[IDX3(i + 1, j, k)], tym = T[IDX3(i - 1, j, k)];
    const double tzp = T[IDX3(i, j, k + 1)], tzm = T[IDX3(i, j, k - 1)];
    if (tc >= ORC_INF || txp >= ORC_INF || txm >= ORC_INF || typ >= ORC_INF ||
        tym >= ORC_INF || tzp >= ORC_INF || tzm >= ORC_INF)
        return 0.0; /* (:45-49) */
    /* upwind coefficients (:51-56) */
    double a1 = posp(-(txp - tc) / dx), a2 = negp(-(tc - txm) / dx);
    double b1 = posp(-(typ - tc) / dy), b2 = negp(-(tc - tym) / dy);
    double c1 = posp(-(tzp - tc) / dz), c2 = negp(-(tc - tzm) / dz);
    /* neighbour terms (:58-63) */
    double La1 = posp(-(tc - txm) / dx) * U[IDX3(i, j - 1, k)];
    double La2 = negp(-(txp - tc) / dx) * U[IDX3(i, j + 1, k)];
    double Lb1 = posp(-(tc - tym) / dy) * U[IDX3(i - 1, j, k)];
    double Lb2 = negp(-(typ - tc) / dy) * U[IDX3(i + 1, j, k)];
    double Lc1 = posp(-(tc - tzm) / dz) * U[IDX3(i, j, k - 1)];
    double Lc2 = negp(-(tzp - tc) / dz) * U[IDX3(i, j, k + 1)];
    double den = (a1 - a2) / dx + (b1 - b2) / dy + (c1 - c2) / dz;
    if (fabs(den) > 1e-9) /* (:65-68) */
        return ((La1 - La2) / dx + (Lb1 - Lb2) / dy + (Lc1 - Lc2) / dz) / den;
    return 0.0;
}

/* ======================================================================= */
/* 3D adjoint state: src/T_padjoint3d.h:73-172                              */
/* ======================================================================= */
double orc_adjoint_fsm3d(double *U, const double *T, const double *T_obs, const int *rec,
                         const double *nxyz, const int *src, const int *srnum, int srcid,
                         int MAXSwp, double dx, double dy, double dz,
                         int nrows, int ncols, int nz, int *ncycles)
{
    const long N = (long)nrows * ncols * nz;
    double *p = (double *)malloc(sizeof(double) * N);
    unsigned char *tag = (unsigned char *)calloc(N, 1);
    const int si = src[srcid * 3 + 0], sj = src[srcid * 3 + 1], sk = src[srcid * 3 + 2];
    double EE = 0.0;
    long ch0 = 0;
    int count = 0;
    for (long n = 0; n < N; n++) { U[n] = 0.0; p[n] = 0.0; } /* (:90) */
    for (int s = 0; s < srcid; s++) ch0 += srnum[s];           /* (:91) */
    /* receiver boundary condition (:94-127) */
    for (int idx = 0; idx < srnum[srcid]; idx++) {
        const long r = idx + ch0;
        if (!(T_obs[r] < 100)) continue; /* only picked transmission times (:95) */
        const int i = rec[r * 3 + 0], j = rec[r * 3 + 1], k = rec[r * 3 + 2];
        const long n = IDX3(i, j, k);
        tag[n] = 1;
        EE += (T_obs[r] - T[n]) * (T_obs[r] - T[n]) / 2.0;
        if (i == si && j == sj && k == sk) continue; /* receiver at the source (:100) */
        const double nny = nxyz[r * 3 + 0], nnx = nxyz[r * 3 + 1], nnz = nxyz[r * 3 + 2];
        double Tx, Ty, Tz;
        if (k + 1 >= nz) Tz = (T[n] - T[IDX3(i, j, k - 1)]) / dz;
        else Tz = (T[IDX3(i, j, k + 1)] - T[n]) / dz;
        if (i + 1 >= nrows) Ty = (T[n] - T[IDX3(i - 1, j, k)]) / dy;
        else Ty = (T[IDX3(i + 1, j, k)] - T[n]) / dy;
        if (j + 1 >= ncols) Tx = (T[n] - T[IDX3(i, j - 1, k)]) / dx;
        else Tx = (T[IDX3(i, j + 1, k)] - T[n]) / dx;
        if (fabs(T_obs[r] - T[n]) > 1e-9 && fabs(nnx * Tx + nny * Ty + nnz * Tz) > 1e-6)
            U[n] = (T_obs[r] - T[n]) / (nnx * Tx + nny * Ty + nnz * Tz);
        else
            U[n] = 0.0;
    }
    /* `while (count++ < MAXSwp && !isconv)` (:135): interior nodes only (:143-145) */
    for (;;) {
        if (!(count < MAXSwp)) { count++; break; }
        count++;
        if (converged(p, U, N)) break;
        memcpy(p, U, sizeof(double) * N);
        for (int s3 = -1; s3 <= 1; s3 += 2)
            for (int s1 = -1; s1 <= 1; s1 += 2)
                for (int s2 = -1; s2 <= 1; s2 += 2)
                    for (int kk = 1; kk < nz - 1; kk++) {
                        int k = s3 < 0 ? nz - 1 - kk : kk;
                        for (int ii = 1; ii < nrows - 1; ii++) {
                            int i = s1 < 0 ? nrows - 1 - ii : ii;
                            for (int jj = 1; jj < ncols - 1; jj++) {
                                int j = s2 < 0 ? ncols - 1 - jj : jj;
                                if (tag[IDX3(i, j, k)]) continue; /* (:149) */
                                U[IDX3(i, j, k)] = orc_adjoint_slv3d(U, T, i, j, k, dx, dy, dz,
                                                                     nrows, ncols, nz);
                            }
                        }
                    }
    }
    /* cycles actually swept: count-1 when stopped by convergence, else MAXSwp */
    if (ncycles) *ncycles = count > MAXSwp ? MAXSwp : count - 1;
    free(p);
    free(tag);
    return EE;
}

/* ======================================================================= */
/* 2D local Godunov solve: src/RT_pfsm2d.h:71-108                           */
/* ======================================================================= */
double orc_eikslv2d(const double *U, const double *vp, int i, int j,
                    double dx, double dz, int nrows, int ncols)
{
    double s = 1.0 / vp[IDX2(i, j)];
    double ha = dx, hb = dz, a, b, q, ta, tb, tc, t;
    /* b: row (depth) neighbours (:75-80); a: column neighbours (:82-87) */
    if (i - 1 < 0) b = U[IDX2(i + 1, j)];
    else if (i + 1 >= nrows) b = U[IDX2(i - 1, j)];
    else b = dmin(U[IDX2(i - 1, j)], U[IDX2(i + 1, j)]);
    if (j - 1 < 0) a = U[IDX2(i, j + 1)];
    else if (j + 1 >= ncols) a = U[IDX2(i, j - 1)];
    else a = dmin(U[IDX2(i, j + 1)], U[IDX2(i, j - 1)]);
    if (b > a) { t = a; a = b; b = t; t = ha; ha = hb; hb = t; } /* (:89) */
    q = b + hb * s;                                            /* (:93-94) */
    if (q <= a) return q;
    ta = 1.0 / ha / ha + 1.0 / hb / hb;                        /* (:95) */
    tb = -2. * (a / ha / ha + b / hb / hb);
    tc = a * a / ha / ha + b * b / hb / hb - s * s;
    if ((tb * tb - 4. * ta * tc) >= 0) q = (-tb + sqrt(tb * tb - 4 * ta * tc)) / 2. / ta;
    else q = ORC_INF;
    return q;
}

/* One 2D FSM solve with an update mask, shared by the three solves of RT_fsm2d
 * (src/RT_pfsm2d.h:160-182, :297-317, :343-365).  mode 0: skip the source node;
 * mode 1: skip the source node and nodes below the reflector depth Zd[j] (:174);
 * mode 2: update only i <= k2vec[j] (:309).  Returns cycles swept. */
static int fsm2d_masked(const double *vp, double *U, double *p, int si, int sj, int mode,
                        const double *Z, int ir, const int *k2vec, int MAXSwp,
                        double dx, double dz, int nrows, int ncols)
{
    const long N = (long)nrows * ncols;
    int count = 0;
    for (;;) {
        if (converged(p, U, N)) break;
        if (!(count < MAXSwp)) { count++; break; }
        count++;
        memcpy(p, U, sizeof(double) * N);
        for (int s1 = -1; s1 <= 1; s1 += 2)
            for (int s2 = -1; s2 <= 1; s2 += 2)
                for (int ii = 0; ii < nrows; ii++) {
                    int i = s1 < 0 ? nrows - 1 - ii : ii;
                    for (int jj = 0; jj < ncols; jj++) {
                        int j = s2 < 0 ? ncols - 1 - jj : jj;
                        int upd;
                        if (mode == 0) upd = !(i == si && j == sj);
                        else if (mode == 1)
                            upd = !(i == si && j == sj) && (i + 1) * dz <= Z[((long)ir * ncols + j) * 3 + 0];
                        else upd = i <= k2vec[j];
                        if (!upd) continue;
                        double q = orc_eikslv2d(U, vp, i, j, dx, dz, nrows, ncols);
                        U[IDX2(i, j)] = dmin(U[IDX2(i, j)], q);
                    }
                }
    }
    return count > MAXSwp ? MAXSwp : count;
}

/* Reflected-wave seed at one column from k2 (src/RT_pfsm2d.h:220-223, :248-251,
 * :269-272): given qx chosen by the caller. */
static double reflect_seed(const double *U, const double *Z, int ir, int jj, int k2, double qx,
                           double dz, int ncols)
{
    const long zr = ((long)ir * ncols + jj) * 3;
    double qz = (U[IDX2(k2 + 1, jj)] - U[IDX2(k2, jj)]) / dz;
    double dnx = Z[zr + 2], dnz = Z[zr + 1];
    double rz = qz - 2 * (qx * dnx + qz * dnz) * dnz;
    return U[IDX2(k2, jj)] + qz * (Z[zr] - k2 * dz - dz) + rz * ((k2 + 1) * dz + dz - Z[zr]);
}

/* ======================================================================= */
/* 2D transmission + reflection FSM: src/RT_pfsm2d.h:116-380                */
/* ======================================================================= */
int orc_rt_fsm2d(const double *vp, double *U, double *TTd, double *TTu, const double *Z,
                 const int *src, int srcid, int NumofR, int MAXSwp,
                 double dx, double dz, int nrows, int ncols, int *ncycles, int *k2vec_out)
{
    const long N = (long)nrows * ncols;
    const int si = src[srcid * 2 + 0], sj = src[srcid * 2 + 1];
    double *p = (double *)malloc(sizeof(double) * N);
    int *k2vec = (int *)malloc(sizeof(int) * ncols);
    int status = 0;
    for (int j = 0; j < ncols; j++) k2vec[j] = -1; /* (:144) */
    for (int ir = 0; ir < NumofR && status == 0; ir++) {
        const double *Zr = Z + (long)ir * ncols * 3;
        /* down-going solve above the reflector (:151-185) */
        for (long n = 0; n < N; n++) { U[n] = ORC_INF; p[n] = 0.0; }
        U[IDX2(si, sj)] = 0.0;
        int c = fsm2d_masked(vp, U, p, si, sj, 1, Z, ir, k2vec, MAXSwp, dx, dz, nrows, ncols);
        if (ncycles) ncycles[1 + 2 * ir] = c;
        memcpy(TTd + (long)ir * N, U, sizeof(double) * N);
        /* interior columns, in increasing column order (:195-234) */
        for (int jj = 1; jj < ncols - 1; jj++) {
            int k2 = -1;
            for (int kk = 0; kk < nrows - 1; kk++)
                if ((kk + 1) * dz + dz <= Zr[jj * 3] &&
                    (kk * dz + dz <= Zr[(jj - 1) * 3] || kk * dz + dz <= Zr[(jj + 1) * 3]))
                    if (kk >= k2) k2 = kk;
            if (Zr[jj * 3] <= nrows * dz && k2 > 0) {
                double qx;
                if (k2 * dz + dz <= Zr[(jj - 1) * 3] && k2 * dz + dz <= Zr[(jj + 1) * 3])
                    qx = (U[IDX2(k2, jj + 1)] - U[IDX2(k2, jj - 1)]) / 2.0 / dx;
                else if (k2 * dz + dz <= Zr[(jj - 1) * 3])
                    qx = (U[IDX2(k2, jj)] - U[IDX2(k2, jj - 1)]) / dx;
                else
                    qx = (U[IDX2(k2, jj + 1)] - U[IDX2(k2, jj)]) / dx;
                U[IDX2(k2 + 1, jj)] = reflect_seed(U, Z, ir, jj, k2, qx, dz, ncols);
                k2vec[jj] = k2;
                if (U[IDX2(k2 + 1, jj)] < 0) { status = 2; break; }
            } else if (Zr[jj * 3] > nrows * dz) {
                k2vec[jj] = nrows - 1;
            } else { status = 1; break; }
        }
        if (status) break;
        /* left edge column (:238-257) */
        {
            int jj = 0, k2 = -1;
            for (int kk = 0; kk < nrows - 1; kk++)
                if ((kk + 1) * dz + dz <= Zr[jj * 3] && kk * dz + dz <= Zr[(jj + 1) * 3])
                    if (kk >= k2) k2 = kk;
            if (Zr[jj * 3] <= nrows * dz && k2 > 0) {
                double qx = (U[IDX2(k2, jj + 1)] - U[IDX2(k2, jj)]) / dx;
                U[IDX2(k2 + 1, jj)] = reflect_seed(U, Z, ir, jj, k2, qx, dz, ncols);
                k2vec[jj] = k2;
                if (U[IDX2(k2 + 1, jj)] < 0) { status = 2; break; }
            } else if (Zr[jj * 3] > nrows * dz) k2vec[jj] = nrows - 1;
            else { status = 1; break; }
        }
        /* right edge column (:259-281) */
        {
            int jj = ncols - 1, k2 = -1;
            for (int kk = 0; kk < nrows - 1; kk++)
                if ((kk + 1) * dz + dz <= Zr[jj * 3] && kk * dz + dz <= Zr[(jj - 1) * 3])
                    if (kk >= k2) k2 = kk;
            if (Zr[jj * 3] <= nrows * dz && k2 > 0) {
                double qx = (U[IDX2(k2, jj)] - U[IDX2(k2, jj - 1)]) / dx;
                U[IDX2(k2 + 1, jj)] = reflect_seed(U, Z, ir, jj, k2, qx, dz, ncols);
                k2vec[jj] = k2;
                if (U[IDX2(k2 + 1, jj)] < 0) { status = 2; break; }
            } else if (Zr[jj * 3] > nrows * dz) k2vec[jj] = nrows - 1;
            else { status = 1; break; }
        }
        /* reset the region above the reflection points (:289-295) */
        for (int i = 0; i < nrows; i++)
            for (int j = 0; j < ncols; j++) {
                if (i <= k2vec[j]) U[IDX2(i, j)] = ORC_INF;
                p[IDX2(i, j)] = 0.0;
            }
        /* up-going solve (:299-320) */
        c = fsm2d_masked(vp, U, p, si, sj, 2, Z, ir, k2vec, MAXSwp, dx, dz, nrows, ncols);
        if (ncycles) ncycles[2 + 2 * ir] = c;
        memcpy(TTu + (long)ir * N, U, sizeof(double) * N);
    }
    if (status == 0) {
        /* transmission first arrivals (:334-365) */
        for (long n = 0; n < N; n++) { U[n] = ORC_INF; p[n] = 0.0; }
        U[IDX2(si, sj)] = 0.0;
        int c = fsm2d_masked(vp, U, p, si, sj, 0, Z, 0, k2vec, MAXSwp, dx, dz, nrows, ncols);
        if (ncycles) ncycles[0] = c;
    }
    if (k2vec_out) memcpy(k2vec_out, k2vec, sizeof(int) * ncols);
    free(p);
    free(k2vec);
    return status;
}

/* ======================================================================= */
/* 2D adjoint local solve: src/RT_padjoint2d.h:42-65                        */
/* ======================================================================= */
double orc_adjoint_slv2d(const double *U, const double *T, int i, int j,
                         double dx, double dz, int nrows, int ncols)
{
    const double tc = T[IDX2(i, j)];
    const double txp = T[IDX2(i, j + 1)], txm = T[IDX2(i, j - 1)];
    const double tzp = T[IDX2(i + 1, j)], tzm = T[IDX2(i - 1, j)];
    if (txp >= ORC_INF || tc >= ORC_INF || txm >= ORC_INF || tzp >= ORC_INF || tzm >= ORC_INF)
        return 0.0;
    double a1 = posp(-(txp - tc) / dx), a2 = negp(-(tc - txm) / dx);
    double b1 = posp(-(tzp - tc) / dz), b2 = negp(-(tc - tzm) / dz);
    double La1 = posp(-(tc - txm) / dx) * U[IDX2(i, j - 1)];
    double La2 = negp(-(txp - tc) / dx) * U[IDX2(i, j + 1)];
    double Lb1 = posp(-(tc - tzm) / dz) * U[IDX2(i - 1, j)];
    double Lb2 = negp(-(tzp - tc) / dz) * U[IDX2(i + 1, j)];
    double den = (a1 - a2) / dx + (b1 - b2) / dz;
    if (fabs(den) > 1e-9) return ((La1 - La2) / dx + (Lb1 - Lb2) / dz) / den;
    return 0.0;
}

/* Interior 2D adjoint sweeps shared by Tadjoint/Radjoint (src/RT_padjoint2d.h:
 * 124-148 and :214-238); `count` is carried by the caller (Radjoint does not reset
 * it between reflectors, :160-161, :214). */
static int adjoint_sweeps2d(double *U, double *p, const double *T, const unsigned char *tag,
                            int *count, int MAXSwp, double dx, double dz, int nrows, int ncols)
{
    const long N = (long)nrows * ncols;
    int swept = 0;
    for (;;) {
        if (!(*count < MAXSwp)) { (*count)++; break; }
        (*count)++;
        if (converged(p, U, N)) break;
        swept++;
        memcpy(p, U, sizeof(double) * N);
        for (int s1 = -1; s1 <= 1; s1 += 2)
            for (int s2 = -1; s2 <= 1; s2 += 2)
                for (int ii = 1; ii < nrows - 1; ii++) {
                    int i = s1 < 0 ? nrows - 1 - ii : ii;
                    for (int jj = 1; jj < ncols - 1; jj++) {
                        int j = s2 < 0 ? ncols - 1 - jj : jj;
                        if (tag[IDX2(i, j)]) continue;
                        U[IDX2(i, j)] = orc_adjoint_slv2d(U, T, i, j, dx, dz, nrows, ncols);
                    }
                }
    }
    return swept;
}

/* ======================================================================= */
/* 2D transmission adjoint: src/RT_padjoint2d.h:288-380 (Cr = 1) and the    */
/* WR overload :67-157 (Cr = 1-WR)                                          */
/* ======================================================================= */
double orc_tadjoint_fsm2d(double *U, const double *T, const double *T_obs, double Cr,
                          const int *rec, const double *nxyz, const int *src,
                          const int *srnum, int srcid, int MAXSwp, double dx, double dz,
                          int nrows, int ncols, int *ncycles)
{
    const long N = (long)nrows * ncols;
    double *p = (double *)malloc(sizeof(double) * N);
    unsigned char *tag = (unsigned char *)calloc(N, 1);
    const int si = src[srcid * 2 + 0], sj = src[srcid * 2 + 1];
    double EE = 0.0;
    long ch0 = 0;
    int count = 0;
    for (long n = 0; n < N; n++) { U[n] = 0.0; p[n] = 0.0; }
    for (int s = 0; s < srcid; s++) ch0 += srnum[s];
    for (int idx = 0; idx < srnum[srcid]; idx++) {
        const long r = idx + ch0;
        if (!(T_obs[r] < 100)) continue;
        const int i = rec[r * 2 + 0], j = rec[r * 2 + 1];
        const long n = IDX2(i, j);
        tag[n] = 1;
        EE += Cr * (T_obs[r] - T[n]) * (T_obs[r] - T[n]) / 2.0;
        if (i == si && j == sj) continue;
        const double nny = nxyz[r * 2 + 0], nnx = nxyz[r * 2 + 1];
        double Tx, Tz;
        if (i + 1 >= nrows) Tz = (T[n] - T[IDX2(i - 1, j)]) / dz;
        else Tz = (T[IDX2(i + 1, j)] - T[n]) / dz;
        if (j + 1 >= ncols) Tx = (T[n] - T[IDX2(i, j - 1)]) / dx;
        else Tx = (T[IDX2(i, j + 1)] - T[n]) / dx;
        if (fabs(T_obs[r] - T[n]) > 1e-9 && fabs(nnx * Tx + nny * Tz) > 1e-6)
            U[n] = Cr * (T_obs[r] - T[n]) / (nnx * Tx + nny * Tz);
        else
            U[n] = 0.0;
    }
    int swept = adjoint_sweeps2d(U, p, T, tag, &count, MAXSwp, dx, dz, nrows, ncols);
    if (ncycles) *ncycles = swept;
    free(p);
    free(tag);
    return EE;
}

/* ======================================================================= */
/* 2D reflection adjoint: src/RT_padjoint2d.h:382-473 (Cr = 1) and the WR   */
/* overload :159-285 (Cr = WR).  Tu holds NumofR up-going fields.           */
/* ======================================================================= */
double orc_radjoint_fsm2d(double *UR, const double *Tu, const double *TR_obs, double Cr,
                          const int *rec, const double *nxyz, const int *srnum, int srcid,
                          int NumofR, int MAXSwp, double dx, double dz,
                          int nrows, int ncols, int *ncycles)
{
    const long N = (long)nrows * ncols;
    double *p = (double *)malloc(sizeof(double) * N);
    double *U = (double *)malloc(sizeof(double) * N);
    unsigned char *tag = (unsigned char *)calloc(N, 1); /* not reset per reflector (:398) */
    double EE = 0.0;
    long ch0 = 0;
    int count = 0;
    for (long n = 0; n < N; n++) UR[n] = 0;
    for (int s = 0; s < srcid; s++) ch0 += srnum[s];
    for (int ir = 0; ir < NumofR; ir++) {
        const double *T = Tu + (long)ir * N;
        for (long n = 0; n < N; n++) { U[n] = 0.0; p[n] = 0.0; }
        for (int idx = 0; idx < srnum[srcid]; idx++) {
            const long r = idx + ch0;
            const double obs = TR_obs[r * NumofR + ir];
            if (!(obs != 0)) continue; /* picked reflection (:406) */
            const int i = rec[r * 2 + 0], j = rec[r * 2 + 1];
            const long n = IDX2(i, j);
            tag[n] = 1;
            EE += Cr * (obs - T[n]) * (obs - T[n]) / 2.0;
            const double nny = nxyz[r * 2 + 0], nnx = nxyz[r * 2 + 1];
            double Tx, Tz;
            if (i + 1 >= nrows) Tz = (T[n] - T[IDX2(i - 1, j)]) / dz;
            else Tz = (T[IDX2(i + 1, j)] - T[n]) / dz;
            if (j + 1 >= ncols) Tx = (T[n] - T[IDX2(i, j - 1)]) / dx;
            else Tx = (T[IDX2(i, j + 1)] - T[n]) / dx;
            if (fabs(obs - T[n]) > 1e-9 && fabs(nnx * Tx + nny * Tz) > 1e-6)
                U[n] = Cr * (obs - T[n]) / (nnx * Tx + nny * Tz);
            else
                U[n] = 0.0;
        }
        int swept = adjoint_sweeps2d(U, p, T, tag, &count, MAXSwp, dx, dz, nrows, ncols);
        if (ncycles) ncycles[ir] = swept;
        for (long n = 0; n < N; n++) UR[n] = UR[n] + U[n];
    }
    free(p);
    free(U);
    free(tag);
    return EE;
}

/* ======================================================================= */
/* Geometry (src/T_PFAST3D.cpp:250-252, src/RT_PFAST2D.cpp:300-301) and     */
/* gradient accumulation (src/T_PFAST3D.cpp:393, src/RT_PFAST2D.cpp:425)     */
/* ======================================================================= */
void orc_map_index(const double *table, int n, int stride, int col, double h,
                   int *out, int out_stride)
{
    for (int r = 0; r < n; r++) out[(long)r * out_stride] = (int)ceil(table[(long)r * stride + col] / h) - 1;
}

void orc_grad_accum(double *g, const double *lam, const double *vp, long N)
{
    for (long n = 0; n < N; n++) g[n] = g[n] - lam[n] / (vp[n] * vp[n] * vp[n]);
}

/* ======================================================================= */
/* CPU baseline: the reference's rank-per-contiguous-source-block layout    */
/* (src/T_PFAST3D.cpp:221-238, :387-399) with threads in place of ranks.    */
/* ======================================================================= */
typedef struct {
    const double *vp; const int *src, *srnum, *rec; const double *nxyz, *T_obs;
    int b0, b1, MAXSwp; double dx, dy, dz; int nrows, ncols, nz;
    double *grad, misfit; int *fcyc, *acyc;
} orc_rank;

static void *orc_rank_main(void *arg)
{
    orc_rank *R = (orc_rank *)arg;
    const long N = (long)R->nrows * R->ncols * R->nz;
    double *T = (double *)malloc(sizeof(double) * N);
    double *lam = (double *)malloc(sizeof(double) * N);
    R->misfit = 0.0;
    for (int s = R->b0; s < R->b1; s++) {
        int fc = orc_fsm3d(R->vp, T, R->src, s, R->MAXSwp, R->dx, R->dy, R->dz,
                           R->nrows, R->ncols, R->nz);
        int ac = 0;
        R->misfit += orc_adjoint_fsm3d(lam, T, R->T_obs, R->rec, R->nxyz, R->src, R->srnum, s,
                                       R->MAXSwp, R->dx, R->dy, R->dz, R->nrows, R->ncols,
                                       R->nz, &ac);
        orc_grad_accum(R->grad, lam, R->vp, N);
        if (R->fcyc) R->fcyc[s] = fc;
        if (R->acyc) R->acyc[s] = ac;
    }
    free(T);
    free(lam);
    return NULL;
}

int orc_fwdadj3d_parallel(int P, const double *vp, const int *src, const int *srnum,
                          const int *rec, const double *nxyz, const double *T_obs,
                          int s0, int s1, int MAXSwp, double dx, double dy, double dz,
                          int nrows, int ncols, int nz, double *grad, double *misfit,
                          int *fcycles, int *acycles)
{
    const long N = (long)nrows * ncols * nz;
    const int ns = s1 - s0;
    if (P < 1 || ns < 0) return -1;
    if (P > ns && ns > 0) P = ns;
    orc_rank *R = (orc_rank *)calloc(P, sizeof(orc_rank));
    pthread_t *th = (pthread_t *)calloc(P, sizeof(pthread_t));
    /* contiguous blocks, remainder to the lowest ranks (src/T_PFAST3D.cpp:221-236) */
    int rmd = ns % P, per = ns / P, start = s0;
    for (int r = 0; r < P; r++) {
        int cnt = per + (r < rmd ? 1 : 0);
        R[r].vp = vp; R[r].src = src; R[r].srnum = srnum; R[r].rec = rec;
        R[r].nxyz = nxyz; R[r].T_obs = T_obs; R[r].b0 = start; R[r].b1 = start + cnt;
        R[r].MAXSwp = MAXSwp; R[r].dx = dx; R[r].dy = dy; R[r].dz = dz;
        R[r].nrows = nrows; R[r].ncols = ncols; R[r].nz = nz;
        R[r].grad = (double *)calloc(N, sizeof(double));
        R[r].fcyc = fcycles; R[r].acyc = acycles;
        start += cnt;
        pthread_create(&th[r], NULL, orc_rank_main, &R[r]);
    }
    double m = 0.0;
    for (long n = 0; n < N; n++) grad[n] = 0.0;
    for (int r = 0; r < P; r++) {
        pthread_join(th[r], NULL);
        for (long n = 0; n < N; n++) grad[n] += R[r].grad[n]; /* MPI_SUM in rank order */
        m += R[r].misfit;
        free(R[r].grad);
    }
    *misfit = m;
    free(R);
    free(th);
    return 0;
}
