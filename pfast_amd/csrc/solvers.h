// solvers.h -- per-node local solves shared by the kernels.
//
// Two families:
//  * *_ref: the reference's FP64 arithmetic, operation for operation, so that the
//    single-source FP64 path is bit-identical to the reference (the library is
//    built with -ffp-contract=off; division and sqrt are IEEE correctly rounded).
//  * *_f32: FP32 forms of the same Godunov / adjoint updates written in
//    difference form (values are offsets from the largest neighbour), which
//    avoids the catastrophic cancellation of the reference's quadratic in FP32.
#pragma once
#include <hip/hip_runtime.h>

#define PFAST_INF_F 9999999.0f
#define PFAST_INF_D 9999999.0

// ------------------------------------------------------------------ FP64 (reference)
// Godunov update at one node, src/T_pfsm3d.h:73-120. a/b/c are the already-formed
// x/y/z neighbour minima, s = 1/vp.
__device__ __forceinline__ double eik3d_ref(double a, double b, double c, double s, double dx,
                                            double dy, double dz)
{
    double ha = dx, hb = dy, hc = dz, t, q, ta, tb, tc;
    if (b > a) { t = a; a = b; b = t; t = ha; ha = hb; hb = t; }
    if (c > a) { t = a; a = c; c = t; t = ha; ha = hc; hc = t; }
    if (c > b) { t = b; b = c; c = t; t = hb; hb = hc; hc = t; }
    if (!(c <= b && b <= a)) return 0;
    q = hc * s + c;
    if (q <= b) return q;
    ta = hc * hc + hb * hb;
    tb = -2. * (b * hc * hc + c * hb * hb);
    tc = hc * hc * b * b + hb * hb * c * c - hb * hb * hc * hc * s * s;
    double disc = tb * tb - 4. * ta * tc;
    if (disc >= 0) q = (-tb + sqrt(disc)) / 2. / ta;
    else q = PFAST_INF_D;
    if (q <= a) return q;
    ta = 1.0 / ha / ha + 1.0 / hb / hb + 1.0 / hc / hc;
    tb = -2.0 * (a / ha / ha + b / hb / hb + c / hc / hc);
    tc = a * a / ha / ha + b * b / hb / hb + c * c / hc / hc - s * s;
    disc = tb * tb - 4. * ta * tc;
    if (disc >= 0) q = (-tb + sqrt(disc)) / 2. / ta;
    else q = PFAST_INF_D;
    return q;
}

// 2D Godunov update, src/RT_pfsm2d.h:71-108. b = row (depth) neighbour min,
// a = column neighbour min.
__device__ __forceinline__ double eik2d_ref(double a, double b, double s, double dx, double dz)
{
    double ha = dx, hb = dz, t, q, ta, tb, tc;
    if (b > a) { t = a; a = b; b = t; t = ha; ha = hb; hb = t; }
    q = b + hb * s;
    if (q <= a) return q;
    ta = 1.0 / ha / ha + 1.0 / hb / hb;
    tb = -2. * (a / ha / ha + b / hb / hb);
    tc = a * a / ha / ha + b * b / hb / hb - s * s;
    double disc = tb * tb - 4. * ta * tc;
    if (disc >= 0) q = (-tb + sqrt(disc)) / 2. / ta;
    else q = PFAST_INF_D;
    return q;
}

__device__ __forceinline__ double posp(double x) { return (x + fabs(x)) / 2.0; }
__device__ __forceinline__ double negp(double x) { return (x - fabs(x)) / 2.0; }

// Adjoint local solve, src/T_padjoint3d.h:42-71 (3D) with the 7 T values and 6
// lambda neighbours already gathered.
__device__ __forceinline__ double adj3d_ref(double tc, double txp, double txm, double typ,
                                            double tym, double tzp, double tzm, double lxp,
                                            double lxm, double lyp, double lym, double lzp,
                                            double lzm, double dx, double dy, double dz)
{
    if (tc >= PFAST_INF_D || txp >= PFAST_INF_D || txm >= PFAST_INF_D || typ >= PFAST_INF_D ||
        tym >= PFAST_INF_D || tzp >= PFAST_INF_D || tzm >= PFAST_INF_D)
        return 0.0;
    double a1 = posp(-(txp - tc) / dx), a2 = negp(-(tc - txm) / dx);
    double b1 = posp(-(typ - tc) / dy), b2 = negp(-(tc - tym) / dy);
    double c1 = posp(-(tzp - tc) / dz), c2 = negp(-(tc - tzm) / dz);
    double La1 = posp(-(tc - txm) / dx) * lxm;
    double La2 = negp(-(txp - tc) / dx) * lxp;
    double Lb1 = posp(-(tc - tym) / dy) * lym;
    double Lb2 = negp(-(typ - tc) / dy) * lyp;
    double Lc1 = posp(-(tc - tzm) / dz) * lzm;
    double Lc2 = negp(-(tzp - tc) / dz) * lzp;
    double den = (a1 - a2) / dx + (b1 - b2) / dy + (c1 - c2) / dz;
    if (fabs(den) > 1e-9) return ((La1 - La2) / dx + (Lb1 - Lb2) / dy + (Lc1 - Lc2) / dz) / den;
    return 0.0;
}

// 2D adjoint local solve, src/RT_padjoint2d.h:42-65 (x = columns, z = rows).
__device__ __forceinline__ double adj2d_ref(double tc, double txp, double txm, double tzp,
                                            double tzm, double lxp, double lxm, double lzp,
                                            double lzm, double dx, double dz)
{
    if (txp >= PFAST_INF_D || tc >= PFAST_INF_D || txm >= PFAST_INF_D || tzp >= PFAST_INF_D ||
        tzm >= PFAST_INF_D)
        return 0.0;
    double a1 = posp(-(txp - tc) / dx), a2 = negp(-(tc - txm) / dx);
    double b1 = posp(-(tzp - tc) / dz), b2 = negp(-(tc - tzm) / dz);
    double La1 = posp(-(tc - txm) / dx) * lxm;
    double La2 = negp(-(txp - tc) / dx) * lxp;
    double Lb1 = posp(-(tc - tzm) / dz) * lzm;
    double Lb2 = negp(-(tzp - tc) / dz) * lzp;
    double den = (a1 - a2) / dx + (b1 - b2) / dz;
    if (fabs(den) > 1e-9) return ((La1 - La2) / dx + (Lb1 - Lb2) / dz) / den;
    return 0.0;
}

// ------------------------------------------------------------------ FP32 forms
// Hardware v_sqrt_f32 / v_rcp_f32 (about 1 ulp): the FP32 sweeps are checked
// against the FP64 oracle at 1e-5, far above this rounding.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Uniform spacing h: f = h/v (travel time across one cell).  Neighbour minima
// a, b, c in any order.  Same case structure as src/T_pfsm3d.h:106-119.
__device__ __forceinline__ float eik3d_f32_uniform(float a, float b, float c, float f)
{
    const float lo = fminf(fminf(a, b), c);
    const float hi = fmaxf(fmaxf(a, b), c);
    const float mid = __builtin_amdgcn_fmed3f(a, b, c);
    const float q1 = lo + f;
    const float d = mid - lo;
    const float ff = f * f;
    // ff + ff is 2 f f rounded once, as (2 f) f would be (scaling by 2 is exact)
    const float q2 = fmaf(0.5f, fsqrt(fmaf(-d, d, ff + ff)) - d, mid);
    const float db = hi - mid, dc = hi - lo;
    const float P = db + dc;
    const float disc3 = fmaf(P, P, -3.0f * (fmaf(db, db, dc * dc) - ff));
    const float q3 = disc3 >= 0.0f ? fmaf(fsqrt(disc3) - P, 1.0f / 3.0f, hi) : PFAST_INF_F;
    return q1 <= mid ? q1 : (q2 <= hi ? q2 : q3);
}

// General spacing: s = 1/v, w* = 1/h*^2, h* the spacings of axes x (a), y (b), z (c).
__device__ __forceinline__ float eik3d_f32_general(float a, float b, float c, float s, float hx,
                                                   float hy, float hz, float wx, float wy,
                                                   float wz)
{
    float ha = hx, hb = hy, hc = hz, wa = wx, wb = wy, wc = wz, t;
    if (b > a) { t = a; a = b; b = t; t = ha; ha = hb; hb = t; t = wa; wa = wb; wb = t; }
    if (c > a) { t = a; a = c; c = t; t = ha; ha = hc; hc = t; t = wa; wa = wc; wc = t; }
    if (c > b) { t = b; b = c; c = t; t = hb; hb = hc; hc = t; t = wb; wb = wc; wc = t; }
    const float q1 = fmaf(hc, s, c);
    const float d = b - c;
    const float W2 = wb + wc;
    const float disc2 = fmaf(W2, s * s, -wb * wc * d * d);
    const float q2 = b + (fsqrt(disc2) - wc * d) * frcp(W2);
    const float db = a - b, dc = a - c;
    const float W3 = wa + wb + wc;
    const float P = fmaf(wb, db, wc * dc);
    const float Q = fmaf(wb * db, db, wc * dc * dc);
    const float disc3 = fmaf(P, P, -W3 * (Q - s * s));
    const float q3 = disc3 >= 0.0f ? a + (fsqrt(disc3) - P) * frcp(W3) : PFAST_INF_F;
    return q1 <= b ? q1 : (q2 <= a ? q2 : q3);
}

// 2D FP32: a = column-neighbour min (spacing dx), b = row-neighbour min (dz).
__device__ __forceinline__ float eik2d_f32(float a, float b, float s, float dx, float dz,
                                           float wx, float wz)
{
    const bool sw = b > a;
    const float lo = sw ? a : b, hi = sw ? b : a;
    const float hlo = sw ? dx : dz;  // spacing of the smaller neighbour's axis
    const float wlo = sw ? wx : wz;
    const float q1 = fmaf(hlo, s, lo);
    const float d = hi - lo;
    const float W = wx + wz;
    const float disc = fmaf(W, s * s, -wx * wz * d * d);
    const float q2 = disc >= 0.0f ? hi + (fsqrt(disc) - wlo * d) * frcp(W) : PFAST_INF_F;
    return q1 <= hi ? q1 : q2;
}

// FP32 adjoint accumulation (same upwind coefficients as src/T_padjoint3d.h:51-68 and
// src/RT_padjoint2d.h:42-65: den = sum over neighbours of max(tc - tn, 0) / h^2, num =
// sum of max(tn - tc, 0) lambda_n / h^2) from the difference d = tn - tc of one
// neighbour: den += w max(-d, 0), num += w max(d, 0) ln, the product added by one fma.
// T values are non-negative magnitudes.  num and den start at +0: a downwind
// neighbour's product is a zero whose sign is that of ln, and a signed zero added to +0
// (or to a non-zero sum) leaves it unchanged, so the result never depends on downwind
// lambdas, not even through the sign of a zero (a -0 there would couple a node to its
// downwind neighbours and let its sign bit oscillate between sweeps).  Uniform spacing
// (UNI): every w is the same and cancels in num / den; it is left out.
// The first term sets den (a zero den of either sign fails the threshold alike) and
// adds num's product to +0.
template <bool UNI>
__device__ __forceinline__ void adj_first(float d, float ln, float w, float &num, float &den)
{
    if (UNI) {
        den = fmaxf(-d, 0.0f);
        num = fmaf(fmaxf(d, 0.0f), ln, 0.0f);
    } else {
        den = w * fmaxf(-d, 0.0f);
        num = fmaf(w * fmaxf(d, 0.0f), ln, 0.0f);
    }
}
template <bool UNI>
__device__ __forceinline__ void adj_acc(float d, float ln, float w, float &num, float &den)
{
    if (UNI) {
        den += fmaxf(-d, 0.0f);
        num = fmaf(fmaxf(d, 0.0f), ln, num);
    } else {
        den = fmaf(w, fmaxf(-d, 0.0f), den);
        num = fmaf(w * fmaxf(d, 0.0f), ln, num);
    }
}
