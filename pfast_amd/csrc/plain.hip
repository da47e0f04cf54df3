// plain.hip -- single-source FP64 kernels (bit-identical to the reference's FP64).
//
// 3D: one launch per anti-diagonal hyperplane of an ordering.  In ordering
// (s3,s1,s2) the reference visits nodes lexicographically; a node's update reads
// only its six face neighbours, and both in lexicographic and in hyperplane order
// every face neighbour "behind" it (in each axis) is already updated in this
// ordering and every neighbour "ahead" is not.  Nodes of one hyperplane are never
// face neighbours, so a hyperplane is one parallel step and the results are the
// reference's bit for bit.
//
// 2D: one work-group per source runs the whole solve (all cycles, orderings and
// hyperplanes, with work-group barriers between hyperplanes), for both the FP64
// reference arithmetic and the FP32 batched form.
#include "pfast_internal.h"
#include "plain_kernels.h"
#include "solvers.h"

// --------------------------------------------------------------------- 3D FP64
__global__ void init3d_f64(double *U, long N, long srcnode)
{
    long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) U[n] = n == srcnode ? 0.0 : PFAST_INF_D;
}

// Plane d of ordering (s3,s1,s2): transformed coords kk,ii,jj with kk+ii+jj = d.
// Thread t covers (kk, ii) in [k0,k0+nk) x [i0,i0+ni); jj is implied.
__global__ void fsm3d_plane_f64(const double *__restrict__ vp, double *__restrict__ U,
                                double *__restrict__ p, int *__restrict__ flag, int si, int sj,
                                int sk, int s1, int s2, int s3, int d, int k0, int nk, int i0,
                                int ni, double dx, double dy, double dz, int nrows, int ncols,
                                int nz, int first, int last)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk * ni) return;
    int kk = k0 + t / ni, ii = i0 + t % ni, jj = d - kk - ii;
    if (jj < 0 || jj >= ncols) return;
    int k = s3 > 0 ? kk : nz - 1 - kk, i = s1 > 0 ? ii : nrows - 1 - ii,
        j = s2 > 0 ? jj : ncols - 1 - jj;
    if (i == si && j == sj && k == sk) return;
    const long plane = (long)nrows * ncols;
    const long n = k * plane + (long)i * ncols + j;
    double a, b, c;
    if (j - 1 < 0) a = U[n + 1];
    else if (j + 1 >= ncols) a = U[n - 1];
    else { double x1 = U[n + 1], x0 = U[n - 1]; a = x1 < x0 ? x1 : x0; }
    if (i - 1 < 0) b = U[n + ncols];
    else if (i + 1 >= nrows) b = U[n - ncols];
    else { double y0 = U[n - ncols], y1 = U[n + ncols]; b = y0 < y1 ? y0 : y1; }
    if (k - 1 < 0) c = U[n + plane];
    else if (k + 1 >= nz) c = U[n - plane];
    else { double z0 = U[n - plane], z1 = U[n + plane]; c = z0 < z1 ? z0 : z1; }
    const double s = 1.0 / vp[n];
    const double q = eik3d_ref(a, b, c, s, dx, dy, dz);
    const double old = U[n];
    const double nw = old < q ? old : q;
    U[n] = nw;
    if (first) p[n] = old;
    if (last && fabs(p[n] - nw) > 1e-9) *flag = 1;
}

// Receiver boundary condition of src/T_padjoint3d.h:90-127 for one source, in
// receiver order (one thread: duplicates overwrite and EE accumulates exactly as
// the reference does).  Also computes the first convergence test isconv(0, U).
__global__ void adj3d_bc_f64(double *__restrict__ U, unsigned char *__restrict__ tag,
                             const double *__restrict__ T, const double *__restrict__ T_obs,
                             const int *__restrict__ rec, const double *__restrict__ nxyz,
                             long ch0, int nrec, int si, int sj, int sk, double dx, double dy,
                             double dz, int nrows, int ncols, int nz, double *__restrict__ EE_out,
                             int *__restrict__ flag)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const long plane = (long)nrows * ncols;
    double EE = 0.0;
    for (int idx = 0; idx < nrec; idx++) {
        const long r = idx + ch0;
        if (!(T_obs[r] < 100)) continue;
        const int i = rec[r * 3 + 0], j = rec[r * 3 + 1], k = rec[r * 3 + 2];
        const long n = k * plane + (long)i * ncols + j;
        tag[n] = 1;
        EE += (T_obs[r] - T[n]) * (T_obs[r] - T[n]) / 2.0;
        if (i == si && j == sj && k == sk) continue;
        const double nny = nxyz[r * 3 + 0], nnx = nxyz[r * 3 + 1], nnz = nxyz[r * 3 + 2];
        double Tx, Ty, Tz;
        if (k + 1 >= nz) Tz = (T[n] - T[n - plane]) / dz;
        else Tz = (T[n + plane] - T[n]) / dz;
        if (i + 1 >= nrows) Ty = (T[n] - T[n - ncols]) / dy;
        else Ty = (T[n + ncols] - T[n]) / dy;
        if (j + 1 >= ncols) Tx = (T[n] - T[n - 1]) / dx;
        else Tx = (T[n + 1] - T[n]) / dx;
        if (fabs(T_obs[r] - T[n]) > 1e-9 && fabs(nnx * Tx + nny * Ty + nnz * Tz) > 1e-6)
            U[n] = (T_obs[r] - T[n]) / (nnx * Tx + nny * Ty + nnz * Tz);
        else
            U[n] = 0.0;
    }
    int nc = 0;
    for (int idx = 0; idx < nrec && !nc; idx++) {
        const long r = idx + ch0;
        if (!(T_obs[r] < 100)) continue;
        const long n = rec[r * 3 + 2] * plane + (long)rec[r * 3 + 0] * ncols + rec[r * 3 + 1];
        if (fabs(U[n]) > 1e-9) nc = 1;
    }
    *EE_out = EE;
    *flag = nc;
}

// Adjoint hyperplane over the interior box [1,n-2]^3 (src/T_padjoint3d.h:143-155).
__global__ void adj3d_plane_f64(double *__restrict__ U, const double *__restrict__ T,
                                const unsigned char *__restrict__ tag, double *__restrict__ p,
                                int *__restrict__ flag, int s1, int s2, int s3, int d, int k0,
                                int nk, int i0, int ni, double dx, double dy, double dz,
                                int nrows, int ncols, int nz, int first, int last)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nk * ni) return;
    const int nci = ncols - 2;
    int kk = k0 + t / ni, ii = i0 + t % ni, jj = d - kk - ii;
    if (jj < 0 || jj >= nci) return;
    int k = s3 > 0 ? 1 + kk : nz - 2 - kk, i = s1 > 0 ? 1 + ii : nrows - 2 - ii,
        j = s2 > 0 ? 1 + jj : ncols - 2 - jj;
    const long plane = (long)nrows * ncols;
    const long n = k * plane + (long)i * ncols + j;
    if (tag[n]) return;
    const double old = U[n];
    const double nw = adj3d_ref(T[n], T[n + 1], T[n - 1], T[n + ncols], T[n - ncols],
                                T[n + plane], T[n - plane], U[n + 1], U[n - 1], U[n + ncols],
                                U[n - ncols], U[n + plane], U[n - plane], dx, dy, dz);
    U[n] = nw;
    if (first) p[n] = old;
    if (last && fabs(p[n] - nw) > 1e-9) *flag = 1;
}

// -------------------------------------------------------------------------- 2D
// Policies: Ref64 = the reference's FP64 arithmetic and its |p-U| <= 1e-9 test;
// Fast32 = FP32 difference-form updates with a bitwise-stationarity test.
struct Ref64 {
    using real = double;
    static constexpr bool exact = true;
    __device__ static real eik(real a, real b, real s, const Grid2 &g)
    {
        return eik2d_ref(a, b, s, g.dx, g.dz);
    }
    __device__ static real adj(real tc, real txp, real txm, real tzp, real tzm, real lxp,
                               real lxm, real lzp, real lzm, const Grid2 &g)
    {
        return adj2d_ref(tc, txp, txm, tzp, tzm, lxp, lxm, lzp, lzm, g.dx, g.dz);
    }
    __device__ static real slow(const double *vp, long n) { return 1.0 / vp[n]; }
};

struct Fast32 {
    using real = float;
    static constexpr bool exact = false;
    __device__ static real eik(real a, real b, real s, const Grid2 &g)
    {
        return eik2d_f32(a, b, s, (float)g.dx, (float)g.dz, (float)g.wx, (float)g.wz);
    }
    // the batched 3D adjoint's form (adj_acc): each difference formed once, fma
    // accumulation from +0, any stencil traveltime at INF by one compare of the maximum,
    // computed in every lane and selected without a branch
    __device__ static real adj(real tc, real txp, real txm, real tzp, real tzm, real lxp,
                               real lxm, real lzp, real lzm, const Grid2 &g)
    {
        const float wx = (float)g.wx, wz = (float)g.wz;
        const float tmax = fmaxf(fmaxf(fmaxf(tc, txp), fmaxf(txm, tzp)), tzm);
        float num, den;
        adj_first<false>(txp - tc, lxp, wx, num, den);
        adj_acc<false>(txm - tc, lxm, wx, num, den);
        adj_acc<false>(tzp - tc, lzp, wz, num, den);
        adj_acc<false>(tzm - tc, lzm, wz, num, den);
        float r = num * frcp(den);
        asm volatile("" : "+v"(r));
        return tmax < PFAST_INF_F && den > 1e-9f ? r : 0.0f;
    }
    __device__ static real slow(const double *vp, long n) { return (float)(1.0 / vp[n]); }
};

// ---------------------------------------------------- 2D FP32 row window
// One ordering of the batched FP32 form without a memory round trip per plane.
// Thread t owns sweep-ordered row t and walks it: at step p it updates walk column
// w = p - t (interior forms: the t-th interior row, w = p - t + 1), so node (t, w) is
// updated at step t + w, the plane of the plane loop, with the same operands and
// arithmetic: the same results bit for bit.  A thread keeps columns w-1 (updated),
// w and w+1 (not yet) of its row in registers.  The row behind it in sweep order is
// at column w+1 and holds this column's updated value as its "w-1"; the row ahead is
// at w-1 and holds the old value as its "w+1": every step each thread publishes those
// two to LDS (double-buffered by step parity, one barrier per step).  Memory traffic
// is batched over W2L steps: one vmcnt(0) wait, the stores of the previous batch's
// changed nodes, then the loads of the next batch (own row two columns ahead and the
// node inputs: the model's FP32 slowness, (float)(1.0 / vp) as the plane loop forms it,
// and the reflector depth), so the steps themselves touch registers and LDS only.  The loads are
// raw buffer loads issued for every column of the batch: one outside the row or the
// grid reads 0 or a neighbouring row's value, never faults, and is never used (edge
// columns and rows take the reference's one-sided rules, and only in-range nodes are
// updated).  Needs one thread per row (blockDim.x == W2NT >= rows).
constexpr int W2NT = 256, W2L = 8;

__device__ bool fsm2d_window(const double *__restrict__ vp, float *__restrict__ U, int si, int sj,
                             int mode, const double *__restrict__ Zr, const int *__restrict__ k2vec,
                             const Grid2 &g, int s1, int s2)
{
    constexpr int NT = W2NT, L = W2L, PB = 2 * NT + 1, INFSLOT = 2 * NT;
    __shared__ float pub[2 * PB];
    const int nrows = g.nrows, ncols = g.ncols, t = threadIdx.x;
    const float INF = PFAST_INF_F;
    const bool valid = t < nrows;
    const int i = s1 > 0 ? t : nrows - 1 - t;
    const long rb = valid ? (long)i * ncols : 0;
    const int rbehind = valid && t > 0 ? t - 1 : INFSLOT;
    const int rahead = t + 1 < nrows ? NT + t + 1 : INFSLOT;
    const bool top = s1 > 0 ? t == 0 : t == nrows - 1;     // i - 1 < 0
    const bool bottom = s1 > 0 ? t == nrows - 1 : t == 0;  // i + 1 >= nrows
    const float fdx = (float)g.dx, fdz = (float)g.dz, fwx = (float)g.wx, fwz = (float)g.wz;
    const double zlim = (i + 1) * g.dz;
    const unsigned N = (unsigned)nrows * ncols;
    const rsrc_t RU = make_rsrc(U, 4u * N), RS = make_rsrc(g.slow, 4u * N);
    const rsrc_t RZ = make_rsrc(Zr, mode == 1 ? 24u * ncols : 0u);
    auto col = [&](int w) { return s2 > 0 ? w : ncols - 1 - w; };
    auto in = [&](int w) { return valid && w >= 0 && w < ncols; };
    auto ld = [&](int w) { return bload(RU, 4u * (unsigned)(rb + col(w))); };
    // raw inputs of the batch's nodes: own row two columns ahead, slowness, and the
    // reflector depth of the column (mode 1).  The reflection rows of mode 2 (k2vec)
    // may live in LDS and are read at the batch start through their own pointer.
    struct Q {
        float u[L], s[L];
        double z[L];
    };
    auto load_batch = [&](int pb, Q &q) {
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int w = pb + u - t;
            const int j = col(w);
            q.u[u] = ld(w + 2);
            q.s[u] = bload(RS, 4u * (unsigned)(rb + j));
            if (mode == 1) q.z[u] = bload_d(RZ, 24u * (unsigned)j);
        }
    };
    float prev = INF, cur = ld(-t), next = ld(1 - t);
    Q qa, qb;
    load_batch(0, qa);
    if (t == 0) pub[INFSLOT] = pub[PB + INFSLOT] = INF;
    __syncthreads();
    bool changed = false;
    float sval[L];
    bool chg[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        sval[u] = 0.0f;
        chg[u] = false;
    }
    auto store_batch = [&](int pb) {  // changed nodes of the batch starting at step pb
#pragma unroll
        for (int u = 0; u < L; u++)
            if (chg[u]) bstore(RU, 4u * (unsigned)(rb + col(pb + u - t)), sval[u]);
    };
    auto batch = [&](int pb, const Q &q, Q &nq) {
        wait_vm_all();
        // node flags (src/RT_pfsm2d.h:174, :309) from the arrived inputs, before this
        // batch's stores and loads are in flight
        bool upd[L];
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int w = pb + u - t, j = col(w);
            const bool notsrc = !(i == si && j == sj);
            bool up;
            if (mode == 2) {
                const int jc = j < 0 ? 0 : j >= ncols ? ncols - 1 : j;
                up = in(w) && i <= k2vec[jc];
            } else {
                up = in(w) && (mode == 0 ? notsrc : notsrc && zlim <= q.z[u]);
            }
            // materialised here: left to the compiler, the FP64 compare sinks into each
            // step's update and lengthens the step's dependent chain
            int upi = up;
            asm volatile("" : "+v"(upi));
            upd[u] = upi != 0;
        }
        store_batch(pb - L);
        load_batch(pb + L, nq);
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int w = pb + u - t;
            float *P = pub + (u & 1) * PB;
            P[t] = prev;
            P[NT + t] = next;
            lds_barrier();
            const float vb = P[rbehind], va = P[rahead];
            float nw = cur;
            if (upd[u]) {
                // physical neighbours U[n-ncols], U[n+ncols], U[n-1], U[n+1] and the
                // reference's edge rules (src/RT_pfsm2d.h:60-75)
                const float um = s1 > 0 ? vb : va, up = s1 > 0 ? va : vb;
                const float lm = s2 > 0 ? prev : next, lp = s2 > 0 ? next : prev;
                const bool left = s2 > 0 ? w == 0 : w == ncols - 1;
                const bool right = s2 > 0 ? w == ncols - 1 : w == 0;
                const float b = top ? up : bottom ? um : (um < up ? um : up);
                const float a = left ? lp : right ? lm : (lp < lm ? lp : lm);
                const float qv = eik2d_f32(a, b, q.s[u], fdx, fdz, fwx, fwz);
                nw = cur < qv ? cur : qv;
                changed |= nw != cur;
            }
            chg[u] = __float_as_uint(nw) != __float_as_uint(cur);
            sval[u] = nw;
            prev = nw;
            cur = next;
            next = q.u[u];
        }
    };
    const int pend = nrows + ncols - 1;
    int pl = 0;  // start of the last batch run
    for (int pb = 0;; pb += 2 * L) {
        batch(pb, qa, qb);
        pl = pb;
        if (pb + L >= pend) break;
        batch(pb + L, qb, qa);
        pl = pb + L;
        if (pb + 2 * L >= pend) break;
    }
    wait_vm_all();
    store_batch(pl);
    return changed;
}

// Adjoint form of the row window: interior rows and columns only; (lambda, T) pairs
// travel together (T read-only).  The boundary rows, never updated, are read from
// memory by the first and last interior row.
__device__ bool adj2d_window(float *__restrict__ U, const float *__restrict__ T,
                             const unsigned char *__restrict__ tag, const Grid2 &g, int s1,
                             int s2)
{
    constexpr int NT = W2NT, L = W2L, PB = 2 * NT + 1, INFSLOT = 2 * NT;
    __shared__ float2 pub[2 * PB];
    const int nrows = g.nrows, ncols = g.ncols, nri = nrows - 2, t = threadIdx.x;
    const float2 NONE = make_float2(0.0f, PFAST_INF_F);
    const bool valid = t < nri;
    const int i = s1 > 0 ? 1 + t : nrows - 2 - t;
    const long rb = valid ? (long)i * ncols : 0;
    const bool hbehind = valid && t == 0, hahead = valid && t == nri - 1;
    const long rbh = (long)(s1 > 0 ? i - 1 : i + 1) * ncols;  // rows behind / ahead
    const long rah = (long)(s1 > 0 ? i + 1 : i - 1) * ncols;
    const long rh = hbehind ? rbh : rah;                       // the halo row, if any
    const bool halo = hbehind || hahead;
    const int rbehind = valid && t > 0 ? t - 1 : INFSLOT;
    const int rahead = t + 1 < nri ? NT + t + 1 : INFSLOT;
    const unsigned N = (unsigned)nrows * ncols;
    const rsrc_t RU = make_rsrc(U, 4u * N), RT = make_rsrc(T, 4u * N), RG = make_rsrc(tag, N);
    auto col = [&](int w) { return s2 > 0 ? w : ncols - 1 - w; };
    // a lane without a halo row reads past the arrays (0, no memory touched)
    auto ld2 = [&](long r, int w, bool use) {
        const unsigned e = use ? (unsigned)(r + col(w)) : 0x3FFFFFFFu;
        return make_float2(bload(RU, 4u * e), bload(RT, 4u * e));
    };
    auto in = [&](int w) { return valid && w >= 1 && w <= ncols - 2; };
    struct Q {
        float2 u[L], h[L];
        unsigned char tg[L];
    };
    auto load_batch = [&](int pb, Q &q) {
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int w = pb + u - t + 1;
            q.u[u] = ld2(rb, w + 2, true);
            q.tg[u] = (unsigned char)bload_u8(RG, (unsigned)(rb + col(w)));
            q.h[u] = ld2(rh, w, halo);
        }
    };
    float2 prev = ld2(rb, -t, true), cur = ld2(rb, 1 - t, true), next = ld2(rb, 2 - t, true);
    Q qa, qb;
    load_batch(0, qa);
    if (t == 0) pub[INFSLOT] = pub[PB + INFSLOT] = NONE;
    __syncthreads();
    bool changed = false;
    float sval[L];
    bool chg[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        sval[u] = 0.0f;
        chg[u] = false;
    }
    auto store_batch = [&](int pb) {
#pragma unroll
        for (int u = 0; u < L; u++)
            if (chg[u]) bstore(RU, 4u * (unsigned)(rb + col(pb + u - t + 1)), sval[u]);
    };
    auto batch = [&](int pb, const Q &q, Q &nq) {
        wait_vm_all();
        store_batch(pb - L);
        load_batch(pb + L, nq);
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int w = pb + u - t + 1;
            float2 *P = pub + (u & 1) * PB;
            P[t] = prev;
            P[NT + t] = next;
            lds_barrier();
            const float2 vb = hbehind ? q.h[u] : P[rbehind];
            const float2 va = hahead ? q.h[u] : P[rahead];
            float nw = cur.x;
            if (in(w) && !q.tg[u]) {
                // physical neighbours n-ncols, n+ncols, n-1, n+1 (src/RT_padjoint2d.h)
                const float2 rm = s1 > 0 ? vb : va, rp = s1 > 0 ? va : vb;
                const float2 cm = s2 > 0 ? prev : next, cp = s2 > 0 ? next : prev;
                nw = Fast32::adj(cur.y, cp.y, cm.y, rp.y, rm.y, cp.x, cm.x, rp.x, rm.x, g);
                changed |= nw != cur.x;
            }
            chg[u] = __float_as_uint(nw) != __float_as_uint(cur.x);
            sval[u] = nw;
            prev = make_float2(nw, cur.y);
            cur = next;
            next = q.u[u];
        }
    };
    const int pend = nri + ncols - 3;
    int pl = 0;
    for (int pb = 0;; pb += 2 * L) {
        batch(pb, qa, qb);
        pl = pb;
        if (pb + L >= pend) break;
        batch(pb + L, qb, qa);
        pl = pb + L;
        if (pb + 2 * L >= pend) break;
    }
    wait_vm_all();
    store_batch(pl);
    return changed;
}

// One 2D FSM solve (all cycles) by the calling work-group.  mode 0: all nodes but
// the source; 1: also only (i+1)*dz <= Z depth of column j (down-going,
// src/RT_pfsm2d.h:174); 2: only i <= k2vec[j] (up-going, :309).
// Returns the cycles swept (reference `count` capped at MAXSwp).
template <class P>
__device__ int wg_fsm2d(const double *__restrict__ vp, typename P::real *__restrict__ U,
                        typename P::real *__restrict__ pbuf, int si, int sj, int mode,
                        const double *__restrict__ Zr, const int *k2vec, int MAXSwp,
                        const Grid2 &g, int *s_flag)
{
    using real = typename P::real;
    const int nrows = g.nrows, ncols = g.ncols;
    const int nplanes = nrows + ncols - 1;
    int count = 0;
    for (;;) {
        if (!(count < MAXSwp)) break;
        count++;
        if (threadIdx.x == 0) *s_flag = 0;
        __syncthreads();
        for (int o = 0; o < 4; o++) {
            const int s1 = (o < 2) ? -1 : 1, s2 = (o & 1) ? 1 : -1;
            const bool first = o == 0, last = o == 3;
            if constexpr (!P::exact) {
                if (g.window && g.slow && blockDim.x == W2NT && nrows <= W2NT) {
                    if (fsm2d_window(vp, U, si, sj, mode, Zr, k2vec, g, s1, s2)) *s_flag = 1;
                    __syncthreads();
                    continue;
                }
            }
            for (int d = 0; d < nplanes; d++) {
                const int ilo = d - (ncols - 1) > 0 ? d - (ncols - 1) : 0;
                const int ihi = d < nrows - 1 ? d : nrows - 1;
                for (int ii = ilo + (int)threadIdx.x; ii <= ihi; ii += blockDim.x) {
                    const int jj = d - ii;
                    const int i = s1 > 0 ? ii : nrows - 1 - ii, j = s2 > 0 ? jj : ncols - 1 - jj;
                    bool upd;
                    if (mode == 0) upd = !(i == si && j == sj);
                    else if (mode == 1) upd = !(i == si && j == sj) && (i + 1) * g.dz <= Zr[j * 3];
                    else upd = i <= k2vec[j];
                    if (!upd) continue;
                    const long n = (long)i * ncols + j;
                    real a, b;
                    if (i - 1 < 0) b = U[n + ncols];
                    else if (i + 1 >= nrows) b = U[n - ncols];
                    else { real u0 = U[n - ncols], u1 = U[n + ncols]; b = u0 < u1 ? u0 : u1; }
                    if (j - 1 < 0) a = U[n + 1];
                    else if (j + 1 >= ncols) a = U[n - 1];
                    else { real u1 = U[n + 1], u0 = U[n - 1]; a = u1 < u0 ? u1 : u0; }
                    const real q = P::eik(a, b, P::slow(vp, n), g);
                    const real old = U[n];
                    const real nw = old < q ? old : q;
                    U[n] = nw;
                    if (P::exact) {
                        if (first) pbuf[n] = old;
                        if (last && fabs((double)pbuf[n] - (double)nw) > 1e-9) *s_flag = 1;
                    } else if (nw != old) {
                        *s_flag = 1;
                    }
                }
                __syncthreads();
            }
        }
        const int changed = *s_flag;
        __syncthreads();
        if (!changed) break;
    }
    return count;
}

template <class P>
__device__ double reflect_seed_dev(const typename P::real *U, const double *Zr, int jj, int k2,
                                   double qx, double dz, int ncols)
{
    // src/RT_pfsm2d.h:220-223 (FP64 arithmetic on the stored times)
    const double u0 = (double)U[(long)k2 * ncols + jj], u1 = (double)U[(long)(k2 + 1) * ncols + jj];
    double qz = (u1 - u0) / dz;
    double dnx = Zr[jj * 3 + 2], dnz = Zr[jj * 3 + 1];
    double rz = qz - 2 * (qx * dnx + qz * dnz) * dnz;
    return u0 + qz * (Zr[jj * 3] - k2 * dz - dz) + rz * ((k2 + 1) * dz + dz - Zr[jj * 3]);
}

// Whole RT_fsm2d (src/RT_pfsm2d.h:116-380) for source blockIdx.x of a batch.
// Fields are per-source contiguous: U [b][N], TTd/TTu [b][NumofR][N].
template <class P>
__global__ __launch_bounds__(256) void rt_fsm2d_wg(const double *__restrict__ vp,
                                                   typename P::real *__restrict__ Uall,
                                                   typename P::real *__restrict__ TTd,
                                                   typename P::real *__restrict__ TTu,
                                                   typename P::real *__restrict__ pall,
                                                   const double *__restrict__ Z,
                                                   const int *__restrict__ src, int NumofR,
                                                   int MAXSwp, Grid2 g, int *__restrict__ cycles,
                                                   int *__restrict__ status,
                                                   int *__restrict__ k2out)
{
    using real = typename P::real;
    extern __shared__ int s_k2[];  // ncols ints, a flag, ncols ints
    int *s_flag = s_k2 + g.ncols;
    int *s_kc = s_flag + 1;  // per column: last row above the reflector (candidate k2)
    const int b = blockIdx.x;
    // roles (grid y = 2, FP32 form with reflectors): block y 0 runs the reflection
    // solves in the up-going field itself, block y 1 the transmission solve, which
    // does not depend on them; with one block per source both run in turn in U
    const int role = gridDim.y > 1 ? (int)blockIdx.y : -1;
    const int nrows = g.nrows, ncols = g.ncols;
    const long N = (long)nrows * ncols;
    real *U = Uall + b * N;
    real *pbuf = pall ? pall + b * N : nullptr;
    const int si = src[b * 2 + 0], sj = src[b * 2 + 1];
    const double dz = g.dz, dx = g.dx;
    int st = 0;
    for (int j = threadIdx.x; j < ncols; j += blockDim.x) s_k2[j] = -1;
    __syncthreads();
    for (int ir = 0; ir < NumofR && st == 0 && role != 1; ir++) {
        // working field of this reflector's solves: U, or the up-going field itself
        real *W = role == 0 ? TTu + ((long)b * NumofR + ir) * N : U;
        const double *Zr = Z + (long)ir * ncols * 3;
        for (long n = threadIdx.x; n < N; n += blockDim.x) {
            W[n] = PFAST_INF_D;
            if (pbuf) pbuf[n] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 0) W[(long)si * ncols + sj] = 0;
        __syncthreads();
        int c = wg_fsm2d<P>(vp, W, pbuf, si, sj, 1, Zr, s_k2, MAXSwp, g, s_flag);
        if (threadIdx.x == 0 && cycles) cycles[b * (1 + 2 * NumofR) + 1 + 2 * ir] = c;
        real *Td = TTd + ((long)b * NumofR + ir) * N;
        for (long n = threadIdx.x; n < N; n += blockDim.x) Td[n] = W[n];
        __syncthreads();
        // reflection points and seeds (src/RT_pfsm2d.h:195-281).  A column's k2 (the
        // last row kk passing the reference's test) depends on the reflector alone, so
        // the columns are searched in parallel; the seeding pass stays sequential in
        // the reference's column order because a seed can feed the next column's qx.
        for (int jj = threadIdx.x; jj < ncols; jj += blockDim.x) {
            const bool left = jj == 0, right = jj == ncols - 1;
            int k2 = -1;
            for (int kk = 0; kk < nrows - 1; kk++) {
                bool cnd;
                if (left) cnd = (kk + 1) * dz + dz <= Zr[jj * 3] && kk * dz + dz <= Zr[(jj + 1) * 3];
                else if (right) cnd = (kk + 1) * dz + dz <= Zr[jj * 3] && kk * dz + dz <= Zr[(jj - 1) * 3];
                else cnd = (kk + 1) * dz + dz <= Zr[jj * 3] &&
                           (kk * dz + dz <= Zr[(jj - 1) * 3] || kk * dz + dz <= Zr[(jj + 1) * 3]);
                if (cnd && kk >= k2) k2 = kk;
            }
            s_kc[jj] = k2;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int step = 0; step < ncols && st == 0; step++) {
                // columns 1..ncols-2, then 0, then ncols-1
                const int jj = step < ncols - 2 ? step + 1 : (step == ncols - 2 ? 0 : ncols - 1);
                const bool left = jj == 0, right = jj == ncols - 1;
                const int k2 = s_kc[jj];
                if (Zr[jj * 3] <= nrows * dz && k2 > 0) {
                    const long r0 = (long)k2 * ncols;
                    double qx;
                    if (left) qx = ((double)W[r0 + jj + 1] - (double)W[r0 + jj]) / dx;
                    else if (right) qx = ((double)W[r0 + jj] - (double)W[r0 + jj - 1]) / dx;
                    else if (k2 * dz + dz <= Zr[(jj - 1) * 3] && k2 * dz + dz <= Zr[(jj + 1) * 3])
                        qx = ((double)W[r0 + jj + 1] - (double)W[r0 + jj - 1]) / 2.0 / dx;
                    else if (k2 * dz + dz <= Zr[(jj - 1) * 3])
                        qx = ((double)W[r0 + jj] - (double)W[r0 + jj - 1]) / dx;
                    else
                        qx = ((double)W[r0 + jj + 1] - (double)W[r0 + jj]) / dx;
                    const double v = reflect_seed_dev<P>(W, Zr, jj, k2, qx, dz, ncols);
                    W[(long)(k2 + 1) * ncols + jj] = (real)v;
                    s_k2[jj] = k2;
                    if (v < 0) st = 2;
                } else if (Zr[jj * 3] > nrows * dz) {
                    s_k2[jj] = nrows - 1;
                } else {
                    st = 1;
                }
            }
            *s_flag = st;
        }
        __syncthreads();
        st = *s_flag;
        __syncthreads();
        if (st) break;
        // reset above the reflection points (src/RT_pfsm2d.h:289-295)
        for (long n = threadIdx.x; n < N; n += blockDim.x) {
            const int i = (int)(n / ncols), j = (int)(n % ncols);
            if (i <= s_k2[j]) W[n] = PFAST_INF_D;
            if (pbuf) pbuf[n] = 0;
        }
        __syncthreads();
        c = wg_fsm2d<P>(vp, W, pbuf, si, sj, 2, Zr, s_k2, MAXSwp, g, s_flag);
        if (threadIdx.x == 0 && cycles) cycles[b * (1 + 2 * NumofR) + 2 + 2 * ir] = c;
        real *Tu = TTu + ((long)b * NumofR + ir) * N;
        if (W != Tu)
            for (long n = threadIdx.x; n < N; n += blockDim.x) Tu[n] = W[n];
        __syncthreads();
    }
    if (st == 0 && role != 0) {
        for (long n = threadIdx.x; n < N; n += blockDim.x) {
            U[n] = PFAST_INF_D;
            if (pbuf) pbuf[n] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 0) U[(long)si * ncols + sj] = 0;
        __syncthreads();
        int c = wg_fsm2d<P>(vp, U, pbuf, si, sj, 0, Z, s_k2, MAXSwp, g, s_flag);
        if (threadIdx.x == 0 && cycles) cycles[b * (1 + 2 * NumofR)] = c;
    }
    if (role == 1) return;  // the reflection block reports k2 and the status
    if (k2out)
        for (int j = threadIdx.x; j < ncols; j += blockDim.x) k2out[(long)b * ncols + j] = s_k2[j];
    if (threadIdx.x == 0 && status) status[b] = st;
}

// Interior adjoint sweeps (src/RT_padjoint2d.h:124-148) with the reference's
// carried `count` (Radjoint_fsm2d does not reset it between reflectors).
// first_nc: result of the first convergence test isconv(0, U) negated.
template <class P>
__device__ int wg_adj2d(typename P::real *__restrict__ U, typename P::real *__restrict__ pbuf,
                        const typename P::real *__restrict__ T,
                        const unsigned char *__restrict__ tag, int &count, int first_nc,
                        int MAXSwp, const Grid2 &g, int *s_flag)
{
    using real = typename P::real;
    const int nrows = g.nrows, ncols = g.ncols;
    const int nri = nrows - 2, nci = ncols - 2;
    const int nplanes = nri + nci - 1;
    int notconv = first_nc, swept = 0;
    for (;;) {
        if (!(count < MAXSwp)) { count++; break; }
        count++;
        if (!notconv) break;
        swept++;
        if (threadIdx.x == 0) *s_flag = 0;
        __syncthreads();
        for (int o = 0; o < 4; o++) {
            const int s1 = (o < 2) ? -1 : 1, s2 = (o & 1) ? 1 : -1;
            const bool first = o == 0, last = o == 3;
            if constexpr (!P::exact) {
                // the window keeps one halo-row queue per thread, so it needs two
                // interior rows at least (one row would be both first and last)
                if (g.window && blockDim.x == W2NT && nri >= 2 && nri <= W2NT && nci >= 1) {
                    if (adj2d_window(U, T, tag, g, s1, s2)) *s_flag = 1;
                    __syncthreads();
                    continue;
                }
            }
            for (int d = 0; d < nplanes; d++) {
                const int ilo = d - (nci - 1) > 0 ? d - (nci - 1) : 0;
                const int ihi = d < nri - 1 ? d : nri - 1;
                for (int ii = ilo + (int)threadIdx.x; ii <= ihi; ii += blockDim.x) {
                    const int jj = d - ii;
                    const int i = s1 > 0 ? 1 + ii : nrows - 2 - ii, j = s2 > 0 ? 1 + jj : ncols - 2 - jj;
                    const long n = (long)i * ncols + j;
                    if (tag[n]) continue;
                    const real old = U[n];
                    const real nw = P::adj(T[n], T[n + 1], T[n - 1], T[n + ncols], T[n - ncols],
                                           U[n + 1], U[n - 1], U[n + ncols], U[n - ncols], g);
                    U[n] = nw;
                    if (P::exact) {
                        if (first) pbuf[n] = old;
                        if (last && fabs((double)pbuf[n] - (double)nw) > 1e-9) *s_flag = 1;
                    } else if (nw != old) {
                        *s_flag = 1;
                    }
                }
                __syncthreads();
            }
        }
        notconv = *s_flag;
        __syncthreads();
    }
    return swept;
}

// Receiver BC for one source (all threads of the block), src/RT_padjoint2d.h:298-330
// (transmission, checkSrc=1, pick T_obs<100) and :402-435 (reflection, checkSrc=0,
// pick TR_obs != 0).  A receiver's boundary value and misfit term read only T, so the
// threads compute them for 256 receivers at a time; thread 0 then applies them in
// receiver order -- a node picked twice keeps the last value and EE (thread 0's) adds
// the terms in the reference's order.  Sets *s_nc to !isconv(0,U).
template <class P>
__device__ void bc2d(typename P::real *U, unsigned char *tag, const typename P::real *T,
                     const double *obs, int obs_stride, int obs_off, bool transmission,
                     double Cr, const int *rec, const double *nxy, long ch0, int nrec, int si,
                     int sj, const Grid2 &g, int *s_nc, double &EE)
{
    using real = typename P::real;
    constexpr int NB = 256;  // receivers per round (>= blockDim.x)
    __shared__ double s_term[NB];
    __shared__ real s_lam[NB];
    __shared__ int s_node[NB], s_kind[NB];  // kind 0: not picked, 1: picked, 2: with lambda
    const int nrows = g.nrows, ncols = g.ncols, tid = threadIdx.x;
    auto picked = [&](double o) { return transmission ? o < 100 : o != 0; };
    for (int base = 0; base < nrec; base += (int)blockDim.x) {
        const int idx = base + tid;
        if (idx < nrec) {
            const long r = idx + ch0;
            const double o = obs[r * obs_stride + obs_off];
            int kind = 0;
            long n = 0;
            double term = 0.0;
            real lam = 0;
            if (picked(o)) {
                const int i = rec[r * 2 + 0], j = rec[r * 2 + 1];
                n = (long)i * ncols + j;
                const double Tn = (double)T[n];
                term = Cr * (o - Tn) * (o - Tn) / 2.0;
                kind = 1;
                if (!(transmission && i == si && j == sj)) {
                    const double nny = nxy[r * 2 + 0], nnx = nxy[r * 2 + 1];
                    double Tx, Tz;
                    if (i + 1 >= nrows) Tz = (Tn - (double)T[n - ncols]) / g.dz;
                    else Tz = ((double)T[n + ncols] - Tn) / g.dz;
                    if (j + 1 >= ncols) Tx = (Tn - (double)T[n - 1]) / g.dx;
                    else Tx = ((double)T[n + 1] - Tn) / g.dx;
                    if (fabs(o - Tn) > 1e-9 && fabs(nnx * Tx + nny * Tz) > 1e-6)
                        lam = (real)(Cr * (o - Tn) / (nnx * Tx + nny * Tz));
                    else
                        lam = 0;
                    kind = 2;
                }
            }
            s_term[tid] = term;
            s_lam[tid] = lam;
            s_node[tid] = (int)n;
            s_kind[tid] = kind;
        }
        __syncthreads();
        if (tid == 0) {
            const int m = nrec - base < (int)blockDim.x ? nrec - base : (int)blockDim.x;
            for (int k = 0; k < m; k++) {
                if (!s_kind[k]) continue;
                const long n = s_node[k];
                tag[n] = 1;
                EE += s_term[k];
                if (s_kind[k] == 2) U[n] = s_lam[k];
            }
        }
        __syncthreads();
    }
    // some picked receiver's lambda above 1e-9 in magnitude (after all the writes)
    if (tid == 0) *s_nc = 0;
    __syncthreads();
    for (int idx = tid; idx < nrec; idx += (int)blockDim.x) {
        const long r = idx + ch0;
        if (!picked(obs[r * obs_stride + obs_off])) continue;
        const long n = (long)rec[r * 2 + 0] * ncols + rec[r * 2 + 1];
        if (fabs((double)U[n]) > 1e-9) *s_nc = 1;
    }
    __syncthreads();
}

// Tadjoint_fsm2d (src/RT_padjoint2d.h:288-380, Cr-weighted overload :67-157) for
// source b of a batch, by the calling block: U [b][N], tag scratch [b][N].
template <class P>
__device__ void tadj2d_src(typename P::real *__restrict__ Uall, typename P::real *__restrict__ pall,
                           unsigned char *__restrict__ tagall,
                           const typename P::real *__restrict__ Tall,
                           const double *__restrict__ T_obs, double Cr,
                           const int *__restrict__ rec, const double *__restrict__ nxy,
                           const long *__restrict__ ch0, const int *__restrict__ srnum,
                           const int *__restrict__ src, int MAXSwp, const Grid2 &g,
                           double *__restrict__ EE, int *__restrict__ cycles, int b)
{
    __shared__ int s_flag[2];
    const long N = (long)g.nrows * g.ncols;
    auto *U = Uall + b * N;
    auto *pbuf = pall ? pall + b * N : nullptr;
    auto *tag = tagall + b * N;
    const auto *T = Tall + b * N;
    for (long n = threadIdx.x; n < N; n += blockDim.x) {
        U[n] = 0;
        tag[n] = 0;
        if (pbuf) pbuf[n] = 0;
    }
    __syncthreads();
    double e = 0.0;
    bc2d<P>(U, tag, T, T_obs, 1, 0, true, Cr, rec, nxy, ch0[b], srnum[b], src[b * 2],
            src[b * 2 + 1], g, &s_flag[1], e);
    if (threadIdx.x == 0) EE[b] = e;
    int count = 0;
    const int nc = s_flag[1];
    int sw = wg_adj2d<P>(U, pbuf, T, tag, count, nc, MAXSwp, g, s_flag);
    if (threadIdx.x == 0 && cycles) cycles[b] = sw;
}

template <class P>
__global__ __launch_bounds__(256) void tadj2d_wg(typename P::real *__restrict__ Uall,
                                                 typename P::real *__restrict__ pall,
                                                 unsigned char *__restrict__ tagall,
                                                 const typename P::real *__restrict__ Tall,
                                                 const double *__restrict__ T_obs, double Cr,
                                                 const int *__restrict__ rec,
                                                 const double *__restrict__ nxy,
                                                 const long *__restrict__ ch0,
                                                 const int *__restrict__ srnum,
                                                 const int *__restrict__ src, int MAXSwp, Grid2 g,
                                                 double *__restrict__ EE, int *__restrict__ cycles)
{
    tadj2d_src<P>(Uall, pall, tagall, Tall, T_obs, Cr, rec, nxy, ch0, srnum, src, MAXSwp, g, EE,
                  cycles, blockIdx.x);
}

// Radjoint_fsm2d (src/RT_padjoint2d.h:382-473, Cr-weighted overload :159-285) for
// source b, by the calling block.  Tu [b][NumofR][N]; UR [b][N]; U/p/tag scratch [b][N].
template <class P>
__device__ void radj2d_src(typename P::real *__restrict__ URall, typename P::real *__restrict__ Uall,
                           typename P::real *__restrict__ pall, unsigned char *__restrict__ tagall,
                           const typename P::real *__restrict__ Tuall,
                           const double *__restrict__ TR_obs, double Cr,
                           const int *__restrict__ rec, const double *__restrict__ nxy,
                           const long *__restrict__ ch0, const int *__restrict__ srnum,
                           const int *__restrict__ src, int NumofR, int MAXSwp, const Grid2 &g,
                           double *__restrict__ EE, int *__restrict__ cycles, int b)
{
    __shared__ int s_flag[2];
    const long N = (long)g.nrows * g.ncols;
    auto *UR = URall + b * N;
    auto *U = Uall + b * N;
    auto *pbuf = pall ? pall + b * N : nullptr;
    auto *tag = tagall + b * N;
    for (long n = threadIdx.x; n < N; n += blockDim.x) {
        UR[n] = 0;
        tag[n] = 0;
    }
    int count = 0;
    double EEacc = 0.0;
    for (int ir = 0; ir < NumofR; ir++) {
        const auto *T = Tuall + ((long)b * NumofR + ir) * N;
        for (long n = threadIdx.x; n < N; n += blockDim.x) {
            U[n] = 0;
            if (pbuf) pbuf[n] = 0;
        }
        __syncthreads();
        bc2d<P>(U, tag, T, TR_obs, NumofR, ir, false, Cr, rec, nxy, ch0[b], srnum[b],
                src[b * 2], src[b * 2 + 1], g, &s_flag[1], EEacc);
        const int nc = s_flag[1];
        int sw = wg_adj2d<P>(U, pbuf, T, tag, count, nc, MAXSwp, g, s_flag);
        if (threadIdx.x == 0 && cycles) cycles[b * NumofR + ir] = sw;
        for (long n = threadIdx.x; n < N; n += blockDim.x) UR[n] = UR[n] + U[n];
        __syncthreads();
    }
    if (threadIdx.x == 0) EE[b] = EEacc;
}

template <class P>
__global__ __launch_bounds__(256) void radj2d_wg(typename P::real *__restrict__ URall,
                                                 typename P::real *__restrict__ Uall,
                                                 typename P::real *__restrict__ pall,
                                                 unsigned char *__restrict__ tagall,
                                                 const typename P::real *__restrict__ Tuall,
                                                 const double *__restrict__ TR_obs, double Cr,
                                                 const int *__restrict__ rec,
                                                 const double *__restrict__ nxy,
                                                 const long *__restrict__ ch0,
                                                 const int *__restrict__ srnum,
                                                 const int *__restrict__ src, int NumofR,
                                                 int MAXSwp, Grid2 g, double *__restrict__ EE,
                                                 int *__restrict__ cycles)
{
    radj2d_src<P>(URall, Uall, pall, tagall, Tuall, TR_obs, Cr, rec, nxy, ch0, srnum, src, NumofR,
                  MAXSwp, g, EE, cycles, blockIdx.x);
}

// Both adjoints of a batch in one launch (FP32 form): block y 0 runs source x's
// transmission adjoint, block y 1 its reflection adjoints; they share no buffer (the
// reflection adjoint has its own tag scratch).
__global__ __launch_bounds__(256) void adj2d_pair_wg(
    float *__restrict__ LT, unsigned char *__restrict__ tagT, const float *__restrict__ T,
    const double *__restrict__ T_obs, float *__restrict__ LR, float *__restrict__ Uw,
    unsigned char *__restrict__ tagR, const float *__restrict__ Tu,
    const double *__restrict__ TR_obs, double Cr, const int *__restrict__ rec,
    const double *__restrict__ nxy, const long *__restrict__ ch0, const int *__restrict__ srnum,
    const int *__restrict__ src, int NumofR, int MAXSwp, Grid2 g, double *__restrict__ EEt,
    int *__restrict__ tcyc, double *__restrict__ EEr, int *__restrict__ rcyc)
{
    if (blockIdx.y == 0)
        tadj2d_src<Fast32>(LT, nullptr, tagT, T, T_obs, Cr, rec, nxy, ch0, srnum, src, MAXSwp, g,
                           EEt, tcyc, blockIdx.x);
    else
        radj2d_src<Fast32>(LR, Uw, nullptr, tagR, Tu, TR_obs, Cr, rec, nxy, ch0, srnum, src,
                           NumofR, MAXSwp, g, EEr, rcyc, blockIdx.x);
}

// ------------------------------------------------------------ launch wrappers
namespace pfast {

hipError_t launch_rt_fsm2d_f64(hipStream_t st, const double *vp, double *U, double *TTd,
                               double *TTu, double *p, const double *Z, const int *src, int nb,
                               int NumofR, int MAXSwp, Grid2 g, int *cycles, int *status,
                               int *k2out)
{
    size_t shm = sizeof(int) * (2 * g.ncols + 1);
    hipLaunchKernelGGL(rt_fsm2d_wg<Ref64>, dim3(nb), dim3(256), shm, st, vp, U, TTd, TTu, p, Z,
                       src, NumofR, MAXSwp, g, cycles, status, k2out);
    return hipGetLastError();
}

hipError_t launch_rt_fsm2d_f32(hipStream_t st, const double *vp, float *U, float *TTd,
                               float *TTu, const double *Z, const int *src, int nb, int NumofR,
                               int MAXSwp, Grid2 g, int *cycles, int *status, int *k2out)
{
    size_t shm = sizeof(int) * (2 * g.ncols + 1);
    hipLaunchKernelGGL(rt_fsm2d_wg<Fast32>, dim3(nb, NumofR > 0 && g.window ? 2 : 1), dim3(256),
                       shm, st, vp,
                       U, TTd, TTu,
                       (float *)nullptr, Z, src, NumofR, MAXSwp, g, cycles, status, k2out);
    return hipGetLastError();
}

hipError_t launch_tadj2d_f64(hipStream_t st, double *U, double *p, unsigned char *tag,
                             const double *T, const double *T_obs, double Cr, const int *rec,
                             const double *nxy, const long *ch0, const int *srnum,
                             const int *src, int nb, int MAXSwp, Grid2 g, double *EE,
                             int *cycles)
{
    hipLaunchKernelGGL(tadj2d_wg<Ref64>, dim3(nb), dim3(256), 0, st, U, p, tag, T, T_obs, Cr,
                       rec, nxy, ch0, srnum, src, MAXSwp, g, EE, cycles);
    return hipGetLastError();
}

hipError_t launch_tadj2d_f32(hipStream_t st, float *U, unsigned char *tag, const float *T,
                             const double *T_obs, double Cr, const int *rec, const double *nxy,
                             const long *ch0, const int *srnum, const int *src, int nb,
                             int MAXSwp, Grid2 g, double *EE, int *cycles)
{
    hipLaunchKernelGGL(tadj2d_wg<Fast32>, dim3(nb), dim3(256), 0, st, U, (float *)nullptr, tag,
                       T, T_obs, Cr, rec, nxy, ch0, srnum, src, MAXSwp, g, EE, cycles);
    return hipGetLastError();
}

hipError_t launch_radj2d_f64(hipStream_t st, double *UR, double *U, double *p,
                             unsigned char *tag, const double *Tu, const double *TR_obs,
                             double Cr, const int *rec, const double *nxy, const long *ch0,
                             const int *srnum, const int *src, int nb, int NumofR, int MAXSwp,
                             Grid2 g, double *EE, int *cycles)
{
    hipLaunchKernelGGL(radj2d_wg<Ref64>, dim3(nb), dim3(256), 0, st, UR, U, p, tag, Tu, TR_obs,
                       Cr, rec, nxy, ch0, srnum, src, NumofR, MAXSwp, g, EE, cycles);
    return hipGetLastError();
}

hipError_t launch_radj2d_f32(hipStream_t st, float *UR, float *U, unsigned char *tag,
                             const float *Tu, const double *TR_obs, double Cr, const int *rec,
                             const double *nxy, const long *ch0, const int *srnum,
                             const int *src, int nb, int NumofR, int MAXSwp, Grid2 g,
                             double *EE, int *cycles)
{
    hipLaunchKernelGGL(radj2d_wg<Fast32>, dim3(nb), dim3(256), 0, st, UR, U, (float *)nullptr,
                       tag, Tu, TR_obs, Cr, rec, nxy, ch0, srnum, src, NumofR, MAXSwp, g, EE,
                       cycles);
    return hipGetLastError();
}

hipError_t launch_adj2d_pair_f32(hipStream_t st, float *LT, unsigned char *tagT, const float *T,
                                 const double *T_obs, float *LR, float *Uw, unsigned char *tagR,
                                 const float *Tu, const double *TR_obs, double Cr, const int *rec,
                                 const double *nxy, const long *ch0, const int *srnum,
                                 const int *src, int nb, int NumofR, int MAXSwp, Grid2 g,
                                 double *EEt, int *tcyc, double *EEr, int *rcyc)
{
    hipLaunchKernelGGL(adj2d_pair_wg, dim3(nb, 2), dim3(256), 0, st, LT, tagT, T, T_obs, LR, Uw,
                       tagR, Tu, TR_obs, Cr, rec, nxy, ch0, srnum, src, NumofR, MAXSwp, g, EEt,
                       tcyc, EEr, rcyc);
    return hipGetLastError();
}

hipError_t launch_init3d_f64(hipStream_t st, double *U, long N, long srcnode)
{
    hipLaunchKernelGGL(init3d_f64, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, U, N,
                       srcnode);
    return hipGetLastError();
}

// One ordering's worth of hyperplane launches (forward).
hipError_t launch_fsm3d_ordering_f64(hipStream_t st, const double *vp, double *U, double *p,
                                     int *flag, int si, int sj, int sk, int s1, int s2, int s3,
                                     double dx, double dy, double dz, int nrows, int ncols,
                                     int nz, int first, int last)
{
    const int nplanes = nz + nrows + ncols - 2;
    for (int d = 0; d < nplanes; d++) {
        int k0 = d - (nrows - 1) - (ncols - 1);
        if (k0 < 0) k0 = 0;
        int k1 = d < nz - 1 ? d : nz - 1;
        int i0 = d - k1 - (ncols - 1);
        if (i0 < 0) i0 = 0;
        int i1 = d - k0 < nrows - 1 ? d - k0 : nrows - 1;
        const int nk = k1 - k0 + 1, ni = i1 - i0 + 1;
        const int nt = nk * ni;
        hipLaunchKernelGGL(fsm3d_plane_f64, dim3((nt + 255) / 256), dim3(256), 0, st, vp, U, p,
                           flag, si, sj, sk, s1, s2, s3, d, k0, nk, i0, ni, dx, dy, dz, nrows,
                           ncols, nz, first, last);
    }
    return hipGetLastError();
}

hipError_t launch_adj3d_bc_f64(hipStream_t st, double *U, unsigned char *tag, const double *T,
                               const double *T_obs, const int *rec, const double *nxyz, long ch0,
                               int nrec, int si, int sj, int sk, double dx, double dy, double dz,
                               int nrows, int ncols, int nz, double *EE, int *flag)
{
    hipLaunchKernelGGL(adj3d_bc_f64, dim3(1), dim3(64), 0, st, U, tag, T, T_obs, rec, nxyz, ch0,
                       nrec, si, sj, sk, dx, dy, dz, nrows, ncols, nz, EE, flag);
    return hipGetLastError();
}

hipError_t launch_adj3d_ordering_f64(hipStream_t st, double *U, const double *T,
                                     const unsigned char *tag, double *p, int *flag, int s1,
                                     int s2, int s3, double dx, double dy, double dz, int nrows,
                                     int ncols, int nz, int first, int last)
{
    const int nzi = nz - 2, nri = nrows - 2, nci = ncols - 2;
    if (nzi <= 0 || nri <= 0 || nci <= 0) return hipSuccess;
    const int nplanes = nzi + nri + nci - 2;
    for (int d = 0; d < nplanes; d++) {
        int k0 = d - (nri - 1) - (nci - 1);
        if (k0 < 0) k0 = 0;
        int k1 = d < nzi - 1 ? d : nzi - 1;
        int i0 = d - k1 - (nci - 1);
        if (i0 < 0) i0 = 0;
        int i1 = d - k0 < nri - 1 ? d - k0 : nri - 1;
        const int nk = k1 - k0 + 1, ni = i1 - i0 + 1;
        const int nt = nk * ni;
        hipLaunchKernelGGL(adj3d_plane_f64, dim3((nt + 255) / 256), dim3(256), 0, st, U, T, tag,
                           p, flag, s1, s2, s3, d, k0, nk, i0, ni, dx, dy, dz, nrows, ncols, nz,
                           first, last);
    }
    return hipGetLastError();
}

}  // namespace pfast
