// batch3d.hip -- batched FP32 3D sweeps (every source of a shot set on one GPU).
//
// Layout in HBM: fields are stored per chunk of C sources, node-major with the C
// sources innermost: F[chunk][k][i][j][c] (k = depth, i = row, j = column).  A
// velocity-derived cell time f[k][i][j] (h/v for uniform spacing, else 1/v) is
// shared by every chunk.
//
// Schedule ("slab ring"): the (i,j) plane is cut into TI x TJ column tiles that
// run the whole depth.  For one ordering (s3,s1,s2) tile (ti',tj') (coordinates
// taken in sweep direction) depends only on tiles (ti'-1,tj') and (ti',tj'-1), so
// one launch per tile anti-diagonal D = ti'+tj' keeps the reference's upwind
// order with no inter-block waiting.  Inside a block, thread (li,lj,c) owns one
// column of one source; hyperplane p of the tile processes level pk = p - pi - pj
// (pi,pj its sweep-ordered position), so every node sees its behind neighbours
// updated and its ahead neighbours not yet, as in the reference's lexicographic
// sweep.  Depth slabs (the tile plus its four side halos) stream through an LDS
// ring of W levels: each hyperplane the block stores the slab it will need next
// into LDS (loaded L hyperplanes earlier into registers, so HBM latency overlaps
// the sweep), updates its nodes in place, and writes back the slab that no
// thread reads any more (only nodes that changed).
#include "batch3d.h"
#include "pfast_internal.h"
#include "solvers.h"

namespace {

template <int TI, int TJ, int C>
struct Geo {
    static constexpr int NT = TI * TJ * C;
    static constexpr int HI = TI + 2, HJ = TJ + 2;
    static constexpr int W = TI + TJ + 3;  // ring depth >= span + 5
    static constexpr int SLAB = HI * HJ * C;
    static constexpr int NH = 2 * (TI + TJ) * C;  // side-halo elements per level
    static_assert(NH <= NT, "halo must fit one element per thread");
    static_assert(64 % C == 0, "C must divide the wavefront");
};

struct TileCtx {
    int chunk, i0, j0, TIe, TJe, span;
    bool live;
};

template <int TI, int TJ>
__device__ __forceinline__ TileCtx decode_tile(const SlabArgs &a, int D, int s1, int s2)
{
    TileCtx t;
    const int slot = blockIdx.x % a.nact_p8;
    const int tt = blockIdx.x / a.nact_p8;
    t.live = slot < a.nact;
    t.chunk = t.live ? a.active[slot] : 0;
    const int tlo = D - (a.ntj - 1) > 0 ? D - (a.ntj - 1) : 0;
    const int ti_ = tlo + tt, tj_ = D - ti_;
    const int ti = s1 > 0 ? ti_ : a.nti - 1 - ti_;
    const int tj = s2 > 0 ? tj_ : a.ntj - 1 - tj_;
    t.i0 = ti * TI;
    t.j0 = tj * TJ;
    t.TIe = min(TI, a.nrows - t.i0);
    t.TJe = min(TJ, a.ncols - t.j0);
    t.span = t.TIe + t.TJe - 2;
    return t;
}

// Side-halo element h of a level: LDS position and global (i,j), or !use.
template <int TI, int TJ, int C>
struct HaloSlot {
    int lds;      // offset inside one ring slot
    long gnode;   // i*ncols + j of the halo node (valid if inside)
    int c;
    bool use;     // this thread owns a halo element that is read
    bool inside;  // halo node is inside the domain (else it reads INF)
};

template <int TI, int TJ, int C>
__device__ __forceinline__ HaloSlot<TI, TJ, C> halo_slot(int h, const TileCtx &t, int nrows,
                                                         int ncols)
{
    using G = Geo<TI, TJ, C>;
    HaloSlot<TI, TJ, C> r;
    r.use = false;
    r.inside = false;
    r.lds = 0;
    r.gnode = 0;
    r.c = h % C;
    if (h >= G::NH) return r;
    int q = h / C;  // node index along the side list
    int hi, hj, gi, gj;
    if (q < TJ) {  // top side, row i0-1
        if (q >= t.TJe) return r;
        hi = 0; hj = q + 1; gi = t.i0 - 1; gj = t.j0 + q;
    } else if (q < 2 * TJ) {  // bottom side, row i0+TIe
        q -= TJ;
        if (q >= t.TJe) return r;
        hi = t.TIe + 1; hj = q + 1; gi = t.i0 + t.TIe; gj = t.j0 + q;
    } else if (q < 2 * TJ + TI) {  // left side, column j0-1
        q -= 2 * TJ;
        if (q >= t.TIe) return r;
        hi = q + 1; hj = 0; gi = t.i0 + q; gj = t.j0 - 1;
    } else {  // right side, column j0+TJe
        q -= 2 * TJ + TI;
        if (q >= t.TIe) return r;
        hi = q + 1; hj = t.TJe + 1; gi = t.i0 + q; gj = t.j0 + t.TJe;
    }
    r.use = true;
    r.inside = gi >= 0 && gi < nrows && gj >= 0 && gj < ncols;
    r.lds = (hi * G::HJ + hj) * C + r.c;
    r.gnode = (long)gi * ncols + gj;
    return r;
}

__device__ __forceinline__ unsigned chunk_bits(unsigned long long m, int C)
{
    // bit c set iff some lane with (lane % C) == c voted
    unsigned bits = 0;
    for (int c = 0; c < C; c++) {
        unsigned long long pat = 0;
        for (int l = c; l < 64; l += C) pat |= 1ull << l;
        if (m & pat) bits |= 1u << c;
    }
    return bits;
}

// ------------------------------------------------------------------- forward
template <int TI, int TJ, int C, int L, bool UNI>
__global__ __launch_bounds__(TI *TJ *C) void fwd3d_slab(SlabArgs a, int D, int s1, int s2,
                                                        int s3)
{
    using G = Geo<TI, TJ, C>;
    constexpr int W = G::W, SLAB = G::SLAB;
    __shared__ float ring[W * SLAB];
    __shared__ float fring[W * TI * TJ];

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const int tid = threadIdx.x;
    const int c = tid % C, lj = (tid / C) % TJ, li = tid / (C * TJ);
    const bool valid = li < t.TIe && lj < t.TJe;
    const int pi = s1 > 0 ? li : t.TIe - 1 - li;
    const int pj = s2 > 0 ? lj : t.TJe - 1 - lj;
    const int nz = a.nz;
    const long plane = (long)a.nrows * a.ncols;
    float *__restrict__ F = a.field + (size_t)t.chunk * plane * nz * C;
    const long mynode = (long)(t.i0 + li) * a.ncols + (t.j0 + lj);
    const int me = (((li + 1) * G::HJ) + (lj + 1)) * C + c;
    const int src = a.srcnode[t.chunk * C + c];
    const HaloSlot<TI, TJ, C> hs = halo_slot<TI, TJ, C>(tid, t, a.nrows, a.ncols);
    const bool fown = tid < TI * TJ && (tid / TJ) < t.TIe && (tid % TJ) < t.TJe;
    const long fnode = (long)(t.i0 + tid / TJ) * a.ncols + (t.j0 + tid % TJ);
    const float *__restrict__ slow = a.slow;

    auto kof = [&](int lev) { return s3 > 0 ? lev : nz - 1 - lev; };
    auto ld_int = [&](int lev) -> float {
        return valid ? F[((size_t)kof(lev) * plane + mynode) * C + c] : 0.0f;
    };
    auto ld_halo = [&](int lev) -> float {
        return (hs.use && hs.inside) ? F[((size_t)kof(lev) * plane + hs.gnode) * C + hs.c]
                                     : PFAST_INF_F;
    };
    auto ld_f = [&](int lev) -> float { return fown ? slow[(size_t)kof(lev) * plane + fnode] : 0.0f; };
    auto st_lds = [&](int lev, float vi, float vh, float vf) {
        float *r = ring + (lev % W) * SLAB;
        if (valid) r[me] = vi;
        if (hs.use) r[hs.lds] = vh;
        if (fown) fring[(lev % W) * TI * TJ + tid] = vf;
    };

    // prologue: level 0 into LDS, levels 1..L into registers
    st_lds(0, ld_int(0), ld_halo(0), ld_f(0));
    float rI[L], rH[L], rF[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        const int lev = u + 1;
        rI[u] = lev < nz ? ld_int(lev) : 0.0f;
        rH[u] = lev < nz ? ld_halo(lev) : 0.0f;
        rF[u] = lev < nz ? ld_f(lev) : 0.0f;
    }

    bool changed = false;
    unsigned dirty = 0;  // bit (level % 32): this thread changed that level
    const int pend = nz + t.span + 2;
    for (int pb = 0; pb < pend; pb += L) {
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            // (a) slab p+1 from registers into the ring; (b) prefetch slab p+1+L
            if (p + 1 < nz) st_lds(p + 1, rI[u], rH[u], rF[u]);
            if (p + 1 + L < nz) {
                rI[u] = ld_int(p + 1 + L);
                rH[u] = ld_halo(p + 1 + L);
                rF[u] = ld_f(p + 1 + L);
            }
            lds_barrier();
            // (d) update this thread's node at level pk
            const int pk = p - pi - pj;
            if (valid && pk >= 0 && pk < nz) {
                const float *r0 = ring + (pk % W) * SLAB;
                const float self = r0[me];
                const float xa = fminf(r0[me - C], r0[me + C]);
                const float ya = fminf(r0[me - G::HJ * C], r0[me + G::HJ * C]);
                const float zm = pk > 0 ? ring[((pk - 1) % W) * SLAB + me] : PFAST_INF_F;
                const float zp = pk + 1 < nz ? ring[((pk + 1) % W) * SLAB + me] : PFAST_INF_F;
                const float za = fminf(zm, zp);
                const float fv = fring[(pk % W) * TI * TJ + li * TJ + lj];
                float q;
                if (UNI) q = eik3d_f32_uniform(xa, ya, za, fv);
                else q = eik3d_f32_general(xa, ya, za, fv, a.hx, a.hy, a.hz, a.wx, a.wy, a.wz);
                const long node = (long)kof(pk) * plane + mynode;
                if (q < self && node != src) {
                    ring[(pk % W) * SLAB + me] = q;
                    changed = true;
                    dirty |= 1u << (pk & 31);
                }
            }
            // (e) retire level p - span - 2 (no thread reads it any more)
            const int lr = p - t.span - 2;
            if (valid && lr >= 0 && lr < nz && (dirty >> (lr & 31)) & 1u) {
                F[((size_t)kof(lr) * plane + mynode) * C + c] = ring[(lr % W) * SLAB + me];
                dirty &= ~(1u << (lr & 31));
            }
        }
    }
    const unsigned long long m = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && m) atomicOr(&a.changed[t.chunk], chunk_bits(m, C));
}

// ------------------------------------------------------------------- adjoint
// lambda in a.field, traveltimes in a.tfield (read-only; a negative value marks a
// receiver node of that source, whose lambda is fixed by the boundary condition).
template <int TI, int TJ, int C, int L>
__global__ __launch_bounds__(TI *TJ *C) void adj3d_slab(SlabArgs a, int D, int s1, int s2,
                                                        int s3)
{
    using G = Geo<TI, TJ, C>;
    constexpr int W = G::W, SLAB = G::SLAB;
    __shared__ float lring[W * SLAB];
    __shared__ float tring[W * SLAB];

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const int tid = threadIdx.x;
    const int c = tid % C, lj = (tid / C) % TJ, li = tid / (C * TJ);
    const bool valid = li < t.TIe && lj < t.TJe;
    const int pi = s1 > 0 ? li : t.TIe - 1 - li;
    const int pj = s2 > 0 ? lj : t.TJe - 1 - lj;
    const int nz = a.nz;
    const long plane = (long)a.nrows * a.ncols;
    const size_t cb = (size_t)t.chunk * plane * nz * C;
    float *__restrict__ Lm = a.field + cb;
    const float *__restrict__ Tm = a.tfield + cb;
    const int gi = t.i0 + li, gj = t.j0 + lj;
    const long mynode = (long)gi * a.ncols + gj;
    const bool col_interior = gi >= 1 && gi <= a.nrows - 2 && gj >= 1 && gj <= a.ncols - 2;
    const int me = (((li + 1) * G::HJ) + (lj + 1)) * C + c;
    const HaloSlot<TI, TJ, C> hs = halo_slot<TI, TJ, C>(tid, t, a.nrows, a.ncols);

    auto kof = [&](int lev) { return s3 > 0 ? lev : nz - 1 - lev; };
    auto gaddr = [&](int lev, long node, int cc) { return ((size_t)kof(lev) * plane + node) * C + cc; };
    auto st_lds = [&](int lev, float li_, float lh, float ti_, float th) {
        const int o = (lev % W) * SLAB;
        if (valid) { lring[o + me] = li_; tring[o + me] = ti_; }
        if (hs.use) { lring[o + hs.lds] = lh; tring[o + hs.lds] = th; }
    };
    // Halo nodes outside the domain are never read by an update (only interior
    // nodes are updated), so their value is irrelevant.
    auto ldLi = [&](int lev) { return valid ? Lm[gaddr(lev, mynode, c)] : 0.0f; };
    auto ldTi = [&](int lev) { return valid ? Tm[gaddr(lev, mynode, c)] : 0.0f; };
    auto ldLh = [&](int lev) { return (hs.use && hs.inside) ? Lm[gaddr(lev, hs.gnode, hs.c)] : 0.0f; };
    auto ldTh = [&](int lev) { return (hs.use && hs.inside) ? Tm[gaddr(lev, hs.gnode, hs.c)] : PFAST_INF_F; };

    st_lds(0, ldLi(0), ldLh(0), ldTi(0), ldTh(0));
    float rL[L], rLH[L], rT[L], rTH[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        const int lev = u + 1;
        const bool ok = lev < nz;
        rL[u] = ok ? ldLi(lev) : 0.0f;
        rLH[u] = ok ? ldLh(lev) : 0.0f;
        rT[u] = ok ? ldTi(lev) : 0.0f;
        rTH[u] = ok ? ldTh(lev) : 0.0f;
    }
    const float wx = a.wx, wy = a.wy, wz = a.wz;
    bool changed = false;
    unsigned dirty = 0;
    const int pend = nz + t.span + 2;
    for (int pb = 0; pb < pend; pb += L) {
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            if (p + 1 < nz) st_lds(p + 1, rL[u], rLH[u], rT[u], rTH[u]);
            if (p + 1 + L < nz) {
                rL[u] = ldLi(p + 1 + L);
                rLH[u] = ldLh(p + 1 + L);
                rT[u] = ldTi(p + 1 + L);
                rTH[u] = ldTh(p + 1 + L);
            }
            lds_barrier();
            const int pk = p - pi - pj;
            const int gk = kof(pk);
            if (valid && col_interior && pk >= 0 && pk < nz && gk >= 1 && gk <= nz - 2) {
                const int o = (pk % W) * SLAB + me;
                const float tcr = tring[o];
                if (!signbit(tcr)) {
                    const int om = ((pk - 1) % W) * SLAB + me;  // pk >= 1 here
                    const int op = ((pk + 1) % W) * SLAB + me;
                    const float tc = tcr;
                    const float txm = fabsf(tring[o - C]), txp = fabsf(tring[o + C]);
                    const float tym = fabsf(tring[o - G::HJ * C]), typ = fabsf(tring[o + G::HJ * C]);
                    const float tzm = fabsf(tring[om]), tzp = fabsf(tring[op]);
                    float nw = 0.0f;
                    const float I = PFAST_INF_F;
                    if (!(tc >= I || txm >= I || txp >= I || tym >= I || typ >= I || tzm >= I ||
                          tzp >= I)) {
                        float den = 0.0f;
                        float num = adj_term(tc, txp, lring[o + C], wx, den);
                        num += adj_term(tc, txm, lring[o - C], wx, den);
                        num += adj_term(tc, typ, lring[o + G::HJ * C], wy, den);
                        num += adj_term(tc, tym, lring[o - G::HJ * C], wy, den);
                        num += adj_term(tc, tzp, lring[op], wz, den);
                        num += adj_term(tc, tzm, lring[om], wz, den);
                        nw = fabsf(den) > 1e-9f ? num / den : 0.0f;
                    }
                    const float old = lring[o];
                    if (nw != old) {
                        lring[o] = nw;
                        changed = true;
                        dirty |= 1u << (pk & 31);
                    }
                }
            }
            const int lr = p - t.span - 2;
            if (valid && lr >= 0 && lr < nz && (dirty >> (lr & 31)) & 1u) {
                Lm[gaddr(lr, mynode, c)] = lring[(lr % W) * SLAB + me];
                dirty &= ~(1u << (lr & 31));
            }
        }
    }
    const unsigned long long m = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && m) atomicOr(&a.changed[t.chunk], chunk_bits(m, C));
}

// ------------------------------------------------------------------- helpers
// T = INF everywhere, 0 at each source's node (padding sources have node -1).
template <int C>
__global__ void init_fwd(float *__restrict__ F, const int *__restrict__ srcnode, long N,
                         int nchunk)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;  // (chunk, node)
    if (e >= N * nchunk) return;
    const int chunk = (int)(e / N);
    const long n = e % N;
#pragma unroll
    for (int c = 0; c < C; c++)
        F[e * C + c] = srcnode[chunk * C + c] == n ? 0.0f : PFAST_INF_F;
}

// Mark receiver nodes of each source as fixed (negative traveltime) and form the
// per-source misfit terms.  One thread per picked receiver entry.
__global__ void adj_mark(float *__restrict__ T, const long *__restrict__ pnode,
                         const int *__restrict__ psrc, int npick, long N, int C)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npick) return;
    const int s = psrc[e];
    const size_t a = ((size_t)(s / C) * N + pnode[e]) * C + (s % C);
    T[a] = -fabsf(T[a]);
}

__global__ void adj_unmark(float *__restrict__ T, const long *__restrict__ pnode,
                           const int *__restrict__ psrc, int npick, long N, int C)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= npick) return;
    const int s = psrc[e];
    const size_t a = ((size_t)(s / C) * N + pnode[e]) * C + (s % C);
    T[a] = fabsf(T[a]);
}

// Receiver boundary condition lambda = r / (n . grad T) with the reference's one-sided
// differences (src/T_padjoint3d.h:102-122), FP64 from the FP32 traveltimes.
// One thread per boundary entry (the last picked receiver of each node, receivers
// at the source excluded -- resolved on the host).
__global__ void adj_bc(float *__restrict__ Lm, const float *__restrict__ T,
                       const long *__restrict__ bnode, const int *__restrict__ bsrc,
                       const double *__restrict__ bobs, const double *__restrict__ bnrm, int nb,
                       int nrows, int ncols, int nz, double dx, double dy, double dz, int C)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb) return;
    const int s = bsrc[e];
    const long N = (long)nrows * ncols * nz, plane = (long)nrows * ncols;
    const long n = bnode[e];
    const int k = (int)(n / plane), i = (int)((n % plane) / ncols), j = (int)(n % ncols);
    const float *Ts = T + (size_t)(s / C) * N * C + (s % C);
    auto tv = [&](long m) { return (double)fabsf(Ts[(size_t)m * C]); };
    const double Tn = tv(n);
    const double Tz = k + 1 >= nz ? (Tn - tv(n - plane)) / dz : (tv(n + plane) - Tn) / dz;
    const double Ty = i + 1 >= nrows ? (Tn - tv(n - ncols)) / dy : (tv(n + ncols) - Tn) / dy;
    const double Tx = j + 1 >= ncols ? (Tn - tv(n - 1)) / dx : (tv(n + 1) - Tn) / dx;
    const double r = bobs[e] - Tn;
    const double dot = bnrm[e * 3 + 1] * Tx + bnrm[e * 3 + 0] * Ty + bnrm[e * 3 + 2] * Tz;
    double lam = 0.0;
    if (fabs(r) > 1e-9 && fabs(dot) > 1e-6) lam = r / dot;
    Lm[(size_t)(s / C) * N * C + (size_t)n * C + (s % C)] = (float)lam;
}

// Per-source misfit EE = sum 1/2 (T_obs - T)^2 over picked receivers, Kahan-summed
// in FP64.  One wavefront per source; picks are grouped by source.
__global__ void adj_misfit(const float *__restrict__ T, const long *__restrict__ pnode,
                           const double *__restrict__ pobs, const long *__restrict__ poff,
                           int nsrc, long N, int C, double *__restrict__ EE)
{
    const int s = blockIdx.x;
    if (s >= nsrc) return;
    const float *Ts = T + (size_t)(s / C) * N * C + (s % C);
    double sum = 0.0, comp = 0.0;
    for (long e = poff[s] + threadIdx.x; e < poff[s + 1]; e += blockDim.x) {
        const double r = pobs[e] - (double)fabsf(Ts[(size_t)pnode[e] * C]);
        const double y = r * r / 2.0 - comp;
        const double tt = sum + y;
        comp = (tt - sum) - y;
        sum = tt;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(sum, off);
        const double oc = __shfl_down(comp, off);
        // combine two compensated partial sums
        const double y = o - (comp + oc);
        const double tt = sum + y;
        comp = (tt - sum) - y;
        sum = tt;
    }
    if (threadIdx.x == 0) EE[s] = sum;
}

// grad[n] = -(sum_s lambda_s[n]) / vp[n]^3, Kahan-summed in FP64 over sources.
__global__ void grad_reduce(const float *__restrict__ Lm, const double *__restrict__ vp,
                            double *__restrict__ g, long N, int nchunk, int C, int nsrc)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double sum = 0.0, comp = 0.0;
    for (int ch = 0; ch < nchunk; ch++) {
        const float *p = Lm + ((size_t)ch * N + n) * C;
        for (int c = 0; c < C; c++) {
            if (ch * C + c >= nsrc) break;
            const double y = (double)p[c] - comp;
            const double tt = sum + y;
            comp = (tt - sum) - y;
            sum = tt;
        }
    }
    const double v = vp[n];
    g[n] = -(sum / (v * v * v));
}

__global__ void gather_picks(const float *__restrict__ T, const long *__restrict__ rnode,
                             const int *__restrict__ rsrc, int nrec, long N, int C,
                             double *__restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrec) return;
    const int s = rsrc[e];
    out[e] = (double)fabsf(T[((size_t)(s / C) * N + rnode[e]) * C + (s % C)]);
}

__global__ void gather_field(const float *__restrict__ F, int s, long N, int C,
                             double *__restrict__ out)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    out[n] = (double)fabsf(F[((size_t)(s / C) * N + n) * C + (s % C)]);
}

__global__ void cell_time(const double *__restrict__ vp, float *__restrict__ f, long N,
                          double h, int uniform)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    f[n] = uniform ? (float)(h / vp[n]) : (float)(1.0 / vp[n]);
}

}  // namespace

namespace pfast {

#define PFAST_TILE_TI 8
#define PFAST_TILE_TJ 8
#define PFAST_PREFETCH 2

hipError_t launch_sweep3d(hipStream_t st, const SlabArgs &a, int adjoint, int uniform, int D,
                          int s1, int s2, int s3, int ntiles_diag)
{
    constexpr int TI = PFAST_TILE_TI, TJ = PFAST_TILE_TJ, C = PFAST_CHUNK, L = PFAST_PREFETCH;
    const dim3 grid((unsigned)(ntiles_diag * a.nact_p8)), block(TI * TJ * C);
    if (adjoint)
        hipLaunchKernelGGL((adj3d_slab<TI, TJ, C, L>), grid, block, 0, st, a, D, s1, s2, s3);
    else if (uniform)
        hipLaunchKernelGGL((fwd3d_slab<TI, TJ, C, L, true>), grid, block, 0, st, a, D, s1, s2, s3);
    else
        hipLaunchKernelGGL((fwd3d_slab<TI, TJ, C, L, false>), grid, block, 0, st, a, D, s1, s2,
                           s3);
    return hipGetLastError();
}

int tile_ti() { return PFAST_TILE_TI; }
int tile_tj() { return PFAST_TILE_TJ; }

hipError_t launch_init_fwd(hipStream_t st, float *F, const int *srcnode, long N, int nchunk)
{
    const long tot = N * nchunk;
    hipLaunchKernelGGL(init_fwd<PFAST_CHUNK>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       st, F, srcnode, N, nchunk);
    return hipGetLastError();
}

hipError_t launch_adj_mark(hipStream_t st, float *T, const long *pnode, const int *psrc,
                           int npick, long N, int unmark)
{
    if (npick <= 0) return hipSuccess;
    if (unmark)
        hipLaunchKernelGGL(adj_unmark, dim3((npick + 255) / 256), dim3(256), 0, st, T, pnode, psrc,
                           npick, N, PFAST_CHUNK);
    else
        hipLaunchKernelGGL(adj_mark, dim3((npick + 255) / 256), dim3(256), 0, st, T, pnode, psrc,
                           npick, N, PFAST_CHUNK);
    return hipGetLastError();
}

hipError_t launch_adj_bc(hipStream_t st, float *Lm, const float *T, const long *bnode,
                         const int *bsrc, const double *bobs, const double *bnrm, int nb,
                         int nrows, int ncols, int nz, double dx, double dy, double dz)
{
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_bc, dim3((nb + 255) / 256), dim3(256), 0, st, Lm, T, bnode, bsrc, bobs,
                       bnrm, nb, nrows, ncols, nz, dx, dy, dz, PFAST_CHUNK);
    return hipGetLastError();
}

hipError_t launch_adj_misfit(hipStream_t st, const float *T, const long *pnode,
                             const double *pobs, const long *poff, int nsrc, long N, double *EE)
{
    hipLaunchKernelGGL(adj_misfit, dim3(nsrc), dim3(64), 0, st, T, pnode, pobs, poff, nsrc, N,
                       PFAST_CHUNK, EE);
    return hipGetLastError();
}

hipError_t launch_grad_reduce(hipStream_t st, const float *Lm, const double *vp, double *g,
                              long N, int nchunk, int nsrc)
{
    hipLaunchKernelGGL(grad_reduce, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, Lm, vp,
                       g, N, nchunk, PFAST_CHUNK, nsrc);
    return hipGetLastError();
}

hipError_t launch_gather_picks(hipStream_t st, const float *T, const long *rnode,
                               const int *rsrc, int nrec, long N, double *out)
{
    if (nrec <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_picks, dim3((nrec + 255) / 256), dim3(256), 0, st, T, rnode, rsrc,
                       nrec, N, PFAST_CHUNK, out);
    return hipGetLastError();
}

hipError_t launch_gather_field(hipStream_t st, const float *F, int s, long N, double *out)
{
    hipLaunchKernelGGL(gather_field, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, F, s, N,
                       PFAST_CHUNK, out);
    return hipGetLastError();
}

hipError_t launch_cell_time(hipStream_t st, const double *vp, float *f, long N, double h,
                            int uniform)
{
    hipLaunchKernelGGL(cell_time, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, vp, f, N,
                       h, uniform);
    return hipGetLastError();
}

}  // namespace pfast
