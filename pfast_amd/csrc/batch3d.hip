// batch3d.hip -- batched FP32 3D sweeps (every source of a shot set on one GPU).
//
// Layout in HBM: fields are stored per chunk of C sources, node-major with the C
// sources innermost: F[k][i][j][c] (k = depth, i = row, j = column).  Each chunk
// names its traveltime and lambda fields and its model's velocity-derived cell
// time f[k][i][j] (h/v for uniform spacing, else 1/v) through a ChunkRef, so one
// launch can sweep chunks of several shot sets and models of one grid shape.
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

// Per-thread constants of a tile task shared by the forward and adjoint sweeps.
// Offsets are 32-bit element offsets inside one chunk's field (the chunk base
// pointer is uniform), advanced by one level per hyperplane.
template <int TI, int TJ, int C>
struct Lane {
    int c, li, lj, pipj, me;
    bool valid, hreal, huse;
    int hl;                 // halo LDS position (if huse)
    unsigned oI, oH;        // element offsets of interior / halo node at level 0
    int stepC;              // signed element step per level
    unsigned node;          // i*ncols+j of this column (tile origin when !valid)
};

template <int TI, int TJ, int C>
__device__ __forceinline__ Lane<TI, TJ, C> make_lane(const SlabArgs &a, const TileCtx &t, int s1,
                                                     int s2, int s3)
{
    using G = Geo<TI, TJ, C>;
    Lane<TI, TJ, C> l;
    const int tid = threadIdx.x;
    l.c = tid % C;
    l.lj = (tid / C) % TJ;
    l.li = tid / (C * TJ);
    l.valid = l.li < t.TIe && l.lj < t.TJe;
    l.pipj = (s1 > 0 ? l.li : t.TIe - 1 - l.li) + (s2 > 0 ? l.lj : t.TJe - 1 - l.lj);
    l.me = ((l.li + 1) * G::HJ + (l.lj + 1)) * C + l.c;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    const unsigned pC = plane * C;
    const unsigned org = (unsigned)t.i0 * a.ncols + t.j0;
    l.node = l.valid ? (unsigned)(t.i0 + l.li) * a.ncols + (t.j0 + l.lj) : org;
    const HaloSlot<TI, TJ, C> hs = halo_slot<TI, TJ, C>(tid, t, a.nrows, a.ncols);
    l.huse = hs.use;
    l.hreal = hs.use && hs.inside;
    l.hl = hs.lds;
    const unsigned hnode = l.hreal ? (unsigned)hs.gnode : org;
    const unsigned k0 = s3 > 0 ? 0u : (unsigned)(a.nz - 1);
    l.oI = k0 * pC + l.node * C + l.c;
    l.oH = k0 * pC + hnode * C + hs.c;
    l.stepC = s3 > 0 ? (int)pC : -(int)pC;
    return l;
}

__device__ __forceinline__ int wrap_inc(int s, int W) { return s == W - 1 ? 0 : s + 1; }

// ------------------------------------------------------------------- forward
template <int TI, int TJ, int C, int L, bool UNI>
__global__ __launch_bounds__(TI *TJ *C) void fwd3d_slab(SlabArgs a, int D, int s1, int s2,
                                                        int s3)
{
    using G = Geo<TI, TJ, C>;
    constexpr int W = G::W, SLAB = G::SLAB, FS = TI * TJ, SLOT = SLAB + FS, HJC = G::HJ * C;
    __shared__ float ring[W * SLOT];  // per level: halo-padded slab, then the cell times

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const Lane<TI, TJ, C> l = make_lane<TI, TJ, C>(a, t, s1, s2, s3);
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    float *__restrict__ Fc = a.chunks[t.chunk].field[0];
    const float *__restrict__ slow = a.chunks[t.chunk].slow;
    // cell times: threads 0..FS-1 stream one node each
    const bool fown = tid < FS && (tid / TJ) < t.TIe && (tid % TJ) < t.TJe;
    const unsigned org = (unsigned)t.i0 * a.ncols + t.j0;
    const unsigned fnode = fown ? (unsigned)(t.i0 + tid / TJ) * a.ncols + (t.j0 + tid % TJ) : org;
    const int stepF = s3 > 0 ? (int)plane : -(int)plane;
    unsigned oF = (s3 > 0 ? 0u : (unsigned)(nz - 1)) * plane + fnode;
    unsigned oI = l.oI, oH = l.oH, oS = l.oI;
    const int fme = SLAB + l.li * TJ + l.lj;
    // level (in sweep order) at which this thread's column holds its own source
    int srclev = -1;
    {
        const int sn = a.chunks[t.chunk].srcnode[l.c];
        if (l.valid && sn >= 0 && (unsigned)sn % plane == l.node) {
            const int sk = (int)((unsigned)sn / plane);
            srclev = s3 > 0 ? sk : nz - 1 - sk;
        }
    }
    auto put = [&](int slot, float vi, float vh, float vf) {
        float *r = ring + slot * SLOT;
        if (l.valid) r[l.me] = vi;
        if (l.huse) r[l.hl] = l.hreal ? vh : PFAST_INF_F;
        if (fown) r[SLAB + tid] = vf;
    };

    // prologue: INF sentinel for level -1 (slot W-1), level 0 into slot 0, levels
    // 1..L into registers
    if (l.valid) ring[(W - 1) * SLOT + l.me] = PFAST_INF_F;
    put(0, Fc[oI], Fc[oH], slow[oF]);
    float rI[L], rH[L], rF[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        rI[u] = rH[u] = rF[u] = 0.0f;
        if (u + 1 < nz) {
            oI += l.stepC;
            oH += l.stepC;
            oF += stepF;
            rI[u] = Fc[oI];
            rH[u] = Fc[oH];
            rF[u] = slow[oF];
        }
    }

    bool changed = false;
    unsigned dirty = 0;  // bit (level % 32): this thread changed that level
    const int pend = nz + t.span + 2;
    int vs = l.pipj == 0 ? 0 : W - l.pipj;  // ring slot of this thread's level p - pipj
    int ss = 1;                              // slot of level p+1
    int sr = W - (t.span + 2);               // slot of the retired level p-span-2
    for (int pb = 0; pb < pend; pb += L) {
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            // (a) slab p+1 into the ring (INF sentinel past the last level);
            // (b) prefetch slab p+1+L into registers
            if (p + 1 < nz) put(ss, rI[u], rH[u], rF[u]);
            else if (p + 1 == nz && l.valid) ring[ss * SLOT + l.me] = PFAST_INF_F;
            if (p + 1 + L < nz) {
                oI += l.stepC;
                oH += l.stepC;
                oF += stepF;
                rI[u] = Fc[oI];
                rH[u] = Fc[oH];
                rF[u] = slow[oF];
            }
            lds_barrier();
            // (d) update this thread's node at level pk
            const int pk = p - l.pipj;
            if (l.valid && pk >= 0 && pk < nz) {
                float *r0 = ring + vs * SLOT;
                const float self = r0[l.me];
                const float xa = fminf(r0[l.me - C], r0[l.me + C]);
                const float ya = fminf(r0[l.me - HJC], r0[l.me + HJC]);
                const int vm = vs == 0 ? W - 1 : vs - 1, vp = wrap_inc(vs, W);
                const float za = fminf(ring[vm * SLOT + l.me], ring[vp * SLOT + l.me]);
                const float fv = r0[fme];
                float q;
                if (UNI) q = eik3d_f32_uniform(xa, ya, za, fv);
                else q = eik3d_f32_general(xa, ya, za, fv, a.hx, a.hy, a.hz, a.wx, a.wy, a.wz);
                if (q < self && pk != srclev) {
                    r0[l.me] = q;
                    changed = true;
                    dirty |= 1u << (pk & 31);
                }
            }
            vs = wrap_inc(vs, W);
            // (e) retire level p-span-2 (no thread reads it any more): write back
            // the nodes this thread changed
            const int lr = p - t.span - 2;
            if (lr >= 0 && lr < nz) {
                if (l.valid && ((dirty >> (lr & 31)) & 1u)) {
                    Fc[oS] = ring[sr * SLOT + l.me];
                    dirty &= ~(1u << (lr & 31));
                }
                oS += l.stepC;
            }
            sr = wrap_inc(sr, W);
            ss = wrap_inc(ss, W);
        }
    }
    const unsigned long long m = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && m) atomicOr(&a.changed[t.chunk], chunk_bits(m, C));
}

// ------------------------------------------------------------------- adjoint
// lambda in field[1] of each chunk, traveltimes in field[0] (read-only; a negative value marks a
// receiver node of that source, whose lambda is fixed by the boundary condition).
// Only interior nodes are updated, so no edge sentinels are needed.
template <int TI, int TJ, int C, int L>
__global__ __launch_bounds__(TI *TJ *C) void adj3d_slab(SlabArgs a, int D, int s1, int s2,
                                                        int s3)
{
    using G = Geo<TI, TJ, C>;
    constexpr int W = G::W, SLAB = G::SLAB, HJC = G::HJ * C;
    __shared__ float lring[W * SLAB];
    __shared__ float tring[W * SLAB];

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const Lane<TI, TJ, C> l = make_lane<TI, TJ, C>(a, t, s1, s2, s3);
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    float *__restrict__ Lm = a.chunks[t.chunk].field[1];
    const float *__restrict__ Tm = a.chunks[t.chunk].field[0];
    const int gi = t.i0 + l.li, gj = t.j0 + l.lj;
    const bool col_interior =
        l.valid && gi >= 1 && gi <= a.nrows - 2 && gj >= 1 && gj <= a.ncols - 2;
    unsigned oI = l.oI, oH = l.oH, oS = l.oI;
    auto put = [&](int slot, float vl, float vlh, float vt, float vth) {
        const int o = slot * SLAB;
        if (l.valid) {
            lring[o + l.me] = vl;
            tring[o + l.me] = vt;
        }
        if (l.huse) {
            lring[o + l.hl] = vlh;
            tring[o + l.hl] = l.hreal ? vth : PFAST_INF_F;
        }
    };
    put(0, Lm[oI], Lm[oH], Tm[oI], Tm[oH]);
    float rL[L], rLH[L], rT[L], rTH[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        rL[u] = rLH[u] = rT[u] = rTH[u] = 0.0f;
        if (u + 1 < nz) {
            oI += l.stepC;
            oH += l.stepC;
            rL[u] = Lm[oI];
            rLH[u] = Lm[oH];
            rT[u] = Tm[oI];
            rTH[u] = Tm[oH];
        }
    }
    const float wx = a.wx, wy = a.wy, wz = a.wz;
    bool changed = false;
    unsigned dirty = 0;
    const int pend = nz + t.span + 2;
    int vs = l.pipj == 0 ? 0 : W - l.pipj;
    int ss = 1;
    int sr = W - (t.span + 2);
    for (int pb = 0; pb < pend; pb += L) {
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            if (p + 1 < nz) put(ss, rL[u], rLH[u], rT[u], rTH[u]);
            if (p + 1 + L < nz) {
                oI += l.stepC;
                oH += l.stepC;
                rL[u] = Lm[oI];
                rLH[u] = Lm[oH];
                rT[u] = Tm[oI];
                rTH[u] = Tm[oH];
            }
            lds_barrier();
            const int pk = p - l.pipj;
            if (col_interior && pk >= 1 && pk <= nz - 2) {
                const int o = vs * SLAB + l.me;
                const float tc = tring[o];
                if (!signbit(tc)) {
                    // ring levels pk-1 / pk+1 are depths k-s3 / k+s3: name them by
                    // absolute depth so that every ordering sums the terms in the
                    // same order (another order rounds differently and makes a node
                    // flip by an ulp from one ordering to the next, so the sweeps
                    // would never become stationary)
                    const int lo_ = (vs == 0 ? W - 1 : vs - 1) * SLAB + l.me;
                    const int hi_ = wrap_inc(vs, W) * SLAB + l.me;
                    const int om = s3 > 0 ? lo_ : hi_;  // depth k-1
                    const int op = s3 > 0 ? hi_ : lo_;  // depth k+1
                    const float txm = fabsf(tring[o - C]), txp = fabsf(tring[o + C]);
                    const float tym = fabsf(tring[o - HJC]), typ = fabsf(tring[o + HJC]);
                    const float tzm = fabsf(tring[om]), tzp = fabsf(tring[op]);
                    float nw = 0.0f;
                    const float I = PFAST_INF_F;
                    if (!(tc >= I || txm >= I || txp >= I || tym >= I || typ >= I || tzm >= I ||
                          tzp >= I)) {
                        float den = 0.0f;
                        float num = adj_term(tc, txp, lring[o + C], wx, den);
                        num += adj_term(tc, txm, lring[o - C], wx, den);
                        num += adj_term(tc, typ, lring[o + HJC], wy, den);
                        num += adj_term(tc, tym, lring[o - HJC], wy, den);
                        num += adj_term(tc, tzp, lring[op], wz, den);
                        num += adj_term(tc, tzm, lring[om], wz, den);
                        nw = fabsf(den) > 1e-9f ? num * frcp(den) : 0.0f;
                    }
                    if (nw != lring[o]) {
                        lring[o] = nw;
                        changed = true;
                        dirty |= 1u << (pk & 31);
                    }
                }
            }
            vs = wrap_inc(vs, W);
            const int lr = p - t.span - 2;
            if (lr >= 0 && lr < nz) {
                if (l.valid && ((dirty >> (lr & 31)) & 1u)) {
                    Lm[oS] = lring[sr * SLAB + l.me];
                    dirty &= ~(1u << (lr & 31));
                }
                oS += l.stepC;
            }
            sr = wrap_inc(sr, W);
            ss = wrap_inc(ss, W);
        }
    }
    const unsigned long long m = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && m) atomicOr(&a.changed[t.chunk], chunk_bits(m, C));
}

// ============================================================ column sweeps
// Second schedule for the same sweep ("column window"): thread (li,lj,c) keeps its
// own column of one source in registers -- levels m-1 (updated), m, m+1 (not yet
// updated) of the level m = p - pi - pj it works on at step p -- so depth needs no
// exchange at all.  Only the in-plane neighbours do: at every step each thread
// publishes two values to LDS, its level m-1 (wanted by the neighbour one step
// ahead of it, i.e. upstream in the sweep, which now works on level m) and its
// level m+1 (wanted by the neighbour one step behind, downstream, which now works
// on level m ... from its own point of view: the upstream neighbour of a thread on
// level m is on level m+1 and holds the thread's level-m value as its level "m-1
// updated"; the downstream neighbour is on level m-1 and holds it as "m+1 not yet
// updated").  Tile-edge neighbours come from the neighbouring tiles in HBM, which
// are either finished for this ordering (upstream tiles: earlier launches) or not
// started (downstream: later launches).  The published array is double-buffered by
// step parity, so one barrier per step suffices.  Memory traffic is batched: every
// L steps one vmcnt(0) wait, then the stores of the previous batch's changed nodes
// and the loads of the next batch, so every load has a whole batch to arrive.

// Raw buffer access: 32-bit byte offsets from a uniform base, and an offset at or
// past num_records (a level outside the column, where the unsigned offset wraps)
// reads 0 without touching memory and never faults.
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             0x00020000);
}
__device__ __forceinline__ float bload(rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, unsigned off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}

// vmcnt(0) (expcnt and lgkmcnt left at their maxima).  Issued through the builtin
// so that the compiler's own wait insertion knows nothing is outstanding after it.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Lane Q of this lane's quad (DPP quad_perm [Q,Q,Q,Q]).  The intrinsic is convergent,
// so it stays where every lane of the quad is active.
template <int Q>
__device__ __forceinline__ float quad_bcast(float v)
{
    constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl, 0xF, 0xF, true));
}
__device__ __forceinline__ float quad_bcast(float v, int q)
{
    switch (q & 3) {  // q is a constant after unrolling
    case 0: return quad_bcast<0>(v);
    case 1: return quad_bcast<1>(v);
    case 2: return quad_bcast<2>(v);
    default: return quad_bcast<3>(v);
    }
}


// Per-thread constants of a column-window tile task.  Level lev (sweep order) of
// the thread's column sits at byte offset oI + lev*step.
template <int TI, int TJ, int C>
struct ColLane {
    int c, li, lj, pipj;
    bool valid;
    bool uxm, uxp, uym, uyp;   // that neighbour is a halo node (read from HBM)
    int rxm, rxp, rym, ryp;    // LDS indices of the neighbours' published values
    unsigned oI, oHX, oHY;     // byte offsets of level 0: own column, x halo, y halo
    unsigned oF;               // byte offset of level 0 in the cell-time array
    unsigned step, stepF;      // byte step per level (two's complement when s3 < 0)
    unsigned node;
};

template <int TI, int TJ, int C>
__device__ __forceinline__ ColLane<TI, TJ, C> make_col_lane(const SlabArgs &a, const TileCtx &t,
                                                            int s1, int s2, int s3)
{
    constexpr int NT = TI * TJ * C, INFSLOT = 2 * NT;  // INF slot of each parity block
    ColLane<TI, TJ, C> l;
    const int tid = threadIdx.x;
    l.c = tid % C;
    l.lj = (tid / C) % TJ;
    l.li = tid / (C * TJ);
    l.valid = l.li < t.TIe && l.lj < t.TJe;
    l.pipj = l.valid ? (s1 > 0 ? l.li : t.TIe - 1 - l.li) + (s2 > 0 ? l.lj : t.TJe - 1 - l.lj) : 0;
    const bool exm = l.lj == 0, exp_ = l.lj == t.TJe - 1;
    const bool eym = l.li == 0, eyp = l.li == t.TIe - 1;
    const int gi = t.i0 + l.li, gj = t.j0 + l.lj;
    // a tile-edge neighbour is a halo node inside the domain, or the domain edge
    // (the INF slot); a one-column tile is the last of its row, so x+ is the edge
    l.uxm = l.valid && exm && t.j0 > 0;
    l.uxp = l.valid && !l.uxm && exp_ && t.j0 + t.TJe < a.ncols;
    l.uym = l.valid && eym && t.i0 > 0;
    l.uyp = l.valid && !l.uym && eyp && t.i0 + t.TIe < a.nrows;
    // upstream neighbour (one step ahead): its "m-1" slot [0, NT); downstream: its
    // "m+1" slot [NT, 2NT).  Idle threads read their own slot (unused).
    const bool ok = l.valid;
    l.rxm = !ok ? tid : exm ? INFSLOT : (s2 > 0 ? 0 : NT) + tid - C;
    l.rxp = !ok ? tid : exp_ ? INFSLOT : (s2 > 0 ? NT : 0) + tid + C;
    l.rym = !ok ? tid : eym ? INFSLOT : (s1 > 0 ? 0 : NT) + tid - C * TJ;
    l.ryp = !ok ? tid : eyp ? INFSLOT : (s1 > 0 ? NT : 0) + tid + C * TJ;
    const unsigned plane = (unsigned)a.nrows * a.ncols, pC = plane * C;
    l.node = l.valid ? (unsigned)gi * a.ncols + gj : 0u;
    const int hxj = l.uxm ? t.j0 - 1 : t.j0 + t.TJe;
    const int hyi = l.uym ? t.i0 - 1 : t.i0 + t.TIe;
    const bool hx = l.uxm || l.uxp, hy = l.uym || l.uyp;
    const unsigned k0 = s3 > 0 ? 0u : (unsigned)(a.nz - 1);
    const unsigned NONE = 0x80000000u;  // out of range at every level (never loaded)
    l.oI = 4u * (k0 * pC + l.node * C + l.c);
    l.oHX = hx ? 4u * (k0 * pC + ((unsigned)gi * a.ncols + hxj) * C + l.c) : NONE;
    l.oHY = hy ? 4u * (k0 * pC + ((unsigned)hyi * a.ncols + gj) * C + l.c) : NONE;
    l.oF = 4u * (k0 * plane + l.node);
    l.step = s3 > 0 ? 4u * pC : 0u - 4u * pC;
    l.stepF = s3 > 0 ? 4u * plane : 0u - 4u * plane;
    return l;
}

// Tile index of (chunk, tile) in the skip records, or -1 without skipping.
template <int TI, int TJ>
__device__ __forceinline__ int skip_index(const SlabArgs &a, const TileCtx &t)
{
    return a.tmod ? (t.chunk * a.nti + t.i0 / TI) * a.ntj + t.j0 / TJ : -1;
}

// The sweep of tile ix in this ordering would change nothing (see SlabArgs): its
// clean sweep is newer than every change of the tile and of the neighbour faces
// it reads (left tile's x+ face, right's x-, upper's y+, lower's y-).
template <int TI, int TJ>
__device__ __forceinline__ bool skip_tile(const SlabArgs &a, const TileCtx &t, int ix)
{
    if (ix < 0) return false;
    const int cl = a.tclean[ix * 8 + a.o];
    if (cl <= 0) return false;
    const int ti = t.i0 / TI, tj = t.j0 / TJ;
    const int *M = a.tmod;
    int m = M[ix * 5];
    if (tj > 0) m = max(m, M[(ix - 1) * 5 + 2]);
    if (tj < a.ntj - 1) m = max(m, M[(ix + 1) * 5 + 1]);
    if (ti > 0) m = max(m, M[(ix - a.ntj) * 5 + 4]);
    if (ti < a.nti - 1) m = max(m, M[(ix + a.ntj) * 5 + 3]);
    return m < cl;
}

// Faces of the tile that this thread's column lies on (bit 0 x-, 1 x+, 2 y-, 3 y+).
__device__ __forceinline__ unsigned col_faces(bool valid, int li, int lj, int TIe, int TJe)
{
    if (!valid) return 0u;
    return (lj == 0 ? 1u : 0u) | (lj == TJe - 1 ? 2u : 0u) | (li == 0 ? 4u : 0u) |
           (li == TIe - 1 ? 8u : 0u);
}

// Record the outcome of this tile sweep: `what` = bit 4 (some node changed) | the
// faces on which a node changed, OR-ed over the block.
__device__ __forceinline__ void record_tile(const SlabArgs &a, int ix, unsigned what, int *any)
{
    if (ix < 0) return;
    unsigned w = 0;
#pragma unroll
    for (int b = 0; b < 5; b++)
        if (__ballot((what >> b) & 1u)) w |= 1u << b;
    if ((threadIdx.x & 63) == 0 && w) atomicOr(any, (int)w);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int v = *any;
        if (v) {
            a.tmod[ix * 5] = a.epoch;
#pragma unroll
            for (int f = 0; f < 4; f++)
                if ((v >> f) & 1) a.tmod[ix * 5 + 1 + f] = a.epoch;
            a.tclean[ix * 8 + a.o] = 0;
        } else {
            a.tclean[ix * 8 + a.o] = a.epoch;
        }
    }
}

template <int TI, int TJ, int C, int L, bool UNI>
__global__ __launch_bounds__(TI *TJ *C) void fwd3d_col(SlabArgs a, int D, int s1, int s2, int s3)
{
    constexpr int NT = TI * TJ * C;
    static_assert(L % 2 == 0, "the step parity must be static inside a batch");
    static_assert(C == 4, "a node's sources fill one DPP quad");
    // per step parity: [m-1 | m+1][thread], then an INF slot (domain-edge neighbour)
    constexpr int PB = 2 * NT + 1;
    __shared__ float pub[2 * PB];

    __shared__ int any_change;

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const int six = skip_index<TI, TJ>(a, t);
    if (skip_tile<TI, TJ>(a, t, six)) return;
    const ColLane<TI, TJ, C> l = make_col_lane<TI, TJ, C>(a, t, s1, s2, s3);
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    const unsigned fbytes = 4u * plane * (unsigned)nz * C;
    const rsrc_t RF = make_rsrc(a.chunks[t.chunk].field[0], fbytes);
    // 8x8 tiles read the tile-major cell times: a tile's level is 256 contiguous bytes
    // (two cache lines) instead of eight 32-byte pieces of grid rows
    constexpr bool F8 = TI == 8 && TJ == 8;
    const unsigned tplane = (unsigned)a.nti * a.ntj * 64u;
    const rsrc_t RS = F8 ? make_rsrc(a.chunks[t.chunk].slow8, 4u * tplane * (unsigned)nz)
                         : make_rsrc(a.chunks[t.chunk].slow, 4u * plane * (unsigned)nz);
    const float INF = PFAST_INF_F;
    if (tid == 0) {
        pub[2 * NT] = pub[PB + 2 * NT] = INF;
        any_change = 0;
    }
    // steps p at which this thread updates its node: pipj <= p < pipj + nz, except
    // the level of its own source
    const int plo = l.valid ? l.pipj : 0x3fffffff, phi = l.pipj + nz;
    int psrc = -1;
    {
        const int sn = a.chunks[t.chunk].srcnode[l.c];
        if (l.valid && sn >= 0 && (unsigned)sn % plane == l.node) {
            const int sk = (int)((unsigned)sn / plane);
            psrc = l.pipj + (s3 > 0 ? sk : nz - 1 - sk);
        }
    }
    const unsigned st = l.step;
    unsigned oF = l.oF, stF = l.stepF;
    if (F8) {
        const unsigned k0 = s3 > 0 ? 0u : (unsigned)(nz - 1);
        const unsigned tix = (unsigned)(t.i0 / TI) * a.ntj + (unsigned)(t.j0 / TJ);
        oF = 4u * (k0 * tplane + tix * 64u + (unsigned)(l.li * TJ + l.lj));
        stF = s3 > 0 ? 4u * tplane : 0u - 4u * tplane;
    }
    // Vector-memory instructions are what the sweep's throughput is bound by (each
    // costs the address unit the same whatever its lanes hold), so: halo loads run
    // only in the lanes that use them (a wave without such a lane skips them), and
    // the cell time, one value per node, is loaded for four levels at once, lane c
    // of the node's four lanes taking level 4g + c; step u broadcasts it by DPP.
    constexpr int FQ = (L + 3) / 4;
    const bool hxuse = l.uxm || l.uxp, hyuse = l.uym || l.uyp;
    const unsigned oFc = oF + (unsigned)l.c * stF;
    // queue of one batch: own column two levels ahead (m+2, INF outside the column),
    // halos at level m itself, cell time of levels m .. m+4*FQ-1
    auto load_batch = [&](int pb, float *ow, float *hx, float *hy, float *fv) {
        const int mb = pb - l.pipj;
        const unsigned dm = (unsigned)mb * st, dmf = (unsigned)mb * stF;
#pragma unroll
        for (int u = 0; u < L; u++) {
            const float o = bload(RF, l.oI + dm + (unsigned)(u + 2) * st);
            ow[u] = (unsigned)(mb + u + 2) < (unsigned)nz ? o : INF;
        }
        if (hxuse) {
#pragma unroll
            for (int u = 0; u < L; u++) hx[u] = bload(RF, l.oHX + dm + (unsigned)u * st);
        } else {
#pragma unroll
            for (int u = 0; u < L; u++) hx[u] = INF;
        }
        if (hyuse) {
#pragma unroll
            for (int u = 0; u < L; u++) hy[u] = bload(RF, l.oHY + dm + (unsigned)u * st);
        } else {
#pragma unroll
            for (int u = 0; u < L; u++) hy[u] = INF;
        }
#pragma unroll
        for (int g = 0; g < FQ; g++) fv[g] = bload(RS, oFc + dmf + (unsigned)(4 * g) * stF);
    };
    // window at step 0 (level m0 = -pipj); level m0-1 is outside the column
    const int m0 = -l.pipj;
    float vprev = INF;
    float vcur = l.valid && m0 == 0 ? bload(RF, l.oI) : INF;
    float vnext = l.valid && m0 >= -1 ? bload(RF, l.oI + (unsigned)(m0 + 1) * st) : INF;
    float qa[4][L], qb[4][L];
    load_batch(0, qa[0], qa[1], qa[2], qa[3]);
    __syncthreads();  // INF slot visible

    bool changed = false;
    float sval[L];
    bool chg[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        sval[u] = 0.0f;
        chg[u] = false;
    }
    // one batch of L steps from queue q (loaded a batch earlier); issues the loads
    // of the following batch into nq first
    auto batch = [&](int pb, float (*q)[L], float (*nq)[L]) {
        wait_vm_all();
        // an opaque copy of the batch start for the addresses: otherwise the compiler
        // keeps every per-step address of the queues in a register of its own
        int pbo = pb;
        asm volatile("" : "+s"(pbo));
        {  // changed nodes of the previous batch (steps pb-L .. pb-1)
            const unsigned d = (unsigned)(pbo - L - l.pipj) * st;
#pragma unroll
            for (int u = 0; u < L; u++)
                if (chg[u]) bstore(RF, l.oI + d + (unsigned)u * st, sval[u]);
        }
        load_batch(pbo + L, nq[0], nq[1], nq[2], nq[3]);
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            float *P = pub + (u & 1) * PB;
            P[tid] = vprev;
            P[NT + tid] = vnext;
            lds_barrier();
            const float pxm = P[l.rxm], pxp = P[l.rxp], pym = P[l.rym], pyp = P[l.ryp];
            const float xa = fminf(l.uxm ? q[1][u] : pxm, l.uxp ? q[1][u] : pxp);
            const float ya = fminf(l.uym ? q[2][u] : pym, l.uyp ? q[2][u] : pyp);
            const float za = fminf(vprev, vnext);
            const float fval = quad_bcast(q[3][u >> 2], u);
            float nv = vcur;
            bool c = false;
            if (p >= plo && p < phi && p != psrc) {
                float qv;
                if (UNI) qv = eik3d_f32_uniform(xa, ya, za, fval);
                else qv = eik3d_f32_general(xa, ya, za, fval, a.hx, a.hy, a.hz, a.wx, a.wy, a.wz);
                c = qv < vcur;
                nv = c ? qv : vcur;
            }
            chg[u] = c;
            changed |= c;
            sval[u] = nv;
            vprev = nv;
            vcur = vnext;
            vnext = q[0][u];
        }
    };
    const int pend = nz + t.span;
    for (int pb = 0; pb < pend; pb += 2 * L) {
        batch(pb, qa, qb);
        batch(pb + L, qb, qa);
    }
    {
        const int pb = (pend + 2 * L - 1) / (2 * L) * (2 * L);
        const unsigned d = (unsigned)(pb - L - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++)
            if (chg[u]) bstore(RF, l.oI + d + (unsigned)u * st, sval[u]);
    }
    const unsigned long long bm = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && bm) atomicOr(&a.changed[t.chunk], chunk_bits(bm, C));
    record_tile(a, six, changed ? 16u | col_faces(l.valid, l.li, l.lj, t.TIe, t.TJe) : 0u,
                &any_change);
}

// Adjoint, column window: lambda and traveltime travel together as float2 (the
// traveltimes are read-only; a negative value marks a fixed receiver node).  Only
// interior nodes are updated, and every neighbour they read is inside the domain.
template <int TI, int TJ, int C, int L, bool S3POS>
__global__ __launch_bounds__(TI *TJ *C) void adj3d_col(SlabArgs a, int D, int s1, int s2)
{
    constexpr int NT = TI * TJ * C;
    constexpr int s3 = S3POS ? 1 : -1;
    static_assert(L % 2 == 0, "the step parity must be static inside a batch");
    constexpr int PB = 2 * NT + 1;  // per step parity: [m-1 | m+1][thread], INF slot
    __shared__ float2 pub[2 * PB];

    __shared__ int any_change;

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const int six = skip_index<TI, TJ>(a, t);
    if (skip_tile<TI, TJ>(a, t, six)) return;
    const ColLane<TI, TJ, C> l = make_col_lane<TI, TJ, C>(a, t, s1, s2, s3);
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    const unsigned fbytes = 4u * plane * (unsigned)nz * C;
    const rsrc_t RL = make_rsrc(a.chunks[t.chunk].field[1], fbytes);
    const rsrc_t RT = make_rsrc(a.chunks[t.chunk].field[0], fbytes);
    const int gi = t.i0 + l.li, gj = t.j0 + l.lj;
    const bool col_interior =
        l.valid && gi >= 1 && gi <= a.nrows - 2 && gj >= 1 && gj <= a.ncols - 2;
    // interior levels 1 .. nz-2 of interior columns
    const int plo = col_interior ? l.pipj + 1 : 0x3fffffff, phi = l.pipj + nz - 1;
    const unsigned st = l.step;
    const float wx = a.wx, wy = a.wy, wz = a.wz;
    if (tid == 0) {
        pub[2 * NT] = pub[PB + 2 * NT] = make_float2(0.0f, PFAST_INF_F);
        any_change = 0;
    }

    // halo loads only in the lanes that use them (vector-memory instructions bound
    // the sweep; a wave without a halo lane skips them)
    const bool hxuse = l.uxm || l.uxp, hyuse = l.uym || l.uyp;
    auto ld2 = [&](unsigned o) { return make_float2(bload(RL, o), bload(RT, o)); };
    auto load_batch = [&](int pb, float2 *ow, float2 *hx, float2 *hy) {
        const unsigned dm = (unsigned)(pb - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++) ow[u] = ld2(l.oI + dm + (unsigned)(u + 2) * st);
        if (hxuse) {
#pragma unroll
            for (int u = 0; u < L; u++) hx[u] = ld2(l.oHX + dm + (unsigned)u * st);
        } else {
#pragma unroll
            for (int u = 0; u < L; u++) hx[u] = make_float2(0.0f, 0.0f);
        }
        if (hyuse) {
#pragma unroll
            for (int u = 0; u < L; u++) hy[u] = ld2(l.oHY + dm + (unsigned)u * st);
        } else {
#pragma unroll
            for (int u = 0; u < L; u++) hy[u] = make_float2(0.0f, 0.0f);
        }
    };
    const int m0 = -l.pipj;
    float2 vprev = make_float2(0.0f, 0.0f);
    float2 vcur = l.valid && m0 == 0 ? ld2(l.oI) : make_float2(0.0f, 0.0f);
    float2 vnext = l.valid && m0 >= -1 ? ld2(l.oI + (unsigned)(m0 + 1) * st) : make_float2(0.0f, 0.0f);
    float2 qa[3][L], qb[3][L];
    load_batch(0, qa[0], qa[1], qa[2]);
    __syncthreads();

    bool changed = false;
    float sval[L];
    bool chg[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        sval[u] = 0.0f;
        chg[u] = false;
    }
    auto batch = [&](int pb, float2 (*q)[L], float2 (*nq)[L]) {
        wait_vm_all();
        // an opaque copy of the batch start for the addresses: otherwise the compiler
        // keeps every per-step address of the queues in a register of its own
        int pbo = pb;
        asm volatile("" : "+s"(pbo));
        {
            const unsigned d = (unsigned)(pbo - L - l.pipj) * st;
#pragma unroll
            for (int u = 0; u < L; u++)
                if (chg[u]) bstore(RL, l.oI + d + (unsigned)u * st, sval[u]);
        }
        load_batch(pbo + L, nq[0], nq[1], nq[2]);
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            float2 *P = pub + (u & 1) * PB;
            P[tid] = vprev;
            P[NT + tid] = vnext;
            lds_barrier();
            // branch-free: the (lambda, T) pairs of the neighbours come in one LDS read
            // each, the update is computed in every lane and selected afterwards
            const float2 pxm = P[l.rxm], pxp = P[l.rxp], pym = P[l.rym], pyp = P[l.ryp];
            const float2 xm = l.uxm ? q[1][u] : pxm;
            const float2 xp = l.uxp ? q[1][u] : pxp;
            const float2 ym = l.uym ? q[2][u] : pym;
            const float2 yp = l.uyp ? q[2][u] : pyp;
            // depth k-1 / k+1 by absolute depth (levels m-1 / m+1 are depths
            // k-s3 / k+s3): every ordering sums the terms in one order, so a
            // stationary node stays bitwise stationary
            const float2 zm = S3POS ? vprev : vnext;
            const float2 zp = S3POS ? vnext : vprev;
            const float tc = vcur.y;
            const float txm = fabsf(xm.y), txp = fabsf(xp.y), tym = fabsf(ym.y),
                        typ = fabsf(yp.y), tzm = fabsf(zm.y), tzp = fabsf(zp.y);
            const float I = PFAST_INF_F;
            const bool anyinf = tc >= I || txm >= I || txp >= I || tym >= I || typ >= I ||
                                tzm >= I || tzp >= I;
            float den = 0.0f;
            float num = adj_term(tc, txp, xp.x, wx, den);
            num += adj_term(tc, txm, xm.x, wx, den);
            num += adj_term(tc, typ, yp.x, wy, den);
            num += adj_term(tc, tym, ym.x, wy, den);
            num += adj_term(tc, tzp, zp.x, wz, den);
            num += adj_term(tc, tzm, zm.x, wz, den);
            const float nw = !anyinf && fabsf(den) > 1e-9f ? num * frcp(den) : 0.0f;
            const bool act = p >= plo && p < phi && !signbit(vcur.y);
            const bool c = act && nw != vcur.x;
            const float nl = act ? nw : vcur.x;
            chg[u] = c;
            changed |= c;
            sval[u] = nl;
            vprev = make_float2(nl, vcur.y);
            vcur = vnext;
            vnext = q[0][u];
        }
    };
    const int pend = nz + t.span;
    for (int pb = 0; pb < pend; pb += 2 * L) {
        batch(pb, qa, qb);
        batch(pb + L, qb, qa);
    }
    {
        const int pb = (pend + 2 * L - 1) / (2 * L) * (2 * L);
        const unsigned d = (unsigned)(pb - L - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++)
            if (chg[u]) bstore(RL, l.oI + d + (unsigned)u * st, sval[u]);
    }
    const unsigned long long bm = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && bm) atomicOr(&a.changed[t.chunk], chunk_bits(bm, C));
    record_tile(a, six, changed ? 16u | col_faces(l.valid, l.li, l.lj, t.TIe, t.TJe) : 0u,
                &any_change);
}

// ============================================================ vector-lane sweeps
// Fourth schedule ("vector lanes"): the column-window sweep with one lane per
// column and the chunk's C = 4 sources of a node (its 16-byte record) in that lane.
// Loads, the LDS exchange and stores move whole records; the four updates of a node
// run as independent instruction chains, and what depends only on the node
// (addresses, masks, neighbour selection, the cell time) is done once for the
// four.  An 8x8 tile is one wave.  Same steps and the same per-node arithmetic as
// fwd3d_col / adj3d_col, so the same results bit for bit; a record is stored when
// any of its sources changed (the others rewrite their own value).

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 bload4(rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ void bstore4(rsrc_t r, unsigned off, f32x4 v)
{
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, off, 0, 0);
}

// Column lane of a vector tile: indices and LDS slots as a one-source column lane,
// field offsets in 16-byte node records (cell-time offsets stay 4-byte).
template <int TI, int TJ>
__device__ __forceinline__ ColLane<TI, TJ, 1> make_vec_lane(const SlabArgs &a, const TileCtx &t,
                                                            int s1, int s2, int s3)
{
    ColLane<TI, TJ, 1> l = make_col_lane<TI, TJ, 1>(a, t, s1, s2, s3);
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    const unsigned k0 = s3 > 0 ? 0u : (unsigned)(a.nz - 1);
    const unsigned NONE = 0x80000000u;  // out of range at every level (never loaded)
    const int gi = t.i0 + l.li, gj = t.j0 + l.lj;
    const int hxj = l.uxm ? t.j0 - 1 : t.j0 + t.TJe;
    const int hyi = l.uym ? t.i0 - 1 : t.i0 + t.TIe;
    l.oI = 16u * (k0 * plane + l.node);
    l.oHX = l.uxm || l.uxp ? 16u * (k0 * plane + (unsigned)gi * a.ncols + hxj) : NONE;
    l.oHY = l.uym || l.uyp ? 16u * (k0 * plane + (unsigned)hyi * a.ncols + gj) : NONE;
    l.step = s3 > 0 ? 16u * plane : 0u - 16u * plane;
    return l;
}

// per-chunk changed bits of a wave: bit c iff some lane changed source c
__device__ __forceinline__ void vote_changed(const SlabArgs &a, int chunk, unsigned changed)
{
    unsigned bits = 0;
#pragma unroll
    for (int c = 0; c < PFAST_CHUNK; c++)
        if (__ballot((changed >> c) & 1u)) bits |= 1u << c;
    if ((threadIdx.x & 63) == 0 && bits) atomicOr(&a.changed[chunk], bits);
}

template <int TI, int TJ, int L, bool UNI>
__global__ __launch_bounds__(TI *TJ) void fwd3d_vec(SlabArgs a, int D, int s1, int s2, int s3)
{
    constexpr int NT = TI * TJ, C = PFAST_CHUNK;
    static_assert(C == 4, "one 16-byte record per lane");
    static_assert(L % 2 == 0, "the step parity must be static inside a batch");
    constexpr int PB = 2 * NT + 1;  // per step parity: [m-1 | m+1][lane], INF slot
    __shared__ f32x4 pub[2 * PB];

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const ColLane<TI, TJ, 1> l = make_vec_lane<TI, TJ>(a, t, s1, s2, s3);
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    const rsrc_t RF = make_rsrc(a.chunks[t.chunk].field[0], 16u * plane * (unsigned)nz);
    const rsrc_t RS = make_rsrc(a.chunks[t.chunk].slow, 4u * plane * (unsigned)nz);
    const float INF = PFAST_INF_F;
    const f32x4 INF4 = {INF, INF, INF, INF};
    if (tid == 0) pub[2 * NT] = pub[PB + 2 * NT] = INF4;
    const int plo = l.valid ? l.pipj : 0x3fffffff, phi = l.pipj + nz;
    int psrc[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int sn = a.chunks[t.chunk].srcnode[c];
        psrc[c] = -1;
        if (l.valid && sn >= 0 && (unsigned)sn % plane == l.node) {
            const int sk = (int)((unsigned)sn / plane);
            psrc[c] = l.pipj + (s3 > 0 ? sk : nz - 1 - sk);
        }
    }
    const unsigned st = l.step, stF = l.stepF;
    auto load_batch = [&](int pb, f32x4 *ow, f32x4 *hx, f32x4 *hy, float *fv) {
        const int mb = pb - l.pipj;
        const unsigned dm = (unsigned)mb * st, dmf = (unsigned)mb * stF;
#pragma unroll
        for (int u = 0; u < L; u++) {
            const f32x4 o = bload4(RF, l.oI + dm + (unsigned)(u + 2) * st);
            ow[u] = (unsigned)(mb + u + 2) < (unsigned)nz ? o : INF4;
            hx[u] = bload4(RF, l.oHX + dm + (unsigned)u * st);
            hy[u] = bload4(RF, l.oHY + dm + (unsigned)u * st);
            fv[u] = bload(RS, l.oF + dmf + (unsigned)u * stF);
        }
    };
    const int m0 = -l.pipj;
    f32x4 vprev = INF4;
    f32x4 vcur = l.valid && m0 == 0 ? bload4(RF, l.oI) : INF4;
    f32x4 vnext = l.valid && m0 >= -1 ? bload4(RF, l.oI + (unsigned)(m0 + 1) * st) : INF4;
    f32x4 qa[3][L], qb[3][L];
    float fa[L], fb[L];
    load_batch(0, qa[0], qa[1], qa[2], fa);
    __syncthreads();  // INF slot visible

    unsigned changed = 0;
    f32x4 sval[L];
    bool chg[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        sval[u] = INF4;
        chg[u] = false;
    }
    auto batch = [&](int pb, f32x4 (*q)[L], const float *qf, f32x4 (*nq)[L], float *nqf) {
        wait_vm_all();
        {  // records changed in the previous batch (steps pb-L .. pb-1)
            const unsigned d = (unsigned)(pb - L - l.pipj) * st;
#pragma unroll
            for (int u = 0; u < L; u++)
                if (chg[u]) bstore4(RF, l.oI + d + (unsigned)u * st, sval[u]);
        }
        load_batch(pb + L, nq[0], nq[1], nq[2], nqf);
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            f32x4 *P = pub + (u & 1) * PB;
            P[tid] = vprev;
            P[NT + tid] = vnext;
            lds_barrier();
            const f32x4 pxm = P[l.rxm], pxp = P[l.rxp], pym = P[l.rym], pyp = P[l.ryp];
            const f32x4 xm = l.uxm ? q[1][u] : pxm, xp = l.uxp ? q[1][u] : pxp;
            const f32x4 ym = l.uym ? q[2][u] : pym, yp = l.uyp ? q[2][u] : pyp;
            const bool live = p >= plo && p < phi;
            f32x4 nv = vcur;
            unsigned cm = 0;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const float xa = fminf(xm[c], xp[c]), ya = fminf(ym[c], yp[c]);
                const float za = fminf(vprev[c], vnext[c]);
                float qv;
                if (UNI) qv = eik3d_f32_uniform(xa, ya, za, qf[u]);
                else qv = eik3d_f32_general(xa, ya, za, qf[u], a.hx, a.hy, a.hz, a.wx, a.wy, a.wz);
                const bool cc = live && p != psrc[c] && qv < vcur[c];
                nv[c] = cc ? qv : vcur[c];
                cm |= cc ? 1u << c : 0u;
            }
            chg[u] = cm != 0;
            changed |= cm;
            sval[u] = nv;
            vprev = nv;
            vcur = vnext;
            vnext = q[0][u];
        }
    };
    const int pend = nz + t.span;
    for (int pb = 0; pb < pend; pb += 2 * L) {
        batch(pb, qa, fa, qb, fb);
        batch(pb + L, qb, fb, qa, fa);
    }
    {
        const int pb = (pend + 2 * L - 1) / (2 * L) * (2 * L);
        const unsigned d = (unsigned)(pb - L - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++)
            if (chg[u]) bstore4(RF, l.oI + d + (unsigned)u * st, sval[u]);
    }
    vote_changed(a, t.chunk, changed);
}

template <int TI, int TJ, int L, bool S3POS>
__global__ __launch_bounds__(TI *TJ) void adj3d_vec(SlabArgs a, int D, int s1, int s2)
{
    constexpr int NT = TI * TJ, C = PFAST_CHUNK;
    constexpr int s3 = S3POS ? 1 : -1;
    static_assert(C == 4, "one 16-byte record per lane");
    static_assert(L % 2 == 0, "the step parity must be static inside a batch");
    constexpr int PB = 2 * NT + 1;
    __shared__ f32x4 pubL[2 * PB];  // lambda records
    __shared__ f32x4 pubT[2 * PB];  // traveltime records (negative: receiver node)

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const ColLane<TI, TJ, 1> l = make_vec_lane<TI, TJ>(a, t, s1, s2, s3);
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    const unsigned fbytes = 16u * plane * (unsigned)nz;
    const rsrc_t RL = make_rsrc(a.chunks[t.chunk].field[1], fbytes);
    const rsrc_t RT = make_rsrc(a.chunks[t.chunk].field[0], fbytes);
    const int gi = t.i0 + l.li, gj = t.j0 + l.lj;
    const bool col_interior =
        l.valid && gi >= 1 && gi <= a.nrows - 2 && gj >= 1 && gj <= a.ncols - 2;
    const int plo = col_interior ? l.pipj + 1 : 0x3fffffff, phi = l.pipj + nz - 1;
    const unsigned st = l.step;
    const float wx = a.wx, wy = a.wy, wz = a.wz;
    const float I = PFAST_INF_F;
    const f32x4 Z4 = {0.0f, 0.0f, 0.0f, 0.0f}, INF4 = {I, I, I, I};
    if (tid == 0) {
        pubL[2 * NT] = pubL[PB + 2 * NT] = Z4;
        pubT[2 * NT] = pubT[PB + 2 * NT] = INF4;
    }
    auto load_batch = [&](int pb, f32x4 (*q)[L]) {
        const unsigned dm = (unsigned)(pb - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++) {
            const unsigned o = l.oI + dm + (unsigned)(u + 2) * st;
            const unsigned ox = l.oHX + dm + (unsigned)u * st, oy = l.oHY + dm + (unsigned)u * st;
            q[0][u] = bload4(RL, o);
            q[1][u] = bload4(RT, o);
            q[2][u] = bload4(RL, ox);
            q[3][u] = bload4(RT, ox);
            q[4][u] = bload4(RL, oy);
            q[5][u] = bload4(RT, oy);
        }
    };
    const int m0 = -l.pipj;
    f32x4 pL = Z4, pT = Z4;  // level m-1
    f32x4 cL = l.valid && m0 == 0 ? bload4(RL, l.oI) : Z4;
    f32x4 cT = l.valid && m0 == 0 ? bload4(RT, l.oI) : Z4;
    const unsigned on = l.oI + (unsigned)(m0 + 1) * st;
    f32x4 nL = l.valid && m0 >= -1 ? bload4(RL, on) : Z4;
    f32x4 nT = l.valid && m0 >= -1 ? bload4(RT, on) : Z4;
    f32x4 qa[6][L], qb[6][L];
    load_batch(0, qa);
    __syncthreads();

    unsigned changed = 0;
    f32x4 sval[L];
    bool chg[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        sval[u] = Z4;
        chg[u] = false;
    }
    auto batch = [&](int pb, f32x4 (*q)[L], f32x4 (*nq)[L]) {
        wait_vm_all();
        {
            const unsigned d = (unsigned)(pb - L - l.pipj) * st;
#pragma unroll
            for (int u = 0; u < L; u++)
                if (chg[u]) bstore4(RL, l.oI + d + (unsigned)u * st, sval[u]);
        }
        load_batch(pb + L, nq);
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            f32x4 *PL = pubL + (u & 1) * PB, *PT = pubT + (u & 1) * PB;
            PL[tid] = pL;
            PT[tid] = pT;
            PL[NT + tid] = nL;
            PT[NT + tid] = nT;
            lds_barrier();
            const f32x4 xmL = l.uxm ? q[2][u] : PL[l.rxm], xmT = l.uxm ? q[3][u] : PT[l.rxm];
            const f32x4 xpL = l.uxp ? q[2][u] : PL[l.rxp], xpT = l.uxp ? q[3][u] : PT[l.rxp];
            const f32x4 ymL = l.uym ? q[4][u] : PL[l.rym], ymT = l.uym ? q[5][u] : PT[l.rym];
            const f32x4 ypL = l.uyp ? q[4][u] : PL[l.ryp], ypT = l.uyp ? q[5][u] : PT[l.ryp];
            // depth k-1 / k+1 by absolute depth, as in adj3d_col
            const f32x4 zmL = S3POS ? pL : nL, zmT = S3POS ? pT : nT;
            const f32x4 zpL = S3POS ? nL : pL, zpT = S3POS ? nT : pT;
            const bool live = p >= plo && p < phi;
            f32x4 nl = cL;
            unsigned cm = 0;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const float tc = cT[c];
                const float txm = fabsf(xmT[c]), txp = fabsf(xpT[c]), tym = fabsf(ymT[c]),
                            typ = fabsf(ypT[c]), tzm = fabsf(zmT[c]), tzp = fabsf(zpT[c]);
                const bool anyinf = tc >= I || txm >= I || txp >= I || tym >= I || typ >= I ||
                                    tzm >= I || tzp >= I;
                float den = 0.0f;
                float num = adj_term(tc, txp, xpL[c], wx, den);
                num += adj_term(tc, txm, xmL[c], wx, den);
                num += adj_term(tc, typ, ypL[c], wy, den);
                num += adj_term(tc, tym, ymL[c], wy, den);
                num += adj_term(tc, tzp, zpL[c], wz, den);
                num += adj_term(tc, tzm, zmL[c], wz, den);
                const float nw = !anyinf && fabsf(den) > 1e-9f ? num * frcp(den) : 0.0f;
                const bool act = live && !signbit(tc);
                const bool cc = act && nw != cL[c];
                nl[c] = act ? nw : cL[c];
                cm |= cc ? 1u << c : 0u;
            }
            chg[u] = cm != 0;
            changed |= cm;
            sval[u] = nl;
            pL = nl;
            pT = cT;
            cL = nL;
            cT = nT;
            nL = q[0][u];
            nT = q[1][u];
        }
    };
    const int pend = nz + t.span;
    for (int pb = 0; pb < pend; pb += 2 * L) {
        batch(pb, qa, qb);
        batch(pb + L, qb, qa);
    }
    {
        const int pb = (pend + 2 * L - 1) / (2 * L) * (2 * L);
        const unsigned d = (unsigned)(pb - L - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++)
            if (chg[u]) bstore4(RL, l.oI + d + (unsigned)u * st, sval[u]);
    }
    vote_changed(a, t.chunk, changed);
}

// ============================================================ loader-wave sweeps
// Third schedule ("loader wave"): the column-window sweep with its HBM reads moved
// out of the compute waves.  One extra wave per block only issues loads: at step p
// it streams the data every compute thread needs at step p + A (own column two
// levels ahead, tile-edge halos, cell time) from HBM straight into a step-indexed
// LDS ring (buffer -> LDS DMA; one dwordx4 lane carries the C = 4 sources of a
// column), then waits until step p + 1's data has landed and joins the step
// barrier.  Because that wave issues loads only, vmcnt counts them in order and
// the wait is exact, so A steps of HBM latency are hidden.  The compute waves read
// everything from LDS and store changed nodes as they go (never waiting on them).
typedef __attribute__((address_space(3))) void lds_void_t;

__device__ __forceinline__ void dma16(rsrc_t r, void *lds, unsigned off)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void_t *)lds, 16, off, 0, 0, 0);
}
__device__ __forceinline__ void dma4(rsrc_t r, void *lds, unsigned off)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void_t *)lds, 4, off, 0, 0, 0);
}
// s_waitcnt vmcnt(n), n <= 63 (expcnt / lgkmcnt at their maxima)
template <int N>
__device__ __forceinline__ void wait_vm()
{
    static_assert(N >= 0 && N <= 63, "vmcnt is 6 bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
}

// The consumer (compute thread) of a tile-edge halo entry h of a step slot, and
// the halo node it needs.  Entries 0 .. 2TI-1: x halos (row h/2, side h%2: 0 = x-,
// 1 = x+); then 2TJ y halos (column, side 0 = y-, 1 = y+).
template <int TI, int TJ>
__device__ __forceinline__ bool halo_entry(int h, const TileCtx &t, int nrows, int ncols, int &cli,
                                           int &clj, unsigned &hnode)
{
    int gi, gj;
    if (h < 2 * TI) {
        cli = h / 2;
        const int side = h % 2;
        clj = side ? t.TJe - 1 : 0;
        gi = t.i0 + cli;
        gj = side ? t.j0 + t.TJe : t.j0 - 1;
        if (cli >= t.TIe) return false;
    } else {
        clj = (h - 2 * TI) / 2;
        const int side = (h - 2 * TI) % 2;
        cli = side ? t.TIe - 1 : 0;
        gi = side ? t.i0 + t.TIe : t.i0 - 1;
        gj = t.j0 + clj;
        if (clj >= t.TJe) return false;
    }
    if (gi < 0 || gi >= nrows || gj < 0 || gj >= ncols) return false;
    hnode = (unsigned)gi * ncols + gj;
    return true;
}

template <int TI, int TJ>
__device__ __forceinline__ int pipj_of(int li, int lj, const TileCtx &t, int s1, int s2)
{
    return (s1 > 0 ? li : t.TIe - 1 - li) + (s2 > 0 ? lj : t.TJe - 1 - lj);
}

template <int TI, int TJ, int C, int A, bool UNI>
__global__ __launch_bounds__(TI *TJ *C + 64) void fwd3d_dma(SlabArgs a, int D, int s1, int s2,
                                                            int s3)
{
    static_assert(C == 4, "one dwordx4 lane per column");
    constexpr int NT = TI * TJ * C, NCOL = TI * TJ, NH = 2 * (TI + TJ);
    static_assert(NCOL <= 64 && NH <= 64, "one loader lane per column / halo node");
    constexpr int R = A + 1;  // ring slots: step p + A reuses the slot of step p - 1
    constexpr int S_HALO = NT, S_F = NT + NH * C, SLOT = S_F + NCOL;  // floats per slot
    static_assert(SLOT % 4 == 0, "16-byte aligned slots");
    constexpr int PB = 2 * NT + 1;
    __shared__ float pub[2 * PB];
    __shared__ __attribute__((aligned(16))) float ring[R * SLOT];

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols, pC = plane * C;
    const rsrc_t RF = make_rsrc(a.chunks[t.chunk].field[0], 4u * pC * nz);
    const rsrc_t RS = make_rsrc(a.chunks[t.chunk].slow, 4u * plane * nz);
    const unsigned k0 = s3 > 0 ? 0u : (unsigned)(nz - 1);
    const unsigned st = s3 > 0 ? 4u * pC : 0u - 4u * pC;
    const unsigned stF = s3 > 0 ? 4u * plane : 0u - 4u * plane;
    const unsigned NONE = 0x80000000u;
    const float INF = PFAST_INF_F;
    const int pend = nz + t.span;
    if (tid == 0) pub[2 * NT] = pub[PB + 2 * NT] = INF;
    __syncthreads();

    if (tid >= NT) {  // ------------------------------------------ loader wave
        const int l = tid - NT;
        // own column: level p - pipj + 2 for step p
        unsigned oOwn = NONE, oF = NONE, oH = NONE;
        if (l < NCOL) {
            const int li = l / TJ, lj = l % TJ;
            if (li < t.TIe && lj < t.TJe) {
                const int pp = pipj_of<TI, TJ>(li, lj, t, s1, s2);
                const unsigned node = (unsigned)(t.i0 + li) * a.ncols + t.j0 + lj;
                oOwn = 4u * (k0 * pC + node * C) + (unsigned)(2 - pp) * st;
                oF = 4u * (k0 * plane + node) + (unsigned)(-pp) * stF;
            }
        }
        if (l < NH) {
            int cli, clj;
            unsigned hn;
            if (halo_entry<TI, TJ>(l, t, a.nrows, a.ncols, cli, clj, hn))
                oH = 4u * (k0 * pC + hn * C) + (unsigned)(-pipj_of<TI, TJ>(cli, clj, t, s1, s2)) * st;
        }
        // the data of step q into slot q % R; each DMA only on the lanes that own an
        // entry (a lane writes its 16 / 4 bytes at lane * size from the base even
        // when its load is out of range), lane 0 always among them, so every issue
        // is exactly three VMEM instructions
        auto issue = [&](int q) {
            float *S = ring + (q % R) * SLOT;
            if (l < NCOL) {
                dma16(RF, S, oOwn);
                dma4(RS, S + S_F, oF);
            }
            if (l < NH) dma16(RF, S + S_HALO, oH);
            oOwn += st;
            oH += st;
            oF += stF;
        };
        for (int q = 0; q < A; q++) issue(q);
        for (int p = 0; p < pend; p++) {
            wait_vm<3 * (A - 1)>();  // step p's three loads have landed
            __builtin_amdgcn_s_barrier();
            issue(p + A);  // slot of step p - 1, whose readers passed this barrier
        }
        wait_vm<0>();
        return;
    }

    // ------------------------------------------------------------ compute waves
    const ColLane<TI, TJ, C> l = make_col_lane<TI, TJ, C>(a, t, s1, s2, s3);
    const int plo = l.valid ? l.pipj : 0x3fffffff, phi = l.pipj + nz;
    int psrc = -1;
    {
        const int sn = a.chunks[t.chunk].srcnode[l.c];
        if (l.valid && sn >= 0 && (unsigned)sn % plane == l.node) {
            const int sk = (int)((unsigned)sn / plane);
            psrc = l.pipj + (s3 > 0 ? sk : nz - 1 - sk);
        }
    }
    // ring indices of this thread's entries
    const int iOwn = tid;
    const int col = l.li * TJ + l.lj;
    const int iF = S_F + col;
    const int iHX = S_HALO + (l.li * 2 + (l.uxp ? 1 : 0)) * C + l.c;
    const int iHY = S_HALO + (2 * TI + l.lj * 2 + (l.uyp ? 1 : 0)) * C + l.c;
    const int m0 = -l.pipj;
    float vprev = INF;
    float vcur = l.valid && m0 == 0 ? bload(RF, l.oI) : INF;
    float vnext = l.valid && m0 >= -1 ? bload(RF, l.oI + (unsigned)(m0 + 1) * l.step) : INF;
    unsigned oS = l.oI + (unsigned)m0 * l.step;  // byte offset of this thread's level at step p
    bool changed = false;
    int slot = 0;
    for (int p = 0; p < pend; p++) {
        float *P = pub + (p & 1) * PB;
        P[tid] = vprev;
        P[NT + tid] = vnext;
        lds_barrier();
        const float *S = ring + slot * SLOT;
        const float own = S[iOwn], hx = S[iHX], hy = S[iHY], fv = S[iF];
        const float pxm = P[l.rxm], pxp = P[l.rxp], pym = P[l.rym], pyp = P[l.ryp];
        const float xa = fminf(l.uxm ? hx : pxm, l.uxp ? hx : pxp);
        const float ya = fminf(l.uym ? hy : pym, l.uyp ? hy : pyp);
        const float za = fminf(vprev, vnext);
        float nv = vcur;
        if (p >= plo && p < phi && p != psrc) {
            float qv;
            if (UNI) qv = eik3d_f32_uniform(xa, ya, za, fv);
            else qv = eik3d_f32_general(xa, ya, za, fv, a.hx, a.hy, a.hz, a.wx, a.wy, a.wz);
            if (qv < vcur) {
                nv = qv;
                bstore(RF, oS, nv);
                changed = true;
            }
        }
        vprev = nv;
        vcur = vnext;
        vnext = (unsigned)(p + 2 - l.pipj) < (unsigned)nz ? own : INF;
        oS += l.step;
        slot = slot == R - 1 ? 0 : slot + 1;
    }
    const unsigned long long bm = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && bm) atomicOr(&a.changed[t.chunk], chunk_bits(bm, C));
}

template <int TI, int TJ, int C, int A, bool S3POS>
__global__ __launch_bounds__(TI *TJ *C + 64) void adj3d_dma(SlabArgs a, int D, int s1, int s2)
{
    static_assert(C == 4, "one dwordx4 lane per column");
    constexpr int s3 = S3POS ? 1 : -1;
    constexpr int NT = TI * TJ * C, NCOL = TI * TJ, NH = 2 * (TI + TJ);
    static_assert(NCOL <= 64 && NH <= 64, "one loader lane per column / halo node");
    constexpr int R = A + 1;
    // per slot: own lambda [NT] | own T [NT] | halo lambda [NH*C] | halo T [NH*C]
    constexpr int S_OT = NT, S_HL = 2 * NT, S_HT = 2 * NT + NH * C, SLOT = 2 * NT + 2 * NH * C;
    constexpr int PB = 2 * NT + 1;
    __shared__ float2 pub[2 * PB];
    __shared__ __attribute__((aligned(16))) float ring[R * SLOT];

    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live) return;
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols, pC = plane * C;
    const rsrc_t RL = make_rsrc(a.chunks[t.chunk].field[1], 4u * pC * nz);
    const rsrc_t RT = make_rsrc(a.chunks[t.chunk].field[0], 4u * pC * nz);
    const unsigned k0 = s3 > 0 ? 0u : (unsigned)(nz - 1);
    const unsigned st = s3 > 0 ? 4u * pC : 0u - 4u * pC;
    const unsigned NONE = 0x80000000u;
    const int pend = nz + t.span;
    if (tid == 0) pub[2 * NT] = pub[PB + 2 * NT] = make_float2(0.0f, PFAST_INF_F);
    __syncthreads();

    if (tid >= NT) {  // ------------------------------------------ loader wave
        const int l = tid - NT;
        unsigned oOwn = NONE, oH = NONE;
        if (l < NCOL) {
            const int li = l / TJ, lj = l % TJ;
            if (li < t.TIe && lj < t.TJe) {
                const int pp = pipj_of<TI, TJ>(li, lj, t, s1, s2);
                const unsigned node = (unsigned)(t.i0 + li) * a.ncols + t.j0 + lj;
                oOwn = 4u * (k0 * pC + node * C) + (unsigned)(2 - pp) * st;
            }
        }
        if (l < NH) {
            int cli, clj;
            unsigned hn;
            if (halo_entry<TI, TJ>(l, t, a.nrows, a.ncols, cli, clj, hn))
                oH = 4u * (k0 * pC + hn * C) + (unsigned)(-pipj_of<TI, TJ>(cli, clj, t, s1, s2)) * st;
        }
        auto issue = [&](int q) {  // four VMEM instructions (lane 0 is in both groups)
            float *S = ring + (q % R) * SLOT;
            if (l < NCOL) {
                dma16(RL, S, oOwn);
                dma16(RT, S + S_OT, oOwn);
            }
            if (l < NH) {
                dma16(RL, S + S_HL, oH);
                dma16(RT, S + S_HT, oH);
            }
            oOwn += st;
            oH += st;
        };
        for (int q = 0; q < A; q++) issue(q);
        for (int p = 0; p < pend; p++) {
            wait_vm<4 * (A - 1)>();
            __builtin_amdgcn_s_barrier();
            issue(p + A);
        }
        wait_vm<0>();
        return;
    }

    const ColLane<TI, TJ, C> l = make_col_lane<TI, TJ, C>(a, t, s1, s2, s3);
    const int gi = t.i0 + l.li, gj = t.j0 + l.lj;
    const bool col_interior =
        l.valid && gi >= 1 && gi <= a.nrows - 2 && gj >= 1 && gj <= a.ncols - 2;
    const int plo = col_interior ? l.pipj + 1 : 0x3fffffff, phi = l.pipj + nz - 1;
    const float wx = a.wx, wy = a.wy, wz = a.wz;
    const int iHX = (l.li * 2 + (l.uxp ? 1 : 0)) * C + l.c;
    const int iHY = (2 * TI + l.lj * 2 + (l.uyp ? 1 : 0)) * C + l.c;
    auto ld2 = [&](unsigned o) { return make_float2(bload(RL, o), bload(RT, o)); };
    const int m0 = -l.pipj;
    float2 vprev = make_float2(0.0f, 0.0f);
    float2 vcur = l.valid && m0 == 0 ? ld2(l.oI) : make_float2(0.0f, 0.0f);
    float2 vnext = l.valid && m0 >= -1 ? ld2(l.oI + (unsigned)(m0 + 1) * l.step) : make_float2(0.0f, 0.0f);
    unsigned oS = l.oI + (unsigned)m0 * l.step;
    bool changed = false;
    int slot = 0;
    for (int p = 0; p < pend; p++) {
        float2 *P = pub + (p & 1) * PB;
        P[tid] = vprev;
        P[NT + tid] = vnext;
        lds_barrier();
        const float *S = ring + slot * SLOT;
        const float2 own = make_float2(S[tid], S[S_OT + tid]);
        const float2 hx = make_float2(S[S_HL + iHX], S[S_HT + iHX]);
        const float2 hy = make_float2(S[S_HL + iHY], S[S_HT + iHY]);
        const float2 pxm = P[l.rxm], pxp = P[l.rxp], pym = P[l.rym], pyp = P[l.ryp];
        float nl = vcur.x;
        if (p >= plo && p < phi && !signbit(vcur.y)) {
            const float2 xm = l.uxm ? hx : pxm;
            const float2 xp = l.uxp ? hx : pxp;
            const float2 ym = l.uym ? hy : pym;
            const float2 yp = l.uyp ? hy : pyp;
            // depth k-1 / k+1 by absolute depth (see adj3d_col)
            const float2 zm = S3POS ? vprev : vnext;
            const float2 zp = S3POS ? vnext : vprev;
            const float tc = vcur.y;
            const float txm = fabsf(xm.y), txp = fabsf(xp.y), tym = fabsf(ym.y),
                        typ = fabsf(yp.y), tzm = fabsf(zm.y), tzp = fabsf(zp.y);
            float nw = 0.0f;
            const float I = PFAST_INF_F;
            if (!(tc >= I || txm >= I || txp >= I || tym >= I || typ >= I || tzm >= I || tzp >= I)) {
                float den = 0.0f;
                float num = adj_term(tc, txp, xp.x, wx, den);
                num += adj_term(tc, txm, xm.x, wx, den);
                num += adj_term(tc, typ, yp.x, wy, den);
                num += adj_term(tc, tym, ym.x, wy, den);
                num += adj_term(tc, tzp, zp.x, wz, den);
                num += adj_term(tc, tzm, zm.x, wz, den);
                nw = fabsf(den) > 1e-9f ? num * frcp(den) : 0.0f;
            }
            if (nw != vcur.x) {
                bstore(RL, oS, nw);
                changed = true;
            }
            nl = nw;
        }
        vprev = make_float2(nl, vcur.y);
        vcur = vnext;
        vnext = own;  // out-of-column levels read 0 and are never used by interior nodes
        oS += l.step;
        slot = slot == R - 1 ? 0 : slot + 1;
    }
    const unsigned long long bm = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && bm) atomicOr(&a.changed[t.chunk], chunk_bits(bm, C));
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

// Widen one source's field to FP64.  Traveltimes (strip_sign) may carry the
// receiver marks of an adjoint in progress; adjoint fields are signed.
__global__ void gather_field(const float *__restrict__ F, int s, long N, int C, int strip_sign,
                             double *__restrict__ out)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float v = F[((size_t)(s / C) * N + n) * C + (s % C)];
    out[n] = (double)(strip_sign ? fabsf(v) : v);
}

__global__ void cell_time(const double *__restrict__ vp, float *__restrict__ f, long N,
                          double h, int uniform)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    f[n] = uniform ? (float)(h / vp[n]) : (float)(1.0 / vp[n]);
}

// 8x8-tile-major copy of a cell-time array: element (k, ti, tj, li, lj) of ft is
// f[k][8ti+li][8tj+lj], 0 where an edge tile runs past the grid.  One thread per
// element of ft.
__global__ void tile8_copy(const float *__restrict__ f, float *__restrict__ ft, int nrows,
                           int ncols, long total)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int nti = (nrows + 7) / 8, ntj = (ncols + 7) / 8;
    const int lj = (int)(e & 7), li = (int)((e >> 3) & 7);
    const long t = e >> 6;
    const int tj = (int)(t % ntj);
    const long r = t / ntj;
    const int ti = (int)(r % nti);
    const long k = r / nti;
    const int i = 8 * ti + li, j = 8 * tj + lj;
    ft[e] = i < nrows && j < ncols ? f[(k * nrows + i) * ncols + j] : 0.0f;
}

}  // namespace

namespace pfast {

// Sweep kernel variants (schedule, tile rows x columns, batch/prefetch depth).
// Variant 0 is the default; the others stay selectable for measurement.
static const SweepVariant kVariants[] = {
    {0, 8, 16, 6, "column-window 8x16 L6"},
    {1, 8, 8, 2, "slab-ring 8x8 L2"},
    {0, 8, 8, 4, "column-window 8x8 L4"},
    {0, 8, 8, 2, "column-window 8x8 L2"},
    {0, 8, 4, 4, "column-window 8x4 L4"},
    {2, 8, 8, 12, "loader-wave 8x8 A12/8"},
    {0, 4, 4, 6, "column-window 4x4 L6"},
    {0, 4, 4, 4, "column-window 4x4 L4"},
    {0, 4, 8, 6, "column-window 4x8 L6"},
    {0, 8, 4, 6, "column-window 8x4 L6"},
    {0, 8, 8, 6, "column-window 8x8 L6"},
    {0, 16, 8, 6, "column-window 16x8 L6"},
    {0, 8, 16, 4, "column-window 8x16 L4"},
    {0, 8, 8, 8, "column-window 8x8 L8"},
    {3, 8, 8, 2, "vector-lane 8x8 L2"},
    {3, 8, 8, 4, "vector-lane 8x8 L4"},
    {3, 8, 16, 2, "vector-lane 8x16 L2"},
    {0, 8, 32, 6, "column-window 8x32 L6"},
    {0, 16, 16, 6, "column-window 16x16 L6"},
    {0, 4, 16, 6, "column-window 4x16 L6"},
};

int sweep_variant_count() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }

const SweepVariant &sweep_variant(int v) { return kVariants[v]; }

namespace {
template <int TI, int TJ, int L>
void launch_col(hipStream_t st, const SlabArgs &a, int adjoint, int uniform, int D, int s1,
                int s2, int s3, dim3 grid)
{
    constexpr int C = PFAST_CHUNK;
    const dim3 block(TI * TJ * C);
    if (adjoint) {
        if (s3 > 0)
            hipLaunchKernelGGL((adj3d_col<TI, TJ, C, L, true>), grid, block, 0, st, a, D, s1, s2);
        else
            hipLaunchKernelGGL((adj3d_col<TI, TJ, C, L, false>), grid, block, 0, st, a, D, s1, s2);
    } else if (uniform) {
        hipLaunchKernelGGL((fwd3d_col<TI, TJ, C, L, true>), grid, block, 0, st, a, D, s1, s2, s3);
    } else {
        hipLaunchKernelGGL((fwd3d_col<TI, TJ, C, L, false>), grid, block, 0, st, a, D, s1, s2, s3);
    }
}

// loader-wave kernels: forward lookahead AF steps, adjoint AA steps
template <int TI, int TJ, int L>
void launch_vec(hipStream_t st, const SlabArgs &a, int adjoint, int uniform, int D, int s1,
                int s2, int s3, dim3 grid)
{
    const dim3 block(TI * TJ);
    if (adjoint) {
        if (s3 > 0)
            hipLaunchKernelGGL((adj3d_vec<TI, TJ, L, true>), grid, block, 0, st, a, D, s1, s2);
        else
            hipLaunchKernelGGL((adj3d_vec<TI, TJ, L, false>), grid, block, 0, st, a, D, s1, s2);
    } else if (uniform) {
        hipLaunchKernelGGL((fwd3d_vec<TI, TJ, L, true>), grid, block, 0, st, a, D, s1, s2, s3);
    } else {
        hipLaunchKernelGGL((fwd3d_vec<TI, TJ, L, false>), grid, block, 0, st, a, D, s1, s2, s3);
    }
}

template <int TI, int TJ, int AF, int AA>
void launch_dma(hipStream_t st, const SlabArgs &a, int adjoint, int uniform, int D, int s1,
                int s2, int s3, dim3 grid)
{
    constexpr int C = PFAST_CHUNK;
    const dim3 block(TI * TJ * C + 64);
    if (adjoint) {
        if (s3 > 0)
            hipLaunchKernelGGL((adj3d_dma<TI, TJ, C, AA, true>), grid, block, 0, st, a, D, s1, s2);
        else
            hipLaunchKernelGGL((adj3d_dma<TI, TJ, C, AA, false>), grid, block, 0, st, a, D, s1, s2);
    } else if (uniform) {
        hipLaunchKernelGGL((fwd3d_dma<TI, TJ, C, AF, true>), grid, block, 0, st, a, D, s1, s2, s3);
    } else {
        hipLaunchKernelGGL((fwd3d_dma<TI, TJ, C, AF, false>), grid, block, 0, st, a, D, s1, s2, s3);
    }
}
}  // namespace

hipError_t launch_sweep3d(hipStream_t st, const SlabArgs &a, int variant, int adjoint,
                          int uniform, int D, int s1, int s2, int s3, int ntiles_diag)
{
    constexpr int C = PFAST_CHUNK;
    const dim3 grid((unsigned)(ntiles_diag * a.nact_p8));
    switch (variant) {
    case 0: launch_col<8, 16, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 1: {
        constexpr int TI = 8, TJ = 8, L = 2;
        const dim3 block(TI * TJ * C);
        if (adjoint)
            hipLaunchKernelGGL((adj3d_slab<TI, TJ, C, L>), grid, block, 0, st, a, D, s1, s2, s3);
        else if (uniform)
            hipLaunchKernelGGL((fwd3d_slab<TI, TJ, C, L, true>), grid, block, 0, st, a, D, s1, s2,
                               s3);
        else
            hipLaunchKernelGGL((fwd3d_slab<TI, TJ, C, L, false>), grid, block, 0, st, a, D, s1,
                               s2, s3);
        break;
    }
    case 2: launch_col<8, 8, 4>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 3: launch_col<8, 8, 2>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 4: launch_col<8, 4, 4>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 5: launch_dma<8, 8, 12, 8>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 6: launch_col<4, 4, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 7: launch_col<4, 4, 4>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 8: launch_col<4, 8, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 9: launch_col<8, 4, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 10: launch_col<8, 8, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 11: launch_col<16, 8, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 12: launch_col<8, 16, 4>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 13: launch_col<8, 8, 8>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 14: launch_vec<8, 8, 2>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 15: launch_vec<8, 8, 4>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 16: launch_vec<8, 16, 2>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 17: launch_col<8, 32, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 18: launch_col<16, 16, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    case 19: launch_col<4, 16, 6>(st, a, adjoint, uniform, D, s1, s2, s3, grid); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

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

hipError_t launch_gather_field(hipStream_t st, const float *F, int s, long N, int strip_sign,
                               double *out)
{
    hipLaunchKernelGGL(gather_field, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, F, s, N,
                       PFAST_CHUNK, strip_sign, out);
    return hipGetLastError();
}

long tile8_size(int nrows, int ncols, int nz)
{
    return (long)nz * ((nrows + 7) / 8) * ((ncols + 7) / 8) * 64;
}

hipError_t launch_tile8_copy(hipStream_t st, const float *f, float *ft, int nrows, int ncols,
                             int nz)
{
    const long total = tile8_size(nrows, ncols, nz);
    hipLaunchKernelGGL(tile8_copy, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, f, ft,
                       nrows, ncols, total);
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
