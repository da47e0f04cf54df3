// batch3d.hip -- batched FP32 3D sweeps (every source of a shot set on one GPU).
//
// Layout in HBM: fields are stored per chunk of C sources, node-major with the C
// sources innermost: F[k][i][j][c] (k = depth, i = row, j = column).  The adjoint
// field pairs every lambda with a copy of its traveltime, LT[k][i][j][c] = (lambda, T),
// so that one 8-byte load brings both values an adjoint update reads of a neighbour.
// Each chunk names its traveltime and adjoint fields and its model's velocity-derived
// cell time f[k][i][j] (h/v for uniform spacing, else 1/v) through a ChunkRef, so
// one launch can sweep chunks of several shot sets and models of one grid shape.
//
// Schedule: the (i,j) plane is cut into TI x TJ column tiles that run the whole
// depth.  For one ordering (s3,s1,s2) tile (ti',tj') (coordinates taken in sweep
// direction) depends only on tiles (ti'-1,tj') and (ti',tj'-1), so one launch per
// tile anti-diagonal D = ti'+tj' keeps the reference's upwind order with no
// inter-block waiting.  Inside a block, thread (li,lj,c) owns one column of one
// source; hyperplane p of the tile processes level pk = p - pi - pj (pi,pj its
// sweep-ordered position), so every node sees its behind neighbours updated and
// its ahead neighbours not yet, as in the reference's lexicographic sweep.
#include "batch3d.h"
#include "flow.h"
#include "pfast_internal.h"
#include "solvers.h"

#include <algorithm>
#include <map>
#include <mutex>

namespace {

struct TileCtx {
    int chunk, i0, j0, TIe, TJe, span;
    bool live;
};

template <int TI, int TJ>
__device__ __forceinline__ TileCtx decode_tile(SweepArgs &a, int D, int s1, int &s2)
{
    TileCtx t;
    const int slot = blockIdx.x % a.nact_p8;
    int tt = blockIdx.x / a.nact_p8;
    if (tt >= a.ntA) {  // the second ordering of a paired launch
        tt -= a.ntA;
        D = a.DB;
        s2 = -s2;
        a.epoch = a.epochB;
        a.o = a.oB;
    }
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

// Device phase state (SweepArgs::phase): is the chunk still sweeping; the launch's
// tile-skip epoch from the cycle number.
__device__ __forceinline__ bool phase_live(const SweepArgs &a, int chunk)
{
    return !a.live || a.live[chunk] != 0;
}
__device__ __forceinline__ void phase_epoch(SweepArgs &a)
{
    if (a.phase) a.epoch += (a.phase[0] - 1) * a.ep_cycle;
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

// ============================================================ column sweeps
// The column window: thread (li,lj,c) keeps its own column of one source in
// registers -- levels m-1 (updated), m, m+1 (not yet updated) of the level
// m = p - pi - pj it works on at step p -- so depth needs no exchange at all.  Only the in-plane neighbours do: at every step each thread
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
    int rxm, rxp, rym, ryp;    // LDS indices of the neighbours' values in a publish block
    unsigned oI;               // byte offset of level 0 of the own column
    unsigned oF;               // byte offset of level 0 in the cell-time array
    unsigned step, stepF;      // byte step per level (two's complement when s3 < 0)
    unsigned node;
};

template <int TI, int TJ, int C>
__device__ __forceinline__ ColLane<TI, TJ, C> make_col_lane(const SweepArgs &a, const TileCtx &t,
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
    // "m+1" slot [NT, 2NT); a halo node: its slot among the halo values published for
    // the step after the INF slot, faces x- (TI rows), x+ (TI), y- (TJ columns), y+ (TJ),
    // C sources per face position (HaloStage).  Idle threads read their own slot (unused).
    constexpr int HALO = INFSLOT + 1;
    const bool ok = l.valid;
    l.rxm = !ok ? tid : l.uxm ? HALO + l.li * C + l.c : exm ? INFSLOT : (s2 > 0 ? 0 : NT) + tid - C;
    l.rxp = !ok ? tid : l.uxp ? HALO + (TI + l.li) * C + l.c
                      : exp_ ? INFSLOT : (s2 > 0 ? NT : 0) + tid + C;
    l.rym = !ok ? tid : l.uym ? HALO + (2 * TI + l.lj) * C + l.c
                      : eym ? INFSLOT : (s1 > 0 ? 0 : NT) + tid - C * TJ;
    l.ryp = !ok ? tid : l.uyp ? HALO + (2 * TI + TJ + l.lj) * C + l.c
                      : eyp ? INFSLOT : (s1 > 0 ? NT : 0) + tid + C * TJ;
    const unsigned plane = (unsigned)a.nrows * a.ncols, pC = plane * C;
    l.node = l.valid ? (unsigned)gi * a.ncols + gj : 0u;
    const unsigned k0 = s3 > 0 ? 0u : (unsigned)(a.nz - 1);
    l.oI = 4u * (k0 * pC + l.node * C + l.c);
    l.oF = 4u * (k0 * plane + l.node);
    l.step = s3 > 0 ? 4u * pC : 0u - 4u * pC;
    l.stepF = s3 > 0 ? 4u * plane : 0u - 4u * plane;
    return l;
}

// Tile index of (chunk, tile) in the skip records, or -1 without skipping.
template <int TI, int TJ>
__device__ __forceinline__ int skip_index(const SweepArgs &a, const TileCtx &t)
{
    return a.tmod ? (t.chunk * a.nti + t.i0 / TI) * a.ntj + t.j0 / TJ : -1;
}

// The sweep of tile ix in this ordering would change nothing (see SweepArgs): its
// clean sweep is newer than every change of the tile and of the neighbour faces
// it reads (left tile's x+ face, right's x-, upper's y+, lower's y-).  DEV:
// device-scope reads, for records written earlier in the same launch.
template <int TI, int TJ, bool DEV = false>
__device__ __forceinline__ bool skip_tile(const SweepArgs &a, const TileCtx &t, int ix)
{
    if (ix < 0) return false;
    auto ld = [](const int *p) { return DEV ? ld_agent(p) : *p; };
    const int cl = ld(&a.tclean[ix * kTileClean + a.o]);
    if (cl <= 0) return false;
    const int ti = t.i0 / TI, tj = t.j0 / TJ;
    const int *M = a.tmod;
    int m = ld(&M[ix * 5]);
    if (tj > 0) m = max(m, ld(&M[(ix - 1) * 5 + 2]));
    if (tj < a.ntj - 1) m = max(m, ld(&M[(ix + 1) * 5 + 1]));
    if (ti > 0) m = max(m, ld(&M[(ix - a.ntj) * 5 + 4]));
    if (ti < a.nti - 1) m = max(m, ld(&M[(ix + a.ntj) * 5 + 3]));
    return m < cl;
}

// Faces of the tile that this thread's column lies on (bit 0 x-, 1 x+, 2 y-, 3 y+).
__device__ __forceinline__ unsigned col_faces(bool valid, int li, int lj, int TIe, int TJe)
{
    if (!valid) return 0u;
    return (lj == 0 ? 1u : 0u) | (lj == TJe - 1 ? 2u : 0u) | (li == 0 ? 4u : 0u) |
           (li == TIe - 1 ? 8u : 0u);
}

// Profiling: the node-sweeps this tile sweep performed (columns x depth x the chunk's
// real sources; padding sources have node -1).
__device__ __forceinline__ void count_swept(const SweepArgs &a, const TileCtx &t, int levels)
{
    if (!a.swept || threadIdx.x != 0) return;
    int nreal = 0;
    for (int c = 0; c < PFAST_CHUNK; c++) nreal += a.chunks[t.chunk].srcnode[c] >= 0 ? 1 : 0;
    atomicAdd(a.swept, (unsigned long long)t.TIe * t.TJe * levels * nreal);
}

// Record the outcome of this tile sweep: `what` = bit 4 (some node changed) | the
// faces on which a node changed, OR-ed over the block.
__device__ __forceinline__ void record_tile(const SweepArgs &a, int ix, unsigned what, int *any)
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
            st_agent(&a.tmod[ix * 5], a.epoch);
#pragma unroll
            for (int f = 0; f < 4; f++)
                if ((v >> f) & 1) st_agent(&a.tmod[ix * 5 + 1 + f], a.epoch);
            st_agent(&a.tclean[ix * kTileClean + a.o], 0);
        } else {
            st_agent(&a.tclean[ix * kTileClean + a.o], a.epoch);
        }
    }
}

// ------------------------------------------------------------ start levels
// (SweepArgs::starts.)  A block's LDS record array: rec[0] the change word of
// record_tile, rec[1 + 2r] / rec[2 + 2r] the lowest / highest level (sweep order)
// changed in region r: 0 the tile, 1..4 its faces x-, x+, y-, y+; rec[11 + b % 3]
// whether batch b changed a node (kStartRec ints).
constexpr int kStartRec = 14;

__device__ __forceinline__ void start_init(int *rec)  // one thread, before a barrier
{
    rec[0] = 0;
    for (int r = 0; r < 5; r++) {
        rec[1 + 2 * r] = 0x7fffffff;
        rec[2 + 2 * r] = -1;
    }
    rec[11] = rec[12] = rec[13] = 0;
}

// This lane's changed levels lo..hi (none: hi < lo), into the regions it lies in;
// before record_tile's barrier.
__device__ __forceinline__ void start_ranges(int *rec, int lo, int hi, unsigned faces)
{
    if (hi < lo) return;
    atomicMin(&rec[1], lo);
    atomicMax(&rec[2], hi);
#pragma unroll
    for (int f = 0; f < 4; f++)
        if ((faces >> f) & 1u) {
            atomicMin(&rec[3 + 2 * f], lo);
            atomicMax(&rec[4 + 2 * f], hi);
        }
}

// Record ix of region r of the tile (its own for r = 0, else the neighbour across
// face r), or -1 outside the grid.
template <int TI, int TJ>
__device__ __forceinline__ int region_tile(const SweepArgs &a, const TileCtx &t, int ix, int r)
{
    const int ti = t.i0 / TI, tj = t.j0 / TJ;
    if (r == 1) return tj > 0 ? ix - 1 : -1;
    if (r == 2) return tj < a.ntj - 1 ? ix + 1 : -1;
    if (r == 3) return ti > 0 ? ix - a.ntj : -1;
    if (r == 4) return ti < a.nti - 1 ? ix + a.ntj : -1;
    return ix;
}

// After record_tile's barrier, lanes 0..39 (region lane / 8, ordering lane % 8): the
// start levels this sweep leaves for the tile and the neighbours across its changed
// faces.  s3: depth direction of this sweep's ordering a.o.
template <int TI, int TJ>
__device__ __forceinline__ void start_update(const SweepArgs &a, const TileCtx &t, int ix,
                                             const int *rec, int s3)
{
    const int lane = (int)threadIdx.x;
    if (lane >= 40) return;
    const int r = lane >> 3, o2 = lane & 7;
    const int tgt = region_tile<TI, TJ>(a, t, ix, r);
    const int lo = rec[1 + 2 * r], hi = rec[2 + 2 * r];
    const int nz = a.nz;
    int *w = a.tclean + (size_t)max(tgt, 0) * kTileClean;
    if (r == 0 && o2 == a.o) {  // this sweep's changed levels and the ones beside them
        w[8 + o2] = hi < 0 ? nz : max(lo - 1, 0);
        w[16 + o2] = hi < 0 ? nz : nz - 1 - min(hi + 1, nz - 1);
        return;
    }
    if (hi < 0 || tgt < 0) return;
    // the changed levels in depth, then in ordering o2's sweep order (the tile's own
    // nodes widened by the level beside them)
    const int klo = s3 > 0 ? lo : nz - 1 - hi, khi = s3 > 0 ? hi : nz - 1 - lo;
    const int first = o2 >= 4 ? klo : nz - 1 - khi, last = o2 >= 4 ? khi : nz - 1 - klo;
    const int m = r == 0 ? 1 : 0;
    atomicMin(&w[8 + o2], max(first - m, 0));
    atomicMin(&w[16 + o2], nz - 1 - min(last + m, nz - 1));
}

// The start level of this block's sweep (0 without start levels) and its last dirty
// level (nz - 1 without).
__device__ __forceinline__ int start_level(const SweepArgs &a, int ix)
{
    return a.starts && ix >= 0 ? a.tclean[ix * kTileClean + 8 + a.o] : 0;
}
__device__ __forceinline__ int last_dirty_level(const SweepArgs &a, int ix)
{
    return a.nz - 1 - (a.starts && ix >= 0 ? a.tclean[ix * kTileClean + 16 + a.o] : 0);
}

// ------------------------------------------------------------ dataflow waits
// Dependency wait of a dataflow task (flow.h).  Wave 0 of the block watches the
// task's dependencies -- lane d < ndep dependency d -- and publishes in LDS the last
// step whose memory operations they allow; the other waves read that word and never
// poll memory.  `dv` holds each word as read one batch earlier, so the check costs no
// memory latency once the dependencies are ahead.
struct FlowWait {
    rsrc_t R;            // the launch's dataflow state words
    unsigned w;          // this lane's dependency: word index << 6 | requirement code
    unsigned dv;         // the word as last read (wave 0)
    unsigned base;       // launch epoch << 16
    int smax_cap;        // pend - 1: later steps touch no memory
    unsigned abort_off;  // byte offset of the abort word
    volatile int *ok;    // LDS: steps up to *ok are allowed
    bool watcher;        // this wave is wave 0
    unsigned long long timeout;  // wall-clock ticks a wait may take
};

// Requirement codes: pflow::need_add values below kCodeDone, kCodeDone for kDone.
constexpr unsigned kCodeDone = 63;
constexpr int kOkAll = 0x7fffffff;

// The last step up to which every dependency (as last read) allows memory operations:
// a word of an earlier launch allows nothing, a finished task everything, else steps
// completed minus the requirement's offset (flow.h).  Wave 0, wave-uniform result.
__device__ __forceinline__ int flow_ok_until(const FlowWait &f)
{
    const unsigned code = f.w & 63u;
    int v;
    if (f.dv < f.base) v = -1;
    else if (f.dv - f.base >= pflow::kDone) v = kOkAll;
    else if (code == kCodeDone) v = -1;
    else v = (int)(f.dv - f.base) - (int)code;
    int m = __builtin_amdgcn_readlane(v, 0);
#pragma unroll
    for (int l = 1; l < pflow::kMaxDeps; l++) m = min(m, __builtin_amdgcn_readlane(v, l));
    return m;
}

__device__ __forceinline__ void flow_prefetch(FlowWait &f)
{
    if (f.watcher)
        f.dv = __builtin_amdgcn_raw_buffer_load_b32(f.R, (f.w >> 6) * 4u, 0, kCpolDevice);
}

// Until the task's dependencies allow the memory operations of steps up to smax
// (wave-uniform).  A wait longer than the timeout sets the abort word, and a wait that
// sees it set gives up: the task then runs to its end without waiting (the launch's
// results are void and the host reports it), so every block leaves.
__device__ __forceinline__ void flow_wait(FlowWait &f, int smax)
{
    smax = min(smax, f.smax_cap);
    if (f.watcher) {
        int ok = flow_ok_until(f);
        if (ok < smax) {
            const unsigned long long t0 = wall_clock64();
            for (;;) {
                __builtin_amdgcn_s_sleep(4);
                flow_prefetch(f);
                ok = flow_ok_until(f);
                if (ok >= smax) break;
                if (__builtin_amdgcn_raw_buffer_load_b32(f.R, f.abort_off, 0, kCpolDevice)) {
                    ok = kOkAll;
                    break;
                }
                if (wall_clock64() - t0 > f.timeout) {
                    __builtin_amdgcn_raw_buffer_store_b32(1u, f.R, f.abort_off, 0, kCpolDevice);
                    ok = kOkAll;
                    break;
                }
            }
        }
        if ((threadIdx.x & 63) == 0 && ok > *f.ok) *f.ok = ok;
    } else {
        while (*f.ok < smax) __builtin_amdgcn_s_sleep(1);
    }
}

// Field access of the tile sweeps.  The per-diagonal kernels read only what earlier
// launches wrote and use the default cache policy; the dataflow kernels read what
// other work-groups of the same launch wrote, possibly on another XCD, and access the
// fields at device scope.
template <bool DEV>
__device__ __forceinline__ float fload(rsrc_t r, unsigned off)
{
    return DEV ? bload_dev(r, off) : bload(r, off);
}
template <bool DEV>
__device__ __forceinline__ float2 fload2(rsrc_t r, unsigned off)
{
    return DEV ? bload2_dev(r, off) : bload2(r, off);
}
template <bool DEV>
__device__ __forceinline__ void fstore(rsrc_t r, unsigned off, float v)
{
    if (DEV) bstore_dev(r, off, v);
    else bstore(r, off, v);
}

// ------------------------------------------------------------ halo values
// A tile's four halo faces (the columns beside its x edges, the rows beside its y
// edges) are read by the lanes on those edges only: with a load per lane, a wave of
// an edge row would carry a second full row of loads and every wave a sliver of x
// halo per step, and the block waits at each step for its slowest wave.  Instead
// every thread of the block loads a share of the faces for a whole batch of L steps
// (full-wave loads, consecutive lanes on consecutive sources of one face node), a
// batch ahead like the column's own values.  At step u the threads holding that
// step's values write them into the halo slots of the step's publish block, beside
// the levels the lanes publish, and an edge lane reads its halo neighbour there like
// any other neighbour (its neighbour index points at the slot), so no lane selects
// between a published and a halo value.  Value s = u * HSU + face offset + q * C + c
// (step u, face f, position q along the face, source c; faces x- (TI rows), x+ (TI),
// y- (TJ columns), y+ (TJ)) is loaded by thread s % NT and goes to halo slot
// s % HSU.  A value stands for the level the consuming edge lane works on at that
// step, so the loader takes the consumer's sweep position into its address.
template <int TI, int TJ, int C, int L>
struct HaloStage {
    static constexpr int NT = TI * TJ * C;
    static constexpr int HSU = 2 * (TI + TJ) * C;  // halo values per step
    static constexpr int NS = L * HSU;             // per batch
    static constexpr int K = (NS + NT - 1) / NT;   // values loaded per thread
    // publish block of a step parity: [m-1 | m+1][thread], INF slot, HSU halo slots
    static constexpr int PB = 2 * NT + 1 + HSU;
    static_assert(HSU % 64 == 0 && NT % 64 == 0, "the values of a wave are for one step");
    unsigned addr[K];  // byte offset of the value at batch start step 0
    bool use[K];       // the value feeds an edge lane that reads a halo in the domain
    int uk[K];         // step of the batch the value is for (wave-uniform; >= L: none)
    int rk[K];         // its halo slot
};

template <int TI, int TJ, int C, int L>
__device__ __forceinline__ HaloStage<TI, TJ, C, L> make_halo_stage(const SweepArgs &a,
                                                                   const TileCtx &t, int s1,
                                                                   int s2, int s3, unsigned rec)
{
    using H = HaloStage<TI, TJ, C, L>;
    H h;
    const unsigned plane = (unsigned)a.nrows * a.ncols, pC = plane * C;
    const unsigned k0 = s3 > 0 ? 0u : (unsigned)(a.nz - 1);
    const unsigned st = (s3 > 0 ? 4u * pC : 0u - 4u * pC) * rec;
    const bool hxm = t.j0 > 0, hxp = t.j0 + t.TJe < a.ncols;
    const bool hym = t.i0 > 0, hyp = t.i0 + t.TIe < a.nrows;
#pragma unroll
    for (int k = 0; k < H::K; k++) {
        const int s = (int)threadIdx.x + k * H::NT;
        const int u = s / H::HSU, r = s % H::HSU;
        const int c = r % C, f = r / C;  // f: face position index over all four faces
        int li, lj, gi, gj;
        bool in;
        if (f < TI) {                      // x-: row f, read at column j0 - 1
            li = f; lj = 0; gi = t.i0 + f; gj = t.j0 - 1; in = hxm && f < t.TIe;
        } else if (f < 2 * TI) {           // x+: row f - TI, column j0 + TJe
            li = f - TI; lj = t.TJe - 1; gi = t.i0 + li; gj = t.j0 + t.TJe; in = hxp && li < t.TIe;
        } else if (f < 2 * TI + TJ) {      // y-: column f - 2TI, row i0 - 1
            li = 0; lj = f - 2 * TI; gi = t.i0 - 1; gj = t.j0 + lj; in = hym && lj < t.TJe;
        } else {                           // y+: column, row i0 + TIe
            li = t.TIe - 1; lj = f - 2 * TI - TJ; gi = t.i0 + t.TIe; gj = t.j0 + lj;
            in = hyp && lj < t.TJe;
        }
        in = in && s < H::NS;
        const int pipj = (s1 > 0 ? li : t.TIe - 1 - li) + (s2 > 0 ? lj : t.TJe - 1 - lj);
        const unsigned node = in ? (unsigned)gi * a.ncols + (unsigned)gj : 0u;
        h.use[k] = in;
        h.addr[k] = 4u * rec * (k0 * pC + node * C + (unsigned)c) + (unsigned)(u - pipj) * st;
        h.uk[k] = __builtin_amdgcn_readfirstlane(u);
        h.rk[k] = r;
    }
    return h;
}

// ======================================================= column-window tiles
// One tile sweep of the forward column window (the body of fwd3d_col and of each
// task of fwd3d_flow).  FLOW: wait on the task's dependencies before each batch's
// memory operations and publish the task's progress every batch (flow.h).  BOX: the
// chunk's sources hold fixed boxes (half-width R > 0) that the sweep must leave alone;
// with R = 0 the source node alone is fixed, and no local solve can lower its 0 (each
// is at least the smallest neighbour value, and traveltimes are never negative), so
// the sweep needs no test for it.  k0, k1: the start level and the last dirty level
// (SweepArgs::starts; 0 and nz - 1 in the dataflow kernels); rec: the block's LDS
// record array (kStartRec ints).
template <int TI, int TJ, int C, int L, bool UNI, bool FLOW, bool BOX>
__device__ __forceinline__ void fwd_tile(const SweepArgs &a, const TileCtx &t, int s1, int s2,
                                         int s3, int six, int k0, int k1, float *pub, int *rec,
                                         FlowWait *fw, unsigned *self)
{
    constexpr int NT = TI * TJ * C;
    static_assert(L % 2 == 0, "the step parity must be static inside a batch");
    static_assert(C % 4 == 0, "a node's sources fill whole DPP quads");
    // per step parity: [m-1 | m+1][thread], an INF slot (domain-edge neighbour), the
    // step's halo values
    using HS = HaloStage<TI, TJ, C, L>;
    constexpr int PB = HS::PB;
    const ColLane<TI, TJ, C> l = make_col_lane<TI, TJ, C>(a, t, s1, s2, s3);
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    const unsigned fbytes = 4u * plane * (unsigned)nz * C;
    const rsrc_t RF = make_rsrc(uniform_ptr(a.chunks[t.chunk].field[0]), fbytes);
    // 8x8 tiles read the tile-major cell times: a tile's level is 256 contiguous bytes
    // (two cache lines) instead of eight 32-byte pieces of grid rows
    constexpr bool F8 = TI == 8 && TJ == 8;
    const unsigned tplane = (unsigned)a.nti * a.ntj * 64u;
    const rsrc_t RS = F8 ? make_rsrc(uniform_ptr(a.chunks[t.chunk].slow8), 4u * tplane * (unsigned)nz)
                         : make_rsrc(uniform_ptr(a.chunks[t.chunk].slow), 4u * plane * (unsigned)nz);
    const float INF = PFAST_INF_F;
    if (tid == 0) {
        pub[2 * NT] = pub[PB + 2 * NT] = INF;
        if constexpr (FLOW) rec[0] = 0;
        else start_init(rec);
    }
    // steps p at which this thread updates its node: plo = pipj <= p < pipj + nz, except
    // the steps pfl .. pfl + pfn at which its column crosses its source's fixed box
    // (the source node alone when the box half-width R is 0; levels of the box outside
    // the grid fall outside the column's steps, so the range needs no clipping)
    const int plo = l.valid ? l.pipj : 0x3fffffff;
    const int R = a.chunks[t.chunk].srcinit;
    const unsigned pfn = 2u * (unsigned)R;
    int pfl = 0x3fffffff;
    if constexpr (BOX) {
        const int sn = a.chunks[t.chunk].srcnode[l.c];
        if (l.valid && sn >= 0) {
            const unsigned sc = (unsigned)sn % plane;
            const int si = (int)(sc / (unsigned)a.ncols), sj = (int)(sc % (unsigned)a.ncols);
            const int sk = (int)((unsigned)sn / plane);
            if (abs(t.i0 + l.li - si) <= R && abs(t.j0 + l.lj - sj) <= R)
                pfl = l.pipj + (s3 > 0 ? sk - R : nz - 1 - sk - R);
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
    // The vector-memory path bounds the sweep's throughput, so: the halo faces are
    // loaded by the whole block and published step by step (HaloStage), and the cell
    // time, one value per node, is loaded for four levels at once, lane c of each DPP
    // quad of the node's lanes taking level 4g + c%4; step u broadcasts it within the
    // quad.
    constexpr int FQ = (L + 3) / 4;
    const HS hs = make_halo_stage<TI, TJ, C, L>(a, t, s1, s2, s3, 1u);
    const unsigned oFc = oF + (unsigned)(l.c & 3) * stF;
    // queue of one batch: own column two levels ahead (m+2, INF outside the column),
    // cell time of levels m .. m+4*FQ-1; and this thread's halo values of that batch
    auto load_batch = [&](int pb, float *ow, float *fv, float *hl) {
        const int mb = pb - l.pipj;
        const unsigned dm = (unsigned)mb * st, dmf = (unsigned)mb * stF;
#pragma unroll
        for (int u = 0; u < L; u++) {
            const float o = fload<FLOW>(RF, l.oI + dm + (unsigned)(u + 2) * st);
            ow[u] = (unsigned)(mb + u + 2) < (unsigned)nz ? o : INF;
        }
        const unsigned dp = (unsigned)pb * st;
#pragma unroll
        for (int k = 0; k < HS::K; k++) hl[k] = hs.use[k] ? fload<FLOW>(RF, hs.addr[k] + dp) : INF;
#pragma unroll
        for (int g = 0; g < FQ; g++) fv[g] = bload(RS, oFc + dmf + (unsigned)(4 * g) * stF);
    };
    if constexpr (FLOW) {
        flow_wait(*fw, L - 1);
        asm volatile("" ::: "memory");
    }
    // the sweep starts at step pb0 <= k0 (a multiple of the batch depth, where every
    // batch of the column starts), where the window holds levels m0-1 (as this sweep leaves it:
    // below k0 that is the stored value), m0 and m0+1 of level m0 = pb0 - pipj
    // (the dataflow kernels start at step 0: level m0-1 is outside the column)
    const int pb0 = FLOW ? 0 : k0 / L * L;
    const int m0 = pb0 - l.pipj;
    auto lev = [&](int m) { return l.valid && (unsigned)m < (unsigned)nz; };
    float vprev, vcur, vnext;
    if constexpr (FLOW) {
        vprev = INF;
        vcur = l.valid && m0 == 0 ? fload<FLOW>(RF, l.oI) : INF;
        vnext = l.valid && m0 >= -1 ? fload<FLOW>(RF, l.oI + (unsigned)(m0 + 1) * st) : INF;
    } else {
        vprev = lev(m0 - 1) ? fload<FLOW>(RF, l.oI + (unsigned)(m0 - 1) * st) : INF;
        vcur = lev(m0) ? fload<FLOW>(RF, l.oI + (unsigned)m0 * st) : INF;
        vnext = lev(m0 + 1) ? fload<FLOW>(RF, l.oI + (unsigned)(m0 + 1) * st) : INF;
    }
    float qa[2][L], qb[2][L], ha[HS::K], hb[HS::K];
    load_batch(pb0, qa[0], qa[1], ha);
    if constexpr (FLOW) flow_prefetch(*fw);
    __syncthreads();  // INF slot visible

    bool changed = false;
    int mlo = 0x7fffffff, mhi = -1;  // this lane's changed levels, batch granularity
    float sval[L];
    bool chg[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        sval[u] = 0.0f;
        chg[u] = false;
    }
    // one batch of L steps from queue q and halo values h (loaded a batch earlier);
    // issues the loads of the following batch into nq, nh first
    // (per-diagonal kernels) steps from pstop on reproduce the stored values once a whole
    // batch has changed nothing: every lane is then past the last dirty level and the
    // level beside it
    const int pstop = k1 + 2 + t.span;
    auto batch = [&](int pb, float (*q)[L], const float *h, float (*nq)[L], float *nh) -> bool {
        const int b = (pb - pb0) / L;
        if constexpr (!FLOW)
            if (tid == 0) rec[11 + (b + 1) % 3] = 0;
        wait_vm_all();
        if constexpr (FLOW) {
            flow_wait(*fw, pb + 2 * L - 1);
            flow_prefetch(*fw);
        }
        // an opaque copy of the batch start for the addresses: otherwise the compiler
        // keeps every per-step address of the queues in a register of its own
        int pbo = pb;
        if constexpr (FLOW) asm volatile("" : "+s"(pbo) : : "memory");
        else asm volatile("" : "+s"(pbo));
        {  // changed nodes of the previous batch (steps pb-L .. pb-1)
            const unsigned d = (unsigned)(pbo - L - l.pipj) * st;
#pragma unroll
            for (int u = 0; u < L; u++)
                if (chg[u]) fstore<FLOW>(RF, l.oI + d + (unsigned)u * st, sval[u]);
        }
        load_batch(pbo + L, nq[0], nq[1], nh);
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            float *P = pub + (u & 1) * PB;
            P[tid] = vprev;
            P[NT + tid] = vnext;
#pragma unroll
            for (int k = 0; k < HS::K; k++)
                if (hs.uk[k] == u) P[2 * NT + 1 + hs.rk[k]] = h[k];
            lds_barrier();
            if constexpr (!FLOW)
                if (u == 0 && b > 0 && pb >= pstop && rec[11 + (b - 1) % 3] == 0) return true;
            if constexpr (FLOW) {
                // every wave has waited for its stores of steps < pb - L
                if (u == 0 && tid == 0 && pbo > L) st_agent(self, fw->base + (unsigned)(pbo - L));
            }
            const float xa = fminf(P[l.rxm], P[l.rxp]);
            const float ya = fminf(P[l.rym], P[l.ryp]);
            const float za = fminf(vprev, vnext);
            const float fval = quad_bcast(q[1][u >> 2], u);
            float nv = vcur;
            bool c = false;
            if ((unsigned)(p - plo) < (unsigned)nz && (!BOX || (unsigned)(p - pfl) > pfn)) {
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
        if constexpr (!FLOW) {
            bool bc = false;
#pragma unroll
            for (int u = 0; u < L; u++) bc = bc || chg[u];
            if (bc) {
                mlo = min(mlo, pb - l.pipj);
                mhi = pb + L - 1 - l.pipj;
            }
            if (__ballot(bc) && (tid & 63) == 0) atomicOr(&rec[11 + b % 3], 1);
        }
        return false;
    };
    // steps 0 .. pend-1, in batches of L (the queues and stage parities alternate); the
    // loop ends after the batch holding step pend-1 (uniform over the block)
    const int pend = nz + t.span;
    int pdone = pend;  // the step the sweep ended at
    for (int pb = pb0;; pb += 2 * L) {
        if (batch(pb, qa, ha, qb, hb)) {
            pdone = pb;
            break;
        }
        if (pb + L >= pend) break;
        if (batch(pb + L, qb, hb, qa, ha)) {
            pdone = pb + L;
            break;
        }
        if (pb + 2 * L >= pend) break;
    }
    {  // changed nodes of the last batch, the one holding step pend-1
        const int pl = (pend - 1) / L * L;
        const unsigned d = (unsigned)(pl - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++)
            if (chg[u]) fstore<FLOW>(RF, l.oI + d + (unsigned)u * st, sval[u]);
    }
    if constexpr (FLOW) wait_vm_all();
    const unsigned long long bm = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && bm) atomicOr(&a.changed[t.chunk], chunk_bits(bm, C));
    if constexpr (FLOW) {
        count_swept(a, t, nz);
        record_tile(a, six, changed ? 16u | col_faces(l.valid, l.li, l.lj, t.TIe, t.TJe) : 0u,
                    rec);
    } else {
        const unsigned faces = col_faces(l.valid, l.li, l.lj, t.TIe, t.TJe);
        if (a.starts) start_ranges(rec, max(mlo, 0), min(mhi, nz - 1), faces);
        count_swept(a, t, max(min(pdone, nz) - pb0, 0));
        record_tile(a, six, changed ? 16u | faces : 0u, rec);
        if (a.starts && six >= 0) start_update<TI, TJ>(a, t, six, rec, s3);
    }
}

template <int TI, int TJ, int C, int L, bool UNI>
__global__ __launch_bounds__(TI *TJ *C) void fwd3d_col(SweepArgs a, int D, int s1, int s2, int s3)
{
    __shared__ float pub[2 * HaloStage<TI, TJ, C, L>::PB];
    __shared__ int rec[kStartRec];
    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live || !phase_live(a, t.chunk)) return;
    const int six = skip_index<TI, TJ>(a, t);
    if (skip_tile<TI, TJ>(a, t, six)) return;
    const int k0 = start_level(a, six);
    if (k0 >= a.nz) return;
    const int k1 = last_dirty_level(a, six);
    phase_epoch(a);
    if (a.chunks[t.chunk].srcinit > 0)
        fwd_tile<TI, TJ, C, L, UNI, false, true>(a, t, s1, s2, s3, six, k0, k1, pub, rec,
                                                 nullptr, nullptr);
    else
        fwd_tile<TI, TJ, C, L, UNI, false, false>(a, t, s1, s2, s3, six, k0, k1, pub, rec,
                                                  nullptr, nullptr);
}

// Adjoint, column window, on the interleaved (lambda, T) field: every load brings a
// node's pair in one 8-byte access (the traveltimes are read-only; a negative value
// marks a fixed receiver node), and a changed node stores its lambda alone.  Record
// offsets are twice the traveltime field's.  Only interior nodes are updated, and
// every neighbour they read is inside the domain.
template <int TI, int TJ, int C, int L, bool UNI, bool S3POS, bool FLOW>
__device__ __forceinline__ void adj_tile(const SweepArgs &a, const TileCtx &t, int s1, int s2,
                                         int six, int k0, int k1, float2 *pub, int *rec,
                                         FlowWait *fw, unsigned *self)
{
    constexpr int NT = TI * TJ * C;
    constexpr int s3 = S3POS ? 1 : -1;
    static_assert(L % 2 == 0, "the step parity must be static inside a batch");
    // per step parity: [m-1 | m+1][thread], INF slot, the step's halo pairs
    using HS = HaloStage<TI, TJ, C, L>;
    constexpr int PB = HS::PB;
    const ColLane<TI, TJ, C> l = make_col_lane<TI, TJ, C>(a, t, s1, s2, s3);
    const int tid = threadIdx.x;
    const int nz = a.nz;
    const unsigned plane = (unsigned)a.nrows * a.ncols;
    const unsigned fbytes = 8u * plane * (unsigned)nz * C;
    const rsrc_t RLT = make_rsrc(uniform_ptr(a.chunks[t.chunk].field[1]), fbytes);
    // 8-byte records: the traveltime field's offsets doubled, mod 2^32 (the shot set's
    // size bound, 8 (N + 64 nrows ncols) C <= 2^32, keeps a level outside the column
    // outside the field)
    const unsigned oI = 2u * l.oI;
    const int gi = t.i0 + l.li, gj = t.j0 + l.lj;
    const bool col_interior =
        l.valid && gi >= 1 && gi <= a.nrows - 2 && gj >= 1 && gj <= a.ncols - 2;
    // interior levels 1 .. nz-2 of interior columns
    const int plo = col_interior ? l.pipj + 1 : 0x3fffffff, phi = l.pipj + nz - 1;
    const unsigned st = 2u * l.step;
    const float wx = a.wx, wy = a.wy, wz = a.wz;
    if (tid == 0) {
        pub[2 * NT] = pub[PB + 2 * NT] = make_float2(0.0f, PFAST_INF_F);
        if constexpr (FLOW) rec[0] = 0;
        else start_init(rec);
    }

    // the halo faces are loaded by the whole block and published step by step
    // (HaloStage)
    const HS hs = make_halo_stage<TI, TJ, C, L>(a, t, s1, s2, s3, 2u);
    auto ld2 = [&](unsigned o) { return fload2<FLOW>(RLT, o); };
    auto load_batch = [&](int pb, float2 *ow, float2 *hl) {
        const unsigned dm = (unsigned)(pb - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++) ow[u] = ld2(oI + dm + (unsigned)(u + 2) * st);
        const unsigned dp = (unsigned)pb * st;
#pragma unroll
        for (int k = 0; k < HS::K; k++)
            hl[k] = hs.use[k] ? ld2(hs.addr[k] + dp) : make_float2(0.0f, 0.0f);
    };
    if constexpr (FLOW) {
        flow_wait(*fw, L - 1);
        asm volatile("" ::: "memory");
    }
    // the window at step pb0 (as in fwd_tile)
    const int pb0 = FLOW ? 0 : k0 / L * L;
    const int m0 = pb0 - l.pipj;
    auto lev = [&](int m) { return l.valid && (unsigned)m < (unsigned)nz; };
    const float2 zero = make_float2(0.0f, 0.0f);
    float2 vprev, vcur, vnext;
    if constexpr (FLOW) {
        vprev = zero;
        vcur = l.valid && m0 == 0 ? ld2(oI) : zero;
        vnext = l.valid && m0 >= -1 ? ld2(oI + (unsigned)(m0 + 1) * st) : zero;
    } else {
        vprev = lev(m0 - 1) ? ld2(oI + (unsigned)(m0 - 1) * st) : zero;
        vcur = lev(m0) ? ld2(oI + (unsigned)m0 * st) : zero;
        vnext = lev(m0 + 1) ? ld2(oI + (unsigned)(m0 + 1) * st) : zero;
    }
    float2 qa[1][L], qb[1][L], ha[HS::K], hb[HS::K];
    load_batch(pb0, qa[0], ha);
    if constexpr (FLOW) flow_prefetch(*fw);
    __syncthreads();

    bool changed = false;
    int mlo = 0x7fffffff, mhi = -1;  // this lane's changed levels, batch granularity
    float sval[L];
    bool chg[L];
#pragma unroll
    for (int u = 0; u < L; u++) {
        sval[u] = 0.0f;
        chg[u] = false;
    }
    const int pstop = k1 + 2 + t.span;  // as in fwd_tile
    auto batch = [&](int pb, float2 (*q)[L], const float2 *h, float2 (*nq)[L], float2 *nh) -> bool {
        const int b = (pb - pb0) / L;
        if constexpr (!FLOW)
            if (tid == 0) rec[11 + (b + 1) % 3] = 0;
        wait_vm_all();
        if constexpr (FLOW) {
            flow_wait(*fw, pb + 2 * L - 1);
            flow_prefetch(*fw);
        }
        // an opaque copy of the batch start for the addresses: otherwise the compiler
        // keeps every per-step address of the queues in a register of its own
        int pbo = pb;
        if constexpr (FLOW) asm volatile("" : "+s"(pbo) : : "memory");
        else asm volatile("" : "+s"(pbo));
        {
            const unsigned d = (unsigned)(pbo - L - l.pipj) * st;
#pragma unroll
            for (int u = 0; u < L; u++)
                if (chg[u]) fstore<FLOW>(RLT, oI + d + (unsigned)u * st, sval[u]);
        }
        load_batch(pbo + L, nq[0], nh);
#pragma unroll
        for (int u = 0; u < L; u++) {
            const int p = pb + u;
            float2 *P = pub + (u & 1) * PB;
            P[tid] = vprev;
            P[NT + tid] = vnext;
#pragma unroll
            for (int k = 0; k < HS::K; k++)
                if (hs.uk[k] == u) P[2 * NT + 1 + hs.rk[k]] = h[k];
            lds_barrier();
            if constexpr (!FLOW)
                if (u == 0 && b > 0 && pb >= pstop && rec[11 + (b - 1) % 3] == 0) return true;
            if constexpr (FLOW) {
                if (u == 0 && tid == 0 && pbo > L) st_agent(self, fw->base + (unsigned)(pbo - L));
            }
            // branch-free: the (lambda, T) pairs of the neighbours come in one LDS read
            // each (halo pairs too), the update is computed in every lane and selected
            // afterwards
            const float2 xm = P[l.rxm], xp = P[l.rxp], ym = P[l.rym], yp = P[l.ryp];
            // depth k-1 / k+1 by absolute depth (levels m-1 / m+1 are depths
            // k-s3 / k+s3): every ordering sums the terms in one order, so a
            // stationary node stays bitwise stationary
            const float2 zm = S3POS ? vprev : vnext;
            const float2 zp = S3POS ? vnext : vprev;
            const float tc = vcur.y;
            const float txm = fabsf(xm.y), txp = fabsf(xp.y), tym = fabsf(ym.y),
                        typ = fabsf(yp.y), tzm = fabsf(zm.y), tzp = fabsf(zp.y);
            // any stencil traveltime at INF: one compare of the maximum (three max3)
            const float tmax = fmaxf(fmaxf(fmaxf(txm, txp), tym),
                                     fmaxf(fmaxf(typ, tzm), fmaxf(tzp, tc)));
            float num, den;
            adj_first<UNI>(txp - tc, xp.x, wx, num, den);
            adj_acc<UNI>(txm - tc, xm.x, wx, num, den);
            adj_acc<UNI>(typ - tc, yp.x, wy, num, den);
            adj_acc<UNI>(tym - tc, ym.x, wy, num, den);
            adj_acc<UNI>(tzp - tc, zp.x, wz, num, den);
            adj_acc<UNI>(tzm - tc, zm.x, wz, num, den);
            // the reference's threshold applies to the weighted denominator: with
            // uniform spacing den * wx > 1e-9 exactly when den > adj_thr
            const float thr = UNI ? a.adj_thr : 1e-9f;
            // computed in every lane and selected: no branch around the update
            float r = num * frcp(den);
            asm volatile("" : "+v"(r));
            const float nw = tmax < PFAST_INF_F && den > thr ? r : 0.0f;
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
        if constexpr (!FLOW) {
            bool bc = false;
#pragma unroll
            for (int u = 0; u < L; u++) bc = bc || chg[u];
            if (bc) {
                mlo = min(mlo, pb - l.pipj);
                mhi = pb + L - 1 - l.pipj;
            }
            if (__ballot(bc) && (tid & 63) == 0) atomicOr(&rec[11 + b % 3], 1);
        }
        return false;
    };
    // steps 0 .. pend-1, in batches of L (the queues alternate); the loop ends after
    // the batch holding step pend-1 (uniform over the block)
    const int pend = nz + t.span;
    int pdone = pend;  // the step the sweep ended at
    for (int pb = pb0;; pb += 2 * L) {
        if (batch(pb, qa, ha, qb, hb)) {
            pdone = pb;
            break;
        }
        if (pb + L >= pend) break;
        if (batch(pb + L, qb, hb, qa, ha)) {
            pdone = pb + L;
            break;
        }
        if (pb + 2 * L >= pend) break;
    }
    {  // changed nodes of the last batch, the one holding step pend-1
        const int pl = (pend - 1) / L * L;
        const unsigned d = (unsigned)(pl - l.pipj) * st;
#pragma unroll
        for (int u = 0; u < L; u++)
            if (chg[u]) fstore<FLOW>(RLT, oI + d + (unsigned)u * st, sval[u]);
    }
    if constexpr (FLOW) wait_vm_all();
    const unsigned long long bm = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && bm) atomicOr(&a.changed[t.chunk], chunk_bits(bm, C));
    if constexpr (FLOW) {
        count_swept(a, t, nz);
        record_tile(a, six, changed ? 16u | col_faces(l.valid, l.li, l.lj, t.TIe, t.TJe) : 0u,
                    rec);
    } else {
        const unsigned faces = col_faces(l.valid, l.li, l.lj, t.TIe, t.TJe);
        if (a.starts) start_ranges(rec, max(mlo, 0), min(mhi, nz - 1), faces);
        count_swept(a, t, max(min(pdone, nz) - pb0, 0));
        record_tile(a, six, changed ? 16u | faces : 0u, rec);
        if (a.starts && six >= 0) start_update<TI, TJ>(a, t, six, rec, s3);
    }
}

template <int TI, int TJ, int C, int L, bool UNI, bool S3POS>
__global__ __launch_bounds__(TI *TJ *C) void adj3d_col(SweepArgs a, int D, int s1, int s2)
{
    __shared__ float2 pub[2 * HaloStage<TI, TJ, C, L>::PB];
    __shared__ int rec[kStartRec];
    const TileCtx t = decode_tile<TI, TJ>(a, D, s1, s2);
    if (!t.live || !phase_live(a, t.chunk)) return;
    const int six = skip_index<TI, TJ>(a, t);
    if (skip_tile<TI, TJ>(a, t, six)) return;
    const int k0 = start_level(a, six);
    if (k0 >= a.nz) return;
    const int k1 = last_dirty_level(a, six);
    phase_epoch(a);
    adj_tile<TI, TJ, C, L, UNI, S3POS, false>(a, t, s1, s2, six, k0, k1, pub, rec, nullptr,
                                              nullptr);
}

// ======================================================= dataflow kernels
// One launch sweeps every active chunk through the 8 orderings of a cycle, one block
// per task: each block claims the next task number in the order of flow.h as it
// starts and waits on its dependencies batch by batch, so tiles of later diagonals
// and orderings start as soon as the values they read are final instead of at a
// launch boundary.  (A block that runs several tasks in a loop would hold the loop's
// invariants in registers across the tile sweep: one task per block keeps the sweep
// near the register count of the per-diagonal kernels.  The kernels are held to 64
// VGPRs, 8 waves per SIMD; the few values that then spill are set up once per task.)

struct FlowTaskLds {
    unsigned t;  // claimed task number (>= ntasks: none left)
    int o, ti, tj, chunk, D, skip;
    int ok;      // steps whose memory operations the dependencies allow (FlowWait)
    int ndep;
    pflow::Dep d[pflow::kMaxDeps];
};

// A value every lane read from the same LDS word, as a scalar.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Device phase state (FlowArgs::phase): this cycle's launch epoch and cycle number.
__device__ __forceinline__ void flow_phase(FlowArgs &f)
{
    if (f.phase) {
        const int cyc = f.phase[0];
        f.base += (unsigned)(2 * cyc) << pflow::kEpochShift;
        f.cycle = cyc - 1;
    }
}

// Progress word of tile (ti, tj) of a chunk in ordering o.
__device__ __forceinline__ unsigned *flow_word(const SweepArgs &a, const FlowArgs &f, int chunk,
                                               int o, int ti, int tj)
{
    return f.prog + ((unsigned)chunk * 8u + (unsigned)o) * ((unsigned)a.nti * a.ntj) +
           (unsigned)ti * a.ntj + (unsigned)tj;
}

// Thread 0 waits until dependency word w is finished (the exact tile-skip test reads
// the neighbours' final records).
__device__ __forceinline__ void flow_wait_done(const FlowArgs &f, const unsigned *w)
{
    const unsigned long long t0 = wall_clock64();
    while (ld_agent(w) < f.base + pflow::kDone) {
        if (ld_agent(f.abort)) return;
        if (wall_clock64() - t0 > f.timeout) {
            st_agent(f.abort, 1);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Claim this block's task into s (all threads).  Blocks take task numbers from the
// counter in the order they start, not by block index, so a block only ever waits on
// tasks that started before it; the last claim returns the counter to zero for the
// next launch.  A task whose sweep would change nothing (exact tile skipping) is
// finished here (s.skip), as is every task once the launch has aborted.
template <int TI, int TJ>
__device__ __forceinline__ void flow_claim(const SweepArgs &a, const FlowArgs &f, FlowTaskLds &s)
{
    if (threadIdx.x == 0) {
        unsigned t = atomicAdd(f.ticket, 1u);
        if (t == gridDim.x - 1) st_agent(f.ticket, 0u);
        if (ld_agent(f.abort)) t = f.ntasks;
        s.t = t;
        s.ok = -1;
        if (t < f.ntasks) {
            const pflow::Task k = pflow::decode(t, f.o0, a.nact, a.nti, a.ntj);
            s.o = k.o;
            s.ti = k.ti;
            s.tj = k.tj;
            s.D = k.D;
            s.chunk = a.active[k.slot];
            const int n = pflow::deps(k.o, f.o0, k.ti, k.tj, a.nti, a.ntj, TI, TJ, a.nrows, a.ncols,
                                      s.d);
            s.ndep = n;
            s.skip = 0;
            if (!phase_live(a, s.chunk)) {
                s.skip = 2;  // retired chunk: nothing of it runs, nothing waits on it
            } else if (a.tmod) {
                const int ix = (s.chunk * a.nti + k.ti) * a.ntj + k.tj;
                const int cl = ld_agent(&a.tclean[ix * kTileClean + k.o]);
                if (cl > 0 && ld_agent(&a.tmod[ix * 5]) < cl) {
                    // a candidate: the exact test needs the neighbours' final records
                    for (int i = 0; i < n; i++)
                        flow_wait_done(f, flow_word(a, f, s.chunk, s.d[i].o, s.d[i].ti, s.d[i].tj));
                    SweepArgs at = a;
                    at.o = k.o;
                    TileCtx tc;
                    tc.chunk = s.chunk;
                    tc.i0 = k.ti * TI;
                    tc.j0 = k.tj * TJ;
                    s.skip = skip_tile<TI, TJ, true>(at, tc, ix) ? 1 : 0;
                }
            }
            if (s.skip == 1) st_agent(flow_word(a, f, s.chunk, k.o, k.ti, k.tj), f.base + pflow::kDone);
        }
    }
}

// Per-task setup shared by both dataflow kernels: the tile, the launch arguments of
// its ordering, this lane's dependency and the task's own progress word.
template <int TI, int TJ>
__device__ __forceinline__ void flow_task(const SweepArgs &a, const FlowArgs &f,
                                          FlowTaskLds &s, TileCtx &t, SweepArgs &at,
                                          FlowWait &fw, unsigned *&self)
{
    const int o = uni(s.o), ti = uni(s.ti), tj = uni(s.tj), chunk = uni(s.chunk);
    t.live = true;
    t.chunk = chunk;
    t.i0 = ti * TI;
    t.j0 = tj * TJ;
    t.TIe = min(TI, a.nrows - t.i0);
    t.TJe = min(TJ, a.ncols - t.j0);
    t.span = t.TIe + t.TJe - 2;
    at = a;
    at.o = o;
    at.epoch = 1 + ((f.cycle * 8 + o) * f.nd + uni(s.D));
    self = flow_word(a, f, chunk, o, ti, tj);
    const int lane = threadIdx.x & 63;
    const bool use = lane < uni(s.ndep);
    const pflow::Dep d = s.d[use ? lane : 0];
    // a lane without a dependency watches the state's all-ones word
    const unsigned ntile = (unsigned)a.nti * a.ntj;
    const unsigned words = f.nprog + 4u;
    const unsigned idx = use ? ((unsigned)chunk * 8u + (unsigned)d.o) * ntile +
                                   (unsigned)d.ti * a.ntj + (unsigned)d.tj
                             : words - 1u;
    const unsigned add = use ? pflow::need_add(d.kind, d.width, TI, TJ) : 0u;
    fw.R = make_rsrc(f.prog, 4u * words);
    fw.w = idx << 6 | (add >= pflow::kDone ? kCodeDone : add);
    fw.base = f.base;
    fw.dv = 0;
    fw.abort_off = 4u * (f.nprog + 2u);
    fw.ok = &s.ok;
    fw.watcher = uni(threadIdx.x >> 6) == 0;
    fw.timeout = f.timeout;
    fw.smax_cap = a.nz + t.span - 1;
    flow_prefetch(fw);
}

// Finish a task: every wave's stores are complete (each waited before the block's
// last barrier); thread 0's record stores complete before the progress word says so.
__device__ __forceinline__ void flow_finish(const FlowArgs &f, unsigned *self)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        wait_vm_all();
        st_agent(self, f.base + pflow::kDone);
    }
}

template <int TI, int TJ, int C, int L, bool UNI, bool S3POS>
__global__ __launch_bounds__(TI *TJ *C) __attribute__((amdgpu_waves_per_eu(8))) void fwd3d_flow(SweepArgs a, FlowArgs f)
{
    flow_phase(f);
    __shared__ float pub[2 * HaloStage<TI, TJ, C, L>::PB];
    __shared__ int rec[1];
    __shared__ FlowTaskLds s;
    flow_claim<TI, TJ>(a, f, s);
    __syncthreads();
    if ((unsigned)uni((int)s.t) >= f.ntasks || uni(s.skip)) return;
    TileCtx t;
    SweepArgs at;
    FlowWait fw;
    unsigned *self;
    flow_task<TI, TJ>(a, f, s, t, at, fw, self);
    const int o = at.o;
    const int six = a.tmod ? (t.chunk * a.nti + t.i0 / TI) * a.ntj + t.j0 / TJ : -1;
    if (at.chunks[t.chunk].srcinit > 0)
        fwd_tile<TI, TJ, C, L, UNI, true, true>(at, t, pflow::s1_of(o), pflow::s2_of(o),
                                                S3POS ? 1 : -1, six, 0, a.nz - 1, pub, rec, &fw,
                                                self);
    else
        fwd_tile<TI, TJ, C, L, UNI, true, false>(at, t, pflow::s1_of(o), pflow::s2_of(o),
                                                 S3POS ? 1 : -1, six, 0, a.nz - 1, pub, rec, &fw,
                                                 self);
    flow_finish(f, self);
}

template <int TI, int TJ, int C, int L, bool UNI, bool S3POS>
__global__ __launch_bounds__(TI *TJ *C) __attribute__((amdgpu_waves_per_eu(8))) void adj3d_flow(SweepArgs a, FlowArgs f)
{
    flow_phase(f);
    __shared__ float2 pub[2 * HaloStage<TI, TJ, C, L>::PB];
    __shared__ int rec[1];
    __shared__ FlowTaskLds s;
    flow_claim<TI, TJ>(a, f, s);
    __syncthreads();
    if ((unsigned)uni((int)s.t) >= f.ntasks || uni(s.skip)) return;
    TileCtx t;
    SweepArgs at;
    FlowWait fw;
    unsigned *self;
    flow_task<TI, TJ>(a, f, s, t, at, fw, self);
    const int o = at.o;
    const int six = a.tmod ? (t.chunk * a.nti + t.i0 / TI) * a.ntj + t.j0 / TJ : -1;
    adj_tile<TI, TJ, C, L, UNI, S3POS, true>(at, t, pflow::s1_of(o), pflow::s2_of(o), six, 0,
                                             a.nz - 1, pub, rec, &fw, self);
    flow_finish(f, self);
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

// Phase state: cycle 0, every chunk live, no source converged.
__global__ void phase_init(int *__restrict__ phase, int nchunk, int nsrc)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 2) phase[e] = 0;
    else if (e < 2 + nchunk) phase[e] = 1;
    else if (e < 2 + nchunk + nsrc) phase[e] = 0;
}

__global__ void cycle_begin(int *__restrict__ phase, unsigned *__restrict__ changed, int nchunk)
{
    for (int i = threadIdx.x; i < nchunk; i += blockDim.x) changed[i] = 0u;
    if (threadIdx.x == 0) phase[0] += 1;
}

// A live chunk's sources that changed nothing this cycle converge now (the first such
// cycle counts); a chunk none of whose sources changed retires.
__global__ void cycle_end(int *__restrict__ phase, const unsigned *__restrict__ changed, int nchunk,
                          int nsrc)
{
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    int *live = phase + 2, *cyc = phase + 2 + nchunk;
    if (ch >= nchunk || !live[ch]) return;
    const unsigned bits = changed[ch];
    for (int c = 0; c < PFAST_CHUNK; c++) {
        const int src = ch * PFAST_CHUNK + c;
        if (src < nsrc && !((bits >> c) & 1u) && cyc[src] == 0) cyc[src] = phase[0];
    }
    if (!bits) live[ch] = 0;
}

// Fixed box around each source (half-width R > 0 nodes, clipped to the grid): FP64
// traveltimes of the linear velocity model v(x) = v_s + G.(x - x_s) fitted at the
// source node (G from centred differences of vp, one-sided at a grid edge), whose
// rays are circular arcs with the closed-form traveltime
//   T = acosh(1 + |G|^2 d^2 / (2 v_s v(x))) / |G|,   d = |x - x_s|,
// evaluated as log1p(u + sqrt(u (u + 2))) / |G| (no cancellation for small u), and
// d / sqrt(v_s v(x)) = d / v_s where G = 0.  Exact for a model linear in depth.  One
// block per (chunk, source) slot; padding sources (node -1) are left alone.
template <int C>
__global__ void init_src_box(float *__restrict__ F, const double *__restrict__ vp,
                             const int *__restrict__ srcnode, long N, int R, int nrows,
                             int ncols, int nz, double dx, double dy, double dz)
{
    const int slot = blockIdx.x;
    const int sn = srcnode[slot];
    if (sn < 0) return;
    const long plane = (long)nrows * ncols;
    const int sk = (int)(sn / plane), sc = (int)(sn % plane);
    const int si = sc / ncols, sj = sc % ncols;
    const double vs = vp[sn];
    auto slope = [&](int n, int pos, long stride, double h) {
        if (n < 2) return 0.0;
        const int lo = pos > 0 ? pos - 1 : pos, hi = pos < n - 1 ? pos + 1 : pos;
        return (vp[sn + (hi - pos) * stride] - vp[sn + (lo - pos) * stride]) / ((hi - lo) * h);
    };
    const double gx = slope(ncols, sj, 1, dx), gy = slope(nrows, si, ncols, dy),
                 gz = slope(nz, sk, plane, dz);
    const double g2 = gx * gx + gy * gy + gz * gz, g = sqrt(g2);
    const int W = 2 * R + 1;
    float *Fs = F + (long)(slot / C) * N * C + (slot % C);
    for (int e = threadIdx.x; e < W * W * W; e += blockDim.x) {
        const int dj = e % W - R, di = (e / W) % W - R, dk = e / (W * W) - R;
        const int j = sj + dj, i = si + di, k = sk + dk;
        if (j < 0 || j >= ncols || i < 0 || i >= nrows || k < 0 || k >= nz) continue;
        const double X = dj * dx, Y = di * dy, Z = dk * dz;
        const double d2 = X * X + Y * Y + Z * Z;
        const double vr = fmax(vs + gx * X + gy * Y + gz * Z, 0.1 * vs);
        const double u = g2 * d2 / (2.0 * vs * vr);
        const double T = u > 1e-30 ? log1p(u + sqrt(u * (u + 2.0))) / g : sqrt(d2 / (vs * vr));
        Fs[((long)k * plane + (long)i * ncols + j) * C] = (float)T;
    }
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

// Start of an adjoint: LT = (0, T) for every node and source of every chunk (T
// already carries the receiver marks).  One thread per (chunk, node, source) value.
__global__ void adj_pack(float2 *__restrict__ LT, const float *__restrict__ T, long total)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    LT[e] = make_float2(0.0f, T[e]);
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
// at the source excluded -- resolved on the host).  lambda goes to the .x of the
// interleaved (lambda, T) field.
__global__ void adj_bc(float2 *__restrict__ LT, const float *__restrict__ T,
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
    LT[(size_t)(s / C) * N * C + (size_t)n * C + (s % C)].x = (float)lam;
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

// grad[n] = -(sum_s lambda_s[n]) / vp[n]^3, Kahan-summed in FP64 over sources in
// source order.  A node's C (lambda, T) pairs of a chunk are one 16-byte-aligned
// record, read with 16-byte loads (two pairs each).
template <int C>
__global__ void grad_reduce(const float2 *__restrict__ LT, const double *__restrict__ vp,
                            double *__restrict__ g, long N, int nchunk, int nsrc)
{
    static_assert(C % 2 == 0, "whole 16-byte loads per record");
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double sum = 0.0, comp = 0.0;
    for (int ch = 0; ch < nchunk; ch++) {
        const float4 *p = reinterpret_cast<const float4 *>(LT + ((size_t)ch * N + n) * C);
        float v[C];
#pragma unroll
        for (int q = 0; q < C / 2; q++) {
            const float4 w = p[q];
            v[2 * q] = w.x;
            v[2 * q + 1] = w.z;
        }
        const int m = nsrc - ch * C;  // padding sources of the last chunk are left out
#pragma unroll
        for (int c = 0; c < C; c++) {
            if (c < m) {
                const double y = (double)v[c] - comp;
                const double tt = sum + y;
                comp = (tt - sum) - y;
                sum = tt;
            }
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

// Widen one source's field to FP64: a traveltime field (rec = 1) or the lambdas of
// the interleaved (lambda, T) field (rec = 2).  Traveltimes (strip_sign) may carry
// the receiver marks of an adjoint in progress; adjoint fields are signed.
__global__ void gather_field(const float *__restrict__ F, int s, long N, int C, int rec,
                             int strip_sign, double *__restrict__ out)
{
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float v = F[(((size_t)(s / C) * N + n) * C + (s % C)) * rec];
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

// Sweep kernel variants (tile rows x columns, batch depth of the forward / adjoint
// kernel), all of the column-window schedule.  Variant 0 is the default; the others
// stay selectable for measurement and for the autotuner.  The adjoint keeps a
// (lambda, T) pair per queued value, so its register count grows fast with the batch
// depth (8x8 tiles: 52 VGPRs at L2, 71 at L4, 90 at L6, which leaves 7 or 5 waves per
// SIMD instead of 8).  Tiles of 1024 threads (8x16, 16x8) and batch depth 8 measured 1.4-2.8x slower
// on the 400-source workload and are not built.
#define PFAST_VARIANTS(X)                                                                 \
    X(8, 8, 4, 2) X(8, 8, 4, 4) X(8, 8, 6, 6) X(8, 8, 2, 2) X(8, 8, 6, 2) X(8, 4, 4, 4)     \
    X(4, 8, 6, 6) X(4, 4, 4, 4) X(4, 16, 6, 6)
#define PFAST_VARIANT_ENTRY(TI, TJ, L, LA) \
    {0, TI, TJ, L, LA, "column-window " #TI "x" #TJ " L" #L "/" #LA},
static const SweepVariant kVariants[] = {PFAST_VARIANTS(PFAST_VARIANT_ENTRY)};
#undef PFAST_VARIANT_ENTRY

int sweep_variant_count() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }

const SweepVariant &sweep_variant(int v) { return kVariants[v]; }

namespace {
template <int TI, int TJ, int L, int LA>
void launch_col(hipStream_t st, const SweepArgs &a, int adjoint, int uniform, int D, int s1,
                int s2, int s3, dim3 grid)
{
    constexpr int C = PFAST_CHUNK;
    static_assert(TI * TJ * C <= 1024, "one thread per column and source");
    const dim3 block(TI * TJ * C);
    if (adjoint) {
        if (uniform) {
            if (s3 > 0)
                hipLaunchKernelGGL((adj3d_col<TI, TJ, C, LA, true, true>), grid, block, 0, st, a, D, s1, s2);
            else
                hipLaunchKernelGGL((adj3d_col<TI, TJ, C, LA, true, false>), grid, block, 0, st, a, D, s1, s2);
        } else {
            if (s3 > 0)
                hipLaunchKernelGGL((adj3d_col<TI, TJ, C, LA, false, true>), grid, block, 0, st, a, D, s1, s2);
            else
                hipLaunchKernelGGL((adj3d_col<TI, TJ, C, LA, false, false>), grid, block, 0, st, a, D, s1, s2);
        }
    } else if (uniform) {
        hipLaunchKernelGGL((fwd3d_col<TI, TJ, C, L, true>), grid, block, 0, st, a, D, s1, s2, s3);
    } else {
        hipLaunchKernelGGL((fwd3d_col<TI, TJ, C, L, false>), grid, block, 0, st, a, D, s1, s2, s3);
    }
}
}  // namespace

hipError_t launch_sweep3d(hipStream_t st, const SweepArgs &a, int variant, int adjoint,
                          int uniform, int D, int s1, int s2, int s3, int ntiles_diag)
{
    const dim3 grid((unsigned)(ntiles_diag * a.nact_p8));
    int v = 0;
#define PFAST_VARIANT_CASE(TI, TJ, L, LA)                                           \
    if (variant == v++) {                                                            \
        launch_col<TI, TJ, L, LA>(st, a, adjoint, uniform, D, s1, s2, s3, grid);     \
        return hipGetLastError();                                                    \
    }
    PFAST_VARIANTS(PFAST_VARIANT_CASE)
#undef PFAST_VARIANT_CASE
    return hipErrorInvalidValue;
}

namespace {
template <int TI, int TJ, int L, int LA, bool S3POS>
const void *flow_kernel(int adjoint, int uniform)
{
    constexpr int C = PFAST_CHUNK;
    if (adjoint)
        return uniform ? reinterpret_cast<const void *>(&adj3d_flow<TI, TJ, C, LA, true, S3POS>)
                       : reinterpret_cast<const void *>(&adj3d_flow<TI, TJ, C, LA, false, S3POS>);
    if (uniform) return reinterpret_cast<const void *>(&fwd3d_flow<TI, TJ, C, L, true, S3POS>);
    return reinterpret_cast<const void *>(&fwd3d_flow<TI, TJ, C, L, false, S3POS>);
}

// The dataflow kernel of a variant for the orderings starting at o0 (0 or 4).
const void *flow_kernel_of(int variant, int adjoint, int uniform, int o0)
{
    int v = 0;
#define PFAST_VARIANT_CASE(TI, TJ, L, LA)                                               \
    if (variant == v++)                                                                  \
        return o0 < 4 ? flow_kernel<TI, TJ, L, LA, false>(adjoint, uniform)             \
                      : flow_kernel<TI, TJ, L, LA, true>(adjoint, uniform);
    PFAST_VARIANTS(PFAST_VARIANT_CASE)
#undef PFAST_VARIANT_CASE
    return nullptr;
}

// Zero the progress words, the claim counter and the abort word; the last word of
// the state reads all ones (the dependency of lanes that watch none).
__global__ void flow_reset(unsigned *__restrict__ w, long n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) w[e] = e == n - 1 ? ~0u : 0u;
}
}  // namespace

int flow_grid(int variant, int adjoint, int uniform)
{
    // (device, variant, kind) -> resident blocks x CUs, measured once
    static std::mutex mu;
    static std::map<long, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    const int kind = (adjoint ? 2 : 0) + (uniform ? 1 : 0);
    const long key = ((long)dev * 64 + variant) * 4 + kind;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const void *k = flow_kernel_of(variant, adjoint, uniform, 0);
    if (!k) return 0;
    const SweepVariant &sv = kVariants[variant];
    int nb = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, sv.ti * sv.tj * PFAST_CHUNK, 0) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const int g = std::max(1, nb) * std::max(1, cus);
    cache[key] = g;
    return g;
}

hipError_t launch_sweep3d_flow(hipStream_t st, const SweepArgs &a, const FlowArgs &f,
                               int variant, int adjoint, int uniform)
{
    if (variant < 0 || variant >= sweep_variant_count()) return hipErrorInvalidValue;
    if (f.ntasks == 0) return hipSuccess;
    const SweepVariant &sv = kVariants[variant];
    const dim3 grid(f.ntasks);
    const dim3 block((unsigned)(sv.ti * sv.tj * PFAST_CHUNK));
    SweepArgs ac = a;
    FlowArgs fc = f;
    void *args[] = {&ac, &fc};
    return hipLaunchKernel(flow_kernel_of(variant, adjoint, uniform, f.o0), grid, block, args, 0,
                           st);
}

size_t flow_words(int nchunk, int ntile) { return (size_t)nchunk * 8 * ntile + 4; }

hipError_t launch_flow_reset(hipStream_t st, unsigned *state, size_t words)
{
    hipLaunchKernelGGL(flow_reset, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st, state,
                       (long)words);
    return hipGetLastError();
}

namespace {
__global__ void starts_reset(int *__restrict__ tclean, const int *__restrict__ act, int nact,
                             int ntile)
{
    // (active slot, tile, start or end level of an ordering)
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nact * ntile * 16) return;
    const long r = e / 16;
    const int tile = (int)(r % ntile), slot = (int)(r / ntile);
    tclean[((long)act[slot] * ntile + tile) * kTileClean + 8 + (int)(e % 16)] = 0;
}
}  // namespace

hipError_t launch_starts_reset(hipStream_t st, int *tclean, const int *act, int nact, int ntile)
{
    const long n = (long)nact * ntile * 16;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(starts_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, tclean,
                       act, nact, ntile);
    return hipGetLastError();
}

hipError_t launch_init_fwd(hipStream_t st, float *F, const int *srcnode, long N, int nchunk)
{
    const long tot = N * nchunk;
    hipLaunchKernelGGL(init_fwd<PFAST_CHUNK>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       st, F, srcnode, N, nchunk);
    return hipGetLastError();
}

hipError_t launch_phase_init(hipStream_t st, int *phase, int nchunk, int nsrc)
{
    const int n = 2 + nchunk + nsrc;
    hipLaunchKernelGGL(phase_init, dim3((n + 255) / 256), dim3(256), 0, st, phase, nchunk, nsrc);
    return hipGetLastError();
}

hipError_t launch_cycle_begin(hipStream_t st, int *phase, unsigned *changed, int nchunk)
{
    hipLaunchKernelGGL(cycle_begin, dim3(1), dim3(256), 0, st, phase, changed, nchunk);
    return hipGetLastError();
}

hipError_t launch_cycle_end(hipStream_t st, int *phase, const unsigned *changed, int nchunk,
                            int nsrc)
{
    hipLaunchKernelGGL(cycle_end, dim3((nchunk + 255) / 256), dim3(256), 0, st, phase, changed,
                       nchunk, nsrc);
    return hipGetLastError();
}

hipError_t launch_init_src_box(hipStream_t st, float *F, const double *vp, const int *srcnode,
                               long N, int nchunk, int R, int nrows, int ncols, int nz,
                               double dx, double dy, double dz)
{
    if (R <= 0) return hipSuccess;
    hipLaunchKernelGGL(init_src_box<PFAST_CHUNK>, dim3((unsigned)(nchunk * PFAST_CHUNK)),
                       dim3(256), 0, st, F, vp, srcnode, N, R, nrows, ncols, nz, dx, dy, dz);
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

hipError_t launch_adj_pack(hipStream_t st, float *LT, const float *T, long N, int nchunk)
{
    const long tot = N * nchunk * PFAST_CHUNK;
    hipLaunchKernelGGL(adj_pack, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       reinterpret_cast<float2 *>(LT), T, tot);
    return hipGetLastError();
}

hipError_t launch_adj_bc(hipStream_t st, float *LT, const float *T, const long *bnode,
                         const int *bsrc, const double *bobs, const double *bnrm, int nb,
                         int nrows, int ncols, int nz, double dx, double dy, double dz)
{
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(adj_bc, dim3((nb + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<float2 *>(LT), T, bnode, bsrc, bobs,
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

hipError_t launch_grad_reduce(hipStream_t st, const float *LT, const double *vp, double *g,
                              long N, int nchunk, int nsrc)
{
    hipLaunchKernelGGL(grad_reduce<PFAST_CHUNK>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                       st, reinterpret_cast<const float2 *>(LT), vp, g, N, nchunk, nsrc);
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

hipError_t launch_gather_field(hipStream_t st, const float *F, int s, long N, int rec,
                               int strip_sign, double *out)
{
    hipLaunchKernelGGL(gather_field, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, F, s, N,
                       PFAST_CHUNK, rec, strip_sign, out);
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
