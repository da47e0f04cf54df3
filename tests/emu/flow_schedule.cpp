// flow_schedule.cpp -- CPU model of the dataflow schedule of the persistent batched
// 3D sweep kernels (pfast_amd/csrc/batch3d.hip, fwd3d_flow / adj3d_flow), built on
// the same task numbering and dependency rules (pfast_amd/csrc/flow.h).
//
// Test infrastructure.  A seeded discrete-event simulation: `nblocks` resident
// blocks claim task numbers from a counter, and at every tick a randomly chosen
// block advances by one event of the kernels' batch structure:
//   prologue: dependency check for steps < L, then the first loads;
//   batch pb: wait for the block's outstanding memory operations, dependency check
//             for steps up to pb + 2L - 1 (blocked until it holds), stores of the
//             previous batch's steps, loads of the next batch's, progress pb - L;
//   end:      stores of the last batch, wait, progress DONE.
// Lanes follow the kernels' column-window geometry (level m = p - pipj at step p,
// own column read two levels ahead, tile-edge halos from the neighbour tiles).  A
// load samples memory either when issued or when the block next waits, a store lands
// either when issued or at that wait (chosen at random), which is every outcome the
// hardware allows between issue and the wait.  Every node carries the number of the
// last ordering that processed it; each read checks that it sees exactly what the
// lexicographic sweep would: the current ordering for nodes of upstream tiles, the
// previous one for its own column ahead and for downstream tiles.  Every node
// processed is stored (a superset of the kernels' changed-node stores).
//
// Built with -DPLANT_MARGIN, -DPLANT_NO_PREV or -DPLANT_EARLY_PUBLISH it carries a
// deliberate protocol mistake, which the tests expect it to report.
//
// usage: flow_schedule nrows ncols nz TI TJ L nchunk nblocks launches seed
//   (launches alternate between orderings 0-3 and 4-7, as the kernels' do)
// prints "tasks <n> reads <r> mismatches <m> deadlock <0|1>"; exit 1 on any failure.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../pfast_amd/csrc/flow.h"

using namespace pflow;

static int nr, nc, nz, TI, TJ, L, nchunk, nblocks, nti, ntj, ntile;
static std::vector<int> ver;          // [chunk][node] last ordering (global count) processed
static std::vector<unsigned> prog;    // [chunk][8][tile]
static long nreads = 0, nmis = 0;
static std::mt19937_64 rng;

static int rnd(int n) { return (int)(rng() % (unsigned long long)n); }

struct Lane {
    int li, lj, gi, gj, pipj;
    int hx, hy;   // halo node column / row (-1: none)
    bool hxup, hyup;
};

struct PendingLoad {
    long node;
    int chunk, expect;
};
struct PendingStore {
    long node;
    int chunk, value;
};

struct Block {
    bool alive = true, intask = false;
    Task k{};
    int chunk = 0, g = 0, i0 = 0, j0 = 0, TIe = 0, TJe = 0, pend = 0;
    std::vector<Lane> lanes;
    int ndep = 0;
    Dep dep[kMaxDeps];
    int phase = 0;   // 0 prologue, 1 batches, 2 end
    int pb = 0;      // next batch start
    std::vector<PendingLoad> loads;
    std::vector<PendingStore> stores;
    std::vector<long> processed_prev;  // nodes processed in the previous batch (stored now)
};

static unsigned *word(int chunk, int o, int ti, int tj)  // o: 0..7
{
    return &prog[((size_t)chunk * 8 + o) * ntile + (size_t)ti * ntj + tj];
}

static void do_read(const PendingLoad &l)
{
    nreads++;
    const int v = ver[(size_t)l.chunk * nr * nc * nz + l.node];
    if (v != l.expect) {
        if (nmis < 10)
            fprintf(stderr, "mismatch: chunk %d node %ld version %d expected %d\n", l.chunk, l.node, v,
                    l.expect);
        nmis++;
    }
}

static void issue_load(Block &b, long node, int expect)
{
    PendingLoad l{node, b.chunk, expect};
    if (rnd(2)) do_read(l);
    else b.loads.push_back(l);
}

static void issue_store(Block &b, long node)
{
    PendingStore s{node, b.chunk, b.g};
    if (rnd(2)) ver[(size_t)s.chunk * nr * nc * nz + s.node] = s.value;
    else b.stores.push_back(s);
}

static void wait_all(Block &b)
{
    // outstanding operations complete in any order
    while (!b.loads.empty() || !b.stores.empty()) {
        const int nl = (int)b.loads.size(), ns = (int)b.stores.size();
        const int pick = rnd(nl + ns);
        if (pick < nl) {
            do_read(b.loads[pick]);
            b.loads.erase(b.loads.begin() + pick);
        } else {
            const PendingStore s = b.stores[pick - nl];
            ver[(size_t)s.chunk * nr * nc * nz + s.node] = s.value;
            b.stores.erase(b.stores.begin() + (pick - nl));
        }
    }
}

static bool deps_ok(const Block &b, unsigned base, int smax)
{
    if (smax > b.pend - 1) smax = b.pend - 1;
    for (int d = 0; d < b.ndep; d++) {
        unsigned add = need_add(b.dep[d].kind, b.dep[d].width, TI, TJ);
#ifdef PLANT_MARGIN
        if (add < kDone && add >= 4) add -= 4;
#endif
        const unsigned v = *word(b.chunk, b.dep[d].o, b.dep[d].ti, b.dep[d].tj);
        if (v < base + need(add, smax)) return false;
    }
    return true;
}

static long node_of(int k, int i, int j) { return ((long)k * nr + i) * nc + j; }

// Memory loads a lane issues for step s: its own column two levels ahead, its halos
// at its own level (the kernels' load_batch; levels outside the column read nothing).
static void loads_for_step(Block &b, const Lane &ln, int s, int s3)
{
    const int m = s - ln.pipj;
    auto depth = [&](int lev) { return s3 > 0 ? lev : nz - 1 - lev; };
    if (m + 2 >= 0 && m + 2 < nz) issue_load(b, node_of(depth(m + 2), ln.gi, ln.gj), b.g - 1);
    if (m >= 0 && m < nz) {
        if (ln.hx >= 0) issue_load(b, node_of(depth(m), ln.gi, ln.hx), ln.hxup ? b.g : b.g - 1);
        if (ln.hy >= 0) issue_load(b, node_of(depth(m), ln.hy, ln.gj), ln.hyup ? b.g : b.g - 1);
    }
}

int main(int argc, char **argv)
{
    if (argc != 11) {
        fprintf(stderr, "usage: %s nrows ncols nz TI TJ L nchunk nblocks launches seed\n", argv[0]);
        return 2;
    }
    nr = atoi(argv[1]);
    nc = atoi(argv[2]);
    nz = atoi(argv[3]);
    TI = atoi(argv[4]);
    TJ = atoi(argv[5]);
    L = atoi(argv[6]);
    nchunk = atoi(argv[7]);
    nblocks = atoi(argv[8]);
    const int launches = atoi(argv[9]);
    rng.seed(strtoull(argv[10], nullptr, 10));
    nti = (nr + TI - 1) / TI;
    ntj = (nc + TJ - 1) / TJ;
    ntile = nti * ntj;
    ver.assign((size_t)nchunk * nr * nc * nz, 0);
    prog.assign((size_t)nchunk * 8 * ntile, 0);
    const int nd = nti + ntj - 1;
    (void)nd;
    long ntask_total = 0;
    bool deadlock = false;
    for (int launch = 0; launch < launches && !deadlock; launch++) {
        // launches alternate between the two depth directions of a cycle
        const int o0 = (launch % 2) * 4;
        const unsigned base = (unsigned)(launch + 1) << kEpochShift;
        const unsigned ntasks = 4u * ntile * nchunk;
        unsigned ticket = 0;
        std::vector<Block> blocks(nblocks);
        int alive = nblocks;
        long idle_ticks = 0;
        while (alive > 0) {
            Block &b = blocks[rnd(nblocks)];
            if (!b.alive) continue;
            bool progressed = true;
            if (!b.intask) {
                const unsigned t = ticket++;
                if (t >= ntasks) {
                    b.alive = false;
                    alive--;
                    continue;
                }
                ntask_total++;
                b.k = decode(t, o0, nchunk, nti, ntj);
                b.chunk = b.k.slot;
                b.g = (launch / 2) * 8 + b.k.o + 1;
                b.i0 = b.k.ti * TI;
                b.j0 = b.k.tj * TJ;
                b.TIe = nr - b.i0 < TI ? nr - b.i0 : TI;
                b.TJe = nc - b.j0 < TJ ? nc - b.j0 : TJ;
                b.pend = nz + b.TIe + b.TJe - 2;
                b.ndep = deps(b.k.o, o0, b.k.ti, b.k.tj, nti, ntj, TI, TJ, nr, nc, b.dep);
#ifdef PLANT_NO_PREV
                {
                    int n = 0;
                    for (int d = 0; d < b.ndep; d++)
                        if (b.dep[d].kind != kPrev) b.dep[n++] = b.dep[d];
                    b.ndep = n;
                }
#endif
                const int s1 = s1_of(b.k.o), s2 = s2_of(b.k.o);
                b.lanes.clear();
                for (int li = 0; li < b.TIe; li++)
                    for (int lj = 0; lj < b.TJe; lj++) {
                        Lane ln;
                        ln.li = li;
                        ln.lj = lj;
                        ln.gi = b.i0 + li;
                        ln.gj = b.j0 + lj;
                        ln.pipj = (s1 > 0 ? li : b.TIe - 1 - li) + (s2 > 0 ? lj : b.TJe - 1 - lj);
                        const bool uxm = lj == 0 && b.j0 > 0;
                        const bool uxp = !uxm && lj == b.TJe - 1 && b.j0 + b.TJe < nc;
                        const bool uym = li == 0 && b.i0 > 0;
                        const bool uyp = !uym && li == b.TIe - 1 && b.i0 + b.TIe < nr;
                        ln.hx = uxm ? b.j0 - 1 : uxp ? b.j0 + b.TJe : -1;
                        ln.hy = uym ? b.i0 - 1 : uyp ? b.i0 + b.TIe : -1;
                        ln.hxup = uxm ? s2 > 0 : s2 < 0;
                        ln.hyup = uym ? s1 > 0 : s1 < 0;
                        b.lanes.push_back(ln);
                    }
                b.intask = true;
                b.phase = 0;
                b.pb = 0;
                b.processed_prev.clear();
            } else if (b.phase == 0) {
                if (!deps_ok(b, base, L - 1)) {
                    progressed = false;
                } else {
                    const int s3 = s3_of(b.k.o);
                    auto depth = [&](int lev) { return s3 > 0 ? lev : nz - 1 - lev; };
                    for (const Lane &ln : b.lanes) {
                        // the window's first values: level 0 (pipj 0) and level 1 (pipj <= 1)
                        if (ln.pipj == 0) issue_load(b, node_of(depth(0), ln.gi, ln.gj), b.g - 1);
                        if (ln.pipj <= 1 && 1 - ln.pipj < nz)
                            issue_load(b, node_of(depth(1 - ln.pipj), ln.gi, ln.gj), b.g - 1);
                        for (int s = 0; s < L; s++) loads_for_step(b, ln, s, s3);
                    }
                    b.phase = 1;
                }
            } else if (b.phase == 1) {
                const int pb = b.pb;
                wait_all(b);
                if (!deps_ok(b, base, pb + 2 * L - 1)) {
                    progressed = false;  // the wait is repeated, which changes nothing
                } else {
                    const int s3 = s3_of(b.k.o);
                    auto depth = [&](int lev) { return s3 > 0 ? lev : nz - 1 - lev; };
                    for (long n : b.processed_prev) issue_store(b, n);
                    b.processed_prev.clear();
                    for (const Lane &ln : b.lanes)
                        for (int s = pb + L; s < pb + 2 * L; s++) loads_for_step(b, ln, s, s3);
#ifdef PLANT_EARLY_PUBLISH
                    const int P = pb;
#else
                    const int P = pb - L;
#endif
                    unsigned *w = word(b.chunk, b.k.o, b.k.ti, b.k.tj);
                    if (P > 0 && base + (unsigned)P > *w) *w = base + (unsigned)P;
                    // steps pb .. pb+L-1: the lanes process their levels
                    for (const Lane &ln : b.lanes)
                        for (int s = pb; s < pb + L; s++) {
                            const int m = s - ln.pipj;
                            if (m >= 0 && m < nz) b.processed_prev.push_back(node_of(depth(m), ln.gi, ln.gj));
                        }
                    b.pb = pb + L;
                    if (pb + L >= b.pend) b.phase = 2;
                }
            } else {
                for (long n : b.processed_prev) issue_store(b, n);
                b.processed_prev.clear();
                wait_all(b);
                *word(b.chunk, b.k.o, b.k.ti, b.k.tj) = base + kDone;
                b.intask = false;
            }
            if (progressed) {
                idle_ticks = 0;
            } else if (++idle_ticks > 200000L * nblocks) {
                deadlock = true;
                break;
            }
        }
    }
    printf("tasks %ld reads %ld mismatches %ld deadlock %d\n", ntask_total, nreads, nmis,
           deadlock ? 1 : 0);
    return (nmis || deadlock) ? 1 : 0;
}
