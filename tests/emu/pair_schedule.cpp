// pair_schedule.cpp -- CPU check of the per-diagonal launch plan of the batched 3D
// sweeps (pfast_amd/csrc/plan.h, included as is), paired and unpaired.
//
// Test infrastructure.  For one grid cut into TI x TJ column tiles it walks the plan
// of a whole cycle (8 orderings; with pairing, orderings 2p and 2p+1 share launches)
// and checks that
//   * every (ordering, tile) is swept exactly once;
//   * no two tasks of one launch touch each other's tiles: a task writes its tile and
//     reads its tile and the four neighbours;
//   * run launch by launch, with the tasks of a launch in a random order and each
//     task sweeping its tile's nodes in the ordering's lexicographic order, the plan
//     gives the same bits as the plain lexicographic sweeps of the 8 orderings in the
//     reference's loop order (src/T_pfsm3d.h:158-160).
// The node update is a fixed, direction-dependent mix of the node and its six
// neighbours, so that any node read before or after the value the sequential sweep
// gives it changes the result.
//
// usage: pair_schedule nrows ncols nz TI TJ paired seed [lag_delta]
//   lag_delta (default 0) is added to the plan's pair lag: a negative one is a planted
//   mistake the tests expect to be reported.
// prints "launches <n> conflicts <c> mismatches <m>" and exits 1 on any problem.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../pfast_amd/csrc/plan.h"

static int nr, nc, nz, TI, TJ, nti, ntj;

static float upd(const std::vector<float> &U, int k, int i, int j)
{
    auto at = [&](int kk, int ii, int jj, float dflt) {
        if (kk < 0 || kk >= nz || ii < 0 || ii >= nr || jj < 0 || jj >= nc) return dflt;
        return U[((size_t)kk * nr + ii) * nc + jj];
    };
    const float c = at(k, i, j, 0.0f);
    float v = 0.31f * c;
    v += 0.11f * at(k, i, j - 1, c) + 0.13f * at(k, i, j + 1, c);
    v += 0.17f * at(k, i - 1, j, c) + 0.19f * at(k, i + 1, j, c);
    v += 0.07f * at(k - 1, i, j, c) + 0.02f * at(k + 1, i, j, c);
    return v;
}

// Nodes of rows [i0, i1) x columns [j0, j1) x all depths in ordering o's order.
static void sweep_box(std::vector<float> &U, int o, int i0, int i1, int j0, int j1)
{
    const int s1 = pfast::ord_s1(o), s2 = pfast::ord_s2(o), s3 = pfast::ord_s3(o);
    for (int kq = 0; kq < nz; kq++) {
        const int k = s3 < 0 ? nz - 1 - kq : kq;
        for (int iq = 0; iq < i1 - i0; iq++) {
            const int i = s1 < 0 ? i1 - 1 - iq : i0 + iq;
            for (int jq = 0; jq < j1 - j0; jq++) {
                const int j = s2 < 0 ? j1 - 1 - jq : j0 + jq;
                U[((size_t)k * nr + i) * nc + j] = upd(U, k, i, j);
            }
        }
    }
}

struct Task {
    int o, ti, tj;
};

// Tile (ti, tj) at position tt of diagonal D of ordering o, as the kernels decode it.
static Task tile_of(int o, int D, int tt)
{
    const int lo = std::max(0, D - (ntj - 1));
    const int ti_ = lo + tt, tj_ = D - ti_;
    Task t;
    t.o = o;
    t.ti = pfast::ord_s1(o) > 0 ? ti_ : nti - 1 - ti_;
    t.tj = pfast::ord_s2(o) > 0 ? tj_ : ntj - 1 - tj_;
    return t;
}

int main(int argc, char **argv)
{
    if (argc < 8) {
        fprintf(stderr, "usage: pair_schedule nrows ncols nz TI TJ paired seed [lag_delta]\n");
        return 2;
    }
    nr = atoi(argv[1]);
    nc = atoi(argv[2]);
    nz = atoi(argv[3]);
    TI = atoi(argv[4]);
    TJ = atoi(argv[5]);
    const bool paired = atoi(argv[6]) != 0;
    const unsigned seed = (unsigned)atoi(argv[7]);
    const int lag_delta = argc > 8 ? atoi(argv[8]) : 0;
    nti = (nr + TI - 1) / TI;
    ntj = (nc + TJ - 1) / TJ;
    const int nd = pfast::plan_diagonals(nti, ntj);
    const int lag = pfast::pair_lag(nti, ntj) + lag_delta;
    const int ns = pfast::plan_group_steps(paired, nti, ntj);
    const int go = pfast::plan_group_orderings(paired);

    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> uni(0.5f, 2.0f);
    std::vector<float> U0((size_t)nr * nc * nz);
    for (float &v : U0) v = uni(rng);

    // the reference's order: 8 orderings, each a lexicographic sweep of the grid
    std::vector<float> R = U0;
    for (int o = 0; o < 8; o++) sweep_box(R, o, 0, nr, 0, nc);

    std::vector<float> U = U0;
    std::vector<int> seen((size_t)8 * nti * ntj, 0);
    long launches = 0, conflicts = 0;
    for (int o = 0; o < 8; o += go) {
        // steps beyond the plan's own count cover a planted longer lag
        for (int k = 0; k < ns + std::max(0, lag_delta); k++) {
            const pfast::PlanStep p = pfast::plan_step(paired, nd, lag, o, k);
            if (p.nparts == 0) continue;
            launches++;
            std::vector<Task> tasks;
            for (int q = 0; q < p.nparts; q++) {
                if (p.D[q] < 0 || p.D[q] >= nd) continue;
                const int n = pfast::plan_diag_tiles(nti, ntj, p.D[q]);
                for (int tt = 0; tt < n; tt++) tasks.push_back(tile_of(p.o[q], p.D[q], tt));
            }
            for (size_t x = 0; x < tasks.size(); x++)
                for (size_t y = 0; y < tasks.size(); y++) {
                    if (x == y) continue;
                    const int di = std::abs(tasks[x].ti - tasks[y].ti);
                    const int dj = std::abs(tasks[x].tj - tasks[y].tj);
                    if (di + dj <= 1) conflicts++;  // y reads (or writes) x's tile
                }
            std::shuffle(tasks.begin(), tasks.end(), rng);
            for (const Task &t : tasks) {
                seen[((size_t)t.o * nti + t.ti) * ntj + t.tj]++;
                sweep_box(U, t.o, t.ti * TI, std::min(nr, (t.ti + 1) * TI), t.tj * TJ,
                          std::min(nc, (t.tj + 1) * TJ));
            }
        }
    }
    long bad = 0;
    for (size_t n = 0; n < U.size(); n++)
        if (U[n] != R[n]) bad++;
    for (int s : seen)
        if (s != 1) bad++;
    printf("launches %ld conflicts %ld mismatches %ld\n", launches, conflicts, bad);
    return conflicts || bad ? 1 : 0;
}
