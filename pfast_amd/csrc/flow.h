// flow.h -- the dataflow schedule of the batched 3D sweeps: task numbering and the
// dependency rules of the persistent sweep kernels (batch3d.hip, *_flow).  Plain
// host/device C++ so that the CPU model of the schedule (tests/emu/flow_schedule.cpp)
// applies the very same rules.
//
// One launch runs the four orderings of one depth direction of a sweep cycle (the 8
// orderings of src/T_pfsm3d.h:158-179 or src/T_padjoint3d.h:135-163, s3 outermost)
// over every active chunk: orderings 0-3, then 4-7.  (Where the depth direction turns,
// a tile's first level is its neighbours' last, so nothing could overlap there.)  A
// task is one column tile of one chunk in one ordering; tasks are numbered (ordering,
// tile anti-diagonal D in sweep coordinates, tile on the diagonal, chunk slot), which
// orders every task after all the tasks it depends on.  Blocks claim task numbers from a counter in that
// order, so a task only ever waits on tasks already claimed by running blocks, and
// the task holding the lowest unfinished number waits on nothing unfinished: the
// launch always drains.
//
// Inside a task, thread (li, lj, c) handles level m = p - pipj of its column at step p
// (pipj its sweep-ordered position in the tile), as in the per-diagonal kernels.  A
// task publishes its progress P (steps 0 .. P-1 processed and their stores complete)
// every batch of L steps, and DONE when it ends.  Before issuing the loads and stores
// of its steps up to smax, a task requires of each dependency a progress of
//   upstream x neighbour, same ordering:   smax + (that tile's width)
//   upstream y neighbour, same ordering:   smax + (that tile's height)
//   itself and its four neighbours in the previous ordering, same depth direction:
//                                          smax + 3 + (TI + TJ - 2)
//   the same five in the previous ordering, reversed depth direction: DONE
// With these, every node a task reads holds exactly the value it holds at that point
// of the lexicographic sweep (behind neighbours updated in this ordering, ahead
// neighbours not yet), and no task overwrites a node before every reader of its
// previous value has read it; the results are those of the per-diagonal launches.
#pragma once

#if defined(__HIPCC__)
#define PFLOW_HD __host__ __device__ __forceinline__
#else
#define PFLOW_HD inline
#endif

namespace pflow {

constexpr unsigned kDone = 0xFFFFu;    // steps field of a finished task
constexpr unsigned kEpochShift = 16;   // progress word = (launch epoch << 16) | steps
constexpr int kMaxDeps = 7;

enum DepKind { kUpX = 0, kUpY = 1, kPrev = 2, kPrevFlip = 3 };

// Sweep directions of ordering o: s3 (depth) outermost, then s1 (rows), s2 (columns).
PFLOW_HD int s3_of(int o) { return o < 4 ? -1 : 1; }
PFLOW_HD int s1_of(int o) { return (o & 2) ? 1 : -1; }
PFLOW_HD int s2_of(int o) { return (o & 1) ? 1 : -1; }

// First sweep-ordered tile row on anti-diagonal D, and the number of tiles on it.
PFLOW_HD int diag_lo(int D, int ntj) { return D - (ntj - 1) > 0 ? D - (ntj - 1) : 0; }
PFLOW_HD int diag_count(int D, int nti, int ntj)
{
    const int hi = D < nti - 1 ? D : nti - 1;
    return hi - diag_lo(D, ntj) + 1;
}

struct Task {
    int o, D, ti, tj, slot;  // ordering, diagonal, tile (real coordinates), chunk slot
};

// Task number t of a launch over nact chunks whose first ordering is o0.
PFLOW_HD Task decode(unsigned t, int o0, int nact, int nti, int ntj)
{
    Task k;
    const unsigned per_o = (unsigned)nti * ntj * nact;
    k.o = o0 + (int)(t / per_o);
    unsigned r = t % per_o;
    int D = 0;
    for (;; D++) {
        const unsigned n = (unsigned)diag_count(D, nti, ntj) * nact;
        if (r < n) break;
        r -= n;
    }
    k.D = D;
    const int tt = (int)(r / nact);
    k.slot = (int)(r % nact);
    const int ti_ = diag_lo(D, ntj) + tt, tj_ = D - ti_;
    k.ti = s1_of(k.o) > 0 ? ti_ : nti - 1 - ti_;
    k.tj = s2_of(k.o) > 0 ? tj_ : ntj - 1 - tj_;
    return k;
}

struct Dep {
    int o, ti, tj, kind, width;
};

// The tasks that task (o, ti, tj) of one chunk depends on within a launch whose
// first ordering is o0.
PFLOW_HD int deps(int o, int o0, int ti, int tj, int nti, int ntj, int TI, int TJ, int nrows,
                  int ncols, Dep *d)
{
    int n = 0;
    const int uj = tj - s2_of(o), ui = ti - s1_of(o);
    if (uj >= 0 && uj < ntj) {
        const int w = ncols - uj * TJ;
        d[n++] = Dep{o, ti, uj, kUpX, w < TJ ? w : TJ};
    }
    if (ui >= 0 && ui < nti) {
        const int h = nrows - ui * TI;
        d[n++] = Dep{o, ui, tj, kUpY, h < TI ? h : TI};
    }
    if (o > o0) {
        const int kind = s3_of(o - 1) == s3_of(o) ? kPrev : kPrevFlip;
        d[n++] = Dep{o - 1, ti, tj, kind, 0};
        if (tj > 0) d[n++] = Dep{o - 1, ti, tj - 1, kind, 0};
        if (tj < ntj - 1) d[n++] = Dep{o - 1, ti, tj + 1, kind, 0};
        if (ti > 0) d[n++] = Dep{o - 1, ti - 1, tj, kind, 0};
        if (ti < nti - 1) d[n++] = Dep{o - 1, ti + 1, tj, kind, 0};
    }
    return n;
}

// Offset added to smax for a dependency of this kind (kDone: it must be finished).
PFLOW_HD unsigned need_add(int kind, int width, int TI, int TJ)
{
    if (kind == kUpX || kind == kUpY) return (unsigned)width;
    if (kind == kPrev) return 3u + (unsigned)(TI + TJ - 2);
    return kDone;
}

// Steps a dependency must have completed before a task issues the memory
// operations of its steps up to smax (smax >= 0).
PFLOW_HD unsigned need(unsigned add, int smax)
{
    if (add >= kDone) return kDone;
    const unsigned r = (unsigned)smax + add;
    return r < kDone ? r : kDone;
}

}  // namespace pflow
