// plan.h -- launch plan of the per-diagonal batched 3D sweeps (host code, no HIP).
//
// A sweep cycle takes the 8 orderings of the reference's loop nest in its order,
// src/T_pfsm3d.h:158-160 (forward) and src/T_padjoint3d.h:139-141 (adjoint): ordering
// o = 4 (s3 > 0) + 2 (s1 > 0) + (s2 > 0), depth s3 outermost, columns s2 innermost.
// Under the column-tile schedule an ordering is a sequence of launches, one per tile
// anti-diagonal D = ti' + tj' (tile coordinates taken in the ordering's direction).
//
// Pairing.  Orderings 2p and 2p+1 differ only in s2, the column direction.  Ordering
// 2p+1 may sweep tile t as soon as ordering 2p has swept t and its four neighbours:
// every value t then reads is the one the sequential sweep gives it, ordering 2p
// reads none of those tiles again, and tiles of the two orderings in one launch
// neither read nor write each other's tiles.  With the column direction flipped this
// holds for every tile when ordering 2p+1 trails by pair_lag diagonals (the last
// column of tiles of ordering 2p is the first of ordering 2p+1), so launch step k of
// the pair sweeps diagonal k of ordering 2p and diagonal k - pair_lag of ordering 2p+1
// in one launch: nti + 2 ntj launches per pair instead of 2 (nti + ntj - 1), and the
// tail diagonals of one ordering share their launches with the head of the next.
// Orderings 2p+1 and 2p+2 flip the row direction and, at the depth turn, the depth
// too: ordering 2p+2 starts on the tile ordering 2p+1 ends on, so nothing overlaps.
#pragma once
#include <algorithm>

namespace pfast {

inline int ord_s1(int o) { return (o & 2) ? 1 : -1; }  // rows
inline int ord_s2(int o) { return (o & 1) ? 1 : -1; }  // columns
inline int ord_s3(int o) { return o < 4 ? -1 : 1; }    // depth

inline int plan_diagonals(int nti, int ntj) { return nti + ntj - 1; }

// Tiles on anti-diagonal D of an nti x ntj tile grid.
inline int plan_diag_tiles(int nti, int ntj, int D)
{
    const int lo = std::max(0, D - (ntj - 1)), hi = std::min(D, nti - 1);
    return hi - lo + 1;
}

// Diagonals by which ordering 2p+1 trails ordering 2p: the smallest lag such that
// every tile t of ordering 2p+1 comes after ordering 2p's diagonals of t and of t's
// neighbours.  A tile in the last column of ordering 2p's direction (diagonal d) has
// ordering 2p+1 diagonal d - (ntj - 1) and a row neighbour on diagonal d + 1, so
// ntj + 1; with a single row of tiles there is no row neighbour: ntj.
inline int pair_lag(int nti, int ntj) { return nti > 1 ? ntj + 1 : ntj; }

// Orderings per launch group and launch steps per group.
inline int plan_group_orderings(bool paired) { return paired ? 2 : 1; }
inline int plan_group_steps(bool paired, int nti, int ntj)
{
    const int nd = plan_diagonals(nti, ntj);
    return paired ? nd + pair_lag(nti, ntj) : nd;
}

// One launch: up to two (ordering, diagonal) parts, the second (if any) of ordering
// o[0] + 1.
struct PlanStep {
    int nparts = 0;
    int o[2] = {0, 0}, D[2] = {0, 0};
};

// Launch step k of the group starting at ordering o (even when paired) on a grid of
// nd tile diagonals; lag = pair_lag of the grid.
inline PlanStep plan_step(bool paired, int nd, int lag, int o, int k)
{
    PlanStep s;
    if (!paired) {
        s.nparts = 1;
        s.o[0] = o;
        s.D[0] = k;
        return s;
    }
    if (k < nd) {
        s.o[s.nparts] = o;
        s.D[s.nparts++] = k;
    }
    const int kb = k - lag;
    if (kb >= 0 && kb < nd) {
        s.o[s.nparts] = o + 1;
        s.D[s.nparts++] = kb;
    }
    return s;
}

}  // namespace pfast
