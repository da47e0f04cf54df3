// shots3d.h -- the batched 3D model and shot-set objects and the pieces of one
// forward+adjoint solve, shared by the C-ABI entry points and the request server.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "batch3d.h"
#include "context.h"

struct pfast_model3d {
    pfast_ctx *ctx;
    int nrows, ncols, nz;
    double dx, dy, dz;
    bool uniform;
    double *vp = nullptr;  // N doubles
    float *f = nullptr;    // N cell times / slowness
    float *ft = nullptr;   // the same, 8x8-tile-major (ChunkRef::slow8)
    int nshots = 0;        // live shot sets bound to this model
};

struct pfast_shots3d {
    pfast_model3d *m;
    int nsrc, nchunk;
    long nrec;
    // Sources are held in an internal order that puts spatially close sources in the
    // same chunk (their sweeps change the same tiles in the same cycles, so more tiles
    // are skipped and chunks converge together); every per-source table below is in
    // that order.  perm[p] = the caller's index of internal source p, iperm its inverse;
    // rmap[r] = the caller's receiver row of internal row r.
    std::vector<int> perm, iperm;
    std::vector<long> rmap;
    std::vector<int> src;    // 3*nsrc
    std::vector<int> srnum;  // nsrc
    std::vector<long> ch0;   // nsrc+1 prefix
    int *srcnode = nullptr;  // [nchunk*C]
    int srcinit = 0;         // half-width of the fixed FP64 box around each source
    float *T = nullptr;      // [nchunk][N][C] traveltimes
    float *L = nullptr;      // [nchunk][N][C][2] adjoint (lambda, T) pairs
    ChunkRef *chunks = nullptr;          // [nchunk] device table the sweep kernels read
    int *tmod = nullptr, *tclean = nullptr;  // tile skip records, [nchunk][ntile](x8)
    int ntile = 0;
    unsigned *flow = nullptr;            // dataflow state (progress words, counters)
    size_t nflow = 0;
    std::vector<ChunkRef> hchunks;       // host copy
    // receiver table (reference order) for picks and the single-source path
    long *rnode = nullptr;
    int *rsrc = nullptr;
    int *rec3 = nullptr;
    double *nxyz = nullptr, *Tobs = nullptr;
    // picks grouped by source (misfit, fixed-node marks)
    long *pnode = nullptr;
    int *psrc = nullptr;
    double *pobs = nullptr;
    long *poff = nullptr;
    int npick = 0;
    // boundary-condition entries (last picked receiver per node, not at the source)
    long *bnode = nullptr;
    int *bsrc = nullptr;
    double *bobs = nullptr, *bnrm = nullptr;
    int nb = 0;
    double *EE = nullptr, *grad = nullptr, *picks = nullptr, *fieldbuf = nullptr;
    unsigned *changed = nullptr;
    int *active = nullptr;
    // phase state on the device (launch_phase_init) and its host mirror (pinned)
    int *phase = nullptr;
    int *hphase = nullptr;
    // captured sweep cycles, keyed by everything their launch arguments depend on
    struct Graph {
        std::vector<long> key;
        hipGraphExec_t exec;
    };
    std::vector<Graph> graphs;
    std::vector<int> fcyc, acyc;
    bool have_adj_data = false;
};

namespace pfast {

struct Grid3 {
    int nrows, ncols, nz;
    double dx, dy, dz;
    bool uniform;
};

inline Grid3 grid_of(const pfast_model3d *m)
{
    return Grid3{m->nrows, m->ncols, m->nz, m->dx, m->dy, m->dz, m->uniform};
}

// Orderings [o0, o1) of one sweep cycle (8 orderings, each as one launch per tile
// anti-diagonal) on ctx's stream over the chunk ids fwd[0..nfwd) (forward sweeps)
// and adj[0..nadj) (adjoint sweeps) of the chunk table `chunks`, all of grid g;
// device arrays.  changed[id] collects, per chunk id, the sources that changed.
// Exact tile skipping records of one phase (see SweepArgs) and the cycle number that
// makes launch epochs increase; records are indexed by chunk id.
struct TileSkip {
    int *tmod, *tclean;
    int cycle;
};

// st: the stream to issue on (nullptr: ctx->stream); skip: NULL for no skipping.
// dphase: the device phase state (SweepArgs::phase) the launches read their cycle
// number and live chunks from, or NULL (the cycle number in skip, every chunk live).
int issue_cycle3d(pfast_ctx *ctx, const Grid3 &g, const ChunkRef *chunks, const int *fwd,
                  int nfwd, const int *adj, int nadj, unsigned *changed, hipStream_t st = nullptr,
                  const TileSkip *skip = nullptr, const int *dphase = nullptr);
// A cycle's per-diagonal launches come in groups of group_orderings3d orderings (2 when
// ctx pairs them, plan.h), group_steps3d launches each.  issue_steps3d: launch steps
// [k0, k1) of the group starting at ordering o (same arguments otherwise).
int group_orderings3d(const pfast_ctx *ctx);
int group_steps3d(const pfast_ctx *ctx, const Grid3 &g);
int issue_steps3d(pfast_ctx *ctx, const Grid3 &g, const ChunkRef *chunks, const int *fwd,
                  int nfwd, const int *adj, int nadj, unsigned *changed, int o, int k0, int k1,
                  hipStream_t st = nullptr, const TileSkip *skip = nullptr,
                  const int *dphase = nullptr);
// Active chunks x tiles of one ordering up to which the automatic schedule takes the
// dataflow launches: they keep the GPU busy where a launch per tile anti-diagonal
// has too few tiles, while the per-diagonal launches sweep only the depth ranges that
// can change (measured on 201x201x101, 676 tiles: dataflow 1.7x faster with 32 sources,
// 1.5x with 64, 1.1x with 128 (10816 tasks), 0.88x with 200 (16900)).
constexpr long kAutoDataflowTasks = 13000;
// Whether a cycle over nact chunks of a grid of ntile tiles runs as dataflow launches
// under ctx's schedule setting (0 never, 1 always, 2 automatic); words: the size of
// the dataflow state.
bool use_dataflow(const pfast_ctx *ctx, int nact, int ntile, size_t words);
// One sweep cycle as two dataflow launches (flow.h: orderings 0-3, then 4-7) on
// stream st over the chunk ids act[0..nact) (device array) of the chunk table,
// forward or adjoint.  flow: the phase's dataflow state, flow_words(id bound, tiles)
// words, reset at the phase start; epoch: this cycle's number in the phase (1, 2, ...,
// at most 0x7FFF); skip as above.  With dphase the launches take the cycle number
// from the device phase state instead (epoch unused; at most 0x7FFF cycles per reset).
int issue_cycle3d_flow(pfast_ctx *ctx, const Grid3 &g, const ChunkRef *chunks, const int *act,
                       int nact, bool adjoint, unsigned *changed, unsigned *flow, int nids,
                       int epoch, hipStream_t st, const TileSkip *skip,
                       const int *dphase = nullptr);
// Tile anti-diagonals of one ordering under ctx's sweep variant.
int diagonals3d(const pfast_ctx *ctx, const Grid3 &g);

// Stages of pfast_shots3d_run on the model's stream.
int fwd_begin3d(pfast_shots3d *s);  // traveltime fields: 0 at the source, INF elsewhere
int adj_begin3d(pfast_shots3d *s);  // misfit, lambda = 0, receiver conditions, marks
int adj_end3d(pfast_shots3d *s);    // unmark, gradient reduction over sources

// Free a shot set without synchronising: no work on the device may still use it.
void shots3d_free(pfast_shots3d *s);

}  // namespace pfast
