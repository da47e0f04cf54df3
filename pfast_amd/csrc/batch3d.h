// batch3d.h -- batched FP32 3D sweep launchers.
#pragma once
#include <hip/hip_runtime.h>

// Sources per chunk (innermost dimension of the batched field layout).
#define PFAST_CHUNK 8

// One chunk of C sources as the sweep kernels see it.  A launch covers chunks
// through a table of these, so the chunks of one launch may belong to different
// shot sets and models of the same grid shape.
struct ChunkRef {
    float *field[2];           // [0] traveltimes [N][C] (negative marks receiver nodes
                               // during the adjoint); [1] adjoint field [N][C] of
                               // (lambda, T) float pairs
    const float *slow;         // [N] the model's cell time h/v (uniform) or slowness 1/v
    const float *slow8;        // the same values 8x8-tile-major: [nz][nti8][ntj8][8][8]
                               // (tile (ti,tj) row li column lj; edge tiles padded)
    int srcnode[PFAST_CHUNK];  // linear source node, -1 for padding sources
    int srcinit;               // half-width in nodes of the box around each source whose
                               // traveltimes are set at the start and never swept (0:
                               // the source node alone, as the reference seeds it)
};

// tclean ints per (chunk, tile): 8 clean epochs, 8 start levels, 8 end levels
constexpr int kTileClean = 24;

struct SweepArgs {
    const ChunkRef *chunks;  // indexed by chunk id
    const int *active;       // chunk ids still sweeping
    unsigned *changed;    // [nchunk] bit c: source c of the chunk changed this cycle
    int nact, nact_p8;    // active chunks; grid stride (multiple of 8 = XCD count)
    int nrows, ncols, nz, nti, ntj;
    float hx, hy, hz, wx, wy, wz;  // spacings (x=cols, y=rows, z=depth), 1/h^2
    // uniform spacing: the adjoint's unweighted denominator den passes the threshold
    // den * wx > 1e-9 (FP32 product) exactly when den > adj_thr
    float adj_thr;
    // Exact tile skipping (column-window kernels; NULL: off).  Per (chunk, tile):
    // tmod[5] = epochs of the last sweeps that changed a node of the tile ([0]) and
    // of its boundary faces x-, x+, y-, y+ ([1..4]: column j = 0, column j = last,
    // row i = 0, row i = last), tclean[o] = epoch of the last sweep in ordering o
    // that changed nothing.  A tile reads from a neighbour only the face that borders
    // it, so a tile whose clean sweep in this ordering is newer than every change of
    // itself and of those four neighbour faces would see the same inputs again and
    // change nothing: skipped.
    // Start levels (starts != 0, per-diagonal kernels): tclean[8 + o] = the first
    // level (in ordering o's sweep order) from which o's next sweep of the tile can
    // differ from its last one.  Below it every input of every update is the value
    // the last sweep saw, so the sweep reproduces the stored values there and starts
    // at that level; nz or more: nothing differs, skipped.  The last sweep in o sets
    // it to its own first changed level less one (a node's update reads the levels
    // beside it), and every change lowers it for the other orderings of the tile (by
    // the same margin) and, on a face, for all orderings of the neighbour across it
    // (to the changed level: a halo node is read at its own level only).
    // End levels, likewise: tclean[16 + o] = the number of trailing levels (sweep
    // order) whose inputs are the ones o's last sweep saw, so that 0 means none (as
    // the records start).  Beyond the last dirty level a sweep reproduces the stored
    // values once a whole batch of steps has changed nothing, and ends there.  The
    // dataflow kernels do not keep them: a dataflow cycle resets its chunks' start and
    // end levels to 0 afterwards (launch_starts_reset).
    int *tmod, *tclean;
    int epoch, o;         // this launch's epoch (increasing within a phase), ordering
    // A paired launch (plan.h) sweeps two orderings that differ in the column direction
    // only: blocks of tile index >= ntA take diagonal DB of ordering oB, the column
    // direction flipped, at epoch epochB.  ntA = INT_MAX: one ordering.
    int ntA, DB, oB, epochB;
    // Phase state on the device (NULL: the host passes every cycle's values): the
    // cycle number phase[0] (1-based) that the cycle's first kernel advanced; epoch then
    // counts from the cycle's start, to which the kernel adds (phase[0] - 1) * ep_cycle.
    // live[chunk] = 0 once a chunk has converged: its tiles return at once.
    const int *phase;
    const int *live;
    int ep_cycle;
    int starts;  // start levels kept and used (per-diagonal kernels; see tclean)
    // node-sweeps of the tiles actually swept (columns x depth x real sources), added up
    // on the device while profiling (NULL: not counted)
    unsigned long long *swept;
};

// State of a dataflow sweep launch (the four orderings of one depth direction of a
// cycle over every active chunk; flow.h has the schedule).
struct FlowArgs {
    unsigned *prog;      // [chunk id][8][tile] progress words: (epoch << 16) | steps
    unsigned nprog;      // progress words; the counter, abort and all-ones words follow
    unsigned *ticket;    // task claim counter (back to 0 when the launch's last block claims)
    int *abort;          // set when a dependency wait times out (results void)
    unsigned base;       // this launch's epoch << 16 (epochs increase within a phase)
    unsigned ntasks;     // 4 * tiles * active chunks
    int o0;              // first ordering: 0 (depth decreasing) or 4 (increasing)
    int cycle, nd;       // cycle number and tile diagonals, for the tile-skip epochs
    unsigned long long timeout;  // wall-clock ticks a dependency wait may take
    // Phase state on the device (NULL: base and cycle as given): base and cycle then
    // count from cycle 0, and the kernel adds (2 * phase[0]) << 16 to base and sets
    // cycle = phase[0] - 1.
    const int *phase;
};

// A sweep kernel variant: schedule (0: column window, the only one), tile rows x
// columns, batch depth of the forward and of the adjoint kernel.
struct SweepVariant {
    int kind, ti, tj, depth, adepth;
    const char *name;
};

namespace pfast {
int sweep_variant_count();
const SweepVariant &sweep_variant(int v);
// Start levels of the chunk ids act[0..nact) back to 0 (after a dataflow cycle).
hipError_t launch_starts_reset(hipStream_t st, int *tclean, const int *act, int nact, int ntile);
hipError_t launch_sweep3d(hipStream_t st, const SweepArgs &a, int variant, int adjoint,
                          int uniform, int D, int s1, int s2, int s3, int ntiles_diag);
// One dataflow launch (a.active / a.nact: the chunks; f.o0: which half of the cycle;
// variant as launch_sweep3d).
hipError_t launch_sweep3d_flow(hipStream_t st, const SweepArgs &a, const FlowArgs &f,
                               int variant, int adjoint, int uniform);
// Resident work-groups of a dataflow kernel on the current device (the grid size).
int flow_grid(int variant, int adjoint, int uniform);
// Words of the dataflow state of `nchunk` chunks of a grid of `ntile` tiles: progress
// words, then the claim counter, a spare word, the abort word and an all-ones word.
size_t flow_words(int nchunk, int ntile);
// Reset the dataflow state (the start of a phase); a kernel, so that it may sit in a
// captured graph.
hipError_t launch_flow_reset(hipStream_t st, unsigned *state, size_t words);
hipError_t launch_init_fwd(hipStream_t st, float *F, const int *srcnode, long N, int nchunk);
// Phase state on the device (SweepArgs::phase): int[2 + nchunk + nsrc] = cycle number,
// spare, live[nchunk], cycles[nsrc] (the first cycle in which a source changed nothing,
// 0 while it still changes).
hipError_t launch_phase_init(hipStream_t st, int *phase, int nchunk, int nsrc);
// Start of a cycle: advance the cycle number, clear the chunks' change bits.
hipError_t launch_cycle_begin(hipStream_t st, int *phase, unsigned *changed, int nchunk);
// End of a cycle: record each live chunk's converged sources and retire the chunks
// none of whose sources changed.
hipError_t launch_cycle_end(hipStream_t st, int *phase, const unsigned *changed, int nchunk,
                            int nsrc);
hipError_t launch_init_src_box(hipStream_t st, float *F, const double *vp, const int *srcnode,
                               long N, int nchunk, int R, int nrows, int ncols, int nz,
                               double dx, double dy, double dz);
hipError_t launch_adj_mark(hipStream_t st, float *T, const long *pnode, const int *psrc,
                           int npick, long N, int unmark);
// The adjoint field LT interleaves (lambda, T) per node and source: 2 floats per
// traveltime-field value, lambda first.
hipError_t launch_adj_pack(hipStream_t st, float *LT, const float *T, long N, int nchunk);
hipError_t launch_adj_bc(hipStream_t st, float *LT, const float *T, const long *bnode,
                         const int *bsrc, const double *bobs, const double *bnrm, int nb,
                         int nrows, int ncols, int nz, double dx, double dy, double dz);
hipError_t launch_adj_misfit(hipStream_t st, const float *T, const long *pnode,
                             const double *pobs, const long *poff, int nsrc, long N, double *EE);
hipError_t launch_grad_reduce(hipStream_t st, const float *LT, const double *vp, double *g,
                              long N, int nchunk, int nsrc);
hipError_t launch_gather_picks(hipStream_t st, const float *T, const long *rnode,
                               const int *rsrc, int nrec, long N, double *out);
hipError_t launch_gather_field(hipStream_t st, const float *F, int s, long N, int rec,
                               int strip_sign, double *out);
hipError_t launch_cell_time(hipStream_t st, const double *vp, float *f, long N, double h,
                            int uniform);
// Elements of the 8x8-tile-major copy of a cell-time array, and the copy itself.
long tile8_size(int nrows, int ncols, int nz);
hipError_t launch_tile8_copy(hipStream_t st, const float *f, float *ft, int nrows, int ncols,
                             int nz);
}  // namespace pfast
