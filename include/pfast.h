/*
 * pfast.h -- C ABI of libpfast, the MI355X (gfx950) traveltime-and-gradient engine.
 *
 * Plain C: pointers, sizes and int status codes.  Every routine returns
 * PFAST_OK (0) or a negative code; pfast_last_error() describes the last failure
 * on the calling thread.
 *
 * Two families of entry points:
 *
 *  1. Single-source FP64 forms, one source per call, host arrays in and out.
 *     They replace the reference's subprograms one for one, with the same argument
 *     meaning, layouts and results (bit-identical to the reference's FP64 code):
 *       pfast_fsm3d            <- fsm3d            src/T_pfsm3d.h:131
 *       pfast_adjoint_fsm3d    <- adjoint_fsm3d    src/T_padjoint3d.h:73
 *       pfast_rt_fsm2d         <- RT_fsm2d         src/RT_pfsm2d.h:116
 *       pfast_tadjoint_fsm2d   <- Tadjoint_fsm2d   src/RT_padjoint2d.h:288 (Cr = 1)
 *                                                  and :67 (Cr = 1-WR)
 *       pfast_radjoint_fsm2d   <- Radjoint_fsm2d   src/RT_padjoint2d.h:382 (Cr = 1)
 *                                                  and :159 (Cr = WR)
 *     The reference's fatal exits (`exit(1)`) become PFAST_ERR_FATAL.
 *
 *  2. Batched device-resident forms (FP32 sweeps, FP64 misfit/gradient reductions):
 *     every source of a shot set on one GPU, model and fields resident in HBM.
 *     They replace the per-source loops + MPI_Allreduce of the drivers
 *     (src/T_PFAST3D.cpp:387-399, src/RT_PFAST2D.cpp:420-457).
 *
 * Layouts follow the reference: 3D node (i,j,k) = (row, col, depth) is stored at
 * k*nrows*ncols + i*ncols + j; 2D node (i,j) = (row/depth, col) at i*ncols + j.
 * src/rec index arrays are DIM-strided ints (3D: i,j,k; 2D: i,j).
 */
#ifndef PFAST_H
#define PFAST_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFAST_OK 0
#define PFAST_ERR_ARG (-1)   /* invalid argument */
#define PFAST_ERR_HIP (-2)   /* HIP runtime failure / no device */
#define PFAST_ERR_NOMEM (-3) /* device allocation failed */
#define PFAST_ERR_FATAL (-4) /* a condition on which the reference calls exit(1) */

#define PFAST_FWD_ONLY 1
#define PFAST_FWD_ADJ 2

const char *pfast_last_error(void);
int pfast_version(void);
/* Number of visible HIP devices (0 when none). */
int pfast_device_count(int *n);

/* ---------------------------------------------------------------------------------
 * Single-source FP64 forms.  ncycles (may be NULL) receives the sweep cycles run.
 * ------------------------------------------------------------------------------- */

/* fsm3d, src/T_pfsm3d.h:131.  U (nz*nrows*ncols) is written. */
int pfast_fsm3d(const double *vp, double *U, const int *src, int srcid, int MAXSwp,
                double dx, double dy, double dz, int nrows, int ncols, int nz, int pe,
                int *ncycles);

/* Allocation-free fsm3d.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under their own graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * pfast_fsm3d, bit for bit, but it calls no hipMalloc, hipFree or other allocating
 * routine: the device copies of vp and U, the convergence copy and the cycle flag live
 * in a static device array of PFAST_NOALLOC_BYTES that each GPU holds until the process
 * ends, and the work runs on the null stream.  Grids above PFAST_NOALLOC_MAX_NODES
 * nodes are refused (PFAST_ERR_ARG).  Calls on one GPU share that array; this library
 * serializes them with a lock (overlapping calls would otherwise race). */
#define PFAST_NOALLOC_BYTES (128u << 20)
#define PFAST_NOALLOC_MAX_NODES 5000000
int pfast_fsm3d_noalloc(const double *vp, double *U, const int *src, int srcid, int MAXSwp,
                        double dx, double dy, double dz, int nrows, int ncols, int nz, int pe,
                        int *ncycles);

/* adjoint_fsm3d, src/T_padjoint3d.h:73.  Writes lambda into U, misfit into *EE. */
int pfast_adjoint_fsm3d(double *U, const double *T, const double *T_obs, const int *rec,
                        const double *nxyz, const int *src, const int *srnum, int srcid,
                        int MAXSwp, double dx, double dy, double dz, int nrows, int ncols,
                        int nz, int pe, double *EE, int *ncycles);

/* RT_fsm2d, src/RT_pfsm2d.h:116.  ncycles (may be NULL) has 1+2*NumofR entries:
 * transmission, then (down, up) per reflector.  k2vec (may be NULL, ncols ints)
 * receives the reflection row map of the last reflector.  rec/srnum are accepted
 * for signature parity and unused, as in the reference. */
int pfast_rt_fsm2d(const double *vp, double *U, double *TTd, double *TTu, const double *Z,
                   const int *src, const int *rec, const int *srnum, int srcid, int NumofR,
                   int MAXSwp, double dx, double dz, int nrows, int ncols, int pe,
                   int *ncycles, int *k2vec);

/* Tadjoint_fsm2d.  Cr = 1 for the plain overload (src/RT_padjoint2d.h:288),
 * Cr = 1-WR for the WR overload (:67). */
int pfast_tadjoint_fsm2d(double *U, const double *T, const double *T_obs, double Cr,
                         const int *rec, const double *nxyz, const int *src, const int *srnum,
                         int srcid, int MAXSwp, double dx, double dz, int nrows, int ncols,
                         int pe, double *EE, int *ncycles);

/* Radjoint_fsm2d.  Cr = 1 for the plain overload (src/RT_padjoint2d.h:382),
 * Cr = WR for the WR overload (:159).  Td is accepted and unused, as in the
 * reference (its Td adjoint is commented out).  ncycles has NumofR entries. */
int pfast_radjoint_fsm2d(double *UR, const double *Td, const double *Tu, const double *TR_obs,
                         double Cr, const int *rec, const double *nxyz, const int *src,
                         const int *srnum, int srcid, int NumofR, int MAXSwp, double dx,
                         double dz, int nrows, int ncols, int pe, double *EE, int *ncycles);

/* ---------------------------------------------------------------------------------
 * Batched, device-resident 3D path.
 * ------------------------------------------------------------------------------- */
typedef struct pfast_ctx pfast_ctx;
typedef struct pfast_model3d pfast_model3d;
typedef struct pfast_shots3d pfast_shots3d;

int pfast_ctx_create(int device, pfast_ctx **out);
int pfast_ctx_destroy(pfast_ctx *ctx);
/* Blocks until all work queued by this context has finished. */
int pfast_ctx_sync(pfast_ctx *ctx);

/* A velocity model kept in HBM.  vp: nz*nrows*ncols doubles (reference layout). */
int pfast_model3d_create(pfast_ctx *ctx, const double *vp, int nrows, int ncols, int nz,
                         double dx, double dy, double dz, pfast_model3d **out);
/* Replace the velocity values (same grid), e.g. a line-search trial model. */
int pfast_model3d_set_vp(pfast_model3d *m, const double *vp);
int pfast_model3d_destroy(pfast_model3d *m);

/* A shot set bound to a model: nsrc sources (src: 3*nsrc node indices, the
 * reference's srcid order), srnum per-source receiver counts, and the flattened
 * receiver table (rec: 3*nrec node indices, nxyz: 3*nrec normals (n_y,n_x,n_z),
 * T_obs: nrec picks; a pick is used when T_obs < 100, as in the reference).
 * T_obs/nxyz/rec may be NULL for forward-only use.  The sweep kernels address one
 * chunk's field (8 sources) with 32-bit byte offsets: a grid whose (N + 64 planes) x 32
 * bytes (x 64 with picks, the adjoint's (lambda, T) pairs) pass 4 GiB is refused
 * (PFAST_ERR_ARG). */
int pfast_shots3d_create(pfast_model3d *m, int nsrc, const int *src, const int *srnum,
                         const int *rec, const double *nxyz, const double *T_obs,
                         pfast_shots3d **out);
int pfast_shots3d_destroy(pfast_shots3d *s);

/* Run the batched FP32 forward (and, with PFAST_FWD_ADJ, the adjoint and the FP64
 * gradient/misfit reductions) for every source of the set.  MAXSwp caps the sweep
 * cycles per solve.  Results stay on the device until fetched. */
int pfast_shots3d_run(pfast_shots3d *s, int mode, int MAXSwp);
/* Source neighbourhood (no counterpart in the reference, which seeds T = 0 at the
 * source node alone, src/T_pfsm3d.h:139-148; that is the default, halfwidth 0, and
 * the parity mode).  With halfwidth h in 1..PFAST_SRCINIT_MAX, every later forward
 * of the set first sets the nodes within h nodes of each source along every axis
 * (a box clipped to the grid) to FP64 traveltimes of the linear velocity model
 * fitted at the source node -- exact for a model linear in depth -- and the sweeps
 * never change them: the first-order error of the source singularity no longer
 * builds up near the source.  Takes effect at the next pfast_shots3d_run. */
#define PFAST_SRCINIT_MAX 16
int pfast_shots3d_set_source_init(pfast_shots3d *s, int halfwidth);
/* Sweep-variant tuning (no counterpart in the reference, whose sweeps have one fixed
 * loop order): times two forward cycles of this shot set under each listed variant
 * (all when variants is NULL), and two adjoint cycles after them when the set has
 * picks, sets the fastest on the shot set's context (the variant in use stays unless
 * another is at least 2% faster) and returns it in *chosen, with
 * each candidate's time in ms[i] when ms is not NULL.  Every variant gives
 * bit-identical results; the shot set's fields are left mid-solve, so run it again
 * afterwards. */
int pfast_shots3d_autotune(pfast_shots3d *s, const int *variants, int nvariants, int *chosen,
                           double *ms);
/* Schedule tuning (no counterpart in the reference): under the context's sweep variant,
 * times a full solve of this shot set (forward, and adjoint when it has picks) under
 * five candidates (schedule, check interval) = (2,1), (0,1), (1,1), (2,2), (2,4) -- see
 * pfast_ctx_set_dataflow and pfast_ctx_set_check_interval -- after a one-cycle warm-up
 * each, sets the fastest on the context (the setting in use stays unless another is
 * at least 2% faster) and returns it; ms[5] receives the times when
 * not NULL.  Results are the same under every candidate; run the shot set again
 * afterwards (its fields hold the last candidate's solve). */
int pfast_shots3d_autotune_schedule(pfast_shots3d *s, int *schedule, int *interval, double *ms);

/* Fetch results.  grad: N doubles (sum over sources of -lambda/v^3, the
 * reference's global gradient before regularisation).  misfit: sum of EE.
 * picks: sum(srnum) doubles, T at every receiver in receiver-table order.
 * field: the FP32 traveltime (which=0) or adjoint (which=1) of source `src`
 * widened to N doubles.  cycles: 2*nsrc ints (forward, adjoint cycles). */
int pfast_shots3d_gradient(pfast_shots3d *s, double *grad, double *misfit);
int pfast_shots3d_picks(pfast_shots3d *s, double *picks);
int pfast_shots3d_field(pfast_shots3d *s, int which, int src, double *field);
int pfast_shots3d_cycles(pfast_shots3d *s, int *cycles);

/* The single-source FP64 form on the same resident inputs: forward + adjoint of
 * source `srcid` (the body of the reference's per-source loop,
 * src/T_PFAST3D.cpp:388-391), one source per launch sequence, FP64 arithmetic
 * identical to pfast_fsm3d/pfast_adjoint_fsm3d.  EE receives the misfit,
 * cycles2 (may be NULL) the forward and adjoint cycle counts, T_out/lam_out
 * (may be NULL, N doubles) the traveltime and adjoint fields. */
int pfast_shots3d_plain_source(pfast_shots3d *s, int srcid, int MAXSwp, double *EE,
                               int *cycles2, double *T_out, double *lam_out);

/* Profiling: accumulated device time (ms) and launch count of the dominant sweep
 * kernels of this context since the last reset, measured with HIP events on the
 * context's stream.  kind 0 / 1 = 3D forward / adjoint sweeps launched per tile
 * anti-diagonal (fwd3d_col / adj3d_col), 2 = 2D shot-set solves (forward and
 * adjoints), 3 / 4 = 3D forward / adjoint sweeps as dataflow launches (fwd3d_flow /
 * adj3d_flow). */
int pfast_ctx_timing(pfast_ctx *ctx, int kind, double *ms, long *launches);
/* Node-sweeps (grid nodes x real sources x orderings) issued by the batched
 * sweeps of `kind` since the last timing reset: every node of every active chunk,
 * the sweep the reference performs. */
int pfast_ctx_work(pfast_ctx *ctx, int kind, double *node_sweeps);
/* The node-sweeps of the tiles the timed launches of `kind` actually swept (tiles
 * left out by exact tile skipping excluded), counted on the device while timing is
 * enabled. */
int pfast_ctx_work_swept(pfast_ctx *ctx, int kind, double *node_sweeps);
int pfast_ctx_timing_enable(pfast_ctx *ctx, int enable);
int pfast_ctx_timing_reset(pfast_ctx *ctx);
/* Batched sweep kernel variant of this context (0 = default).  Every variant
 * computes the same sweeps in the same per-node order (same results); they differ
 * in tile shape and batch depth.  pfast_sweep_name: a variant's description, NULL
 * past the last one. */
int pfast_ctx_set_sweep(pfast_ctx *ctx, int variant);
/* Batched 2D FP32 sweep form of this context (no counterpart in the reference, whose
 * sweeps have one loop order): 0 (default) = row window, each work-group walking the
 * rows of one source, with a source's independent solves (transmission / reflection,
 * and their adjoints) in work-groups of their own; 1 = one work-group per source
 * stepping the anti-diagonal planes.  Both give bit-identical results. */
int pfast_ctx_set_sweep2d(pfast_ctx *ctx, int variant);
const char *pfast_sweep_name(int variant);
/* Streams (1..4) over which a batched 3D cycle splits its active source chunks, so
 * that the launches of one stream fill the idle tail of another's; results are
 * the same for every count (each chunk's sweeps run in the same order). */
int pfast_ctx_set_streams(pfast_ctx *ctx, int nstreams);
/* Exact tile skipping of the batched 3D sweeps: 0 = off; 1 = whole tiles: a tile
 * whose last sweep in an ordering changed nothing, and whose nodes and neighbour
 * tiles have not changed since, is not swept again in that ordering; 2 (default) =
 * whole tiles and depth ranges: the per-diagonal launches also start a tile's sweep
 * at the first level whose inputs differ from what the ordering's last sweep of the
 * tile saw, and end it once it is past the last such level and a batch of steps
 * changed nothing (the levels outside would be reproduced).  Results are the same. */
int pfast_ctx_set_tile_skip(pfast_ctx *ctx, int mode);
/* Paired per-diagonal launches of the batched 3D sweeps (on by default): orderings 2p
 * and 2p+1 of a cycle, which differ only in the column direction (the innermost loop of
 * src/T_pfsm3d.h:158-160 and src/T_padjoint3d.h:139-141), share launches, the second
 * trailing the first by ntj + 1 tile anti-diagonals (ntj: tiles across the columns), so
 * that the thin tail diagonals of one ordering run beside the head of the next.  Every
 * tile still reads exactly the values of the sequential sweep; results are the same. */
int pfast_ctx_set_paired(pfast_ctx *ctx, int enable);
/* Schedule of the batched 3D sweep cycles (the loops of src/T_pfsm3d.h:152-179 and
 * src/T_padjoint3d.h:135-163, same node updates in the same order under every mode):
 * 0 = one launch per tile anti-diagonal and ordering; 1 = dataflow launches, two per
 * cycle, in which a tile starts as soon as the values it reads are final (tiles of
 * later diagonals and orderings overlap, level by level); 2 (default) = per cycle,
 * dataflow when the active chunks x tiles of an ordering are few enough that the
 * per-diagonal launches would leave the GPU idle.  Results are the same. */
int pfast_ctx_set_dataflow(pfast_ctx *ctx, int mode);
/* Sweep cycles of a batched 3D phase issued between two convergence checks on the
 * host (1 = every cycle, the default; at most PFAST_CHECK_INTERVAL_MAX).  The
 * reference tests convergence after every cycle (src/T_pfsm3d.h:181,
 * src/T_padjoint3d.h:165), and so does the device here whatever the interval: a
 * converged source's cycle count is the reference's, and a chunk whose sources all
 * converged sweeps no further.  A longer interval saves host round trips at the cost
 * of launches that return at once after the last chunk converged. */
#define PFAST_CHECK_INTERVAL_MAX 64
int pfast_ctx_set_check_interval(pfast_ctx *ctx, int cycles);
/* Replay the sweep cycles between two checks from captured hipGraphs (no counterpart
 * in the reference): captured on first use of a chunk list, schedule and interval,
 * kept per shot set (at most 16), then launched as one graph.  Off by default and
 * ignored while timing is on; results are the same. */
int pfast_ctx_set_graphs(pfast_ctx *ctx, int enable);
/* Internal source order of the batched 3D shot sets created afterwards on this context
 * (no counterpart in the reference, which runs one source per MPI rank), see
 * pfast_source_order: 2 (default) = balanced clusters, 1 = Morton order, 0 = the
 * caller's order (group the sources yourself).  Eight consecutive sources share a
 * sweep; neighbours' sweeps change the same tiles in the same cycles, so more tiles
 * are skipped and chunks converge together.  Every per-source result is returned in the
 * caller's order, and each source's fields and cycle counts are the same bit for bit
 * under any order. */
int pfast_ctx_set_source_order(pfast_ctx *ctx, int mode);
/* The internal order itself (host only, no GPU needed): perm[p] = the caller's index of
 * internal source p, for nsrc source nodes src[3*i] = (row, column, depth) on a grid of
 * spacings dx (columns), dy (rows), dz (depth).  mode 0: identity; 1: Morton order of the
 * nodes (stable); 2: clusters of 8 sources (the last holds the remainder) minimising the
 * distance to their centroid, by Lloyd iterations from the Morton chunks with a greedy
 * capacity-bounded assignment (nearest pairs first, each source among its 16 nearest
 * centroids), clusters and members in Morton order; falls back to 1 when nsrc x clusters
 * exceeds 1e8.  Deterministic. */
int pfast_source_order(int nsrc, const int *src, double dx, double dy, double dz, int mode,
                       int *perm);

/* ---------------------------------------------------------------------------------
 * Batched 2D shot sets: every source of a 2D survey resident on one GPU (FP32
 * sweeps, FP64 reductions).  Replaces the per-source loop of the 2D driver,
 * src/RT_PFAST2D.cpp:420-435 (RT_fsm2d, Tadjoint_fsm2d, Radjoint_fsm2d, the
 * -lambda/v^3 accumulations) and its MPI sums.
 * ------------------------------------------------------------------------------- */
typedef struct pfast_model2d pfast_model2d;
typedef struct pfast_shots2d pfast_shots2d;

/* vp: nrows*ncols doubles (reference layout); Z: NumofR*ncols*3 reflector table
 * (depth, n_z, n_x) as read from the .ref file (NULL when NumofR = 0). */
int pfast_model2d_create(pfast_ctx *ctx, const double *vp, int nrows, int ncols, double dx,
                         double dz, const double *Z, int NumofR, pfast_model2d **out);
int pfast_model2d_set_vp(pfast_model2d *m, const double *vp);
int pfast_model2d_destroy(pfast_model2d *m);
/* src: 2*nsrc (row, column); rec: 2*nrec; nxy: 2*nrec normals (n_z, n_x);
 * TT_obs: nrec transmission picks (used where < 100); TR_obs: nrec*NumofR
 * reflection picks (used where != 0).  rec..TR_obs may be NULL for forward-only use. */
int pfast_shots2d_create(pfast_model2d *m, int nsrc, const int *src, const int *srnum,
                         const int *rec, const double *nxy, const double *TT_obs,
                         const double *TR_obs, pfast_shots2d **out);
int pfast_shots2d_destroy(pfast_shots2d *s);
/* RT_fsm2d for every source, then (PFAST_FWD_ADJ) both adjoints.  A source that
 * hits one of the reference's fatal exits makes the call return PFAST_ERR_FATAL. */
int pfast_shots2d_run(pfast_shots2d *s, int mode, int MAXSwp);
/* gT, gR: sum over sources of -lambda/v^3 (transmission, reflection); gRT =
 * gT (1 - WR) + gR WR with the per-node weights WR (NULL = 0); E3 = {Et, Er, E}.
 * Any output may be NULL. */
int pfast_shots2d_gradient(pfast_shots2d *s, const double *WR, double *gT, double *gR,
                           double *gRT, double *E3);
/* which: 0 transmission T, 1 TTd / 2 TTu of reflector ir, 3 transmission adjoint,
 * 4 reflection adjoint (summed over reflectors); N doubles. */
int pfast_shots2d_field(pfast_shots2d *s, int which, int src, int ir, double *field);
/* cycles: per source 1+2R forward counts, 1 transmission-adjoint count, R
 * reflection-adjoint counts; k2vec: per source ncols ints (last reflector). */
int pfast_shots2d_info(pfast_shots2d *s, int *cycles, int *k2vec);

/* ---------------------------------------------------------------------------------
 * Optimiser support (host arrays), used by the reference-style drivers.
 * ------------------------------------------------------------------------------- */
/* FFT smoothing of a gradient, Dj <- Re IFFT[FFT(Dj) / (1 + nux^2 kx^2 + nuy^2 ky^2
 * + nuz^2 kz^2)], air nodes (vp = 330) zeroed, on the GPU (hipFFT); wavenumbers
 * built exactly as the reference builds them (including its odd-n ordering).
 * Replaces regu3d, src/T_padjoint3d.h:205, and regu2d, src/RT_padjoint2d.h:506. */
int pfast_regu3d(double *Dj, const double *vp, double nux, double nuy, double nuz, double dx,
                 double dy, double dz, int nrows, int ncols, int nz);
int pfast_regu2d(double *Dj, const double *vp, double nux, double nuz, double dx, double dz,
                 int nrows, int ncols);
/* beta_HS src/T_padjoint3d.h:264, beta_CGDESCENT :276, gd2sk :289, zoom :306,
 * cubicInterp :316, magnitudeOf :328, meanOf :335 (2D twins in
 * src/RT_padjoint2d.h:557-646).  N = number of grid nodes. */
double pfast_beta_hs(const double *gk1, const double *gk0, const double *dk, long N);
double pfast_beta_cgdescent(const double *gk1, const double *gk0, const double *dk, long N);
double pfast_gd2sk(const double *gk1, const double *gk0, const double *dk, double alpha1,
                   double *rhok, double *sk, double *yk, int Iterk, int lm, long N);
double pfast_zoom(double alpha0, double alpha1, double phi0, double phi1, double dphi0,
                  double dphi1);
double pfast_cubic_interp(double alpha0, double alpha1, double phi0, double phi1, double dphi0,
                          double dphi1);
double pfast_magnitude_of(const double *g, long N);
double pfast_mean_of(const double *g, long N);

/* Device-resident L-BFGS memory: the LM pairs (s, y) and rho of N doubles that the
 * drivers keep in sk / yk / rhoi (src/T_PFAST3D.cpp:430-440, src/RT_PFAST2D.cpp:495-505),
 * on ctx's device, with the two-loop recursion and the line search's dot products run
 * there in FP64, every dot product a compensated (Neumaier) sum. */
typedef struct pfast_lbfgs pfast_lbfgs;
int pfast_lbfgs_create(pfast_ctx *ctx, long N, int LM, pfast_lbfgs **out);
int pfast_lbfgs_destroy(pfast_lbfgs *h);
/* L-BFGS direction at iteration Iterk >= 2, replacing gd2sk and the two-loop of
 * src/T_PFAST3D.cpp:450-484 and src/RT_PFAST2D.cpp:520-545: the pair s = alpha1 dk,
 * y = gk0 - gk1, rho = 1 / s.y goes to slot (Iterk - 2) % LM, then the recursion runs
 * over the last min(LM, Iterk - 1) pairs (slot i % LM for pair i: the reference's ai[i]
 * leaves its array once Iterk >= LM + 2, and i % LM is the evident intent) from
 * H0 = gamma I, gamma = s.y / y.y.  gk1, gk0: host arrays of N doubles; dk: host array,
 * in the previous direction, out the new one; *gamma (may be NULL) receives gamma. */
int pfast_lbfgs_direction(pfast_lbfgs *h, int Iterk, const double *gk1, const double *gk0,
                          double *dk, double alpha1, double *gamma);
/* a . b (host arrays of N doubles) on the device, FP64 compensated: the line search's
 * dk . Dj (src/RT_PFAST2D.cpp:622, src/T_PFAST3D.cpp:520-575). */
int pfast_lbfgs_dot(pfast_lbfgs *h, const double *a, const double *b, double *out);

/* ---------------------------------------------------------------------------------
 * Request server (no counterpart in the reference: its drivers run one survey per
 * MPI job).  Velocity models stay resident on the device; traveltime-table and
 * gradient requests (a source set on a resident model) arrive at any time, are
 * admitted by device-memory footprint against the budget, and run batched: the
 * requests of one 3D grid shape share every sweep launch, joining and leaving at
 * sweep-cycle boundaries, so each result is bit-identical to the same request run
 * alone through pfast_shots3d_run / pfast_shots2d_run.  One host thread drives the
 * server through pfast_server_progress; calls on one server must not overlap.
 * ------------------------------------------------------------------------------- */
typedef struct pfast_server pfast_server;

#define PFAST_REQ_TABLE 1    /* traveltime fields of every source */
#define PFAST_REQ_GRADIENT 2 /* forward + adjoint: gradient and misfit */

#define PFAST_REQ_QUEUED 0
#define PFAST_REQ_RUNNING 1
#define PFAST_REQ_DONE 2
#define PFAST_REQ_FAILED 3

/* budget_bytes bounds the device memory of resident models and admitted requests. */
int pfast_server_create(int device, size_t budget_bytes, pfast_server **out);
int pfast_server_destroy(pfast_server *srv);
int pfast_server_add_model3d(pfast_server *srv, const double *vp, int nrows, int ncols, int nz,
                             double dx, double dy, double dz, int *model_id);
/* Z, NumofR as for pfast_model2d_create. */
int pfast_server_add_model2d(pfast_server *srv, const double *vp, int nrows, int ncols,
                             double dx, double dz, const double *Z, int NumofR, int *model_id);
/* Only once every request on the model is released. */
int pfast_server_remove_model(pfast_server *srv, int model_id);
/* Arrays as for pfast_shots3d_create / pfast_shots2d_create, copied at submission.
 * A table request needs no receivers; a request larger than the budget is refused
 * (PFAST_ERR_NOMEM), and a 3D request on a grid past pfast_shots3d_create's 4 GiB chunk
 * field bound (a gradient request by its (lambda, T) field) with PFAST_ERR_ARG.  ticket
 * receives the request id. */
int pfast_server_submit3d(pfast_server *srv, int model_id, int kind, int nsrc, const int *src,
                          const int *srnum, const int *rec, const double *nxyz,
                          const double *T_obs, int MAXSwp, long *ticket);
int pfast_server_submit2d(pfast_server *srv, int model_id, int kind, int nsrc, const int *src,
                          const int *srnum, const int *rec, const double *nxy,
                          const double *TT_obs, const double *TR_obs, int MAXSwp, long *ticket);
/* Admit, launch and retire work; returns after at least one request completed or
 * wait_us microseconds (0: one pass).  ncompleted (may be NULL) counts completions. */
int pfast_server_progress(pfast_server *srv, int wait_us, int *ncompleted);
/* state: PFAST_REQ_*; times (may be NULL): submission, admission and completion in
 * seconds since the server was created (-1 when not reached). */
int pfast_server_status(pfast_server *srv, long ticket, int *state, double *times);
/* Per source (forward, adjoint) sweep cycles (3D: 2*nsrc ints; 2D: as
 * pfast_shots2d_info). */
int pfast_server_cycles(pfast_server *srv, long ticket, int *cycles);
/* Gradient request: 3D grad (N doubles) and misfit[0]; 2D grad = gT, grad_r = gR
 * and misfit[0..1] = (Et, Er).  Outputs may be NULL.  A failed request returns its
 * error. */
int pfast_server_gradient(pfast_server *srv, long ticket, double *grad, double *grad_r,
                          double *misfit);
/* Traveltime field of source src of a finished request (N doubles). */
int pfast_server_field(pfast_server *srv, long ticket, int src, double *T);
/* Drop a request: a queued one leaves the queue, a finished one frees its device
 * memory, a running one is dropped when it finishes. */
int pfast_server_release(pfast_server *srv, long ticket);
/* Device memory the request is charged against the budget while admitted. */
int pfast_server_footprint(pfast_server *srv, long ticket, size_t *bytes);
/* in_use: device memory of resident models and admitted requests; cycles and
 * node_sweeps: 3D sweep cycles issued and the source node-sweeps they carried. */
int pfast_server_info(pfast_server *srv, size_t *in_use, int *queued, int *running,
                      long *cycles, double *node_sweeps);

#ifdef __cplusplus
}
#endif
#endif
