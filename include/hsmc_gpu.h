/*
 * hsmc_gpu.h -- C ABI of the MI355X (gfx950) hard-sphere Monte Carlo device
 * layer, libhsmc_gpu.so.
 *
 * The reference (fedluc/HSMC) has no plugin or FFI boundary: its hot path is
 * a set of C functions over global state (moves.h, cell_list.h, nvt.h,
 * npt.h, compute_*.h).  This header is the device-side boundary beneath the
 * reference-named C host layer (include/hsmc.h); each entry point names the
 * reference function whose work it performs.  Plain C types only: host
 * buffers in and out, device state owned by the context.
 *
 * Conventions
 *   - Every function returns an int status: HSMC_OK (0) or a negative code
 *     below.  Codes that mirror one of the reference's exit(EXIT_FAILURE)
 *     sites name that site.  hsmc_gpu_strerror() and hsmc_gpu_last_error()
 *     give text.
 *   - A context owns one HIP stream and the device memory of one or more
 *     independent systems ("replicas").  Calls on one context are
 *     stream-ordered and must come from one host thread at a time.
 *   - Configurations cross the boundary in the reference's layout
 *     (sim_info.h:20): N rows of {id, x, y, z} doubles.
 *   - All distances are FP64 with the reference's evaluation order
 *     (moves.c:400-431), so overlap decisions, cell lists, histogram counts
 *     and Widom hits equal the reference's on the same configuration.
 */
#ifndef HSMC_GPU_H
#define HSMC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSMC_OK                 0
#define HSMC_ERR_CELL_OVERFLOW -1  /* more occupants than a cell holds (cell_list.c:229-232) */
#define HSMC_ERR_CELL_TOO_SMALL -2 /* cell side below the sphere diameter (cell_list.c:126-129) */
#define HSMC_ERR_INIT_OVERLAP  -3  /* overlapping starting lattice (sim_info.c:128,143) */
#define HSMC_ERR_HIP           -4  /* a HIP runtime call failed */
#define HSMC_ERR_BAD_ARG       -5  /* argument out of range */
#define HSMC_ERR_NO_MEMORY     -7  /* device or host allocation failed */
#define HSMC_ERR_NOT_SUPPORTED -8  /* request outside what the kernels handle */
#define HSMC_ERR_NO_DEVICE     -9  /* no HIP device */

typedef struct hsmc_gpu_ctx hsmc_gpu_ctx;

/* One independent system.  lx/ly/lz are the box edges (sim_info.c:63-65),
 * dr_max the full width of a trial displacement (moves.c:52-54), seed the
 * 64-bit Philox4x32-10 key of its random stream. */
typedef struct hsmc_gpu_replica {
  int32_t  n_part;
  int32_t  reserved;
  double   lx, ly, lz;
  double   dr_max;
  uint64_t seed;
} hsmc_gpu_replica;

/* Launch shape of the sweep kernels (chosen by hsmc_gpu_autotune or set by
 * hand).  lanes_per_cell: 32 or 64 lanes test one active cell;
 * cells_per_block: active cells per workgroup, 1, 2, 4 or 8 (at most 256
 * lanes); sweeps_per_graph: sweeps captured in one hipGraph launch (0 =
 * direct launches, at most 512).  persistent = 1: instead, one cooperative
 * launch of resident workgroups runs sweeps_per_graph whole sweeps (64 when
 * 0), with a grid barrier between the rebin and each sublattice phase in
 * place of a kernel boundary; the trajectory is the same.  Other values
 * give HSMC_ERR_BAD_ARG. */
typedef struct hsmc_gpu_tuning {
  int32_t lanes_per_cell;
  int32_t cells_per_block;
  int32_t sweeps_per_graph;
  int32_t persistent;
} hsmc_gpu_tuning;

/* ---- life cycle -------------------------------------------------------- */

/* Create a context on HIP device `device` holding n_replicas systems.  The
 * sweep grid of each replica is chosen here: the largest even number of
 * cells per axis whose side exceeds one diameter. */
int hsmc_gpu_create(hsmc_gpu_ctx **out, int device, int n_replicas,
                    const hsmc_gpu_replica *replicas);
int hsmc_gpu_destroy(hsmc_gpu_ctx *ctx);
const char *hsmc_gpu_strerror(int code);
const char *hsmc_gpu_last_error(const hsmc_gpu_ctx *ctx);
int hsmc_gpu_sync(hsmc_gpu_ctx *ctx);

/* ---- configuration (sim_info.c:179-190 part_config_get; io_config.c) ---- */

/* Copy N x {id,x,y,z} into the replica (the reference's part_conf). */
int hsmc_gpu_upload(hsmc_gpu_ctx *ctx, int replica, const double *conf4);
/* Copy the replica's configuration out as N x {id,x,y,z}. */
int hsmc_gpu_download(hsmc_gpu_ctx *ctx, int replica, double *conf4);
/* Overwrite the coordinates of m particles: idx[k] moves to
 * xyz[3k..3k+2] (the part_conf update of an accepted move, moves.c:52-54). */
int hsmc_gpu_set_particles(hsmc_gpu_ctx *ctx, int replica, const int32_t *idx,
                           const double *xyz, int m);
/* Replace the box edges without moving particles. */
int hsmc_gpu_set_box(hsmc_gpu_ctx *ctx, int replica, double lx, double ly, double lz);
/* Accepted volume move (moves.c:129-142): multiply every coordinate by sf,
 * install the new box, then wrap each particle as apply_pbc does
 * (moves.c:215-226).  The sweep grid is re-chosen for the new box. */
int hsmc_gpu_scale(hsmc_gpu_ctx *ctx, int replica, double sf,
                   double lx, double ly, double lz);
/* Set the displacement width of the replica (G_IN.dr_max, changed by the
 * step-size optimizer, optimizer.c:37-53). */
int hsmc_gpu_set_dr_max(hsmc_gpu_ctx *ctx, int replica, double dr_max);
/* Sweep grid of the replica (cells per axis). */
int hsmc_gpu_grid(const hsmc_gpu_ctx *ctx, int replica, int32_t out_g[3]);
/* Box edges of the replica (sim_box_info_get's lx, ly, lz, sim_info.c:179). */
int hsmc_gpu_box(const hsmc_gpu_ctx *ctx, int replica, double out[3]);

/* ---- sweeps (nvt.c:201-209 sweep_nvt; moves.c:27-80 part_move) ---------- */

/* n_sweeps sweeps of every replica: per sweep a random shift of the cell
 * grid and a random order of the 8 sublattices; each active cell makes as
 * many displacement attempts as it has occupants, on uniformly chosen
 * occupants; a move that leaves its cell is rejected.  N attempts per
 * replica per sweep, like sweep_nvt. */
int hsmc_gpu_sweeps(hsmc_gpu_ctx *ctx, int n_sweeps);
/* Same, bracketed by HIP events on the context's stream; *ms = elapsed. */
int hsmc_gpu_time_sweeps(hsmc_gpu_ctx *ctx, int n_sweeps, float *ms);
/* Per-replica counters: attempts, accepted, rejected for leaving the cell,
 * sweeps done (moves.c:22-24 part_moves / acc_part_moves, kept 64-bit). */
int hsmc_gpu_counters(hsmc_gpu_ctx *ctx, int replica, uint64_t out[4]);
int hsmc_gpu_reset_counters(hsmc_gpu_ctx *ctx);
/* Position of the sweep streams (the sweep counter in the Philox counters),
 * as counters() out[3] reports it; restored from a restart file so a resumed
 * run continues the same streams (the restart of io_config.c:28-130 keeps
 * the MT19937 state for the same purpose).  Attempt counts are unchanged. */
int hsmc_gpu_set_sweep_counter(hsmc_gpu_ctx *ctx, uint64_t sweep);
/* NpT sweeps of every replica (sweep_npt, npt.c:177-196, batched): each sweep
 * is one checkerboard sweep followed by the replica's Binomial(N, 1/(N+1))
 * volume moves (vol_move, moves.c:83-153): ln V' = ln V + (r - 0.5) dv_max[r],
 * sf = (V'/V)^(1/3); an overlap after scaling rejects, else r' <
 * exp(-press[r] (V' - V) + (N+1) ln(V'/V)) accepts and scales positions and
 * box by sf.  press[r] is beta P.  The volume draws come from a Philox stream
 * per replica keyed by its seed, so a replica's run does not depend on the
 * batch.  vol_counts (4 per replica, may be NULL) accumulates proposed,
 * accepted, rejected for overlap, rejected by the acceptance rule. */
int hsmc_gpu_npt_sweeps(hsmc_gpu_ctx *ctx, int n_sweeps, const double *press,
                        const double *dv_max, uint64_t *vol_counts);
/* The sweep's occupant pick (rng_get_int, rng.c:34-36: GSL's scale-and-reject
 * uniform_int, here on the Philox stream of replica seed `seed`), evaluated
 * by the same device function the sweep kernel calls: out[t] = pick of
 * attempt t (t < count <= 65536) of cell `cell` in sweep `sweep` among n
 * occupants.  n may be any value in [1, 2^32 - 1], so a caller can check the
 * rule's range, uniformity and redraw path, which a cell of at most 16
 * occupants almost never takes.  No context is needed. */
int hsmc_gpu_occupant_picks(int device, uint64_t seed, uint32_t cell, uint64_t sweep, uint32_t n,
                            int32_t count, uint32_t *out);
int hsmc_gpu_set_tuning(hsmc_gpu_ctx *ctx, const hsmc_gpu_tuning *t);
int hsmc_gpu_get_tuning(const hsmc_gpu_ctx *ctx, hsmc_gpu_tuning *t);
/* Time the candidate launch shapes on this context's own replicas (their
 * state is saved and restored) and keep the fastest.  *best receives it. */
int hsmc_gpu_autotune(hsmc_gpu_ctx *ctx, int sweeps_per_trial, hsmc_gpu_tuning *best);
/* Average duration (ms) of one launch of the sweep's dominant kernel (the
 * sublattice-phase kernel), measured with HIP events around each launch
 * of n_sweeps sweeps on the context's stream. */
int hsmc_gpu_time_phase_kernel(hsmc_gpu_ctx *ctx, int n_sweeps, float *avg_ms,
                               int32_t *launches);

/* ---- cell list (cell_list.c:142-175 cell_list_new, :253-301) ------------- */

/* Fresh cell list on the reference's grid nx*ny*nz (size = L/n per axis):
 * part_cell is C rows of max_part ints, [0] = count, then ascending particle
 * ids, then -1 (the layout cell_list_new builds).  neigh_mat (C x 27, may be
 * NULL) is the 27-cell stencil in the reference's order.  Cells are indexed
 * ix*ny*nz + iy*nz + iz (equal to the reference's formula on cubic grids). */
int hsmc_gpu_cell_list(hsmc_gpu_ctx *ctx, int replica, int nx, int ny, int nz,
                       int max_part, int32_t *part_cell, int32_t *neigh_mat);

/* ---- overlap tests (moves.c:157-212 check_overlap) ----------------------- */

/* out[k] = 1 when particle idx[k], placed at xyz[3k..3k+2], overlaps any
 * other particle (the test part_move makes, moves.c:60). */
int hsmc_gpu_overlap_batch(hsmc_gpu_ctx *ctx, int replica, const int32_t *idx,
                           const double *xyz, int m, uint8_t *out);
/* out[k] = check_overlap(idx[k], sf[k], sf[k], sf[k]) with every coordinate
 * and the box scaled by sf[k] (the test vol_move and presst make). */
int hsmc_gpu_overlap_scaled(hsmc_gpu_ctx *ctx, int replica, const int32_t *idx,
                            const double *sf, int m, uint8_t *out);
/* no_overlap[k] = 1 when no pair overlaps after scaling by sf[k], k < n_sf
 * (n_sf <= 64): vol_move's scan (moves.c:106-112) and presst_compute_hist
 * (compute_press.c:243-273), as one all-particle reduction. */
int hsmc_gpu_any_overlap_scaled(hsmc_gpu_ctx *ctx, int replica, const double *sf,
                                int n_sf, uint8_t *no_overlap);
/* Allocation-free variant of hsmc_gpu_any_overlap_scaled, with no
 * counterpart in the reference: for callers that may not allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results; the neighbour grid, scale
 * table and result mask live in a 16 MiB device array that the library
 * holds on each GPU until the process ends, so no call allocates.  A replica
 * whose scan needs more (about 32 B per sphere plus 12 B per grid cell:
 * N = 256000 needs about 11 MB) gets HSMC_ERR_NOT_SUPPORTED.  Calls on one
 * GPU share the array: calls that overlap in time on one GPU race. */
int hsmc_gpu_any_overlap_scaled_noalloc(hsmc_gpu_ctx *ctx, int replica, const double *sf,
                                        int n_sf, uint8_t *no_overlap);
/* The same scan for every replica at once: no_overlap[r] = 1 when no pair of
 * replica r overlaps after scaling by sf[r] (replicas with sf[r] <= 0 are
 * not scanned and report 1).  One kernel serves the replicas whose sweep-grid
 * cells are at least 1/sf[r] wide, reading the cell list the last sweep left;
 * the others are scanned one by one as above.  The volume-move scan of
 * batched NpT (vol_move, moves.c:104-112). */
int hsmc_gpu_any_overlap_scaled_batch(hsmc_gpu_ctx *ctx, const double *sf, uint8_t *no_overlap);
/* compute_dist (moves.c:400-431) for pairs (i[k], j[k]) with scale sf. */
int hsmc_gpu_pair_dist(hsmc_gpu_ctx *ctx, int replica, const int32_t *i,
                       const int32_t *j, int m, double sf, double *out);

/* ---- observables --------------------------------------------------------- */

/* rdf_hist_compute (compute_rdf.c:110-128): number of pairs per bin
 * (int)((dr-1)/dr_bin) with dr < rmax_eff (the reference's histogram is 2x
 * these counts).  *out_of_range counts pairs whose bin falls outside
 * [0, nbins). */
int hsmc_gpu_rdf_counts(hsmc_gpu_ctx *ctx, int replica, double dr_bin, int nbins,
                        double rmax_eff, int64_t *counts, int64_t *out_of_range);
/* pressv_compute_hist (compute_press.c:123-165): contact pairs, same
 * binning, rmax_eff = dr_bin * nbins + 1. */
int hsmc_gpu_contact_counts(hsmc_gpu_ctx *ctx, int replica, double dr_bin, int nbins,
                            double rmax_eff, int64_t *counts, int64_t *out_of_range);
/* widom_check_overlap (compute_widom_chem_pot.c:82-120) for given trial
 * points: hit[k] = 1 when the point is insertable; *hits = sum. */
int hsmc_gpu_widom_points(hsmc_gpu_ctx *ctx, int replica, const double *xyz, int m,
                          uint8_t *hit, int64_t *hits);
/* widom_insertion (compute_widom_chem_pot.c:44-71) with n uniform points
 * i = first..first+n-1 of Philox stream (seed, sample); *hits = insertable. */
int hsmc_gpu_widom_uniform(hsmc_gpu_ctx *ctx, int replica, uint64_t seed,
                           uint64_t sample, int64_t first, int64_t n, int64_t *hits);

/* Cell-stratified uniform insertion: each cell of the replica's insertion
 * grid (cells of side >= 1, out_grid[3] receives its shape; may be NULL)
 * gets per_cell points uniform inside it, point j of cell c from Philox
 * stream (seed, sample) at counter {c, j}.  The cells have equal volume, so
 * *hits / *n_points estimates the insertion probability without bias, as
 * widom_insertion's uniform points do (compute_widom_chem_pot.c:44-80); the
 * overlap test is the same FP64 one.  One wave per cell stages the 27-cell
 * neighbourhood once, so this is the fast form of uniform Widom. */
int hsmc_gpu_widom_cells(hsmc_gpu_ctx *ctx, int replica, uint64_t seed, uint64_t sample,
                         int32_t per_cell, int64_t *hits, int64_t *n_points, int32_t out_grid[3]);

/* ---- learned Widom proposal (no counterpart in the reference, whose
 * insertions are uniform, compute_widom_chem_pot.c:73-80) -----------------
 * Voxels of edge ~`voxel` tile the box; a voxel's descriptor is the sorted
 * distances (minus one diameter) from its centre to its 12 nearest sphere
 * centres, FP32; a 12->64->64->1 ReLU network with BF16 weights (MFMA,
 * FP32 accumulation) scores it.  Parameters are packed FP32 in the order
 * W1[12][64], b1[64], W2[64][64], b2[64], w3[64], b3. */
#define HSMC_MLP_IN 12
#define HSMC_MLP_HIDDEN 64
#define HSMC_MLP_NPARAM (12 * 64 + 64 + 64 * 64 + 64 + 64 + 1)

typedef struct hsmc_gpu_mlp hsmc_gpu_mlp;

/* Network on ctx's device, He-uniform weights from Philox(seed).  A network
 * uses its context's stream and buffers: destroy it before the context. */
int hsmc_gpu_mlp_create(hsmc_gpu_ctx *ctx, hsmc_gpu_mlp **out, uint64_t seed);
int hsmc_gpu_mlp_destroy(hsmc_gpu_mlp *m);
int hsmc_gpu_mlp_set(hsmc_gpu_mlp *m, const float *params);
int hsmc_gpu_mlp_get(hsmc_gpu_mlp *m, float *params);
/* y[k] = network(x[k][0..11]) */
int hsmc_gpu_mlp_forward(hsmc_gpu_mlp *m, const float *x, int n, float *y);
/* Mean squared error over n rows and its gradient (packed like params). */
int hsmc_gpu_mlp_gradient(hsmc_gpu_mlp *m, const float *x, const float *t, int n, float *loss,
                          float *grad);
/* Descriptors at given points (m x 12). */
int hsmc_gpu_point_descriptors(hsmc_gpu_ctx *ctx, int replica, const double *xyz, int m,
                               float *desc);
/* Train on n_points uniform insertions of Philox stream (seed, sample):
 * input the descriptor of the voxel holding the point, target
 * -beta * max(0, 1 - d_min); Adam, minibatches of `batch`, `epochs` passes. */
int hsmc_gpu_mlp_train(hsmc_gpu_mlp *m, int replica, double voxel, uint64_t seed,
                       uint64_t sample, int64_t n_points, int epochs, int batch, float lr,
                       float beta, float *final_loss);
/* The network's scores of all voxels (cap >= *n_voxels; scores may be NULL
 * to query V). */
int hsmc_gpu_voxel_scores(hsmc_gpu_mlp *m, int replica, double voxel, float *scores,
                          int64_t cap, int64_t *n_voxels);
/* The scores the learned proposal uses: -inf for a voxel whose centre lies
 * closer than 1 - (half the voxel diagonal) to a sphere centre (no point of
 * it can be insertable), the network's score for every other voxel (the
 * free voxels, *n_free of them; n_free may be NULL).  Every voxel is scored
 * when the voxel is too large for that exclusion (half diagonal >= 1).  No
 * counterpart in the reference (its insertions are uniform). */
int hsmc_gpu_proposal_scores(hsmc_gpu_mlp *m, int replica, double voxel, float *scores,
                             int64_t cap, int64_t *n_voxels, int64_t *n_free);
/* Drop the network's cached proposal (scores and cumulative mass) of the
 * last configuration scored, so that the next learned Widom call scores its
 * configuration again (for timing each call with its full cost). */
int hsmc_gpu_mlp_forget(hsmc_gpu_mlp *m);
/* Importance-sampled Widom: n insertions from q = mix * softmax(score)
 * + (1 - mix) / V over the proposal scores above (voxels scored -inf keep
 * the uniform part alone; with no free voxel q is uniform), by systematic
 * sampling over voxels (each voxel gets floor or ceil of n q_v, n q_v on
 * average); *estimate = (1/(V n)) sum over insertable points of 1/q_v, an
 * unbiased estimate of the insertion probability for any q > 0; *n_done
 * insertions made (n), *hits insertable. */
int hsmc_gpu_widom_learned(hsmc_gpu_mlp *m, int replica, double voxel, double mix,
                           uint64_t seed, uint64_t sample, int64_t n, double *estimate,
                           int64_t *n_done, int64_t *hits);

#ifdef __cplusplus
}
#endif
#endif
