/*
 * hsmc_oracle.h -- CPU oracle for the hard-sphere Metropolis hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a clean-room restatement, in
 * plain C, of the reference program's algorithm (fedluc/HSMC, the C files under src/).  It is
 * used by tests/ (as the checker), by __graft_entry__.smoke() (as the checker)
 * and by bench.py's cpu_baseline leg (as the timed CPU path).  No product code
 * links, loads or calls it.
 *
 * Every function names the reference file:line it follows.  The reference's
 * global state is replaced by an explicit orc_sys context; the arithmetic,
 * evaluation order, cell-list layout and RNG semantics are kept.
 *
 * Pinning status (how each group is checked, see tests/test_oracle_*.py):
 *   - MT19937 core and seeding: C++11 std::mt19937 known-answer (10000th
 *     output of seed 5489 is 4123659995), and the raw stream of numpy's own
 *     MT19937 (init_genrand seeding) for several seeds.
 *   - rng_get_double: draw / (2^32 - 1) against that independent stream.
 *   - rng_get_int (GSL uniform_int): against that stream with n chosen so a
 *     quarter to a third of all draws fall in the discarded tail (outputs are
 *     exactly the in-range scaled draws, one draw consumed per discard), and
 *     by monotone, equal-width intervals and chi-square uniformity for small
 *     n.  GSL's seed-0 rule (seed 0 -> 4357) stays "parity unpinned": GSL is
 *     absent and no check here exercises seed 0.
 *   - vol_move's acceptance (moves.c:117-121): the ideal-gas closed form
 *     <V> = (N+1)/beta P of a chain in which no pair can overlap, which an
 *     N or N+2 exponent misses by about 13 standard errors.
 *   - Philox4x32-10: Random123 known-answer vectors.
 *   - box / lattice / cell list / compute_dist / check_overlap / part_move
 *     building blocks / rdf / contact / presst / widom: golden vectors written
 *     by the reference's own translation units compiled from their sources
 *     (oracle/ref/), plus brute-force and closed-form checks.
 *   - checkerboard sweep (orc_cb_*) and cell-stratified insertion
 *     (orc_widom_cells): restate the HIP kernels' own algorithms (no reference
 *     counterpart); pinned by invariants (no overlaps, particle conservation,
 *     detailed-balance statistics), by agreement with the reference-pinned
 *     i.i.d. Widom estimate, and the occupant pick (orc_cb_pick) by range and
 *     uniformity at n where half the draws are redrawn.
 */
#ifndef HSMC_ORACLE_H
#define HSMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (same values as the product's hsmc_gpu.h) */
#define ORC_OK              0
#define ORC_CELL_OVERFLOW  -1   /* cell_list.c:229-232 */
#define ORC_CELL_TOO_SMALL -2   /* cell_list.c:126-129 */
#define ORC_INIT_OVERLAP   -3   /* sim_info.c:128,143 */
#define ORC_BAD_ARG        -5
#define ORC_CELL_DEL       -6   /* cell_list.c:214-218 */
#define ORC_NO_MEMORY      -7
#define ORC_NOT_SUPPORTED  -8

/* ---- MT19937 with GSL semantics (rng.c:21-36) ---- */
typedef struct { uint32_t mt[624]; int mti; } orc_mt;
void     orc_mt_seed(orc_mt *m, unsigned long seed);
uint32_t orc_mt_next(orc_mt *m);
double   orc_mt_double(orc_mt *m);          /* rng.c:29-31 */
int      orc_mt_int(orc_mt *m, int n);      /* rng.c:34-36 */
/* helpers for ctypes: allocate/free a generator */
orc_mt  *orc_mt_new(unsigned long seed);
void     orc_mt_free(orc_mt *m);

/* ---- Philox4x32-10 (Salmon et al. 2011), the GPU generator ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- system (sim_info.h:6-20, cell_list.h:6-19, moves.c:22-24) ---- */
typedef struct orc_sys {
  /* box_info (sim_info.h:10-18) */
  double vol, lx, ly, lz, min_size, cell_size;
  int cell_x, cell_y, cell_z, cell_type;
  /* p_info (sim_info.h:6-9) */
  int N, Ncell;
  double *conf;            /* N x 4 : id, x, y, z (sim_info.h:20) */
  /* cl_info (cell_list.h:6-19) */
  int neigh_num, num_x, num_y, num_z, num_tot;
  double size_x, size_y, size_z;
  int *part_cell;          /* num_tot x max_part, [0] = count */
  int *neigh_mat;          /* num_tot x 27 */
  int max_part;
  double neigh_dr;
  /* G_IN fields used by the path (read_input.h:4-69) */
  double rho, dr_max, dv_max, press;
  /* move counters (moves.c:22-24), kept 64-bit */
  uint64_t part_moves, acc_part_moves, rej_part_moves;
  uint64_t vol_moves, acc_vol_moves, rej_vol_moves;
  orc_mt rng;
} orc_sys;

/* Build box, lattice, RNG and cell list as hs_nvt does (nvt.c:32-55). */
orc_sys *orc_create(int type, int nx, int ny, int nz, double rho,
                    double neigh_dr, int max_part, double dr_max,
                    unsigned long seed, int *status);
void     orc_destroy(orc_sys *s);
int      orc_box_init(orc_sys *s, int type, int nx, int ny, int nz, double rho);
int      orc_part_init(orc_sys *s);
int      orc_cell_list_init(orc_sys *s);           /* cell_list.c:33-58 */
int      orc_cell_list_new(orc_sys *s);            /* cell_list.c:142-175 */
int      orc_cell_part_idx(const orc_sys *s, int id);           /* cell_list.c:237-249 */
int      orc_cell_list_update(orc_sys *s, int cdel, int cadd, int id); /* :180-183 */
int      orc_cell_list_add(orc_sys *s, int c, int id);           /* :185-192 */
int      orc_cell_list_del(orc_sys *s, int c, int id);           /* :194-220 */
void     orc_set_params(orc_sys *s, double dr_max, double dv_max, double press);
void     orc_set_seed(orc_sys *s, unsigned long seed);
int      orc_set_config(orc_sys *s, const double *conf4);  /* copy N x 4, rebuild cells */
void     orc_get_config(const orc_sys *s, double *conf4);
void     orc_get_cell_list(const orc_sys *s, int *part_cell, int *neigh_mat);
void     orc_get_box(const orc_sys *s, double out[8]);       /* vol lx ly lz size_x size_y size_z rho */
void     orc_get_counters(const orc_sys *s, uint64_t out[6]);
void     orc_reset_counters(orc_sys *s);

/* moves.c */
double orc_compute_dist(const orc_sys *s, int i, int j, double sfx, double sfy, double sfz); /* :400-431 */
int    orc_check_overlap(const orc_sys *s, int idx, double sfx, double sfy, double sfz);   /* :157-212 */
void   orc_apply_pbc(orc_sys *s, int idx);                                               /* :215-226 */
int    orc_part_move(orc_sys *s);                                                        /* :27-80 */
int    orc_part_move_scripted(orc_sys *s, int idx, double rx, double ry, double rz);
int    orc_vol_move(orc_sys *s);                                                         /* :83-153 */
int    orc_sweep_nvt(orc_sys *s, int n_sweeps);                                          /* nvt.c:201-209 */
int    orc_sweep_npt(orc_sys *s, int n_sweeps);                                          /* npt.c:177-196 */
/* overlap decision of particle idx placed at (x,y,z), as part_move sees it */
int    orc_overlap_at(orc_sys *s, int idx, double x, double y, double z);

/* observables */
int  orc_rdf_counts(const orc_sys *s, double dr_bin, int nbins, double rmax_eff,
                    int64_t *counts, int64_t *out_of_range);           /* compute_rdf.c:110-128 */
int  orc_contact_counts(const orc_sys *s, double dr_bin, int nbins, double rmax_eff,
                        int64_t *counts, int64_t *out_of_range);       /* compute_press.c:123-165 */
int  orc_presst_flags(const orc_sys *s, const double *sf, int n, uint8_t *no_overlap); /* :243-273 */
int64_t orc_widom_points(const orc_sys *s, const double *xyz, int m, uint8_t *hit);   /* compute_widom_chem_pot.c:82-160 */
int64_t orc_widom_insertion(orc_sys *s, int n);                                     /* :44-71 */
/* Widom on the Philox stream used by the GPU's uniform insertion kernel */
int64_t orc_widom_philox(const orc_sys *s, uint64_t seed, uint64_t sample,
                         int64_t first, int64_t n);
/* cell-stratified uniform insertion (restates hsmc_gpu_widom_cells) */
int64_t orc_widom_cells(const orc_sys *s, uint64_t seed, uint64_t sample, int nx, int ny,
                        int nz, int per_cell);
void orc_widom_philox_point(const orc_sys *s, uint64_t seed, uint64_t sample,
                            int64_t i, double out[3]);

/* timing helper for the CPU baseline: n sweeps, wall and CPU seconds */
int orc_time_sweeps(orc_sys *s, int n_sweeps, double *wall_s, double *cpu_s);

/* ---- checkerboard sweep restatement (the HIP kernel's algorithm) ---- */
typedef struct orc_cb {
  int N;
  double lx, ly, lz;
  int gx, gy, gz;
  double dr_max;
  uint32_t key[2];
  uint64_t sweep;
  double *x, *y, *z;        /* by particle id */
  uint64_t attempts, accepted, left_cell;
} orc_cb;
orc_cb *orc_cb_create(int N, const double *conf4, double lx, double ly, double lz,
                      int gx, int gy, int gz, double dr_max, uint64_t seed);
void    orc_cb_destroy(orc_cb *c);
int     orc_cb_sweeps(orc_cb *c, int n_sweeps);
void    orc_cb_get_config(const orc_cb *c, double *conf4);
void    orc_cb_get_counters(const orc_cb *c, uint64_t out[4]);
/* per-sweep random quantities, exposed so tests can pin them */
void    orc_cb_sweep_randoms(const orc_cb *c, uint64_t sweep, double shift[3], int perm[8]);
/* occupant pick (rng_get_int rule, rng.c:34-36) on the checkerboard's Philox
 * stream; n in [1, 2^32-1] */
uint32_t orc_cb_pick(const uint32_t key[2], uint32_t cid, uint32_t t, uint64_t sweep, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
