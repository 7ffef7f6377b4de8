/*
 * hsmc.h -- C host layer of the MI355X hard-sphere Monte Carlo, libhsmc.so.
 *
 * Drop-in for fedluc/HSMC's hot path: the same global input structure and
 * keyword grammar (read_input.h:4-79), box/particle state (sim_info.h), cell
 * list (cell_list.h), moves (moves.h), NVT/NpT drivers (nvt.h, npt.h),
 * observables (compute_rdf.h, compute_press.h, compute_widom_chem_pot.h),
 * configuration/restart I/O (io_config.h) and step-size optimizer
 * (optimizer.h), with the same names, argument meaning, output files and
 * print-and-exit error behaviour.  Underneath, every sweep, overlap test and
 * observable runs on the GPU through include/hsmc_gpu.h.
 *
 * Extensions (no reference counterpart): keyword "gpu_sweep serial" makes
 * sweep_nvt/sweep_npt replay the reference's own serial Markov chain (random
 * particle, MT19937 stream) move by move through the GPU overlap test;
 * the default "gpu_sweep checkerboard" runs the parallel sweep.  Keyword
 * "gpu_device D" selects the HIP device.
 */
#ifndef HSMC_H
#define HSMC_H

#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* read_input.h:4-69 */
struct input {
  double rho;
  int nx, ny, nz, type;
  double neigh_dr;
  int neigh_max_part;
  double dr_max;
  int sweep_eq, sweep_stat;
  int output_int;
  double press, dv_max;
  int opt_flag, opt_samples, opt_sweeps;
  double opt_part_target, opt_vol_target;
  unsigned long int seed;
  double cavity_pcav, cavity_maxdr, cavity_mindr;
  double cavity_out_dr;
  int cavity_sample_int;
  int cluster_flag, cluster_moves_sweep, cluster_init_step;
  int restart_read;
  char restart_name[100];
  int restart_write;
  int config_write, config_samples;
  double pressv_dr, presst_xi_max, presst_dxi;
  int pressv_sample_int, presst_sample_int;
  int ql_order, ql_sample_int;
  double ql_rmax;
  int mu_insertions, mu_sample_int;
  int rdf_sample_int, rdf_samples;
  double rdf_dr, rdf_rmax;
  /* extensions */
  int gpu_serial;   /* 1: reference serial chain; 0: checkerboard sweep */
  int gpu_device;
};
extern struct input G_IN;

void print_example(void);                               /* read_input.c:18 */
void read_input_file(char *filename);                   /* read_input.c:62 */
void read_input_file_err(int err_id, char *last_string);/* read_input.c:477 */

/* sim_info.h:6-20 */
typedef struct { int NN; int Ncell; } p_info;
typedef struct {
  double vol;
  double lx, ly, lz;
  double min_size;
  double cell_size;
  int cell_x, cell_y, cell_z;
  int cell_type;
} box_info;
typedef double (*config)[4];

void sim_box_init(int cell_type, int nx, int ny, int nz, double rho);
void part_alloc(void);
void part_init(void);
void part_init_sc(void);
void part_init_fcc(void);
void part_init_err(void);
void add_particle(int id, double xx, double yy, double zz);
box_info sim_box_info_get(void);
p_info part_info_get(void);
config part_config_get(void);   /* host copy, brought up to date from the GPU */
void part_free(void);
void print_sim_info(void);
void sim_box_info_write(FILE *fid);
void sim_box_info_read(FILE *fid);
void part_info_write(FILE *fid);
void part_info_read(FILE *fid);
void part_conf_write(FILE *fid);
void part_conf_read(FILE *fid);
/* Tell the layer that the host copy returned by part_config_get() was
 * edited and must be sent to the GPU (no reference counterpart: the
 * reference has a single copy). */
void part_config_changed(void);

/* rng.h: MT19937 with GSL's conventions */
void rng_init(void);
double rng_get_double(void);
int rng_get_int(int MM);
void rng_free(void);
void rng_write(FILE *fid);
void rng_read(FILE *fid);

/* cell_list.h:6-47 */
typedef struct {
  int neigh_num;
  int num_x, num_y, num_z, num_tot;
  double size_x, size_y, size_z;
  int **part_cell;
  int **neigh_mat;
} cl_info;

void cell_list_init(bool alloc);
void cell_list_free(void);
void cell_list_alloc_arr(int ***arr, int rows, int cols);
void cell_list_free_arr(int ***arr, int rows);
void compute_cell_list_info(bool init);
cl_info get_cell_list_info(void);
void cell_list_new(void);
void cell_list_update(int cell_idx_del, int cell_idx_add, int part_idx);
void cell_list_add(int cell_idx, int part_idx);
void cell_list_del(int cell_idx, int part_idx);
void cell_list_check(int cell_idx);
int cell_part_idx(int id);
void cell_neigh_init(void);
void neigh_id(int ref_idx_x, int ref_idx_y, int ref_idx_z);

/* moves.h:7-29 */
void part_move(void);
void vol_move(void);
bool check_overlap(int idx_ref, double sf_x, double sf_y, double sf_z);
double compute_dist(int idx1, int idx2, double sf_x, double sf_y, double sf_z);
void apply_pbc(int idx);
void get_moves_counters(int *pm, int *apm, int *rpm, int *vm, int *avm, int *rvm);
void reset_moves_counters(void);
/* 64-bit counters (the reference's are truncated to int, moves.c:231-241) */
void get_moves_counters64(uint64_t out[6]);

/* nvt.h, npt.h */
void hs_nvt(void);
void run_nvt(bool prod_flag, int sweep_offset);
void sweep_nvt(void);
void hs_npt(void);
void run_npt(bool prod_flag, int sweep_offset);
void sweep_npt(void);

/* compute_rdf.h, compute_press.h, compute_widom_chem_pot.h */
void compute_rdf(bool init, int sweep);
void rdf_hist_free(void);
void compute_pressv(bool init);
void pressv_hist_free(void);
void compute_presst(bool init);
void presst_hist_free(void);
void compute_mu(bool init);

/* io_config.h */
void write_restart(int sweep);
void read_restart(char *restart_file);
void write_config(int sweep);

/* optimizer.h */
void opt_nvt(void);
void get_sample_nvt(double *dr, double *acc_ratio, int sample_iter);
void opt_npt(void);
void get_sample_npt(double *dr, double *acc_ratio_part, double *dv, double *acc_ratio_vol,
                    int sample_iter);

/* GPU context of the layer (created by cell_list_init, freed by
 * cell_list_free); NULL before. */
struct hsmc_gpu_ctx *hsmc_host_gpu(void);

#ifdef __cplusplus
}
#endif
#endif
