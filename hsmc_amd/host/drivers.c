/*
 * drivers.c -- NVT and NpT drivers and the step-size optimizer.
 *
 * Same schedule as the reference (nvt.c:30-209, npt.c:25-196,
 * optimizer.c:23-148): per sweep index ii, screen output, restart file, then
 * (production) configuration, pressure (virial, thermodynamic), Widom and
 * g(r) samples, then the sweep.  Sweeps between two sample points are sent to
 * the GPU as one batch (one hipGraph launch per batch of the sweep kernels).
 *
 * With "gpu_sweep serial" a sweep is the reference's: N part_move calls (and,
 * in NpT, a volume move with probability 1/(N+1) per step), each move decided
 * by the GPU.  With the default checkerboard sweep, a sweep is N attempts of
 * the parallel kernel; in NpT it is followed by the volume moves of that sweep,
 * their number drawn from the reference's Binomial(N, 1/(N+1)).
 */
#include <math.h>
#include <stdlib.h>
#include <time.h>

#include "hsmc.h"
#include "hsmc_gpu.h"
#include "host_internal.h"

/* the ql keyword (compute_order_parameter.c), the cavity and cluster
 * ensembles are outside this build's scope */
static void refuse_out_of_scope(void) {
  if (G_IN.ql_sample_int > 0) {
    printf("ERROR: the ql order parameter is not part of this build\n");
    exit(EXIT_FAILURE);
  }
}

static int every(int ii, int interval) { return interval > 0 && ii % interval == 0; }

/* first index > ii (and < end) at which the driver has something to do */
static int next_event(int ii, int end, int prod, int npt) {
  for (int j = ii + 1; j < end; j++) {
    if (every(j, G_IN.output_int) || every(j, G_IN.restart_write)) return j;
    if (prod) {
      if (every(j, G_IN.config_write) || every(j, G_IN.presst_sample_int)) return j;
      if (!npt && (every(j, G_IN.pressv_sample_int) || every(j, G_IN.mu_sample_int) ||
                   every(j, G_IN.rdf_sample_int)))
        return j;
    }
  }
  return end;
}

/* n checkerboard sweeps on the GPU; move counters updated */
static void gpu_sweeps(int n) {
  hsmc_gpu_ctx *g = hsmc_host_gpu();
  host_sync_device();
  uint64_t before[4], after[4];
  host_gpu_check(hsmc_gpu_counters(g, 0, before), "counters");
  host_gpu_check(hsmc_gpu_sweeps(g, n), "sweeps");
  host_gpu_check(hsmc_gpu_counters(g, 0, after), "counters");
  host_count_sweep_moves(after[0] - before[0], after[1] - before[1]);
  host_device_moved();
}

/* Number of volume-move steps in one sweep of N steps, each a volume move
 * with probability 1/(N+1) (npt.c:185): Binomial(N, 1/(N+1)) by inversion. */
static int volume_steps(int N) {
  const double p = 1.0 / (N + 1.0);
  double u = rng_get_double();
  double pk = pow(1.0 - p, N);   /* P(0) */
  double cdf = pk;
  int k = 0;
  while (u > cdf && k < N) {
    pk *= (double)(N - k) / (double)(k + 1) * p / (1.0 - p);
    k++;
    cdf += pk;
  }
  return k;
}

void sweep_nvt(void) {
  p_info pi = part_info_get();
  if (G_IN.gpu_serial) {
    for (int ii = 0; ii < pi.NN; ii++) part_move();   /* nvt.c:205-207 */
  } else {
    gpu_sweeps(1);
  }
}

void sweep_npt(void) {
  p_info pi = part_info_get();
  if (G_IN.gpu_serial) {
    for (int ii = 0; ii < pi.NN; ii++) {   /* npt.c:183-194 */
      int r = rng_get_int(pi.NN + 1);
      if (r < pi.NN) part_move();
      else vol_move();
    }
  } else {
    gpu_sweeps(1);
    int nv = volume_steps(pi.NN);
    for (int k = 0; k < nv; k++) vol_move();
  }
}

static void run_sweeps(int n, int npt) {
  if (n <= 0) return;
  if (G_IN.gpu_serial || npt) {
    for (int k = 0; k < n; k++) {
      if (npt) sweep_npt(); else sweep_nvt();
    }
  } else {
    gpu_sweeps(n);
  }
}

void run_nvt(bool prod_flag, int sweep_offset) {
  bool pressv_init = true, presst_init = true, mu_ave_init = true, rdf_init = true;
  int n_sweeps = prod_flag ? G_IN.sweep_stat : G_IN.sweep_eq;
  int end = sweep_offset + n_sweeps;
  int ii = sweep_offset;
  while (ii < end) {
    if (ii == sweep_offset) printf("Sweep number\n");
    if (every(ii, G_IN.output_int)) {   /* out <= 0: never (the reference divides by it) */
      printf("%d\n", ii);
      fflush(stdout);
    }
    if (every(ii, G_IN.restart_write)) write_restart(ii);
    if (prod_flag) {
      if (every(ii, G_IN.config_write)) write_config(ii);
      if (every(ii, G_IN.pressv_sample_int)) {
        compute_pressv(pressv_init);
        pressv_init = false;
      }
      if (every(ii, G_IN.presst_sample_int)) {
        compute_presst(presst_init);
        presst_init = false;
      }
      if (every(ii, G_IN.mu_sample_int)) {
        compute_mu(mu_ave_init);
        mu_ave_init = false;
      }
      if (every(ii, G_IN.rdf_sample_int)) {
        compute_rdf(rdf_init, ii);
        rdf_init = false;
      }
    }
    int nxt = next_event(ii, end, prod_flag, 0);
    run_sweeps(nxt - ii, 0);
    ii = nxt;
  }
  if (!pressv_init) pressv_hist_free();
  if (!presst_init) presst_hist_free();
  if (!rdf_init) rdf_hist_free();
}

void run_npt(bool prod_flag, int sweep_offset) {
  bool presst_init = true;
  int n_sweeps = prod_flag ? G_IN.sweep_stat : G_IN.sweep_eq;
  int end = sweep_offset + n_sweeps;
  int ii = sweep_offset;
  while (ii < end) {
    if (ii == 0) printf("Sweep number  Density\n");   /* npt.c:120 prints only at 0 */
    if (every(ii, G_IN.output_int)) {
      printf("%d  %.8f\n", ii, G_IN.rho);
      fflush(stdout);
    }
    if (every(ii, G_IN.restart_write)) write_restart(ii);
    if (prod_flag) {
      if (every(ii, G_IN.config_write)) write_config(ii);
      if (every(ii, G_IN.presst_sample_int)) {
        compute_presst(presst_init);
        presst_init = false;
      }
    }
    int nxt = next_event(ii, end, prod_flag, 1);
    run_sweeps(nxt - ii, 1);
    ii = nxt;
  }
  if (!presst_init) presst_hist_free();
}

static void print_moves(int npt) {
  uint64_t c[6];
  get_moves_counters64(c);
  printf("---------------------------------------------------\n");
  printf("-- Particle moves: %.8e\n", (double)c[0]);
  printf("   Acceptance percentage: %f\n", (double)c[1] / ((double)c[0]));
  printf("   Rejection percentage: %f\n", (double)c[2] / ((double)c[0]));
  if (npt) {
    printf("-- Volume moves: %.8e\n", (double)c[3]);
    printf("   Acceptance percentage: %f\n", (double)c[4] / ((double)c[3]));
    printf("   Rejection percentage: %f\n", (double)c[5] / ((double)c[3]));
  }
}

static double wall_now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

void hs_nvt(void) {   /* nvt.c:30-103 */
  refuse_out_of_scope();
  if (G_IN.restart_read == 0) {
    sim_box_init(G_IN.type, G_IN.nx, G_IN.ny, G_IN.nz, G_IN.rho);
    part_alloc();
    part_init();
    rng_init();
  } else {
    read_restart(G_IN.restart_name);
  }
  print_sim_info();
  cell_list_init(true);
  if (G_IN.opt_flag == 1) {
    opt_nvt();
    part_init();
    cell_list_init(false);
  }
  clock_t start = clock();
  double w0 = wall_now();
  reset_moves_counters();
  printf("---------------------------------------------------\n");
  printf("Equilibration...\n");
  run_nvt(false, 0);
  printf("Equilibration completed.\n");
  printf("---------------------------------------------------\n");
  printf("Production...\n");
  run_nvt(true, G_IN.sweep_eq);
  printf("Production completed.\n");
  clock_t end = clock();
  print_moves(0);
  printf("Elapsed time: %f seconds\n", (double)(end - start) / CLOCKS_PER_SEC);
  printf("Wall time: %f seconds\n", wall_now() - w0);
  part_free();
  cell_list_free();
  rng_free();
}

void hs_npt(void) {   /* npt.c:25-104 */
  refuse_out_of_scope();
  if (G_IN.restart_read == 0) {
    sim_box_init(G_IN.type, G_IN.nx, G_IN.ny, G_IN.nz, G_IN.rho);
    part_alloc();
    part_init();
    rng_init();
  } else {
    read_restart(G_IN.restart_name);
  }
  print_sim_info();
  cell_list_init(true);
  if (G_IN.opt_flag == 1) {
    double rho_start = G_IN.rho;
    opt_npt();
    G_IN.rho = rho_start;
    sim_box_init(G_IN.type, G_IN.nx, G_IN.ny, G_IN.nz, rho_start);
    part_init();
    host_gpu_box_sync();
    cell_list_init(false);
  }
  clock_t start = clock();
  double w0 = wall_now();
  reset_moves_counters();
  printf("---------------------------------------------------\n");
  printf("Equilibration...\n");
  run_npt(false, 0);
  printf("Equilibration completed.\n");
  printf("---------------------------------------------------\n");
  printf("Production...\n");
  run_npt(true, G_IN.sweep_eq);
  printf("Production completed.\n");
  clock_t end = clock();
  print_moves(1);
  printf("Elapsed time: %f seconds\n", (double)(end - start) / CLOCKS_PER_SEC);
  printf("Wall time: %f seconds\n", wall_now() - w0);
  part_free();
  cell_list_free();
  rng_free();
}

/* ---- optimizer (optimizer.c:23-148) ---- */

void get_sample_nvt(double *dr, double *acc_ratio, int sweeps) {
  reset_moves_counters();
  run_sweeps(sweeps, 0);
  uint64_t c[6];
  get_moves_counters64(c);
  *dr = G_IN.dr_max;
  *acc_ratio = (double)c[1] / ((double)c[0]);
}

static void set_gpu_dr_max(void) {
  /* the checkerboard kernels read dr_max from the context's replica */
  hsmc_gpu_ctx *g = hsmc_host_gpu();
  if (g) host_gpu_check(hsmc_gpu_set_dr_max(g, 0, G_IN.dr_max), "set_dr_max");
}

/* One secant step of the optimizer (optimizer.c:47, :117-118) towards the
 * target acceptance.  When two samples give the same acceptance the
 * reference divides by zero and carries an infinite or NaN step into the
 * next moves (an infinite volume step is then accepted through exp(NaN));
 * here the step keeps its last value instead. */
static double secant_step(double x2, double x1, double a2, double a1, double target) {
  const double den = a2 - a1;
  if (den == 0.0) return x2;
  const double x = x2 - (a2 - target) * (x2 - x1) / den;
  return isfinite(x) ? x : x2;
}

void get_sample_npt(double *dr, double *acc_ratio_part, double *dv, double *acc_ratio_vol,
                    int sweeps) {
  reset_moves_counters();
  run_sweeps(sweeps, 1);
  uint64_t c[6];
  get_moves_counters64(c);
  *dr = G_IN.dr_max;
  *dv = G_IN.dv_max;
  *acc_ratio_part = (double)c[1] / ((double)c[0]);
  *acc_ratio_vol = (double)c[4] / ((double)c[3]);
}

/* A step width the optimizer tunes toward a target acceptance: the width
 * itself (G_IN.dr_max or G_IN.dv_max), the last two (width, acceptance)
 * samples, an upper bound (0: none), and the largest width a reflected
 * negative step may keep before it falls back to half the last width. */
typedef struct {
  double *width;
  double target;
  double upper;
  double reflect_max;
  double w_prev, acc_prev, w_last, acc_last;
} tuned_width;

/* Next width from the last two samples (optimizer.c:47-56, :117-130). */
static void advance_width(tuned_width *t) {
  double w = secant_step(t->w_last, t->w_prev, t->acc_last, t->acc_prev, t->target);
  if (t->upper > 0.0 && w > t->upper) {
    w = t->upper;
  } else if (w <= 0.0) {
    w = -w;
    if (w > t->reflect_max) w = t->w_last / 2;
  }
  *t->width = w;
}

/* Run one sample of sweeps at the current widths and record, for each tuned
 * width, the width used and the acceptance it gave.  One width tunes the
 * particle displacement alone (NVT); two add the volume step (NpT). */
static void sample_widths(tuned_width *t, int n, int sweeps) {
  double acc[2];
  if (n == 1)
    get_sample_nvt(&t[0].w_last, &acc[0], sweeps);
  else
    get_sample_npt(&t[0].w_last, &acc[0], &t[1].w_last, &acc[1], sweeps);
  for (int k = 0; k < n; k++) t[k].acc_last = acc[k];
}

/* The optimizer's schedule (optimizer.c:23-148): one sample at the input
 * widths, each width doubled if its acceptance is above target and halved
 * otherwise, a second sample, then opt_samples secant updates, each followed
 * by a sample.  opt_sweeps is split evenly over the samples. */
static void tune_widths(tuned_width *t, int n) {
  const int sweeps = G_IN.opt_sweeps / G_IN.opt_samples;
  printf("---------------------------------------------------\n");
  printf("Maximum displacement optimization started ...\n");
  printf("Sweeps for optimization: %d\n", G_IN.opt_sweeps);
  printf("Number of samples: %d\n", G_IN.opt_samples);
  sample_widths(t, n, sweeps);
  for (int k = 0; k < n; k++)
    *t[k].width = t[k].acc_last > t[k].target ? *t[k].width * 2 : *t[k].width / 2;
  for (int round = 0; round <= G_IN.opt_samples; round++) {
    if (round > 0)
      for (int k = 0; k < n; k++) advance_width(&t[k]);
    for (int k = 0; k < n; k++) {
      t[k].w_prev = t[k].w_last;
      t[k].acc_prev = t[k].acc_last;
    }
    set_gpu_dr_max();
    sample_widths(t, n, sweeps);
  }
}

void opt_nvt(void) {
  tuned_width t[1] = {{&G_IN.dr_max, G_IN.opt_part_target, 1.0, 1.0, 0, 0, 0, 0}};
  tune_widths(t, 1);
  printf("Optimal maximum displacement: %.8f\n", G_IN.dr_max);
  printf("Acceptance ratio: %.8f\n", t[0].acc_last);
  printf("Maximum displacement optimization completed\n");
}

void opt_npt(void) {
  tuned_width t[2] = {{&G_IN.dr_max, G_IN.opt_part_target, 1.0, 1.0, 0, 0, 0, 0},
                      {&G_IN.dv_max, G_IN.opt_vol_target, 0.0, 0.1, 0, 0, 0, 0}};
  tune_widths(t, 2);
  printf("Optimal maximum particle displacement: %.8f\n", G_IN.dr_max);
  printf("Acceptance ratio: %.8f \n", t[0].acc_last);
  printf("Optimal maximum volume deformation: %.8f\n", G_IN.dv_max);
  printf("Acceptance ratio: %.8f \n", t[1].acc_last);
  printf("Maximum displacement optimization completed\n");
}
