/*
 * observables.c -- pressure (virial and thermodynamic routes), Widom chemical
 * potential, g(r), configuration and restart files.
 *
 * Counting runs on the GPU (pair histograms, all-particle overlap
 * reductions, Widom hits); normalisation and file output stay on the host
 * with the reference's arithmetic and formats:
 *   press_virial.dat  compute_press.c:47-187, :277-309
 *   press_thermo.dat  compute_press.c:191-273, :311-359 (+ density.dat, NpT)
 *   chem_pot.dat      compute_widom_chem_pot.c:32-71, :164-182
 *   rdf_%06d.dat.gz   compute_rdf.c:36-205
 *   config_%06d.dat.gz, restart_*.bin  io_config.c:28-191
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "hsmc.h"
#include "hsmc_gpu.h"
#include "host_internal.h"

/* ---- virial pressure: contact histogram ---- */
static int pressv_hist_nn;
static double pressv_rmax = 1.05;   /* compute_press.c:36 */
static double *pressv_rr = NULL, *pressv_hist = NULL;

static void pressv_hist_alloc(void) {   /* compute_press.c:85-95 */
  pressv_hist_nn = (int)((pressv_rmax - 1.0) / G_IN.pressv_dr);
  pressv_rr = (double *)host_alloc(sizeof(double) * pressv_hist_nn);
  pressv_hist = (double *)host_alloc(sizeof(double) * pressv_hist_nn);
  if (pressv_rr == NULL || pressv_hist == NULL) {
    printf("ERROR: Failed histogram allocation\n");
    exit(EXIT_FAILURE);
  }
}

void pressv_hist_free(void) {
  free(pressv_rr);
  free(pressv_hist);
  pressv_rr = pressv_hist = NULL;
}

static void pressv_hist_init(void) {   /* compute_press.c:107-119 */
  for (int ii = 0; ii < pressv_hist_nn; ii++) {
    pressv_rr[ii] = (ii + 1. / 2.) * G_IN.pressv_dr + 1.0;
    pressv_hist[ii] = 0.0;
  }
  pressv_rmax = G_IN.pressv_dr * pressv_hist_nn + 1.0;
}

static void shell_normalise(double *hist, const double *rr, int nn) {
  /* pressv_compute_rdf / rdf_hist_norm (compute_press.c:170-187, compute_rdf.c:133-150) */
  p_info pi = part_info_get();
  double dr = rr[1] - rr[0];
  for (int ii = 0; ii < nn; ii++) {
    double r1 = rr[ii] - dr / 2.;
    double r2 = rr[ii] + dr / 2.;
    double bin_vol = (4. * M_PI / 3.) * (pow(r2, 3.) - pow(r1, 3.));
    hist[ii] = hist[ii] / (bin_vol * G_IN.rho * pi.NN);
  }
}

void compute_pressv(bool init) {
  if (init) pressv_hist_alloc();
  pressv_hist_init();
  if (init) {   /* compute_press.c:58-71: checks the largest cell side */
    cl_info nl = get_cell_list_info();
    double m = nl.size_x;
    if (m < nl.size_y) m = nl.size_y;
    if (m < nl.size_z) m = nl.size_z;
    if (m < pressv_rmax) {
      printf("ERROR: The size of the cells in the neighbor list does not allow a correct calculation of the pressure, increase to %f\n",
             pressv_rmax);
      exit(EXIT_FAILURE);
    }
  }
  host_sync_device();
  int64_t *cnt = (int64_t *)host_alloc(sizeof(int64_t) * (pressv_hist_nn > 0 ? pressv_hist_nn : 1));
  int64_t oor = 0;
  host_gpu_check(hsmc_gpu_contact_counts(hsmc_host_gpu(), 0, G_IN.pressv_dr, pressv_hist_nn,
                                         pressv_rmax, cnt, &oor),
                 "contact_counts");
  for (int ii = 0; ii < pressv_hist_nn; ii++) pressv_hist[ii] = 2.0 * (double)cnt[ii];
  free(cnt);
  if (pressv_hist_nn >= 2) shell_normalise(pressv_hist, pressv_rr, pressv_hist_nn);
  box_info b = sim_box_info_get();
  p_info pi = part_info_get();
  FILE *fid = fopen("press_virial.dat", init ? "w" : "a");
  if (fid == NULL) {
    perror("Error while creating the file for the virial pressure\n");
    exit(EXIT_FAILURE);
  }
  fprintf(fid, "######################################\n");
  fprintf(fid, "# Bins, volume, number of particles\n");
  fprintf(fid, "######################################\n");
  fprintf(fid, "%d %.8e %d\n", pressv_hist_nn, b.vol, pi.NN);
  fprintf(fid, "###############################\n");
  fprintf(fid, "# rr, rdf\n");
  fprintf(fid, "###############################\n");
  for (int ii = 0; ii < pressv_hist_nn; ii++) fprintf(fid, "%.8e %.8e\n", pressv_rr[ii], pressv_hist[ii]);
  fclose(fid);
}

/* ---- thermodynamic pressure: virtual compressions ---- */
static int presst_hist_nn;
static double *presst_xi = NULL, *presst_hist = NULL;

void presst_hist_free(void) {
  free(presst_xi);
  free(presst_hist);
  presst_xi = presst_hist = NULL;
}

void compute_presst(bool init) {
  if (init) {   /* compute_press.c:209-219 */
    presst_hist_nn = (int)(G_IN.presst_xi_max / G_IN.presst_dxi);
    presst_xi = (double *)host_alloc(sizeof(double) * (presst_hist_nn > 0 ? presst_hist_nn : 1));
    presst_hist = (double *)host_alloc(sizeof(double) * (presst_hist_nn > 0 ? presst_hist_nn : 1));
    if (presst_xi == NULL || presst_hist == NULL) {
      printf("ERROR: Failed histogram allocation\n");
      exit(EXIT_FAILURE);
    }
  }
  for (int ii = 0; ii < presst_hist_nn; ii++) {   /* compute_press.c:233-236 */
    presst_xi[ii] = (ii + 1) * G_IN.presst_dxi;
    presst_hist[ii] = 0.0;
  }
  host_sync_device();
  /* one all-particle reduction for up to 64 compressions at a time */
  for (int base = 0; base < presst_hist_nn; base += 64) {
    int n = presst_hist_nn - base < 64 ? presst_hist_nn - base : 64;
    double sf[64];
    uint8_t ok[64];
    for (int k = 0; k < n; k++) sf[k] = pow(1 - presst_xi[base + k], 1. / 3.);   /* :253-255 */
    host_gpu_check(hsmc_gpu_any_overlap_scaled(hsmc_host_gpu(), 0, sf, n, ok), "any_overlap_scaled");
    for (int k = 0; k < n; k++)
      if (ok[k]) presst_hist[base + k] += 1.0;
  }
  box_info b = sim_box_info_get();
  p_info pi = part_info_get();
  FILE *fid = fopen("press_thermo.dat", init ? "w" : "a");
  if (fid == NULL) {
    perror("Error while creating the file for the thermo pressure\n");
    exit(EXIT_FAILURE);
  }
  fprintf(fid, "######################################\n");
  fprintf(fid, "# Bins, volume, number of particles\n");
  fprintf(fid, "######################################\n");
  fprintf(fid, "%d %.8e %d\n", presst_hist_nn, b.vol, pi.NN);
  fprintf(fid, "######################################\n");
  fprintf(fid, "# abs(xi), exp(-beta*U)\n");
  fprintf(fid, "######################################\n");
  for (int ii = 0; ii < presst_hist_nn; ii++) fprintf(fid, "%.8e %.8e\n", presst_xi[ii], presst_hist[ii]);
  fclose(fid);
  if (G_IN.press > 0) {
    fid = fopen("density.dat", init ? "w" : "a");
    if (fid == NULL) {
      perror("Error while creating the file for the density\n");
      exit(EXIT_FAILURE);
    }
    if (init) {
      fprintf(fid, "######################################\n");
      fprintf(fid, "# Density (each line is one sample)\n");
      fprintf(fid, "######################################\n");
    }
    fprintf(fid, "%.8e\n", G_IN.rho);
    fclose(fid);
  }
}

/* ---- Widom insertions ---- */
static uint64_t widom_sample = 0;

void compute_mu(bool init) {
  host_sync_device();
  p_info pi = part_info_get();
  (void)pi;
  box_info b = sim_box_info_get();
  int64_t wtest = 0;
  const int n = G_IN.mu_insertions;
  if (G_IN.gpu_serial) {
    /* the reference's trial points from the MT stream (widom_rand_pos) */
    double *q = (double *)host_alloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
      q[3 * i] = rng_get_double() * b.lx;
      q[3 * i + 1] = rng_get_double() * b.ly;
      q[3 * i + 2] = rng_get_double() * b.lz;
    }
    host_gpu_check(hsmc_gpu_widom_points(hsmc_host_gpu(), 0, q, n, NULL, &wtest), "widom_points");
    free(q);
  } else {
    host_gpu_check(hsmc_gpu_widom_uniform(hsmc_host_gpu(), 0, host_philox_seed() ^ 0x5749444F4DULL,
                                          widom_sample, 0, n, &wtest),
                   "widom_uniform");
  }
  widom_sample++;
  double mu = wtest > 0 ? -log((double)wtest / n) : 0.0;   /* :64-69 */
  FILE *fid = fopen("chem_pot.dat", init ? "w" : "a");
  if (fid == NULL) {
    perror("Error while creating the file for the chemical potential\n");
    exit(EXIT_FAILURE);
  }
  if (init) {
    fprintf(fid, "##################################################################################\n");
    fprintf(fid, "# Chemical potenital (average over %d insertions, Fraction of accepted insertions)\n", n);
    fprintf(fid, "##################################################################################\n");
  }
  fprintf(fid, "%.8e %.8e\n", mu, (double)wtest / n);
  fclose(fid);
}

/* ---- radial distribution function ---- */
static int rdf_hist_nn;
static double *rdf_rr = NULL, *rdf_hist = NULL;
static int rdf_samples = 0, rdf_file_id = 0;

void rdf_hist_free(void) {
  free(rdf_hist);
  free(rdf_rr);
  rdf_hist = rdf_rr = NULL;
}

void compute_rdf(bool init, int sweep) {
  if (init) {   /* compute_rdf.c:39-56: cutoff against the largest box edge */
    box_info b = sim_box_info_get();
    double m = b.lx;
    if (m < b.ly) m = b.ly;
    if (m < b.lz) m = b.lz;
    if (m < 2.0 * G_IN.rdf_rmax) {
      G_IN.rdf_rmax = m / 2.0;
      printf("WARNING: Cutoff for the rdf extraction reduced to %f in order to be consistent with minimum image convention\n",
             G_IN.rdf_rmax);
    }
    rdf_hist_nn = (int)((G_IN.rdf_rmax - 1.0) / G_IN.rdf_dr);
    rdf_rr = (double *)host_alloc(sizeof(double) * rdf_hist_nn);
    rdf_hist = (double *)host_alloc(sizeof(double) * rdf_hist_nn);
    if (rdf_rr == NULL || rdf_hist == NULL) {
      printf("ERROR: Failed histogram allocation\n");
      exit(EXIT_FAILURE);
    }
  }
  for (int ii = 0; ii < rdf_hist_nn; ii++) {   /* rdf_hist_init (:95-106) */
    rdf_rr[ii] = (ii + 1. / 2.) * G_IN.rdf_dr + 1.0;
    rdf_hist[ii] = 0.0;
  }
  G_IN.rdf_rmax = G_IN.rdf_dr * rdf_hist_nn + 1.0;
  host_sync_device();
  int64_t *cnt = (int64_t *)host_alloc(sizeof(int64_t) * rdf_hist_nn);
  int64_t oor = 0;
  host_gpu_check(hsmc_gpu_rdf_counts(hsmc_host_gpu(), 0, G_IN.rdf_dr, rdf_hist_nn, G_IN.rdf_rmax,
                                     cnt, &oor),
                 "rdf_counts");
  for (int ii = 0; ii < rdf_hist_nn; ii++) rdf_hist[ii] = 2.0 * (double)cnt[ii];
  free(cnt);
  shell_normalise(rdf_hist, rdf_rr, rdf_hist_nn);
  if (init) {   /* rdf_output (:155-205) */
    if ((double)(G_IN.sweep_stat + G_IN.sweep_eq) / (G_IN.rdf_sample_int * G_IN.rdf_samples) > 100000)
      printf("ERROR: Too many (> 100000) rdf files will be produced. Consider increasing number of samples per file\n");
  }
  char name[64];
  snprintf(name, sizeof(name), "rdf_%06d.dat.gz", rdf_file_id);
  gzFile fid = Z_NULL;
  if (rdf_samples == 0) fid = gzopen(name, "w");
  else if (rdf_samples > 0 && rdf_samples < G_IN.rdf_samples) fid = gzopen(name, "a");
  if (fid == Z_NULL) {
    perror("Error while creating rdf file");
    exit(EXIT_FAILURE);
  }
  p_info pi = part_info_get();
  box_info b = sim_box_info_get();
  gzprintf(fid, "######################################\n");
  gzprintf(fid, "# Sweep, Bins, volume, number of particles\n");
  gzprintf(fid, "######################################\n");
  gzprintf(fid, "%d %d %.8e %d\n", sweep, rdf_hist_nn, b.vol, pi.NN);
  gzprintf(fid, "###############################\n");
  gzprintf(fid, "# rr, rdf\n");
  gzprintf(fid, "###############################\n");
  for (int ii = 0; ii < rdf_hist_nn; ii++) gzprintf(fid, "%.8e %.8e\n", rdf_rr[ii], rdf_hist[ii]);
  gzclose(fid);
  rdf_samples += 1;
  if (rdf_samples == G_IN.rdf_samples) {
    rdf_samples = 0;
    rdf_file_id += 1;
  }
}

/* ---- configuration and restart files (io_config.c) ---- */
static int config_output_check = 1, restart_output_check = 1;
static int config_samples = 0, config_file_id = 0;

void write_config(int sweep) {
  if (config_output_check) {
    if ((double)(G_IN.sweep_stat + G_IN.sweep_eq) / (G_IN.config_write * G_IN.config_samples) > 100000)
      printf("ERROR: Too many (> 100000) configuration files will be produced. Consider increasing number of samples per file\n");
    config_output_check = 0;
  }
  char name[64];
  snprintf(name, sizeof(name), "config_%06d.dat.gz", config_file_id);
  gzFile fid = Z_NULL;
  if (config_samples == 0) fid = gzopen(name, "w");
  else if (config_samples > 0 && config_samples < G_IN.config_samples) fid = gzopen(name, "a");
  if (fid == Z_NULL) {
    perror("Error while creating configuration file");
    exit(EXIT_FAILURE);
  }
  p_info pi = part_info_get();
  box_info b = sim_box_info_get();
  config c = part_config_get();
  gzprintf(fid, "# Sweep number\n");
  gzprintf(fid, "%d\n", sweep);
  gzprintf(fid, "# Number of particles\n");
  gzprintf(fid, "%d\n", pi.NN);
  gzprintf(fid, "# Simulation box size\n");
  gzprintf(fid, "%.8f\n", b.lx);
  gzprintf(fid, "%.8f\n", b.ly);
  gzprintf(fid, "%.8f\n", b.lz);
  gzprintf(fid, "# Configuration\n");
  for (int ii = 0; ii < pi.NN; ii++)
    gzprintf(fid, "%d %.8f %.8f %.8f\n", (int)c[ii][0], c[ii][1], c[ii][2], c[ii][3]);
  gzclose(fid);
  config_samples += 1;
  if (config_samples == G_IN.config_samples) {
    config_samples = 0;
    config_file_id += 1;
  }
}

/* Restart file: the reference's fields in its order and layout
 * (io_config.c:53-69, sim_info.c:220-249): dr_max, dv_max, box_info,
 * p_info, N x 4 doubles, MT19937 state (GSL's mt19937 state struct); then a
 * trailer with the sweep-stream position of the GPU sweep (an 8-byte tag and
 * the counter), which a reader that stops after the MT state ignores. */
static const uint64_t kRestartTag = 0x3130504753434d48ull;   /* "HMCSGP01" */
void write_restart(int sweep) {
  if (restart_output_check) {
    if ((double)(G_IN.sweep_stat + G_IN.sweep_eq) / (G_IN.restart_write) > 100000)
      printf("ERROR: Too many (> 100000) restart files will be produced. Consider writing less often\n");
    restart_output_check = 0;
  }
  int width = (int)ceil(log10(G_IN.sweep_stat + G_IN.sweep_eq));
  if (width < 1) width = 1;
  char tmpl[32], name[64];
  snprintf(tmpl, sizeof(tmpl), "restart_%%0%dd.bin", width);
  snprintf(name, sizeof(name), tmpl, sweep);
  FILE *fid = fopen(name, "wb");
  if (fid == NULL) {
    perror("Error while creating restart file\n");
    exit(EXIT_FAILURE);
  }
  fwrite(&G_IN.dr_max, sizeof(double), 1, fid);
  fwrite(&G_IN.dv_max, sizeof(double), 1, fid);
  sim_box_info_write(fid);
  part_info_write(fid);
  part_conf_write(fid);
  rng_write(fid);
  uint64_t trailer[2] = {kRestartTag, host_sweep_counter()};
  fwrite(trailer, sizeof(uint64_t), 2, fid);
  fclose(fid);
}

void read_restart(char *restart_file) {
  printf("Reading data from restart file %s...\n", restart_file);
  FILE *fid = fopen(restart_file, "rb");
  if (fid == NULL) {
    perror("Error while reading restart file");
    exit(EXIT_FAILURE);
  }
  if (fread(&G_IN.dr_max, sizeof(double), 1, fid) != 1 || fread(&G_IN.dv_max, sizeof(double), 1, fid) != 1) {
    printf("ERROR: short restart file\n");
    exit(EXIT_FAILURE);
  }
  sim_box_info_read(fid);
  part_info_read(fid);
  part_alloc();
  part_conf_read(fid);
  rng_read(fid);
  uint64_t trailer[2];
  if (fread(trailer, sizeof(uint64_t), 2, fid) == 2 && trailer[0] == kRestartTag)
    host_set_sweep_counter(trailer[1]);
  fclose(fid);
  box_info b = sim_box_info_get();
  p_info pi = part_info_get();
  G_IN.rho = pi.NN / b.vol;
  printf("The following data was initialized via the restart file:\n"
         "- Maximum displacement (override with optimization)\n"
         "- Maximum volume displacement (only for NpT, override with optimization)\n"
         "- Dimensions of the simulation box\n"
         "- Cell list information\n"
         "- Number of particles\n"
         "- Particle's positions\n"
         "- Density\n"
         "- Status of the random number generator\n");
}
