/*
 * harness.c -- drives fedluc/HSMC's own hot-path functions, compiled from
 * the reference sources where they lie (see oracle/ref/Makefile), to write
 * golden vectors.  TEST INFRASTRUCTURE ONLY: it runs where the reference
 * sources are present, never beside the GPU runs, and only its numeric
 * outputs are committed (under
 * tests/golden/).  No reference source is copied; the module-static state of
 * three observable files is reached by compiling them into this translation
 * unit with #include of their path under the reference tree (the -I flag in
 * the Makefile).
 *
 * The reference's RNG module needs GSL, which is absent; nothing here calls
 * into it, and the linker's --gc-sections drops the functions that would.
 *
 * Usage: ref_harness CMD type nx ny nz rho neigh_dr max_part conf.bin ARGS...
 *   conf.bin : N x 4 doubles (id, x, y, z), or "-" for the lattice.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdbool.h>
#include <math.h>

#include "sim_info.h"
#include "read_input.h"
#include "cell_list.h"
#include "moves.h"

/* module statics reached by compiling these reference files here */
#include "compute_rdf.c"
#include "compute_press.c"
#include "compute_widom_chem_pot.c"

static void *slurp(const char *path, long *bytes) {
  FILE *f = fopen(path, "rb");
  if (!f) { perror(path); exit(2); }
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  void *buf = malloc(n > 0 ? n : 1);
  if (n > 0 && fread(buf, 1, n, f) != (size_t)n) { perror(path); exit(2); }
  fclose(f);
  *bytes = n;
  return buf;
}

static void spit(const char *path, const void *buf, long bytes) {
  FILE *f = fopen(path, "wb");
  if (!f) { perror(path); exit(2); }
  if (bytes > 0 && fwrite(buf, 1, bytes, f) != (size_t)bytes) { perror(path); exit(2); }
  fclose(f);
}

int main(int argc, char **argv) {
  if (argc < 10) { fprintf(stderr, "usage: see header\n"); return 2; }
  const char *cmd = argv[1];
  memset(&G_IN, 0, sizeof(G_IN));
  G_IN.type = atoi(argv[2]);
  G_IN.nx = atoi(argv[3]); G_IN.ny = atoi(argv[4]); G_IN.nz = atoi(argv[5]);
  G_IN.rho = atof(argv[6]);
  G_IN.neigh_dr = atof(argv[7]);
  G_IN.neigh_max_part = atoi(argv[8]);
  const char *conf_path = argv[9];
  char **rest = argv + 10;
  int nrest = argc - 10;

  if (strcmp(cmd, "parse") == 0) {
    /* read_input_file on rest[0]; dump the fields of the hot path */
    read_input_file(rest[0]);
    double d[24] = {G_IN.rho, G_IN.nx, G_IN.ny, G_IN.nz, G_IN.type, G_IN.neigh_dr,
                    G_IN.neigh_max_part, G_IN.dr_max, G_IN.sweep_eq, G_IN.sweep_stat,
                    G_IN.output_int, G_IN.press, G_IN.dv_max, G_IN.opt_flag, (double)G_IN.seed,
                    G_IN.config_write, G_IN.config_samples, G_IN.pressv_dr, G_IN.pressv_sample_int,
                    G_IN.presst_dxi, G_IN.presst_xi_max, G_IN.presst_sample_int,
                    G_IN.mu_insertions, G_IN.mu_sample_int};
    double e[4] = {G_IN.rdf_dr, G_IN.rdf_rmax, G_IN.rdf_sample_int, G_IN.rdf_samples};
    double all[28];
    memcpy(all, d, sizeof(d)); memcpy(all + 24, e, sizeof(e));
    spit(rest[1], all, sizeof(all));
    return 0;
  }

  sim_box_init(G_IN.type, G_IN.nx, G_IN.ny, G_IN.nz, G_IN.rho);
  part_alloc();
  part_init();
  p_info pi = part_info_get();
  config conf = part_config_get();
  if (strcmp(conf_path, "-") != 0) {
    long nb;
    double *c = (double *)slurp(conf_path, &nb);
    if (nb != (long)sizeof(double) * 4 * pi.NN) { fprintf(stderr, "conf size mismatch\n"); return 2; }
    memcpy(conf, c, nb);
    free(c);
  }
  cell_list_init(true);
  cl_info cli = get_cell_list_info();

  if (strcmp(cmd, "box") == 0) {
    box_info b = sim_box_info_get();
    double out[12] = {b.vol, b.lx, b.ly, b.lz, b.min_size, b.cell_size, pi.NN,
                      cli.num_x, cli.num_y, cli.num_z, cli.size_x, cli.size_z};
    spit(rest[0], out, sizeof(out));
    spit(rest[1], conf, sizeof(double) * 4 * pi.NN);
  } else if (strcmp(cmd, "celllist") == 0) {
    int C = cli.num_tot, mp = G_IN.neigh_max_part;
    int *pc = (int *)malloc(sizeof(int) * C * mp), *nm = (int *)malloc(sizeof(int) * C * 27);
    for (int c = 0; c < C; c++) {
      for (int j = 0; j < mp; j++) pc[c * mp + j] = cli.part_cell[c][j];
      for (int j = 0; j < 27; j++) nm[c * 27 + j] = cli.neigh_mat[c][j];
    }
    spit(rest[0], pc, sizeof(int) * C * mp);
    spit(rest[1], nm, sizeof(int) * C * 27);
  } else if (strcmp(cmd, "overlap") == 0) {
    /* rest[0]: m x 4 doubles (idx, x, y, z); overlap of idx placed there */
    long nb;
    double *q = (double *)slurp(rest[0], &nb);
    int m = (int)(nb / (4 * sizeof(double)));
    unsigned char *out = (unsigned char *)malloc(m > 0 ? m : 1);
    for (int k = 0; k < m; k++) {
      int idx = (int)q[4 * k];
      double ox = conf[idx][1], oy = conf[idx][2], oz = conf[idx][3];
      conf[idx][1] = q[4 * k + 1]; conf[idx][2] = q[4 * k + 2]; conf[idx][3] = q[4 * k + 3];
      out[k] = check_overlap(idx, 1.0, 1.0, 1.0) ? 1 : 0;
      conf[idx][1] = ox; conf[idx][2] = oy; conf[idx][3] = oz;
    }
    spit(rest[1], out, m);
  } else if (strcmp(cmd, "scaled") == 0) {
    /* rest[0]: m x 2 doubles (idx, sf) -> check_overlap(idx, sf, sf, sf) */
    long nb;
    double *q = (double *)slurp(rest[0], &nb);
    int m = (int)(nb / (2 * sizeof(double)));
    unsigned char *out = (unsigned char *)malloc(m > 0 ? m : 1);
    for (int k = 0; k < m; k++)
      out[k] = check_overlap((int)q[2 * k], q[2 * k + 1], q[2 * k + 1], q[2 * k + 1]) ? 1 : 0;
    spit(rest[1], out, m);
  } else if (strcmp(cmd, "dist") == 0) {
    /* rest[0]: m x 3 doubles (i, j, sf) -> compute_dist */
    long nb;
    double *q = (double *)slurp(rest[0], &nb);
    int m = (int)(nb / (3 * sizeof(double)));
    double *out = (double *)malloc(sizeof(double) * (m > 0 ? m : 1));
    for (int k = 0; k < m; k++)
      out[k] = compute_dist((int)q[3 * k], (int)q[3 * k + 1], q[3 * k + 2], q[3 * k + 2], q[3 * k + 2]);
    spit(rest[1], out, sizeof(double) * m);
  } else if (strcmp(cmd, "rdf") == 0) {
    /* rest[0] = dr, rest[1] = rmax ; output hist (doubles) then rmax_eff */
    G_IN.rdf_dr = atof(rest[0]);
    G_IN.rdf_rmax = atof(rest[1]);
    box_info b = sim_box_info_get();
    double mx = b.lx;                               /* compute_rdf.c:43-51 */
    if (mx < b.ly) mx = b.ly;
    if (mx < b.lz) mx = b.lz;
    if (mx < 2.0 * G_IN.rdf_rmax) G_IN.rdf_rmax = mx / 2.0;
    rdf_hist_alloc();
    rdf_hist_init();
    rdf_hist_compute();
    double *out = (double *)malloc(sizeof(double) * (rdf_hist_nn + 2));
    memcpy(out, rdf_hist, sizeof(double) * rdf_hist_nn);
    out[rdf_hist_nn] = G_IN.rdf_rmax;
    rdf_hist_norm();
    spit(rest[2], out, sizeof(double) * (rdf_hist_nn + 1));
    spit(rest[3], rdf_hist, sizeof(double) * rdf_hist_nn);
  } else if (strcmp(cmd, "contact") == 0) {
    G_IN.pressv_dr = atof(rest[0]);
    pressv_hist_alloc();
    pressv_hist_init();
    pressv_compute_hist();
    double *out = (double *)malloc(sizeof(double) * (pressv_hist_nn + 1));
    memcpy(out, pressv_hist, sizeof(double) * pressv_hist_nn);
    out[pressv_hist_nn] = pressv_rmax;
    pressv_compute_rdf();
    spit(rest[1], out, sizeof(double) * (pressv_hist_nn + 1));
    spit(rest[2], pressv_hist, sizeof(double) * pressv_hist_nn);
  } else if (strcmp(cmd, "presst") == 0) {
    G_IN.presst_dxi = atof(rest[0]);
    G_IN.presst_xi_max = atof(rest[1]);
    presst_hist_alloc();
    presst_hist_init();
    presst_compute_hist();
    spit(rest[2], presst_hist, sizeof(double) * presst_hist_nn);
  } else if (strcmp(cmd, "widom") == 0) {
    /* rest[0]: m x 3 doubles of trial points -> hit bytes */
    long nb;
    double *q = (double *)slurp(rest[0], &nb);
    int m = (int)(nb / (3 * sizeof(double)));
    unsigned char *out = (unsigned char *)malloc(m > 0 ? m : 1);
    nl = get_cell_list_info();
    for (int k = 0; k < m; k++) {
      r_x = q[3 * k]; r_y = q[3 * k + 1]; r_z = q[3 * k + 2];
      out[k] = widom_check_overlap() ? 0 : 1;
    }
    spit(rest[1], out, m);
  } else if (strcmp(cmd, "moves") == 0) {
    /* rest[0]: m x 4 doubles (idx, rx, ry, rz): the steps of part_move
     * (moves.c:44-78) driven with these random numbers, composed from the
     * reference's own apply_pbc, check_overlap, cell_part_idx and
     * cell_list_update.  rest[1] = dr_max. Outputs: accept bytes, final conf,
     * final part_cell. */
    long nb;
    double *q = (double *)slurp(rest[0], &nb);
    int m = (int)(nb / (4 * sizeof(double)));
    G_IN.dr_max = atof(rest[1]);
    unsigned char *acc = (unsigned char *)malloc(m > 0 ? m : 1);
    for (int k = 0; k < m; k++) {
      int r_idx = (int)q[4 * k];
      double xo = conf[r_idx][1], yo = conf[r_idx][2], zo = conf[r_idx][3];
      int c_old = cell_part_idx(r_idx);
      conf[r_idx][1] += (q[4 * k + 1] - 0.5) * G_IN.dr_max;
      conf[r_idx][2] += (q[4 * k + 2] - 0.5) * G_IN.dr_max;
      conf[r_idx][3] += (q[4 * k + 3] - 0.5) * G_IN.dr_max;
      apply_pbc(r_idx);
      if (check_overlap(r_idx, 1.0, 1.0, 1.0)) {
        conf[r_idx][1] = xo; conf[r_idx][2] = yo; conf[r_idx][3] = zo;
        acc[k] = 0;
      } else {
        int c_new = cell_part_idx(r_idx);
        if (c_new != c_old) cell_list_update(c_old, c_new, r_idx);
        acc[k] = 1;
      }
    }
    spit(rest[2], acc, m);
    spit(rest[3], conf, sizeof(double) * 4 * pi.NN);
    cli = get_cell_list_info();
    int C = cli.num_tot, mp = G_IN.neigh_max_part;
    int *pc = (int *)malloc(sizeof(int) * C * mp);
    for (int c = 0; c < C; c++)
      for (int j = 0; j < mp; j++) pc[c * mp + j] = cli.part_cell[c][j];
    spit(rest[4], pc, sizeof(int) * C * mp);
  } else {
    fprintf(stderr, "unknown command %s (%d args)\n", cmd, nrest);
    return 2;
  }
  return 0;
}
