/*
 * input.c -- the global input structure and its keyword grammar.
 *
 * Same defaults (read_input.c:71-115), line grammar (read_input.c:127-448:
 * '#'- or newline-led lines skipped, key = first space-separated token, a key
 * may carry the line's newline) and errors (read_input.c:477-499) as the
 * reference.  Two extension keywords select the GPU sweep ("gpu_sweep
 * serial|checkerboard") and the HIP device ("gpu_device D").
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hsmc.h"

struct input G_IN;

void print_example(void) {
  /* an input file for the NVT virial pressure (read_input.c:18-58 prints an
   * example of the same grammar) */
  static const char example[] =
      "# NVT simulation of hard spheres: pressure from the contact value of g(r),\n"
      "# chemical potential from Widom insertions and the radial distribution.\n"
      "# 500 spheres on an fcc lattice at density 0.5\n\n"
      "# Density\n"
      "rho 0.5\n\n"
      "# Simulation box (lattice cells per axis, 1 = sc, 2 = fcc)\n"
      "cells_x 5\n"
      "cells_y 5\n"
      "cells_z 5\n"
      "type 2\n\n"
      "# Cell list: minimum cell side, particles per cell + 1\n"
      "neigh_list 1.05 10\n\n"
      "# Maximum particle displacement\n"
      "dr_max 0.1\n\n"
      "# Optimizer of the displacement (off)\n"
      "opt 0 1000 10 0.5 0.5\n\n"
      "# Observables\n"
      "press_virial 0.01 10\n"
      "widom 100 10\n"
      "rdf 0.01 5 10 100\n\n"
      "# Seed, sweeps, screen output\n"
      "seed 1\n"
      "sweep_eq 2000\n"
      "sweep_stat 20000\n"
      "out 1000\n\n"
      "# Optional: 'gpu_sweep serial' replays the serial single-particle chain\n"
      "# instead of the default checkerboard sweep\n";
  fputs(example, stdout);
}

void read_input_file_err(int err_id, char *last_string) {
  switch (err_id) {
    case 1: printf("Missing value to key\n"); break;
    case 2: printf("Unknown key\n"); break;
    case 3: printf("Name of restart file is too long, maximum 100 characters\n"); break;
    default: break;
  }
  printf("Last read line in the input file:\n%s\n", last_string);
  exit(EXIT_FAILURE);
}

static void defaults(void) {
  memset(&G_IN, 0, sizeof(G_IN));
  G_IN.rho = 0.5;
  G_IN.nx = 3;
  G_IN.ny = 3;
  G_IN.nz = 3;
  G_IN.type = 1;
  G_IN.neigh_dr = 1.0;
  G_IN.neigh_max_part = 10;
  G_IN.dr_max = 0.05;
  G_IN.sweep_eq = 0;
  G_IN.sweep_stat = 0;
  G_IN.output_int = 0;
  G_IN.cavity_pcav = 0;
  G_IN.cavity_maxdr = 1.2;
  G_IN.cavity_mindr = 0;
  G_IN.cavity_out_dr = 0.01;
  G_IN.cavity_sample_int = 100;
  G_IN.cluster_flag = 0;
  G_IN.cluster_moves_sweep = 1;
  G_IN.cluster_init_step = 10000;
  G_IN.restart_read = 0;
  G_IN.restart_write = 0;
  G_IN.config_write = 0;
  G_IN.config_samples = 128;
  G_IN.pressv_dr = 0.01;
  G_IN.pressv_sample_int = 0;
  G_IN.presst_dxi = 0.0001;
  G_IN.presst_xi_max = 0.002;
  G_IN.presst_sample_int = 0;
  G_IN.press = 0;
  G_IN.dv_max = 0.001;
  G_IN.opt_flag = 1;
  G_IN.opt_sweeps = 1000;
  G_IN.opt_samples = 10;
  G_IN.opt_part_target = 0.5;
  G_IN.opt_vol_target = 0.5;
  G_IN.seed = 0;
  G_IN.ql_order = 6;
  G_IN.ql_rmax = 1.5;
  G_IN.ql_sample_int = 0;
  G_IN.mu_sample_int = 0;
  G_IN.mu_insertions = 100;
  G_IN.rdf_dr = 0.01;
  G_IN.rdf_rmax = 10;
  G_IN.rdf_sample_int = 0;
  G_IN.rdf_samples = 100;
  G_IN.gpu_serial = 0;
  G_IN.gpu_device = 0;
}

/* value kinds of a keyword's fields */
enum { VI, VD, VL, VS };
typedef struct {
  const char *key;
  int n;
  int kind[5];
  void *dst[5];
} keyword;

static int key_is(const char *tok, const char *key) {
  size_t n = strlen(key);
  return strncmp(tok, key, n) == 0 && (tok[n] == '\0' || (tok[n] == '\n' && tok[n + 1] == '\0'));
}

void read_input_file(char *filename) {
  defaults();
  const keyword table[] = {
      {"rho", 1, {VD}, {&G_IN.rho}},
      {"cells_x", 1, {VI}, {&G_IN.nx}},
      {"cells_y", 1, {VI}, {&G_IN.ny}},
      {"cells_z", 1, {VI}, {&G_IN.nz}},
      {"type", 1, {VI}, {&G_IN.type}},
      {"neigh_list", 2, {VD, VI}, {&G_IN.neigh_dr, &G_IN.neigh_max_part}},
      {"dr_max", 1, {VD}, {&G_IN.dr_max}},
      {"sweep_eq", 1, {VI}, {&G_IN.sweep_eq}},
      {"sweep_stat", 1, {VI}, {&G_IN.sweep_stat}},
      {"out", 1, {VI}, {&G_IN.output_int}},
      {"npt", 2, {VD, VD}, {&G_IN.press, &G_IN.dv_max}},
      {"opt", 5, {VI, VI, VI, VD, VD},
       {&G_IN.opt_flag, &G_IN.opt_sweeps, &G_IN.opt_samples, &G_IN.opt_part_target,
        &G_IN.opt_vol_target}},
      {"seed", 1, {VL}, {&G_IN.seed}},
      {"cavity", 5, {VD, VD, VD, VI, VD},
       {&G_IN.cavity_pcav, &G_IN.cavity_maxdr, &G_IN.cavity_mindr, &G_IN.cavity_sample_int,
        &G_IN.cavity_out_dr}},
      {"cluster", 3, {VI, VI, VI},
       {&G_IN.cluster_flag, &G_IN.cluster_moves_sweep, &G_IN.cluster_init_step}},
      {"restart_read", 2, {VI, VS}, {&G_IN.restart_read, G_IN.restart_name}},
      {"restart_write", 1, {VI}, {&G_IN.restart_write}},
      {"config_write", 2, {VI, VI}, {&G_IN.config_write, &G_IN.config_samples}},
      {"press_virial", 2, {VD, VI}, {&G_IN.pressv_dr, &G_IN.pressv_sample_int}},
      {"press_thermo", 3, {VD, VD, VI},
       {&G_IN.presst_dxi, &G_IN.presst_xi_max, &G_IN.presst_sample_int}},
      {"ql", 3, {VI, VD, VI}, {&G_IN.ql_order, &G_IN.ql_rmax, &G_IN.ql_sample_int}},
      {"widom", 2, {VI, VI}, {&G_IN.mu_insertions, &G_IN.mu_sample_int}},
      {"rdf", 4, {VD, VD, VI, VI},
       {&G_IN.rdf_dr, &G_IN.rdf_rmax, &G_IN.rdf_sample_int, &G_IN.rdf_samples}},
  };
  const int nkeys = (int)(sizeof(table) / sizeof(table[0]));

  printf("Reading input data from %s ...\n", filename);
  FILE *in = fopen(filename, "r");
  if (!in) {
    printf("Error! Could not open file %s\n", filename);
    exit(EXIT_FAILURE);
  }
  char *line = NULL;
  size_t cap = 0;
  ssize_t len;
  while ((len = getline(&line, &cap, in)) >= 0) {
    if (line[0] == '#' || line[0] == '\n') continue;
    char *key = strtok(line, " ");
    if (!key) continue;
    int found = 0;
    for (int k = 0; k < nkeys && !found; k++) {
      if (!key_is(key, table[k].key)) continue;
      found = 1;
      for (int f = 0; f < table[k].n; f++) {
        char *v = strtok(NULL, " ");
        if (!v) read_input_file_err(1, line);
        switch (table[k].kind[f]) {
          case VI: *(int *)table[k].dst[f] = atoi(v); break;
          case VD: *(double *)table[k].dst[f] = atof(v); break;
          case VL: *(unsigned long *)table[k].dst[f] = (unsigned long)(long)atoi(v); break;
          case VS: {   /* read_input.c:328-333 */
            int n = (int)strlen(v);
            if (n > 0 && v[n - 1] == '\n') v[n - 1] = '\0';
            if (n < 100) strcpy((char *)table[k].dst[f], v);
            else read_input_file_err(3, line);
            break;
          }
        }
      }
    }
    if (found) continue;
    if (key_is(key, "gpu_sweep")) {
      char *v = strtok(NULL, " \n");
      if (!v) read_input_file_err(1, line);
      if (strcmp(v, "serial") == 0) G_IN.gpu_serial = 1;
      else if (strcmp(v, "checkerboard") == 0) G_IN.gpu_serial = 0;
      else read_input_file_err(2, line);
    } else if (key_is(key, "gpu_device")) {
      char *v = strtok(NULL, " ");
      if (!v) read_input_file_err(1, line);
      G_IN.gpu_device = atoi(v);
    } else {
      read_input_file_err(2, line);
    }
  }
  free(line);
  fclose(in);
  printf("Done\n");
  fflush(stdout);
}
