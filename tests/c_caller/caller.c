/*
 * caller.c -- a driver written the way the reference's own translation units
 * use its global-state API (set-up, per-sweep observables, counters, I/O,
 * clean-up), compiled against include/hsmc.h and linked with libhsmc.so.  It
 * checks that a reference-style caller links and runs against the host layer
 * (the Makefile binding of the integration notes), not only the CLI.
 *
 * Usage: caller IN_FILE.  Runs sweep_stat sweeps with sweep_nvt, sampling the
 * contact histogram and Widom insertions at their intervals, checks every
 * sphere with check_overlap, writes the final configuration and prints one
 * summary line.  Exit status 1 on an overlap or an inconsistent counter.
 */
#include <stdio.h>
#include <stdlib.h>

#include "hsmc.h"

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s IN_FILE\n", argv[0]);
    return 2;
  }
  read_input_file(argv[1]);
  sim_box_init(G_IN.type, G_IN.nx, G_IN.ny, G_IN.nz, G_IN.rho);
  part_alloc();
  part_init();
  rng_init();
  cell_list_init(true);
  reset_moves_counters();

  int pressv_first = 1, mu_first = 1;
  for (int s = 0; s < G_IN.sweep_stat; s++) {
    if (G_IN.pressv_sample_int > 0 && s % G_IN.pressv_sample_int == 0) {
      compute_pressv(pressv_first);
      pressv_first = 0;
    }
    if (G_IN.mu_sample_int > 0 && s % G_IN.mu_sample_int == 0) {
      compute_mu(mu_first);
      mu_first = 0;
    }
    sweep_nvt();
  }

  int pm, apm, rpm, vm, avm, rvm;
  get_moves_counters(&pm, &apm, &rpm, &vm, &avm, &rvm);
  const p_info pi = part_info_get();
  const box_info bi = sim_box_info_get();
  int overlaps = 0;
  for (int i = 0; i < pi.NN; i++) overlaps += check_overlap(i, 1.0, 1.0, 1.0) ? 1 : 0;
  /* the host copy is current: a pair distance from it equals the layer's */
  const config conf = part_config_get();
  int wrapped = 0;
  for (int i = 0; i < pi.NN; i++)
    for (int k = 1; k <= 3; k++) {
      const double L = k == 1 ? bi.lx : (k == 2 ? bi.ly : bi.lz);
      if (!(conf[i][k] >= 0.0 && conf[i][k] <= L)) wrapped++;
    }
  write_config(G_IN.sweep_stat);
  if (!pressv_first) pressv_hist_free();
  cell_list_free();
  part_free();
  rng_free();

  const long want = (long)pi.NN * G_IN.sweep_stat;
  printf("N %d sweeps %d attempts %d accepted %d overlaps %d outside %d\n", pi.NN, G_IN.sweep_stat,
         pm, apm, overlaps, wrapped);
  return (overlaps || wrapped || pm != want || apm <= 0 || apm > pm) ? 1 : 0;
}
