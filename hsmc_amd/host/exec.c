/*
 * exec.c -- the hsmc command line (exec.c:85-146 of the reference): -i IN_FILE
 * (default in.dat), -o OUT_FILE, -e; NpT when the input sets a pressure,
 * NVT otherwise.  The cavity and cluster ensembles are not part of this build.
 */
#include <argp.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#include "hsmc.h"

static char doc[] =
    "hsmc performs Monte Carlo simulations of mono-disperse hard-sphere systems on an AMD GPU.\n"
    "The program is run as ./hsmc -i IN_FILE where IN_FILE is a file describing\n"
    "the parameters of the simulation. An example of input file can be obtained by\n"
    "running ./hsmc -e.";

static struct argp_option options[] = {
    {"input", 'i', "IN_FILE", 0, "Input read from IN_FILE instead of from  in.dat", 0},
    {"example", 'e', 0, 0, "Print example of input file on screen", 0},
    {"output", 'o', "OUT_FILE", 0, "Output to OUT_FILE instead of standard output", 0},
    {0}};

struct arguments {
  char *input_file;
  char *output_file;
  bool example;
};

static error_t parse_opt(int key, char *arg, struct argp_state *state) {
  struct arguments *a = state->input;
  switch (key) {
    case 'i': a->input_file = arg; break;
    case 'o': a->output_file = arg; break;
    case 'e': a->example = true; break;
    case ARGP_KEY_ARG:
      if (state->arg_num > 0) argp_usage(state);
      break;
    default: return ARGP_ERR_UNKNOWN;
  }
  return 0;
}

static struct argp argp = {options, parse_opt, 0, doc, 0, 0, 0};

int main(int argc, char **argv) {
  struct arguments a;
  bool run = true;
  a.input_file = "in.dat";
  a.output_file = "";
  a.example = false;
  argp_parse(&argp, argc, argv, 0, 0, &a);
  if (a.example) {
    print_example();
    run = false;
  }
  if (strcmp(a.output_file, "") != 0) {
    if (!freopen(a.output_file, "w", stdout)) {
      printf("Failed to pipe output to file %s\n", a.output_file);
      exit(EXIT_FAILURE);
    }
  }
  if (run) {
    read_input_file(a.input_file);
    if (G_IN.press > 0) {
      hs_npt();
    } else if (G_IN.cavity_pcav > 0 || G_IN.cluster_flag > 0) {
      printf("ERROR: cavity and cluster simulations are not part of this build\n");
      exit(EXIT_FAILURE);
    } else {
      hs_nvt();
    }
    printf("Simulation complete!\n");
  }
  return 0;
}
