/*
 * exec.c -- the `hsmc` executable: command-line handling and the choice of
 * ensemble driver.
 *
 * Interface kept from the reference's CLI (exec.c:28-35 options, :122-137
 * dispatch), so existing job scripts keep working:
 *   -i, --input FILE    keyword file to read (default: in.dat)
 *   -o, --output FILE   send the screen log to FILE
 *   -e, --example       print a sample keyword file and stop
 * The driver is hs_npt when the input sets a positive pressure, hs_nvt
 * otherwise.  The cavity-biased and cluster-move ensembles belong to the
 * reference only; an input that asks for them is refused.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>

#include "hsmc.h"

enum cli_action { CLI_RUN, CLI_EXAMPLE, CLI_HELP };

struct cli_request {
  enum cli_action action;
  const char *keyword_file;
  const char *log_file; /* NULL: keep stdout */
};

static void show_help(FILE *to, const char *prog) {
  fprintf(to,
          "usage: %s [-i FILE] [-o FILE] [-e] [-h]\n"
          "Hard-sphere Metropolis Monte Carlo (NVT or NpT) on an MI355X GPU.\n"
          "  -i, --input FILE    keyword file describing the run (default in.dat)\n"
          "  -o, --output FILE   write the run log to FILE rather than the terminal\n"
          "  -e, --example       print a sample keyword file, then exit\n"
          "  -h, --help          show this text\n",
          prog);
}

/* Fills *req from argv; returns 0, or 2 after printing a usage error. */
static int read_command_line(int argc, char **argv, struct cli_request *req) {
  static const struct option longopts[] = {
      {"input", required_argument, NULL, 'i'},
      {"output", required_argument, NULL, 'o'},
      {"example", no_argument, NULL, 'e'},
      {"help", no_argument, NULL, 'h'},
      {NULL, 0, NULL, 0}};
  req->action = CLI_RUN;
  req->keyword_file = "in.dat";
  req->log_file = NULL;
  int opt;
  while ((opt = getopt_long(argc, argv, "i:o:eh", longopts, NULL)) != -1) {
    if (opt == 'i') {
      req->keyword_file = optarg;
    } else if (opt == 'o') {
      req->log_file = optarg;
    } else if (opt == 'e') {
      if (req->action == CLI_RUN) req->action = CLI_EXAMPLE;
    } else if (opt == 'h') {
      req->action = CLI_HELP;
    } else {
      show_help(stderr, argv[0]);
      return 2;
    }
  }
  /* one stray operand is tolerated (the reference's parser allows it); more
   * is a usage error */
  if (argc - optind > 1) {
    fprintf(stderr, "%s: unexpected operand '%s'\n", argv[0], argv[optind + 1]);
    show_help(stderr, argv[0]);
    return 2;
  }
  return 0;
}

int main(int argc, char **argv) {
  struct cli_request req;
  int st = read_command_line(argc, argv, &req);
  if (st) return st;
  if (req.action == CLI_HELP) {
    show_help(stdout, argv[0]);
    return 0;
  }
  /* the sample goes to the terminal even with -o (the log file is still
   * created, empty), as in the reference */
  if (req.action == CLI_EXAMPLE) print_example();
  if (req.log_file && !freopen(req.log_file, "w", stdout)) {
    fprintf(stderr, "hsmc: cannot open log file %s for writing\n", req.log_file);
    return EXIT_FAILURE;
  }
  if (req.action == CLI_EXAMPLE) return 0;

  read_input_file((char *)req.keyword_file); /* the reference signature takes char * */
  const int other_ensemble = G_IN.cavity_pcav > 0 || G_IN.cluster_flag > 0;
  if (G_IN.press > 0) {
    hs_npt();
  } else if (other_ensemble) {
    printf("ERROR: this build runs NVT and NpT only (cavity-biased and cluster runs are not provided)\n");
    return EXIT_FAILURE;
  } else {
    hs_nvt();
  }
  printf("Simulation complete!\n");
  return 0;
}
