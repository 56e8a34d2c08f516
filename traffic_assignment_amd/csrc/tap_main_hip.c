/*
 * tap_main_hip.c -- command line of the MI355X Frank-Wolfe solver, with the
 * flags of the reference's tap_frankwolfe_pthread (trunk/tap/tap_main.c:
 * 102-277):
 *
 *   tap_frankwolfe_hip [-n threads] [-w flows] [-i iters] [-r relgap]
 *                      [-t tripmods] [-v] netfile tripsfile
 *
 * Flows go to stdout; "iterations = ... total cost = ..." to stderr.  With
 * neither -i nor -r the iteration limit is 10000.  -n is accepted for
 * compatibility and ignored.
 */
#include <errno.h>
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "tap_hip.h"

static void usage(const char *prog)
{
  fprintf(stderr,
          "Usage: %s [-n numthreads] [-w warmstart_flows_filename] [-i iterations]\n"
          "          [-r relgap] [-t tripmods_filename] [-v] netfilename demandfilename\n"
          "  -n numthreads : accepted and ignored (the GPU does the work)\n"
          "  -w flows      : warm start from a previous solution\n"
          "  -i iterations : stop after this many iterations (default 10000)\n"
          "  -r relgap     : stop when relative gap <= relgap\n"
          "  -t tripmods   : modify O-D demand data\n"
          "  -v            : write status each iteration to stderr\n", prog);
  exit(1);
}

static FILE *open_or_die(const char *name, const char *what)
{
  FILE *fp = fopen(name, "r");
  if (!fp) {
    fprintf(stderr, "error opening %s file %s: %s\n", what, name, strerror(errno));
    exit(1);
  }
  return fp;
}

int main(int argc, char *argv[])
{
  int c, target_iterations = 0, verbose = 0, num_trip_mods = 0;
  double target_relgap = 0;
  const char *warm_name = NULL, *tripmods_name = NULL;
  trip_mod_t *trip_mods = NULL;
  while ((c = getopt(argc, argv, "n:w:i:r:t:v")) != -1) {
    switch (c) {
    case 'n':
      if (atoi(optarg) < 1) usage(argv[0]);
      num_threads = (unsigned)atoi(optarg);
      break;
    case 'w': warm_name = optarg; break;
    case 'i':
      if (atoi(optarg) < 1) {
        fprintf(stderr, "iteration target must be >= 1\n");
        usage(argv[0]);
      }
      target_iterations = atoi(optarg);
      break;
    case 'r':
      if (atof(optarg) <= 0) {
        fprintf(stderr, "relative gap must be positive\n");
        usage(argv[0]);
      }
      target_relgap = atof(optarg);
      break;
    case 't': tripmods_name = optarg; break;
    case 'v': verbose = 1; break;
    default: usage(argv[0]);
    }
  }
  if (argc - optind != 2)
    usage(argv[0]);
  if (target_relgap <= 0 && target_iterations <= 0)
    target_iterations = 10000;

  if (tripmods_name) {
    FILE *fp = open_or_die(tripmods_name, "trip modifications");
    if (parse_trip_mod_file(fp, &trip_mods, &num_trip_mods) != 0) {
      fprintf(stderr, "parse trip mods file %s failed\n", tripmods_name);
      exit(1);
    }
    fclose(fp);
  }
  FILE *net_fp = open_or_die(argv[optind], "net");
  FILE *dem_fp = open_or_die(argv[optind + 1], "trips");
  FILE *warm_fp = warm_name ? open_or_die(warm_name, "flows input") : NULL;

  net_data_t net;
  if (parse_net_file(net_fp, &net) != 0) {
    fprintf(stderr, "error parsing net file\n");
    exit(1);
  }
  fclose(net_fp);
  qsort(net.links, net.num_links, sizeof(link_data_t), link_data_compar);
  demand_data_t **demands = NULL;
  long num_zones = 0;
  if (parse_trips_file(dem_fp, &demands, &num_zones) != 0) {
    fprintf(stderr, "error parsing trips file\n");
    return 1;
  }
  fclose(dem_fp);
  if (num_zones != net.num_zones)
    fprintf(stderr, "warning: %ld zones in net data but %ld in trips file\n", net.num_zones,
            num_zones);
  for (int k = 0; k < num_trip_mods; k++)
    apply_trip_modification(demands, &trip_mods[k], num_zones);

  double total_cost, objvalue, relgap;
  int iterations;
  if (tap_frankwolfe(&net, demands, warm_fp != NULL, warm_fp, target_iterations, target_relgap,
                     stdout, &total_cost, &objvalue, &iterations, &relgap, verbose) != 0) {
    fprintf(stderr, "failed\n");
    exit(1);
  }
  fprintf(stderr, "iterations = %d objective value = %.15f relative gap = %.15f total cost = %.15f\n",
          iterations, objvalue, relgap, total_cost);
  free(net.links);
  free(trip_mods);
  return 0;
}
