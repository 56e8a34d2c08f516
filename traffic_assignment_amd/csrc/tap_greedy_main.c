/*
 * tap_greedy_main.c -- command line of the greedy step-cost assignment on the
 * MI355X engine, with the reference tap_greedy's interface
 * (trunk/tap/tap_greedy.c:196-349):
 *
 *   tap_greedy_hip netfile tripsfile
 *
 * Flows go to stdout; status lines and the total cost to stderr.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tap_hip.h"

int main(int argc, char *argv[])
{
  if (argc != 3) {
    fprintf(stderr, "Usage: %s netfilename demandfilename\n", argv[0]);
    exit(1);
  }
  FILE *net_fp = fopen(argv[1], "r");
  if (!net_fp) {
    fprintf(stderr, "error opening net file %s: %s\n", argv[1], strerror(errno));
    exit(1);
  }
  FILE *dem_fp = fopen(argv[2], "r");
  if (!dem_fp) {
    fprintf(stderr, "error opening trips file %s: %s\n", argv[2], strerror(errno));
    exit(1);
  }
  net_data_t net;
  if (parse_net_file(net_fp, &net) != 0) {
    fprintf(stderr, "error parsing net file %s\n", argv[1]);
    exit(1);
  }
  fclose(net_fp);
  demand_data_t **demands = NULL;
  long num_zones = 0;
  if (parse_trips_file(dem_fp, &demands, &num_zones) != 0) {
    fprintf(stderr, "error parsing trips file %s\n", argv[2]);
    exit(1);
  }
  fclose(dem_fp);
  if (num_zones != net.num_zones)
    fprintf(stderr, "warning: %ld zones in net data but %ld in trips file\n", net.num_zones,
            num_zones);
  double total_cost;
  int iterations;
  const int rc = tap_greedy_assign(&net, demands, stdout, &total_cost, &iterations);
  for (long o = 0; o < num_zones + 1; o++)
    free(demands[o]);
  free(demands);
  free(net.links);
  return rc ? 1 : 0;
}
