/*
 * ref_harness.c -- drives the reference's own compiled functions to write
 * golden vectors (TEST INFRASTRUCTURE ONLY; built into oracle/_ref/ by
 * oracle/Makefile from the reference sources where they lie, never shipped).
 *
 * Linked against objects compiled from trunk/tap/{sssp_pape,tap_functions,
 * parsetapfiles,utils}.c.  All doubles are printed with %a so the fixtures
 * round-trip exactly.
 *
 *   ref_harness sssp NET TRIPS COSTS|- ORIG...   dist/prevlink per origin
 *                                                (Pape-LLL and plain Pape)
 *   ref_harness fw NET TRIPS ITERS RELGAP        per-iteration scalars and
 *                                                final volumes/costs
 *
 * The fw mode calls the reference's line_search, link_directional_derivative,
 * links_objective_function, total_link_cost, update_link_costs,
 * adjlist_to_packed_arrays, demand_sum and sssp_pape_lll in the order of
 * tap_frankwolfe_pthread.c:654-803; the all-or-nothing path walk (a static
 * function in the reference) is the loop of tap_frankwolfe.c:131-163.  The
 * golden script checks this loop against the reference's own
 * tap_frankwolfe_pthread binary before writing any fixture.
 */
#include <assert.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "parsetapfiles.h"
#include "sssp_pape.h"
#include "tap_functions.h"

static void load(const char *netf, const char *tripf, net_data_t *net,
                 demand_data_t ***demands, long *nz)
{
  FILE *fp = fopen(netf, "r");
  if (!fp || parse_net_file(fp, net) != 0) { fprintf(stderr, "net parse failed\n"); exit(2); }
  fclose(fp);
  qsort(net->links, net->num_links, sizeof(link_data_t), link_data_compar);
  fp = fopen(tripf, "r");
  if (!fp || parse_trips_file(fp, demands, nz) != 0) { fprintf(stderr, "trips parse failed\n"); exit(2); }
  fclose(fp);
}

static void fix_capacity(net_data_t *net)
{
  double m = DBL_MAX;
  for (long i = 0; i < net->num_links; i++)
    if (net->links[i].capacity > 0.0 && net->links[i].capacity < m)
      m = net->links[i].capacity;
  for (long i = 0; i < net->num_links; i++)
    if (net->links[i].capacity <= 0.0)
      net->links[i].capacity = m;
}

static void aon(net_data_t *net, demand_data_t **demands, long *Va, long *Ea,
                double *Wa, double *y, double *dist, long *q, long *pred)
{
  for (long k = 0; k < net->num_links; k++)
    y[k] = 0;
  for (long orig = 1; orig < net->num_zones + 1; orig++) {
    sssp_pape_lll(Va, Ea, Wa, net->num_nodes + 1, net->num_links, orig,
                  net->first_thru_node, pred, dist, q);
    long dest;
    for (long i = 0; (dest = demands[orig][i].dest) != 0; i++) {
      double f = demands[orig][i].demand;
      if (orig == dest || f <= 0)
        continue;
      long v = dest;
      while (v != orig) {
        long k = pred[v];
        if (k < 0) { fprintf(stderr, "no path %ld %ld\n", orig, dest); exit(3); }
        y[k] += f;
        v = net->links[k].init_node;
      }
    }
  }
}

int main(int argc, char **argv)
{
  if (argc < 4) { fprintf(stderr, "usage: see source\n"); return 2; }
  net_data_t net;
  demand_data_t **demands;
  long nz;
  load(argv[2], argv[3], &net, &demands, &nz);
  fix_capacity(&net);
  const long N1 = net.num_nodes + 1, E = net.num_links;
  long *Va = malloc(sizeof(long) * (N1 + 1)), *Ea = malloc(sizeof(long) * E);
  double *Wa = malloc(sizeof(double) * E);
  adjlist_to_packed_arrays(net.links, N1, E, Va, Ea, Wa);
  double *dist = malloc(sizeof(double) * N1);
  long *pred = malloc(sizeof(long) * N1), *q = malloc(sizeof(long) * N1);

  if (strcmp(argv[1], "sssp") == 0) {
    if (strcmp(argv[4], "-") != 0) {        /* costs: one %a per line */
      FILE *fp = fopen(argv[4], "r");
      for (long k = 0; k < E; k++)
        if (fscanf(fp, "%la", &Wa[k]) != 1) { fprintf(stderr, "costs\n"); return 2; }
      fclose(fp);
    }
    for (int a = 5; a < argc; a++) {
      long o = atol(argv[a]);
      for (int lll = 1; lll >= 0; lll--) {
        if (lll)
          sssp_pape_lll(Va, Ea, Wa, N1, E, o, net.first_thru_node, pred, dist, q);
        else
          sssp_pape(Va, Ea, Wa, N1, E, o, net.first_thru_node, pred, dist, q);
        printf("origin %ld %s\n", o, lll ? "pape_lll" : "pape");
        for (long v = 1; v < N1; v++)
          printf("%ld %a %ld\n", v, dist[v], pred[v]);
      }
    }
    return 0;
  }

  /* fw */
  int iters = atoi(argv[4]);
  double target = atof(argv[5]);
  double *x = malloc(sizeof(double) * E), *y = malloc(sizeof(double) * E);
  double *c = Wa;
  double total_demand = demand_sum(demands, net.num_zones + 1);
  int iter = 0;
  double relgap = DBL_MAX, best_lb = -DBL_MAX;
  aon(&net, demands, Va, Ea, c, x, dist, q, pred);
  update_link_costs(&net, x, c);
  printf("total_demand %a\n", total_demand);
  while ((iters <= 0 || iter < iters) && (target <= 0 || relgap > target)) {
    aon(&net, demands, Va, Ea, c, y, dist, q, pred);
    for (long k = 0; k < E; k++)
      y[k] = y[k] - x[k];
    double lambda = line_search(&net, x, y);
    double gap = -link_directional_derivative(&net, x, y, 0);
    double obj = links_objective_function(&net, x);
    double lb = obj - gap;
    if (lb > best_lb)
      best_lb = lb;
    if (best_lb != 0)
      relgap = (obj - best_lb) / fabs(best_lb);
    double avg_excess = -total_link_cost(&net, y, c) / total_demand;
    double vht = total_link_cost(&net, x, c);
    for (long k = 0; k < E; k++)
      x[k] += lambda * y[k];
    update_link_costs(&net, x, c);
    iter++;
    printf("iter %d %a %a %a %a %a %a %a %a\n", iter, lambda, gap, obj, lb, best_lb,
           relgap, avg_excess, vht);
  }
  printf("final %d %a\n", iter, total_link_cost(&net, x, c));
  for (long k = 0; k < E; k++)
    printf("flow %ld %ld %a %a\n", net.links[k].init_node, net.links[k].term_node, x[k], c[k]);
  return 0;
}
