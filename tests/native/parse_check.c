/* Host check of the C TNTP readers in libtap_hip.so (TEST INFRASTRUCTURE):
 * parse a net / trips / flows / mods file and print what was read so the
 * Python test can compare it with its own reader.
 *   parse_check net FILE | trips FILE | flows FILE | mods FILE | tripmods FILE */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "tap_hip.h"

int main(int argc, char **argv)
{
  if (argc != 3) return 2;
  FILE *fp = fopen(argv[2], "r");
  if (!fp) return 2;
  if (!strcmp(argv[1], "net")) {
    net_data_t n;
    if (parse_net_file(fp, &n)) return 1;
    printf("%ld %ld %ld %ld\n", n.num_zones, n.num_nodes, n.first_thru_node, n.num_links);
    for (long k = 0; k < n.num_links; k++)
      printf("%ld %ld %a %a %a %a %a %d\n", n.links[k].init_node, n.links[k].term_node,
             n.links[k].capacity, n.links[k].length, n.links[k].free_flow_time, n.links[k].B,
             n.links[k].power, n.links[k].link_type);
  } else if (!strcmp(argv[1], "trips")) {
    demand_data_t **d;
    long z;
    if (parse_trips_file(fp, &d, &z)) return 1;
    printf("%ld\n", z);
    for (long o = 1; o <= z; o++)
      for (long i = 0; d[o][i].dest; i++)
        printf("%ld %ld %a\n", o, d[o][i].dest, d[o][i].demand);
  } else if (!strcmp(argv[1], "flows")) {
    link_flow_t *f;
    long n;
    if (parse_flows_file(fp, &f, &n)) return 1;
    printf("%ld\n", n);
    for (long k = 0; k < n; k++)
      printf("%ld %ld %a %a\n", f[k].init_node, f[k].term_node, f[k].volume, f[k].cost);
  } else if (!strcmp(argv[1], "mods")) {
    net_mod_t *m;
    int n;
    if (parse_net_mod_file(fp, &m, &n)) return 1;
    printf("%d\n", n);
    for (int k = 0; k < n; k++)
      printf("%s %d %ld %ld %a %a\n", m[k].change_id, m[k].modtype, m[k].mod_link.init_node,
             m[k].mod_link.term_node, m[k].mod_link.capacity, m[k].project_cost);
  } else if (!strcmp(argv[1], "tripmods")) {
    trip_mod_t *m;
    int n;
    if (parse_trip_mod_file(fp, &m, &n)) return 1;
    printf("%d\n", n);
    for (int k = 0; k < n; k++)
      printf("%s %d %ld %ld %a\n", m[k].trip_change_id, m[k].trip_modtype, m[k].origin,
             m[k].dest, m[k].multiplier);
  }
  return 0;
}
