/* Host check of the trip-modification path (TEST INFRASTRUCTURE): parse a
 * trips file and a trip-modification file, apply every modification in file
 * order, print each return code and then every demand (%a).  Linked once
 * against libtap_hip.so and once against the reference's own parser and
 * tap_functions objects (oracle/_ref), so the two runs' stdout and stderr
 * can be compared.
 *   tripmod_check TRIPS TRIPMODS */
#include <stdio.h>
#include <stdlib.h>
#include "tap_hip.h"

int main(int argc, char **argv)
{
  if (argc != 3) return 2;
  FILE *tf = fopen(argv[1], "r"), *mf = fopen(argv[2], "r");
  if (!tf || !mf) return 2;
  demand_data_t **d;
  long z;
  trip_mod_t *m;
  int n;
  if (parse_trips_file(tf, &d, &z)) return 1;
  if (parse_trip_mod_file(mf, &m, &n)) return 1;
  for (int k = 0; k < n; k++)
    printf("%d\n", apply_trip_modification(d, &m[k], z));
  for (long o = 1; o <= z; o++)
    for (long i = 0; d[o][i].dest; i++)
      printf("%ld %ld %a\n", o, d[o][i].dest, d[o][i].demand);
  return 0;
}
