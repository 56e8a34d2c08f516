/*
 * tntp_io.c -- Bar-Gera TNTP readers and writers with the reference's
 * calling conventions (include/tap_hip.h).  Behaviour follows
 * trunk/tap/parsetapfiles.c (metadata tags, '~' comment lines, demands
 * terminated by dest == 0, missing origins get an empty list) and
 * trunk/tap/tap_functions.c:431-445,521-598 (flow table, trip changes).
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tap_hip.h"

#define LINE_MAX_LEN 8192

static int blank_or_comment(const char *s)
{
  if (s[0] == '~')
    return 1;
  for (; *s; s++)
    if (!isspace((unsigned char)*s))
      return 0;
  return 1;
}

/* "<TAG> value": returns the value text after '>' or NULL */
static const char *tag_value(const char *line, const char *tag)
{
  size_t n = strlen(tag);
  if (strncmp(line, tag, n) != 0)
    return NULL;
  return line + n;
}

static int grow(void **p, long *cap, long need, size_t elem)
{
  if (need <= *cap)
    return 0;
  long nc = *cap ? *cap : 64;
  while (nc < need)
    nc *= 2;
  void *q = realloc(*p, (size_t)nc * elem);
  if (!q)
    return -1;
  *p = q;
  *cap = nc;
  return 0;
}

int parse_net_file(FILE *fp, net_data_t *net)
{
  char line[LINE_MAX_LEN];
  long cap = 0, count = 0;
  const char *v;
  net->links = NULL;
  net->num_zones = net->num_nodes = net->first_thru_node = net->num_links = 0;
  while (fgets(line, sizeof line, fp)) {
    if (blank_or_comment(line))
      continue;
    if (line[0] == '<') {
      if (!strchr(line, '>')) {
        fprintf(stderr, "error: no closing '>' on metadata tag: %s\n", line);
        return -1;
      }
      if ((v = tag_value(line, "<NUMBER OF ZONES>"))) net->num_zones = atol(v);
      else if ((v = tag_value(line, "<NUMBER OF NODES>"))) net->num_nodes = atol(v);
      else if ((v = tag_value(line, "<FIRST THRU NODE>"))) net->first_thru_node = atol(v);
      else if ((v = tag_value(line, "<NUMBER OF LINKS>"))) net->num_links = atol(v);
      else if (!tag_value(line, "<END OF METADATA>"))
        fprintf(stderr, "warning: unknown metadata tag: %s", line);
      continue;
    }
    if (grow((void **)&net->links, &cap, count + 1, sizeof(link_data_t)))
      return -1;
    link_data_t *L = &net->links[count];
    char *p = line, *end;
    double f[8];
    L->init_node = strtol(p, &end, 10); p = end;
    L->term_node = strtol(p, &end, 10); p = end;
    for (int i = 0; i < 7; i++) { f[i] = strtod(p, &end); p = end; }
    L->capacity = f[0]; L->length = f[1]; L->free_flow_time = f[2]; L->B = f[3];
    L->power = f[4]; L->speed_limit = f[5]; L->toll = f[6];
    L->link_type = (short)strtol(p, &end, 10);
    count++;
  }
  if (count != net->num_links) {
    fprintf(stderr, "error: num_links %ld but %ld links\n", net->num_links, count);
    return -1;
  }
  return 0;
}

int parse_trips_file(FILE *fp, demand_data_t **demands[], long *num_zones)
{
  char line[LINE_MAX_LEN];
  const char *v;
  double total_flow = 0, total_read = 0;
  long cur = 0, *count = NULL, *cap = NULL;
  *demands = NULL;
  *num_zones = 0;
  while (fgets(line, sizeof line, fp)) {
    if (blank_or_comment(line))
      continue;
    if (line[0] == '<') {
      if ((v = tag_value(line, "<NUMBER OF ZONES>"))) *num_zones = atol(v);
      else if ((v = tag_value(line, "<TOTAL OD FLOW>"))) total_flow = atof(v);
      else if (tag_value(line, "<END OF METADATA>")) {
        if (!*demands) {
          *demands = calloc((size_t)*num_zones + 1, sizeof(demand_data_t *));
          count = calloc((size_t)*num_zones + 1, sizeof(long));
          cap = calloc((size_t)*num_zones + 1, sizeof(long));
          if (!*demands || !count || !cap)
            return -1;
        }
      } else
        fprintf(stderr, "warning: unknown metadata tag: %s", line);
      continue;
    }
    if (!*demands) {
      fprintf(stderr, "error: trips data before <END OF METADATA>\n");
      return -1;
    }
    if (strncmp(line, "Origin", 6) == 0) {
      cur = atol(line + 6);
      if (cur < 1 || cur > *num_zones) {
        fprintf(stderr, "error: origin %ld outside 1..%ld\n", cur, *num_zones);
        return -1;
      }
      if ((*demands)[cur]) {
        fprintf(stderr, "error: multiple entries for Origin %ld\n", cur);
        return -1;
      }
      cap[cur] = 16;
      (*demands)[cur] = calloc(16, sizeof(demand_data_t));
      if (!(*demands)[cur])
        return -1;
      continue;
    }
    if (cur < 1) {
      fprintf(stderr, "error: demand before any Origin line\n");
      return -1;
    }
    /* items "dest : value;" */
    for (char *p = line;;) {
      char *colon = strchr(p, ':');
      if (!colon)
        break;
      char *end;
      long d = strtol(p, &end, 10);
      double val = strtod(colon + 1, &end);
      p = end;
      char *semi = strchr(p, ';');
      p = semi ? semi + 1 : p;
      if (d > *num_zones)
        fprintf(stderr, "warning: dest %ld > num_zones %ld\n", d, *num_zones);
      if (grow((void **)&(*demands)[cur], &cap[cur], count[cur] + 2, sizeof(demand_data_t)))
        return -1;
      (*demands)[cur][count[cur]].dest = d;
      (*demands)[cur][count[cur]].demand = val;
      count[cur]++;
      (*demands)[cur][count[cur]].dest = 0;           /* terminator */
      (*demands)[cur][count[cur]].demand = 0;
      total_read += val;
      if (!semi)
        break;
    }
  }
  if (!*demands) {
    fprintf(stderr, "error: no <END OF METADATA> in trips file\n");
    return -1;
  }
  for (long o = 1; o <= *num_zones; o++)
    if (!(*demands)[o]) {
      fprintf(stderr, "warning: no trip data for origin %ld\n", o);
      if (!((*demands)[o] = calloc(1, sizeof(demand_data_t))))
        return -1;
    }
  if (fabs(total_flow - total_read) > 1e-08)
    fprintf(stderr, "warning: total_od_flow %f != %f\n", total_flow, total_read);
  free(count);
  free(cap);
  return 0;
}

int parse_flows_file(FILE *flow_fp, link_flow_t **link_flows, long *num_links)
{
  char line[LINE_MAX_LEN];
  const char *v;
  long cap = 0, count = 0, meta_links = -1;
  *link_flows = NULL;
  while (fgets(line, sizeof line, flow_fp)) {
    if (blank_or_comment(line))
      continue;
    if (line[0] == '<') {
      if ((v = tag_value(line, "<NUMBER OF LINKS>"))) meta_links = atol(v);
      continue;
    }
    char *p = line, *end;
    long a = strtol(p, &end, 10);
    if (end == p) {
      fprintf(stderr, "bad line in flows file: %s", line);
      return -1;
    }
    p = end;
    long b = strtol(p, &end, 10);
    p = strchr(end, ':');
    if (!p) {
      fprintf(stderr, "bad line in flows file: %s", line);
      return -1;
    }
    double vol = strtod(p + 1, &end);
    double cost = strtod(end, &end);
    if (grow((void **)link_flows, &cap, count + 1, sizeof(link_flow_t)))
      return -1;
    (*link_flows)[count].init_node = a;
    (*link_flows)[count].term_node = b;
    (*link_flows)[count].volume = vol;
    (*link_flows)[count].cost = cost;
    count++;
  }
  if (meta_links >= 0 && count != meta_links) {
    fprintf(stderr, "error: num_links %ld but %ld links\n", meta_links, count);
    return -1;
  }
  *num_links = count;
  return 0;
}

static char *next_field(char **save)
{
  return strtok_r(NULL, "\t\n", save);
}

int parse_net_mod_file(FILE *fp, net_mod_t **mods, int *num_mods)
{
  char line[LINE_MAX_LEN];
  long cap = 0, count = 0;
  *mods = NULL;
  while (fgets(line, sizeof line, fp)) {
    if (blank_or_comment(line))
      continue;
    if (grow((void **)mods, &cap, count + 1, sizeof(net_mod_t)))
      return -1;
    net_mod_t *m = &(*mods)[count];
    char *save = NULL, *t = strtok_r(line, "\t\n", &save);
    char *f[13];
    f[0] = t;
    for (int i = 1; i < 13; i++)
      f[i] = next_field(&save);
    for (int i = 0; i < 13; i++)
      if (!f[i]) {
        fprintf(stderr, "error: short line in network modification file\n");
        return -1;
      }
    memset(m, 0, sizeof *m);
    strncpy(m->change_id, f[0], MAX_CHANGE_ID_LEN);
    if (strcmp(f[1], "ADD") == 0)
      m->modtype = MOD_TYPE_ADD_E;
    else if (strcmp(f[1], "CHANGE") == 0)
      m->modtype = MOD_TYPE_CHANGE_E;
    else {
      fprintf(stderr, "invalid change type '%s'\n", f[1]);
      return -1;
    }
    link_data_t *L = &m->mod_link;
    L->init_node = atol(f[2]); L->term_node = atol(f[3]);
    L->capacity = atof(f[4]); L->length = atof(f[5]); L->free_flow_time = atof(f[6]);
    L->B = atof(f[7]); L->power = atof(f[8]); L->speed_limit = atof(f[9]);
    L->toll = atof(f[10]); L->link_type = (short)atoi(f[11]);
    m->project_cost = atof(f[12]);
    count++;
  }
  *num_mods = (int)count;
  return 0;
}

int link_data_compar(const void *ent1, const void *ent2)
{
  const link_data_t *a = ent1, *b = ent2;
  if (a->init_node != b->init_node)
    return a->init_node < b->init_node ? -1 : 1;
  if (a->term_node != b->term_node)
    return a->term_node < b->term_node ? -1 : 1;
  return 0;
}

int parse_trip_mod_file(FILE *fp, trip_mod_t **mods, int *num_mods)
{
  char line[LINE_MAX_LEN];
  long cap = 0, count = 0;
  *mods = NULL;
  while (fgets(line, sizeof line, fp)) {
    if (blank_or_comment(line))
      continue;
    if (grow((void **)mods, &cap, count + 1, sizeof(trip_mod_t)))
      return -1;
    trip_mod_t *m = &(*mods)[count];
    char *save = NULL, *f[5];
    f[0] = strtok_r(line, "\t\n", &save);
    for (int i = 1; i < 5; i++)
      f[i] = next_field(&save);
    for (int i = 0; i < 5; i++)
      if (!f[i]) {
        fprintf(stderr, "error: short line in trip modification file\n");
        return -1;
      }
    memset(m, 0, sizeof *m);
    strncpy(m->trip_change_id, f[0], MAX_CHANGE_ID_LEN);
    if (strcmp(f[1], "P2P") == 0) m->trip_modtype = TRIP_MOD_TYPE_POINT_E;
    else if (strcmp(f[1], "UNIFORM") == 0) m->trip_modtype = TRIP_MOD_TYPE_UNIFORM_E;
    else if (strcmp(f[1], "ALL") == 0) m->trip_modtype = TRIP_MOD_TYPE_ALL_E;
    else {
      fprintf(stderr, "invalid trip data change type '%s'\n", f[1]);
      return -1;
    }
    m->origin = atol(f[2]);
    m->dest = atol(f[3]);
    m->multiplier = atof(f[4]);
    if (m->trip_modtype == TRIP_MOD_TYPE_POINT_E && (m->origin == 0 || m->dest == 0)) {
      fprintf(stderr, "ERROR: origin and dest must be nonzero for point-to-point\n");
      return -1;
    }
    if (m->trip_modtype != TRIP_MOD_TYPE_ALL_E && m->origin == 0 && m->dest == 0) {
      fprintf(stderr, "ERROR: cannot have 0 for both origin and dest except type ALL\n");
      return -1;
    }
    count++;
  }
  *num_mods = (int)count;
  return 0;
}

void print_flow(FILE *fp, const net_data_t *net, const double link_volumes[],
                const double link_costs[])
{
  fprintf(fp, "<NUMBER OF NODES>\t%ld\n", net->num_nodes);
  fprintf(fp, "<NUMBER OF LINKS>\t%ld\n", net->num_links);
  fprintf(fp, "<END OF METADATA>\n\n");
  fprintf(fp, "~\tTail\tHead\t:\tVolume\tCost\t;\n");
  for (long k = 0; k < net->num_links; k++)
    fprintf(fp, "\t%ld\t%ld\t:\t%.15f\t%.15f;\n", net->links[k].init_node,
            net->links[k].term_node, link_volumes[k], link_costs[k]);
}

/* multiply matching demands (point-to-point, one origin, one destination or
   all), trunk/tap/tap_functions.c:521-598 */
int apply_trip_modification(demand_data_t *demands[], trip_mod_t *mod, long num_zones)
{
  if (mod->origin < 0 || mod->dest < 0 || mod->origin > num_zones || mod->dest > num_zones) {
    fprintf(stderr, "ERROR: bad origin %ld or dest %ld in trip modification\n", mod->origin,
            mod->dest);
    return -1;
  }
  switch (mod->trip_modtype) {
  case TRIP_MOD_TYPE_POINT_E:
    fprintf(stderr, "[tripchange %s] %ld -> %ld multiplier %f\n", mod->trip_change_id,
            mod->origin, mod->dest, mod->multiplier);
    for (long i = 0; demands[mod->origin][i].dest != 0; i++)
      if (demands[mod->origin][i].dest == mod->dest) {
        demands[mod->origin][i].demand *= mod->multiplier;
        break;
      }
    return 0;
  case TRIP_MOD_TYPE_UNIFORM_E:
    if (mod->dest == 0 && mod->origin > 0) {
      fprintf(stderr, "[tripchange %s] origin %ld multiplier %f\n", mod->trip_change_id,
              mod->origin, mod->multiplier);
      for (long i = 0; demands[mod->origin][i].dest != 0; i++)
        demands[mod->origin][i].demand *= mod->multiplier;
      return 0;
    }
    if (mod->origin == 0 && mod->dest > 0) {
      fprintf(stderr, "[tripchange %s] dest %ld multiplier %f\n", mod->trip_change_id,
              mod->dest, mod->multiplier);
      for (long o = 1; o <= num_zones; o++)
        for (long i = 0; demands[o][i].dest != 0; i++)
          if (demands[o][i].dest == mod->dest)
            demands[o][i].demand *= mod->multiplier;
      return 0;
    }
    fprintf(stderr, "ERROR: both origin and dest nonzero in uniform trip mod\n");
    return -1;
  case TRIP_MOD_TYPE_ALL_E:
    fprintf(stderr, "[tripchange %s] ALL multiplier %f\n", mod->trip_change_id, mod->multiplier);
    for (long o = 1; o <= num_zones; o++)
      for (long i = 0; demands[o][i].dest != 0; i++)
        demands[o][i].demand *= mod->multiplier;
    return 0;
  default:
    fprintf(stderr, "bad trip modification type %d\n", mod->trip_modtype);
    return -1;
  }
}
