/*
 * tap_host.c -- the reference's solver entry points (include/tap_hip.h) on
 * the MI355X engine (include/tapgpu.h).
 *
 *   tap_frankwolfe()     trunk/tap/tap_frankwolfe_pthread.c:488-825
 *   fw_assign()          trunk/tap/tap_frankwolfe_cuda.cu:176-275
 *   pape_cuda()          trunk/tap/pape_cuda_host.cu:51-195
 *   link_volume_*()      trunk/tap/volume_update_host.cu:55-247
 *   sssp_pape*() etc.    trunk/tap/sssp_pape.h:36-64 (one origin per call)
 *
 * The engine context (graph, link records and demand resident in HBM) is
 * created once and cached; the shims that receive raw arrays every call
 * (fw_assign, pape_cuda, sssp_*) compare them with the arrays the context
 * was built from (one read pass) and only upload the new link costs.
 */
#include <errno.h>
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>
#include <unistd.h>

#include "tap_hip.h"
#include "tapgpu.h"

unsigned int num_cores = 1;
unsigned int num_threads = 1;
int using_fermi_architecture = 0;

static void die(const char *what, tapgpu_ctx *ctx, int rc)
{
  fprintf(stderr, "ERROR: %s failed (%d): %s\n", what, rc, tapgpu_last_error(ctx));
  exit(1);
}

/* pape_cuda's predecessor array (long) is handed to the engine as int64_t */
_Static_assert(sizeof(long) == sizeof(int64_t), "LP64 expected");

/* --- network + demand -> engine arrays --------------------------------------- */

typedef struct {
  int64_t *init, *term, *od_start, *od_dest;
  double *cap, *t0, *B, *pw, *od_val;
  tapgpu_network net;
  tapgpu_demand dem;
} engine_input;

static void free_input(engine_input *in)
{
  free(in->init); free(in->term); free(in->od_start); free(in->od_dest);
  free(in->cap); free(in->t0); free(in->B); free(in->pw); free(in->od_val);
  memset(in, 0, sizeof *in);
}

static int make_input(const net_data_t *net, demand_data_t *const *demands, engine_input *in)
{
  memset(in, 0, sizeof *in);
  const long E = net->num_links, Z = net->num_zones;
  in->init = malloc(sizeof(int64_t) * (E + 1));
  in->term = malloc(sizeof(int64_t) * (E + 1));
  in->cap = malloc(sizeof(double) * (E + 1));
  in->t0 = malloc(sizeof(double) * (E + 1));
  in->B = malloc(sizeof(double) * (E + 1));
  in->pw = malloc(sizeof(double) * (E + 1));
  in->od_start = calloc((size_t)Z + 1, sizeof(int64_t));
  if (!in->init || !in->term || !in->cap || !in->t0 || !in->B || !in->pw || !in->od_start)
    return -1;
  for (long k = 0; k < E; k++) {
    in->init[k] = net->links[k].init_node;
    in->term[k] = net->links[k].term_node;
    in->cap[k] = net->links[k].capacity;
    in->t0[k] = net->links[k].free_flow_time;
    in->B[k] = net->links[k].B;
    in->pw[k] = net->links[k].power;
  }
  long nnz = 0;
  for (long o = 1; o <= Z && demands; o++)
    for (long i = 0; demands[o] && demands[o][i].dest != 0; i++)
      nnz++;
  in->od_dest = malloc(sizeof(int64_t) * (nnz + 1));
  in->od_val = malloc(sizeof(double) * (nnz + 1));
  if (!in->od_dest || !in->od_val)
    return -1;
  nnz = 0;
  for (long o = 1; o <= Z; o++) {
    for (long i = 0; demands && demands[o] && demands[o][i].dest != 0; i++) {
      in->od_dest[nnz] = demands[o][i].dest;
      in->od_val[nnz] = demands[o][i].demand;
      nnz++;
    }
    in->od_start[o] = nnz;
  }
  in->net = (tapgpu_network){Z, net->num_nodes, net->first_thru_node, E, in->init, in->term,
                             in->cap, in->t0, in->B, in->pw};
  in->dem = (tapgpu_demand){Z, in->od_start, in->od_dest, in->od_val};
  return 0;
}

/* --- fw_assign ------------------------------------------------------------------ */

/* The context fw_assign reuses, with the arrays it was built from.  A call
   whose network and demand equal them (one read pass: no allocation, no
   rebuild) only uploads the new costs. */
static struct {
  tapgpu_ctx *ctx;
  engine_input in;
} g_assign;

static int same_input(const engine_input *in, const net_data_t *net, demand_data_t *const *demands)
{
  const long E = net->num_links, Z = net->num_zones;
  if (!in->init || in->net.num_links != E || in->net.num_zones != Z ||
      in->net.num_nodes != net->num_nodes || in->net.first_thru_node != net->first_thru_node)
    return 0;
  for (long k = 0; k < E; k++) {
    const link_data_t *l = &net->links[k];
    if (in->init[k] != l->init_node || in->term[k] != l->term_node ||
        memcmp(&in->cap[k], &l->capacity, sizeof(double)) ||
        memcmp(&in->t0[k], &l->free_flow_time, sizeof(double)) ||
        memcmp(&in->B[k], &l->B, sizeof(double)) || memcmp(&in->pw[k], &l->power, sizeof(double)))
      return 0;
  }
  const int64_t nnz = in->od_start[Z];
  int64_t j = 0;
  for (long o = 1; o <= Z; o++) {
    for (long i = 0; demands && demands[o] && demands[o][i].dest != 0; i++, j++) {
      if (j >= nnz || in->od_dest[j] != demands[o][i].dest ||
          memcmp(&in->od_val[j], &demands[o][i].demand, sizeof(double)))
        return 0;
    }
    if (in->od_start[o] != j) return 0;
  }
  return 1;
}

/* All-or-nothing assignment of demands at costs Wa into link_volumes
   (zeroed first).  Va/Ea must be the CSR of net->links (link index ==
   edge index); dist/predlink are work arrays of the reference and are not
   written. */
void fw_assign(net_data_t *net, demand_data_t *demands[], long Va[], long Ea[], double Wa[],
               double link_volumes[], double dist[], long predlink[])
{
  (void)Va; (void)Ea; (void)dist; (void)predlink;
  int rc;
  if (!g_assign.ctx || !same_input(&g_assign.in, net, demands)) {
    if (g_assign.ctx) tapgpu_destroy(g_assign.ctx);
    g_assign.ctx = NULL;
    free_input(&g_assign.in);
    if (make_input(net, demands, &g_assign.in)) {
      fprintf(stderr, "malloc failed in fw_assign\n");
      exit(1);
    }
    if ((rc = tapgpu_create(&g_assign.ctx, &g_assign.in.net, &g_assign.in.dem)))
      die("tapgpu_create", NULL, rc);
  }
  tapgpu_ctx *ctx = g_assign.ctx;
  if ((rc = tapgpu_set_costs(ctx, 0, Wa))) die("tapgpu_set_costs", ctx, rc);
  rc = tapgpu_aon(ctx, 0, link_volumes);
  if (rc == TAPGPU_E_NOPATH) {
    fprintf(stderr, "ERROR: %s\n", tapgpu_last_error(ctx));
    exit(1);
  }
  if (rc) die("tapgpu_aon", ctx, rc);
}

/* --- single-source shortest paths (sssp_pape.h) ------------------------------- */

/* One engine context for the single-source entry points, built from the CSR
   the caller passes (Va, Ea; topology only) and kept while the topology is
   the same; the costs are uploaded when they change.  For a zone origin the
   trees of every zone are computed in one launch and kept on the host, so
   that fw_assign's loop over the zones (trunk/tap/tap_frankwolfe.c:131-147)
   costs one launch per cost vector; other start nodes run on their own.
   The reference's routines are reentrant; these share the context under a
   lock. */
#define SSSP_PREFETCH_MAX (32L * 1024 * 1024)   /* nodes x zones kept on the host */

static pthread_mutex_t g_sssp_lock = PTHREAD_MUTEX_INITIALIZER;
static struct {
  tapgpu_ctx *ctx;
  long num_nodes, num_edges, first_thru;
  long *Va, *Ea;            /* topology the context was built from */
  double *Wa;               /* costs last uploaded                  */
  int costs_set;
  long zones;               /* zone trees held for these costs (0: none) */
  double *zd;               /* [num_nodes * zones] node-major        */
  int64_t *zp;
} g_sssp;

static void sssp_reset(void)
{
  if (g_sssp.ctx) tapgpu_destroy(g_sssp.ctx);
  free(g_sssp.Va); free(g_sssp.Ea); free(g_sssp.Wa); free(g_sssp.zd); free(g_sssp.zp);
  memset(&g_sssp, 0, sizeof g_sssp);
}

/* context for the CSR (Va, Ea) with num_nodes entries (node ids 1..N, row 0
   unused) */
static void sssp_context(long Va[], long Ea[], long num_nodes, long num_edges,
                         long first_thru_node)
{
  if (g_sssp.ctx && g_sssp.num_nodes == num_nodes && g_sssp.num_edges == num_edges &&
      g_sssp.first_thru == first_thru_node &&
      memcmp(g_sssp.Va, Va, sizeof(long) * (size_t)(num_nodes + 1)) == 0 &&
      memcmp(g_sssp.Ea, Ea, sizeof(long) * (size_t)(num_edges > 0 ? num_edges : 1)) == 0)
    return;
  sssp_reset();
  if (num_nodes < 2 || num_edges < 0 || Va[0] != 0 || Va[1] != 0 || Va[num_nodes] != num_edges) {
    fprintf(stderr, "ERROR: shortest paths: bad graph (num_nodes %ld, num_edges %ld; "
            "node ids are 1-based)\n", num_nodes, num_edges);
    exit(1);
  }
  engine_input in;
  memset(&in, 0, sizeof in);
  const long E1 = num_edges + 1;
  in.init = malloc(sizeof(int64_t) * E1);
  in.term = malloc(sizeof(int64_t) * E1);
  in.cap = malloc(sizeof(double) * E1);
  in.t0 = calloc((size_t)E1, sizeof(double));
  in.B = calloc((size_t)E1, sizeof(double));
  in.pw = malloc(sizeof(double) * E1);
  in.od_start = calloc(1, sizeof(int64_t));
  in.od_dest = malloc(sizeof(int64_t));
  in.od_val = malloc(sizeof(double));
  g_sssp.Va = malloc(sizeof(long) * (size_t)(num_nodes + 1));
  g_sssp.Ea = malloc(sizeof(long) * (size_t)E1);
  g_sssp.Wa = malloc(sizeof(double) * (size_t)E1);
  if (!in.init || !in.term || !in.cap || !in.t0 || !in.B || !in.pw || !in.od_start ||
      !in.od_dest || !in.od_val || !g_sssp.Va || !g_sssp.Ea || !g_sssp.Wa) {
    fprintf(stderr, "malloc failed in shortest paths\n");
    exit(1);
  }
  for (long v = 0; v < num_nodes; v++)
    for (long e = Va[v]; e < Va[v + 1]; e++) {
      in.init[e] = v;
      in.term[e] = Ea[e];
    }
  for (long e = 0; e < num_edges; e++) {
    in.cap[e] = 1.0;
    in.pw[e] = 1.0;
  }
  in.net = (tapgpu_network){0, num_nodes - 1, first_thru_node, num_edges, in.init, in.term,
                            in.cap, in.t0, in.B, in.pw};
  in.dem = (tapgpu_demand){0, in.od_start, in.od_dest, in.od_val};
  int rc = tapgpu_create(&g_sssp.ctx, &in.net, &in.dem);
  free_input(&in);
  if (rc) die("tapgpu_create", NULL, rc);
  memcpy(g_sssp.Va, Va, sizeof(long) * (size_t)(num_nodes + 1));
  memcpy(g_sssp.Ea, Ea, sizeof(long) * (size_t)(num_edges > 0 ? num_edges : 1));
  g_sssp.num_nodes = num_nodes;
  g_sssp.num_edges = num_edges;
  g_sssp.first_thru = first_thru_node;
}

/* labels and predecessor links from start_node into dist[] / prev[]
   (num_nodes entries each; node 0 unreachable) */
static void sssp_run(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
                     long start_node, long first_thru_node, long prev[], double dist[])
{
  if (start_node < 1 || start_node >= num_nodes) {
    fprintf(stderr, "ERROR: shortest paths: start node %ld out of range\n", start_node);
    exit(1);
  }
  pthread_mutex_lock(&g_sssp_lock);
  sssp_context(Va, Ea, num_nodes, num_edges, first_thru_node);
  tapgpu_ctx *ctx = g_sssp.ctx;
  int rc;
  if (!g_sssp.costs_set || memcmp(g_sssp.Wa, Wa, sizeof(double) * (size_t)num_edges) != 0) {
    if ((rc = tapgpu_set_costs(ctx, 0, Wa))) die("tapgpu_set_costs", ctx, rc);
    memcpy(g_sssp.Wa, Wa, sizeof(double) * (size_t)num_edges);
    g_sssp.costs_set = 1;
    g_sssp.zones = 0;
  }
  const long Z = first_thru_node - 1;
  if (start_node <= Z && (long)num_nodes * Z <= SSSP_PREFETCH_MAX) {
    if (g_sssp.zones != Z) {
      free(g_sssp.zd); free(g_sssp.zp);
      g_sssp.zd = malloc(sizeof(double) * (size_t)num_nodes * Z);
      g_sssp.zp = malloc(sizeof(int64_t) * (size_t)num_nodes * Z);
      int64_t *starts = malloc(sizeof(int64_t) * Z);
      if (!g_sssp.zd || !g_sssp.zp || !starts) {
        fprintf(stderr, "malloc failed in shortest paths\n");
        exit(1);
      }
      for (long i = 0; i < Z; i++) starts[i] = i + 1;
      if ((rc = tapgpu_sssp(ctx, 0, starts, Z, g_sssp.zd, g_sssp.zp, Z, 0)))
        die("tapgpu_sssp", ctx, rc);
      free(starts);
      g_sssp.zones = Z;
    }
    for (long v = 0; v < num_nodes; v++) {
      dist[v] = g_sssp.zd[(size_t)v * Z + start_node - 1];
      prev[v] = (long)g_sssp.zp[(size_t)v * Z + start_node - 1];
    }
  } else {
    int64_t s = start_node;
    double *d = malloc(sizeof(double) * (size_t)num_nodes);
    int64_t *p = malloc(sizeof(int64_t) * (size_t)num_nodes);
    if (!d || !p) {
      fprintf(stderr, "malloc failed in shortest paths\n");
      exit(1);
    }
    if ((rc = tapgpu_sssp(ctx, 0, &s, 1, d, p, 1, 0))) die("tapgpu_sssp", ctx, rc);
    for (long v = 0; v < num_nodes; v++) {
      dist[v] = d[v];
      prev[v] = (long)p[v];
    }
    free(d); free(p);
  }
  pthread_mutex_unlock(&g_sssp_lock);
}

/* --- pape_cuda ---------------------------------------------------------------- */

/* All-origin shortest paths over the CSR (Va, Ea, Wa) from start_nodes; the
   reference's output layout: entry [v * num_start_nodes + s] for start node
   s and node v (0 <= v < num_nodes).  The context is the shortest-path
   family's (topology compared, costs uploaded when they change); with start
   nodes 1..num_start_nodes (fw_assign's) the engine writes that layout
   directly. */
void pape_cuda(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
               long start_nodes[], long num_start_nodes, long first_thru_node,
               double *distances, long *predecessors)
{
  pthread_mutex_lock(&g_sssp_lock);
  sssp_context(Va, Ea, num_nodes, num_edges, first_thru_node);
  tapgpu_ctx *ctx = g_sssp.ctx;
  int rc;
  if (!g_sssp.costs_set || memcmp(g_sssp.Wa, Wa, sizeof(double) * (size_t)num_edges) != 0) {
    if ((rc = tapgpu_set_costs(ctx, 0, Wa))) die("tapgpu_set_costs", ctx, rc);
    memcpy(g_sssp.Wa, Wa, sizeof(double) * (size_t)num_edges);
    g_sssp.costs_set = 1;
    g_sssp.zones = 0;
  }
  int64_t *starts = malloc(sizeof(int64_t) * (num_start_nodes + 1));
  if (!starts) {
    fprintf(stderr, "malloc failed in pape_cuda\n");
    exit(1);
  }
  int dense = 1;
  for (long i = 0; i < num_start_nodes; i++) {
    starts[i] = start_nodes[i];
    dense = dense && start_nodes[i] == i + 1;
  }
  if (dense) {
    /* column = start node id: col_base 1, row stride num_start_nodes */
    rc = tapgpu_sssp(ctx, 0, starts, num_start_nodes, distances, (int64_t *)predecessors,
                     num_start_nodes, 1);
    if (rc) die("tapgpu_sssp", ctx, rc);
  } else {
    double *d = malloc(sizeof(double) * (size_t)num_nodes * (num_start_nodes + 1));
    int64_t *p = malloc(sizeof(int64_t) * (size_t)num_nodes * (num_start_nodes + 1));
    if (!d || !p) {
      fprintf(stderr, "malloc failed in pape_cuda\n");
      exit(1);
    }
    rc = tapgpu_sssp(ctx, 0, starts, num_start_nodes, d, p, num_start_nodes, 0);
    if (rc) die("tapgpu_sssp", ctx, rc);
    for (long v = 0; v < num_nodes; v++)
      for (long i = 0; i < num_start_nodes; i++) {
        const size_t o = (size_t)v * num_start_nodes + start_nodes[i];
        distances[o] = d[(size_t)v * num_start_nodes + i];
        predecessors[o] = (long)p[(size_t)v * num_start_nodes + i];
      }
    free(d); free(p);
  }
  free(starts);
  pthread_mutex_unlock(&g_sssp_lock);
}

/* sssp_pape, sssp_pape_lll, sssp_slf, sssp_slf_lll (trunk/tap/sssp_pape.h:36-54):
   one labelling for all four.  The labels are the least fixed point of
   fl(dist[u] + Wa[e]) that every one of the reference's deque policies
   reaches (sssp_slf's float temporaries, trunk/tap/sssp_pape.c:543, are not
   reproduced: its labels here are the FP64 ones).  prev = link index (-1:
   none), dist = DBL_MAX where unreachable; queue_next is the reference's
   work array and is not used. */
void sssp_pape(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
               long start_node, long first_thru_node, long prev[], double dist[],
               long queue_next[])
{
  (void)queue_next;
  sssp_run(Va, Ea, Wa, num_nodes, num_edges, start_node, first_thru_node, prev, dist);
}

void sssp_pape_lll(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
                   long start_node, long first_thru_node, long prev[], double dist[],
                   long queue_next[])
{
  (void)queue_next;
  sssp_run(Va, Ea, Wa, num_nodes, num_edges, start_node, first_thru_node, prev, dist);
}

void sssp_slf(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
              long start_node, long first_thru_node, long prev[], double dist[],
              long queue_next[])
{
  (void)queue_next;
  sssp_run(Va, Ea, Wa, num_nodes, num_edges, start_node, first_thru_node, prev, dist);
}

void sssp_slf_lll(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
                  long start_node, long first_thru_node, long prev[], double dist[],
                  long queue_next[])
{
  (void)queue_next;
  sssp_run(Va, Ea, Wa, num_nodes, num_edges, start_node, first_thru_node, prev, dist);
}

/* sssp_prevnodefirst (trunk/tap/sssp_pape.h:60-65): the same labels and
   links, plus the tail node of each predecessor link (-1: none).
   old_prevnode only steers the reference's queue order, not the result. */
void sssp_prevnodefirst(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
                        long start_node, long first_thru_node, long old_prevnode[],
                        long prevlink[], long prevnode[], double dist[], long queue_next[])
{
  (void)old_prevnode; (void)queue_next;
  sssp_run(Va, Ea, Wa, num_nodes, num_edges, start_node, first_thru_node, prevlink, dist);
  for (long v = 0; v < num_nodes; v++) {
    prevnode[v] = -1;
    const long e = prevlink[v];
    if (e < 0) continue;
    /* the tail of link e: the node whose out-link range holds it */
    long lo = 0, hi = num_nodes - 1;
    while (lo < hi) {
      const long mid = (lo + hi + 1) / 2;
      if (Va[mid] <= e) lo = mid; else hi = mid - 1;
    }
    prevnode[v] = lo;
  }
}

/* get_shortest_path (trunk/tap/sssp_pape.h:56-57, sssp_pape.c:237-253): the
   nodes of the path from dest back to orig along predecessor NODES pred[]
   (dest first, orig last); pathlen 0 if dest has no predecessor. */
void get_shortest_path(long pred[], long orig, long dest, long path[], long *pathlen)
{
  long v = dest;
  *pathlen = 0;
  if (pred[v] == -1) return;
  while (v != orig) {
    if (v == -1) {
      fprintf(stderr, "ERROR: get_shortest_path: no path from %ld to %ld\n", orig, dest);
      exit(1);
    }
    path[(*pathlen)++] = v;
    v = pred[v];
  }
  path[(*pathlen)++] = orig;
}

/* --- link_volume_* ---------------------------------------------------------------- */

static tapgpu_ctx *g_volume_ctx;
static long g_volume_links;

void link_volume_data_setup(link_data_t *links, long num_links, long num_start_nodes,
                            demand_data_t *demands[])
{
  long maxnode = 0;
  for (long k = 0; k < num_links; k++) {
    if (links[k].init_node > maxnode) maxnode = links[k].init_node;
    if (links[k].term_node > maxnode) maxnode = links[k].term_node;
  }
  net_data_t net = {num_start_nodes, maxnode, 1, num_links, links};
  engine_input in;
  if (make_input(&net, demands, &in)) {
    fprintf(stderr, "malloc failed in link_volume_data_setup\n");
    exit(1);
  }
  link_volume_data_cleanup();
  int rc = tapgpu_create(&g_volume_ctx, &in.net, &in.dem);
  if (rc) die("tapgpu_create", NULL, rc);
  g_volume_links = num_links;
  free_input(&in);
}

/* link_volumes := all-or-nothing volumes along the predecessor links
   predlink[v * num_start_nodes + s] of origins s = 1..num_start_nodes */
void link_volume_update(long num_start_nodes, long num_edges, long num_nodes,
                        double link_volumes[], long predlink[])
{
  (void)num_nodes;
  if (!g_volume_ctx || num_edges != g_volume_links) {
    fprintf(stderr, "ERROR: link_volume_update without link_volume_data_setup\n");
    exit(1);
  }
  int rc = tapgpu_load_pred(g_volume_ctx, 0, (const int64_t *)predlink, num_start_nodes, 1,
                            link_volumes);
  if (rc) die("tapgpu_load_pred", g_volume_ctx, rc);
}

void link_volume_data_cleanup(void)
{
  if (g_volume_ctx)
    tapgpu_destroy(g_volume_ctx);
  g_volume_ctx = NULL;
}

/* --- tap_frankwolfe --------------------------------------------------------------- */

static volatile sig_atomic_t g_signal;   /* 0 none, 1 status, 2 checkpoint, 3 stop */

static void on_signal(int sig)
{
  if (sig == SIGUSR1) g_signal = 1;
  else if (sig == SIGUSR2) g_signal = 2;
  else g_signal = 3;
}

static int linkflow_compar(const void *a, const void *b)
{
  const link_flow_t *x = a, *y = b;
  if (x->init_node != y->init_node) return x->init_node < y->init_node ? -1 : 1;
  if (x->term_node != y->term_node) return x->term_node < y->term_node ? -1 : 1;
  return 0;
}

static void write_flows(FILE *fp, const net_data_t *net, tapgpu_ctx *ctx, int iter,
                        double obj, double relgap, double *x, double *c, const char *first)
{
  tapgpu_get_flows(ctx, 0, x, c);
  double vht = 0;
  for (long k = 0; k < net->num_links; k++)
    vht += x[k] * c[k];
  rewind(fp);
  if (first)
    fputs(first, fp);
  fprintf(fp, "~ iterations = %d objvalue = %.15f relgap = %.15f\n", iter, obj, relgap);
  fprintf(fp, "~ totalcost = %.15f\n", vht);
  print_flow(fp, net, x, c);
  fflush(fp);
  /* a rewritten table may be shorter than a checkpoint written before it;
     unlike the reference, cut what is left of the older one (regular files) */
  const long end = ftell(fp);
  if (end >= 0 && ftruncate(fileno(fp), end) != 0)
    errno = 0;
}

/* Frank-Wolfe to target_iterations / target_relgap, writing the flow table
   to flow_fp; frees demands as the reference does.  Returns 0 or 1. */
int tap_frankwolfe(net_data_t *net, demand_data_t **demands, int warm_start_mode,
                   FILE *flows_input_fp, int target_iterations, double target_relgap,
                   FILE *flow_fp, double *total_cost_at_end, double *objvalue_at_end,
                   int *iterations_at_end, double *relgap_at_end, int time_each_iteration)
{
  const long E = net->num_links;
  double *warm = NULL, *warm_cost = NULL;
  if (warm_start_mode) {
    link_flow_t *lf = NULL;
    long nf = 0;
    if (!flows_input_fp || parse_flows_file(flows_input_fp, &lf, &nf) != 0) {
      fprintf(stderr, "error parsing flows input file\n");
      return 1;
    }
    fclose(flows_input_fp);
    qsort(lf, nf, sizeof(link_flow_t), linkflow_compar);
    if (nf > E) {
      fprintf(stderr, "error: %ld links in net file but %ld in flows input file\n", E, nf);
      free(lf);
      return 1;
    }
    /* links added since the flows were written start at zero volume and
       cost (trunk/tap/tap_frankwolfe_pthread.c:564-590) */
    warm = calloc((size_t)E + 1, sizeof(double));
    warm_cost = calloc((size_t)E + 1, sizeof(double));
    if (!warm || !warm_cost) {
      fprintf(stderr, "malloc failed in tap_frankwolfe\n");
      free(lf); free(warm); free(warm_cost);
      return 1;
    }
    long j = 0;
    for (long k = 0; k < E; k++)
      if (j < nf && lf[j].init_node == net->links[k].init_node &&
          lf[j].term_node == net->links[k].term_node) {
        warm_cost[k] = lf[j].cost;
        warm[k] = lf[j++].volume;
      }
    if (j != nf) {
      fprintf(stderr, "error: flows input file does not match the network links\n");
      free(lf); free(warm); free(warm_cost);
      return 1;
    }
    free(lf);
  }

  engine_input in;
  if (make_input(net, demands, &in)) {
    fprintf(stderr, "malloc failed in tap_frankwolfe\n");
    free(warm); free(warm_cost);
    return 1;
  }
  tapgpu_ctx *ctx = NULL;
  int rc = tapgpu_create(&ctx, &in.net, &in.dem);
  if (rc) {
    fprintf(stderr, "ERROR: %s\n", tapgpu_last_error(NULL));
    free_input(&in); free(warm); free(warm_cost);
    return 1;
  }
  double *x = malloc(sizeof(double) * (E + 1)), *c = malloc(sizeof(double) * (E + 1));
  struct sigaction sa, old[7];
  const int sigs[7] = {SIGUSR1, SIGUSR2, SIGHUP, SIGINT, SIGTERM, SIGXCPU, SIGPWR};
  memset(&sa, 0, sizeof sa);
  sa.sa_handler = on_signal;
  sigemptyset(&sa.sa_mask);
  g_signal = 0;
  for (int i = 0; i < 7; i++)
    sigaction(sigs[i], &sa, &old[i]);

  int iter = 0, ret = 0;
  double obj = 0, relgap = DBL_MAX;
  rc = tapgpu_fw_init(ctx, warm);
  if (rc) {
    fprintf(stderr, "ERROR: %s\n", tapgpu_last_error(ctx));
    ret = 1;
    goto done;
  }
  if (warm) {
    /* the file's costs should be BPR(volume): warn where they differ by
       more than EPS (trunk/tap/tap_frankwolfe_pthread.c:682-690) */
    if ((rc = tapgpu_get_flows(ctx, 0, x, c))) {
      fprintf(stderr, "ERROR: %s\n", tapgpu_last_error(ctx));
      ret = 1;
      goto done;
    }
    for (long k = 0; k < E; k++)
      if (fabs(c[k] - warm_cost[k]) > 1e-08)
        fprintf(stderr, "warning: mismatched costs (delta = %g) on link %d from %ld to %ld\n",
                fabs(c[k] - warm_cost[k]), (int)k, net->links[k].init_node,
                net->links[k].term_node);
  }
  /* the device applies the loop test of tap_frankwolfe_pthread.c:712-714
     every iteration; the host looks at signals between chunks of iterations
     (every iteration when the per-iteration report is on) */
  const int chunk = time_each_iteration ? 1 : 16;
  for (;;) {
    if (!((target_iterations <= 0 || iter < target_iterations) &&
          (target_relgap <= 0 || relgap > target_relgap)))
      break;
    struct timeval t0, t1;
    gettimeofday(&t0, NULL);
    tapgpu_result r;
    int upto = iter + chunk;
    if (target_iterations > 0 && upto > target_iterations)
      upto = target_iterations;
    rc = tapgpu_fw_run(ctx, upto, target_relgap, &r);
    if (rc) {
      fprintf(stderr, "ERROR: %s\n", tapgpu_last_error(ctx));
      ret = 1;
      break;
    }
    gettimeofday(&t1, NULL);
    const int ms = (int)((t1.tv_sec - t0.tv_sec) * 1000 + (t1.tv_usec - t0.tv_usec) / 1000);
    tapgpu_iter t;
    memset(&t, 0, sizeof t);
    for (int i = iter; i < r.iterations; i++) {
      if (tapgpu_get_iter(ctx, 0, i, &t))
        break;
      if (time_each_iteration)
        fprintf(stderr, "iter = %d objective value = %.15f relative gap = %.15f average "
                "exccess cost = %.15f total cost = %.15f (%d ms)\n", i + 1, t.objective,
                t.relgap, t.avg_excess, t.vht, ms);
    }
    const int progressed = r.iterations > iter;
    iter = r.iterations;
    if (progressed) {
      obj = t.objective;
      relgap = t.relgap;
    }
    if (!time_each_iteration && (g_signal == 1 || g_signal == 2))
      fprintf(stderr, "iter = %d objective value = %.15f relative gap = %.15f average exccess "
              "cost = %.15f total cost = %.15f\n", iter, obj, relgap, t.avg_excess, t.vht);
    if (!progressed)
      break;
    if (g_signal == 2) {
      char buf[128];
      time_t now = time(NULL);
      snprintf(buf, sizeof buf, "~ checkpoint written at %s\n", ctime(&now));
      fprintf(stderr, "checkpointing to flows output file...");
      write_flows(flow_fp, net, ctx, iter, obj, relgap, x, c, buf);
      fprintf(stderr, "done\n");
    }
    const int stop = g_signal == 3;
    g_signal = 0;
    if (stop)
      break;
  }
  if (!ret) {
    write_flows(flow_fp, net, ctx, iter, obj, relgap, x, c, NULL);
    double vht = 0;
    for (long k = 0; k < E; k++)
      vht += x[k] * c[k];
    *total_cost_at_end = vht;
    *objvalue_at_end = obj;
    *iterations_at_end = iter;
    *relgap_at_end = relgap;
  }
done:
  for (int i = 0; i < 7; i++)
    sigaction(sigs[i], &old[i], NULL);
  tapgpu_destroy(ctx);
  free_input(&in);
  free(warm); free(warm_cost); free(x); free(c);
  if (demands) {
    for (long o = 0; o < net->num_zones + 1; o++)
      free(demands[o]);
    free(demands);
  }
  return ret;
}


/* tap_greedy's main loop and output (trunk/tap/tap_greedy.c:296-335) on the
   engine's greedy assignment (tapgpu_greedy). */
int tap_greedy_assign(net_data_t *net, demand_data_t **demands, FILE *flow_fp,
                      double *total_cost_at_end, int *iterations_at_end)
{
  const long E = net->num_links;
  engine_input in;
  if (make_input(net, demands, &in)) {
    fprintf(stderr, "malloc failed in tap_greedy_assign\n");
    return 1;
  }
  tapgpu_ctx *ctx = NULL;
  int rc = tapgpu_create(&ctx, &in.net, &in.dem);
  if (rc) {
    fprintf(stderr, "ERROR: %s\n", tapgpu_last_error(NULL));
    free_input(&in);
    return 1;
  }
  double initial = 0;
  for (long o = 1; o < net->num_zones + 1; o++)
    for (long i = 0; demands[o][i].dest != 0; i++)
      if (demands[o][i].dest != o)
        initial += demands[o][i].demand;
  fprintf(stderr, "initial unsatifised demand = %f\n", initial);
  const int cap = 100000;
  double *x = malloc(sizeof(double) * (E + 1)), *c = malloc(sizeof(double) * (E + 1));
  double *un = malloc(sizeof(double) * cap), *ms = malloc(sizeof(double) * cap);
  tapgpu_greedy_result res;
  int ret = 0;
  rc = tapgpu_greedy(ctx, 0, cap, x, c, un, ms, cap, &res);
  if (rc) {
    fprintf(stderr, "ERROR: %s\n", tapgpu_last_error(ctx));
    ret = 1;
  } else {
    for (int i = 0; i < res.iterations && i < cap; i++)
      fprintf(stderr, "iter = %d unsatisfied demand = %f (time %d ms)\n", i + 1, un[i],
              (int)ms[i]);
    if (res.stalled)
      fprintf(stderr, "WARNING: unsatisfied demand unchanged, terminating with unsatisfied "
                      "demand = %f\n", res.unsatisfied);
    fprintf(stderr, "total cost = %f\n", res.total_cost);
    print_flow(flow_fp, net, x, c);
    *total_cost_at_end = res.total_cost;
    *iterations_at_end = res.iterations;
  }
  tapgpu_destroy(ctx);
  free_input(&in);
  free(x); free(c); free(un); free(ms);
  return ret;
}
