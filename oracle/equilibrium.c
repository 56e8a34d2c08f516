/*
 * equilibrium.c -- path-based user-equilibrium solver for the gradient pins.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference has no gradient
 * code; the derivatives of the equilibrium are pinned by central finite
 * differences of re-solved equilibria and by a path-space reference gradient
 * (oracle/gradient.py).  Frank-Wolfe converges sublinearly, far too slowly to
 * give differences of re-solved equilibria to 1e-4, so the equilibria here
 * come from path-based gradient projection (Jayakrishnan et al., "A faster
 * path-based algorithm for traffic assignment", TRR 1443, 1994): per OD pair
 * a set of paths with flows; each sweep adds the current shortest path
 * (column generation, d'Esopo-Pape trees from oracle.c) and moves flow from
 * every other path p to the shortest path s by the Newton step
 *     delta = min(h_p, (c_p - c_s) / sum_{a in p xor s} t'_a).
 * It converges linearly, to the rounding level of the path costs.  Costs are
 * the reference's BPR in the oracle's arithmetic (ora_bpr).  Link k of the
 * network is CSR edge k (links sorted by init node).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
  int64_t origin, dest;
  double demand;
  int64_t first, n;     /* its paths: ids pid[first .. first+n) in od_paths */
  int64_t cap;
} od_t;

typedef struct {
  ora_net net;
  double *t0, *cap, *B, *pw;        /* own copies (parameters can change) */
  int64_t *init, *term;
  int64_t N1, E, Z, ft;
  int64_t *Va, *Ea;
  double *x, *t, *dt;               /* volumes, costs, d cost / d volume */
  /* OD pairs with positive demand */
  od_t *od;
  int64_t nod;
  int64_t *od_start;                /* [Z+1] into od */
  /* path pool */
  int64_t *p_off, *p_len;           /* into links */
  double *p_flow;
  int64_t np, np_cap;
  int64_t *links;
  int64_t nl, nl_cap;
  int64_t *od_paths;                /* per-OD slots (cap each) */
  int64_t od_paths_len, od_paths_cap;
  /* scratch */
  int64_t *prev, *qn, *mark, *walk;
  double *dist;
  int64_t stamp;
} gp_t;

static void *xrealloc(void *p, size_t n)
{
  void *q = realloc(p, n ? n : 1);
  if (!q) abort();
  return q;
}

static double link_cost(const gp_t *g, int64_t k)
{
  return ora_bpr(g->t0[k], g->cap[k], g->B[k], g->pw[k], g->x[k]);
}

/* d t_k / d x_k of t0 (1 + B (x/c)^p) */
static double link_deriv(const gp_t *g, int64_t k)
{
  const double p = g->pw[k], x = g->x[k], c = g->cap[k];
  if (g->B[k] == 0 || p == 0 || x <= 0) return (p == 1 && g->B[k] != 0) ? g->t0[k] * g->B[k] / c : 0.0;
  return g->t0[k] * g->B[k] * p * pow(x / c, p - 1) / c;
}

static void refresh(gp_t *g, int64_t k)
{
  g->t[k] = link_cost(g, k);
  g->dt[k] = link_deriv(g, k);
}

static int64_t add_path(gp_t *g, const int64_t *lk, int64_t n, double flow)
{
  if (g->np == g->np_cap) {
    g->np_cap = g->np_cap ? 2 * g->np_cap : 1024;
    g->p_off = xrealloc(g->p_off, sizeof(int64_t) * g->np_cap);
    g->p_len = xrealloc(g->p_len, sizeof(int64_t) * g->np_cap);
    g->p_flow = xrealloc(g->p_flow, sizeof(double) * g->np_cap);
  }
  if (g->nl + n > g->nl_cap) {
    while (g->nl + n > g->nl_cap) g->nl_cap = g->nl_cap ? 2 * g->nl_cap : 4096;
    g->links = xrealloc(g->links, sizeof(int64_t) * g->nl_cap);
  }
  memcpy(g->links + g->nl, lk, sizeof(int64_t) * n);
  g->p_off[g->np] = g->nl;
  g->p_len[g->np] = n;
  g->p_flow[g->np] = flow;
  g->nl += n;
  return g->np++;
}

static void od_push(gp_t *g, int64_t k, int64_t pid)
{
  od_t *o = &g->od[k];
  if (o->n == o->cap) {                      /* move to a bigger slot run */
    const int64_t ncap = o->cap * 2;
    if (g->od_paths_len + ncap > g->od_paths_cap) {
      while (g->od_paths_len + ncap > g->od_paths_cap) g->od_paths_cap *= 2;
      g->od_paths = xrealloc(g->od_paths, sizeof(int64_t) * g->od_paths_cap);
    }
    memcpy(g->od_paths + g->od_paths_len, g->od_paths + o->first, sizeof(int64_t) * o->n);
    o->first = g->od_paths_len;
    o->cap = ncap;
    g->od_paths_len += ncap;
  }
  g->od_paths[o->first + o->n++] = pid;
}

/* shortest path to dest from the tree in g->prev, as links origin..dest in
   g->walk; returns its length or -1 without a path */
static int64_t tree_path(gp_t *g, int64_t origin, int64_t dest)
{
  int64_t n = 0, v = dest;
  while (v != origin) {
    const int64_t e = g->prev[v];
    if (e < 0 || n > g->N1) return -1;
    g->walk[n++] = e;
    v = g->init[e];
  }
  for (int64_t i = 0; i < n / 2; i++) {
    const int64_t tmp = g->walk[i];
    g->walk[i] = g->walk[n - 1 - i];
    g->walk[n - 1 - i] = tmp;
  }
  return n;
}

static void sssp(gp_t *g, int64_t origin)
{
  ora_sssp_pape_lll(g->Va, g->Ea, g->t, g->N1, origin, g->ft, g->prev, g->dist, g->qn);
}

static double path_cost(const gp_t *g, int64_t pid)
{
  double s = 0;
  const int64_t *lk = g->links + g->p_off[pid];
  for (int64_t i = 0; i < g->p_len[pid]; i++) s += g->t[lk[i]];
  return s;
}

/* One gradient-projection update of OD k towards its cheapest path. */
static void shift_od(gp_t *g, int64_t k)
{
  od_t *o = &g->od[k];
  if (o->n < 2) return;
  int64_t *ids = g->od_paths + o->first;
  for (int rep = 0; rep < 2; rep++) {
    int64_t s = 0;
    double cs = path_cost(g, ids[0]);
    for (int64_t i = 1; i < o->n; i++) {
      const double c = path_cost(g, ids[i]);
      if (c < cs) { cs = c; s = i; }
    }
    const int64_t ps = ids[s];
    for (int64_t i = 0; i < o->n; i++) {
      if (i == s) continue;
      const int64_t pp = ids[i];
      if (g->p_flow[pp] <= 0) continue;
      const double cp = path_cost(g, pp);
      cs = path_cost(g, ps);
      if (!(cp > cs)) continue;
      /* second derivative along p - s: t' over the symmetric difference */
      const int64_t st = ++g->stamp;
      const int64_t *ls = g->links + g->p_off[ps], *lp = g->links + g->p_off[pp];
      for (int64_t j = 0; j < g->p_len[ps]; j++) g->mark[ls[j]] = st;
      double d2 = 0;
      for (int64_t j = 0; j < g->p_len[pp]; j++)
        if (g->mark[lp[j]] != st) d2 += g->dt[lp[j]];
      const int64_t st2 = ++g->stamp;
      for (int64_t j = 0; j < g->p_len[pp]; j++) g->mark[lp[j]] = st2;
      for (int64_t j = 0; j < g->p_len[ps]; j++)
        if (g->mark[ls[j]] != st2) d2 += g->dt[ls[j]];
      double delta = g->p_flow[pp];
      if (d2 > 0) {
        const double step = (cp - cs) / d2;
        if (step < delta) delta = step;
      }
      if (!(delta > 0)) continue;
      g->p_flow[pp] -= delta;
      g->p_flow[ps] += delta;
      if (g->p_flow[pp] < 1e-300) g->p_flow[pp] = 0;
      for (int64_t j = 0; j < g->p_len[pp]; j++) { g->x[lp[j]] -= delta; }
      for (int64_t j = 0; j < g->p_len[ps]; j++) { g->x[ls[j]] += delta; }
      for (int64_t j = 0; j < g->p_len[pp]; j++) {
        if (g->x[lp[j]] < 0) g->x[lp[j]] = 0;
        refresh(g, lp[j]);
      }
      for (int64_t j = 0; j < g->p_len[ps]; j++) refresh(g, ls[j]);
    }
  }
  /* drop unused paths other than the cheapest */
  int64_t best = 0;
  double cb = path_cost(g, ids[0]);
  for (int64_t i = 1; i < o->n; i++) {
    const double c = path_cost(g, ids[i]);
    if (c < cb) { cb = c; best = i; }
  }
  int64_t m = 0;
  for (int64_t i = 0; i < o->n; i++)
    if (i == best || g->p_flow[ids[i]] > 0) ids[m++] = ids[i];
  o->n = m;
}

static int same_path(const gp_t *g, int64_t pid, const int64_t *lk, int64_t n)
{
  return g->p_len[pid] == n && memcmp(g->links + g->p_off[pid], lk, sizeof(int64_t) * n) == 0;
}

void *ora_gp_create(const ora_net *net, const ora_demand *dem)
{
  gp_t *g = calloc(1, sizeof *g);
  const int64_t E = net->num_links;
  g->E = E;
  g->N1 = net->num_nodes + 1;
  g->Z = net->num_zones;
  g->ft = net->first_thru_node;
  g->t0 = malloc(sizeof(double) * (E + 1));
  g->cap = malloc(sizeof(double) * (E + 1));
  g->B = malloc(sizeof(double) * (E + 1));
  g->pw = malloc(sizeof(double) * (E + 1));
  g->init = malloc(sizeof(int64_t) * (E + 1));
  g->term = malloc(sizeof(int64_t) * (E + 1));
  double minpos = DBL_MAX;
  for (int64_t k = 0; k < E; k++)
    if (net->capacity[k] > 0 && net->capacity[k] < minpos) minpos = net->capacity[k];
  for (int64_t k = 0; k < E; k++) {
    g->t0[k] = net->fftime[k];
    g->cap[k] = net->capacity[k] > 0 ? net->capacity[k] : minpos;
    g->B[k] = net->B[k];
    g->pw[k] = net->power[k];
    g->init[k] = net->init[k];
    g->term[k] = net->term[k];
  }
  g->net = *net;
  g->Va = calloc(g->N1 + 1, sizeof(int64_t));
  g->Ea = calloc(E + 1, sizeof(int64_t));
  double *Wa = calloc(E + 1, sizeof(double));
  ora_build_csr(net, g->Va, g->Ea, Wa);
  free(Wa);
  g->x = calloc(E + 1, sizeof(double));
  g->t = calloc(E + 1, sizeof(double));
  g->dt = calloc(E + 1, sizeof(double));
  g->prev = calloc(g->N1, sizeof(int64_t));
  g->qn = calloc(g->N1, sizeof(int64_t));
  g->dist = calloc(g->N1, sizeof(double));
  g->mark = calloc(E + 1, sizeof(int64_t));
  g->walk = calloc(g->N1 + 1, sizeof(int64_t));
  /* OD pairs */
  g->od_start = calloc(g->Z + 2, sizeof(int64_t));
  int64_t cnt = 0;
  for (int64_t o = 1; o <= dem->num_zones && o <= g->Z; o++)
    for (int64_t i = dem->start[o - 1]; i < dem->start[o]; i++)
      if (dem->dest[i] != o && dem->demand[i] > 0) cnt++;
  g->od = calloc(cnt + 1, sizeof(od_t));
  g->od_paths_cap = 4 * (cnt + 1);
  g->od_paths = malloc(sizeof(int64_t) * g->od_paths_cap);
  for (int64_t o = 1; o <= g->Z; o++) {
    g->od_start[o] = g->nod;
    if (o > dem->num_zones) continue;
    for (int64_t i = dem->start[o - 1]; i < dem->start[o]; i++) {
      if (dem->dest[i] == o || !(dem->demand[i] > 0)) continue;
      od_t *p = &g->od[g->nod++];
      p->origin = o;
      p->dest = dem->dest[i];
      p->demand = dem->demand[i];
      p->first = g->od_paths_len;
      p->cap = 2;
      g->od_paths_len += 2;
    }
  }
  g->od_start[g->Z + 1] = g->nod;
  /* all-or-nothing start at free flow */
  for (int64_t k = 0; k < E; k++) refresh(g, k);
  for (int64_t o = 1; o <= g->Z; o++) {
    if (g->od_start[o] == g->od_start[o + 1]) continue;
    sssp(g, o);
    for (int64_t k = g->od_start[o]; k < g->od_start[o + 1]; k++) {
      const int64_t n = tree_path(g, o, g->od[k].dest);
      if (n < 0) { ora_gp_destroy(g); return NULL; }
      od_push(g, k, add_path(g, g->walk, n, g->od[k].demand));
      for (int64_t j = 0; j < n; j++) g->x[g->walk[j]] += g->od[k].demand;
    }
  }
  for (int64_t k = 0; k < E; k++) refresh(g, k);
  return g;
}

void ora_gp_destroy(void *h)
{
  gp_t *g = h;
  if (!g) return;
  void *ptrs[] = {g->t0, g->cap, g->B, g->pw, g->init, g->term, g->Va, g->Ea, g->x, g->t,
                  g->dt, g->od, g->od_start, g->p_off, g->p_len, g->p_flow, g->links,
                  g->od_paths, g->prev, g->qn, g->mark, g->walk, g->dist};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
  free(g);
}

/* New link parameters (free-flow time, capacity); flows are kept. */
void ora_gp_set_params(void *h, const double *t0, const double *cap)
{
  gp_t *g = h;
  for (int64_t k = 0; k < g->E; k++) {
    g->t0[k] = t0[k];
    g->cap[k] = cap[k];
    refresh(g, k);
  }
}

/* relative gap (sum x t - sum d u) / sum d u with fresh shortest paths */
double ora_gp_relgap(void *h)
{
  gp_t *g = h;
  double tstt = 0, sptt = 0;
  for (int64_t k = 0; k < g->E; k++) tstt += g->x[k] * g->t[k];
  for (int64_t o = 1; o <= g->Z; o++) {
    if (g->od_start[o] == g->od_start[o + 1]) continue;
    sssp(g, o);
    for (int64_t k = g->od_start[o]; k < g->od_start[o + 1]; k++)
      sptt += g->od[k].demand * g->dist[g->od[k].dest];
  }
  return (tstt - sptt) / sptt;
}

/* Sweeps until the relative gap is <= tol (checked every 4 sweeps) or
   max_sweeps; returns the sweeps done, -1 on a missing path. */
int ora_gp_run(void *h, int max_sweeps, double tol, double *relgap)
{
  gp_t *g = h;
  int sw = 0;
  *relgap = ora_gp_relgap(g);
  while (sw < max_sweeps && !(*relgap <= tol)) {
    for (int64_t o = 1; o <= g->Z; o++) {
      if (g->od_start[o] == g->od_start[o + 1]) continue;
      sssp(g, o);
      for (int64_t k = g->od_start[o]; k < g->od_start[o + 1]; k++) {
        const int64_t n = tree_path(g, o, g->od[k].dest);
        if (n < 0) return -1;
        od_t *p = &g->od[k];
        int found = 0;
        for (int64_t i = 0; i < p->n && !found; i++)
          found = same_path(g, g->od_paths[p->first + i], g->walk, n);
        if (!found) od_push(g, k, add_path(g, g->walk, n, 0.0));
        shift_od(g, k);
      }
    }
    /* volumes rebuilt from the path flows (no drift from the increments) */
    memset(g->x, 0, sizeof(double) * g->E);
    for (int64_t k = 0; k < g->nod; k++) {
      const od_t *p = &g->od[k];
      for (int64_t i = 0; i < p->n; i++) {
        const int64_t pid = g->od_paths[p->first + i];
        const int64_t *lk = g->links + g->p_off[pid];
        for (int64_t j = 0; j < g->p_len[pid]; j++) g->x[lk[j]] += g->p_flow[pid];
      }
    }
    for (int64_t k = 0; k < g->E; k++) refresh(g, k);
    sw++;
    if (sw % 4 == 0 || sw == max_sweeps) *relgap = ora_gp_relgap(g);
  }
  return sw;
}

void ora_gp_flows(void *h, double *x, double *t)
{
  gp_t *g = h;
  memcpy(x, g->x, sizeof(double) * g->E);
  if (t) memcpy(t, g->t, sizeof(double) * g->E);
}

/* Path set: counts first (paths, links) */
void ora_gp_count(void *h, int64_t *num_paths, int64_t *num_links)
{
  gp_t *g = h;
  int64_t np = 0, nl = 0;
  for (int64_t k = 0; k < g->nod; k++)
    for (int64_t i = 0; i < g->od[k].n; i++) {
      np++;
      nl += g->p_len[g->od_paths[g->od[k].first + i]];
    }
  *num_paths = np;
  *num_links = nl;
}

/* path p: OD index od[p] (0-based, origin-major order), flow[p], links
   links[off[p] .. off[p+1]) (0-based link ids, origin to destination) */
void ora_gp_paths(void *h, int64_t *od, double *flow, int64_t *off, int64_t *links,
                  int64_t *od_origin, int64_t *od_dest, double *od_demand)
{
  gp_t *g = h;
  int64_t np = 0, nl = 0;
  off[0] = 0;
  for (int64_t k = 0; k < g->nod; k++) {
    od_origin[k] = g->od[k].origin;
    od_dest[k] = g->od[k].dest;
    od_demand[k] = g->od[k].demand;
    for (int64_t i = 0; i < g->od[k].n; i++) {
      const int64_t pid = g->od_paths[g->od[k].first + i];
      od[np] = k;
      flow[np] = g->p_flow[pid];
      memcpy(links + nl, g->links + g->p_off[pid], sizeof(int64_t) * g->p_len[pid]);
      nl += g->p_len[pid];
      off[++np] = nl;
    }
  }
}

int64_t ora_gp_num_od(void *h) { return ((gp_t *)h)->nod; }
