/*
 * oracle.c -- CPU restatement of the reference's Frank-Wolfe UE path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP product,
 * used by tests/, the smoke check and bench.py's CPU-baseline leg.
 *
 * Build: gcc -O3 -fno-associative-math -ffp-contract=off (the reference's
 * host flags, trunk/tap/common.mk:50, plus an explicit no-contraction flag).
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define UNSET   (-1)   /* never queued / no predecessor          */
#define POPPED  (-2)   /* was in the queue and has been removed   */

/* CSR over 1-based node ids; link index == edge index because links are
   sorted by (init, term).  Restates tap_functions.c:313-339 (initial cost is
   the free-flow time). */
void ora_build_csr(const ora_net *net, int64_t *Va, int64_t *Ea, double *Wa)
{
  const int64_t n_rows = net->num_nodes + 1;
  int64_t e = 0;
  for (int64_t v = 0; v < n_rows; v++) {
    Va[v] = e;
    for (; e < net->num_links && net->init[e] == v; e++) {
      Ea[e] = net->term[e];
      Wa[e] = net->fftime[e];
    }
  }
  Va[n_rows] = e;
}

/* d'Esopo-Pape with Large-Label-Last, restating sssp_pape.c:284-501.
 * Queue discipline: a node relabelled for the first time joins the tail, a
 * node that was queued before re-enters at the head; before each pop, heads
 * whose label exceeds the running mean label of the queue are rotated to the
 * tail (the running sum is plain, not compensated: USE_KAHAN_SUM is off,
 * sssp_pape.c:39).  Zones (< first_thru_node) other than the start node are
 * labelled but never expanded (sssp_pape.c:355). */
void ora_sssp_pape_lll(const int64_t *Va, const int64_t *Ea, const double *Wa,
                       int64_t num_nodes, int64_t start_node, int64_t first_thru_node,
                       int64_t *prevlink, double *dist, int64_t *nxt)
{
  for (int64_t v = 0; v < num_nodes; v++) {
    prevlink[v] = UNSET;
    nxt[v] = UNSET;
    dist[v] = (v == start_node) ? 0.0 : DBL_MAX;
  }
  int64_t head = UNSET, tail = UNSET, count = 0;
  double label_sum = 0.0;
  int64_t u = start_node;

  while (u != UNSET && u != POPPED) {
    if (u >= first_thru_node || u == start_node) {
      const double du = dist[u];
      for (int64_t e = Va[u]; e < Va[u + 1]; e++) {
        const int64_t v = Ea[e];
        const double cand = du + Wa[e];
        if (!(cand < dist[v]))
          continue;
        const double prev = dist[v];
        dist[v] = cand;
        prevlink[v] = e;
        if (nxt[v] == POPPED) {                 /* re-entry: to the head */
          nxt[v] = head;
          head = v;
          if (tail == UNSET)
            tail = v;
          count++;
          label_sum += dist[v];
        } else if (nxt[v] == UNSET && v != tail) {  /* first entry: to the tail */
          if (tail != UNSET) {
            nxt[tail] = v;
            tail = v;
            count++;
            label_sum += dist[v];
          } else {                              /* queue empty: sum restarts */
            head = tail = v;
            nxt[v] = UNSET;
            count = 1;
            label_sum = dist[v];
          }
        } else {                                /* already queued */
          label_sum += cand;
          label_sum += -prev;
        }
      }
    }

    /* pick the next node: rotate large labels to the tail */
    const double mean = label_sum / (double)count;
    int64_t checked = 0;
    u = head;
    while (checked < count) {
      const int64_t after = nxt[u];
      if (count > 1 && dist[u] > mean && checked < count - 1) {
        head = nxt[u];
        nxt[u] = UNSET;
        if (tail != UNSET) {
          nxt[tail] = u;
          tail = u;
        } else {
          head = tail = u;
        }
      } else {
        head = nxt[u];
        nxt[u] = POPPED;
        if (tail == u)
          tail = UNSET;
        count--;
        label_sum += -dist[u];
        break;
      }
      checked++;
      u = after;
    }
  }
}

/* plain d'Esopo-Pape, restating sssp_pape.c:120-220 */
void ora_sssp_pape(const int64_t *Va, const int64_t *Ea, const double *Wa,
                   int64_t num_nodes, int64_t start_node, int64_t first_thru_node,
                   int64_t *prevlink, double *dist, int64_t *nxt)
{
  for (int64_t v = 0; v < num_nodes; v++) {
    prevlink[v] = UNSET;
    nxt[v] = UNSET;
    dist[v] = (v == start_node) ? 0.0 : DBL_MAX;
  }
  int64_t head = UNSET, tail = UNSET;
  int64_t u = start_node;
  for (;;) {
    if (u >= first_thru_node || u == start_node) {
      for (int64_t e = Va[u]; e < Va[u + 1]; e++) {
        const int64_t v = Ea[e];
        const double cand = dist[u] + Wa[e];
        if (!(cand < dist[v]))
          continue;
        dist[v] = cand;
        prevlink[v] = e;
        if (nxt[v] == POPPED) {
          nxt[v] = head;
          head = v;
          if (tail == UNSET)
            tail = v;
        } else if (nxt[v] == UNSET && v != tail) {
          if (tail != UNSET) {
            nxt[tail] = v;
            tail = v;
          } else {
            head = tail = v;
            nxt[v] = UNSET;
          }
        }
      }
    }
    u = head;
    if (u == UNSET || u == POPPED)
      break;
    head = nxt[u];
    nxt[u] = POPPED;
    if (tail == u)
      tail = UNSET;
  }
}

/* ---- all-origin SSSP (node-major output, the pape_cuda layout) ---------- */

typedef struct {
  const int64_t *Va, *Ea;
  const double *Wa;
  int64_t num_nodes, num_zones, first_thru, o_lo, o_hi;
  int lll;
  double *dist;
  int64_t *pred;
} sssp_job;

static void *sssp_worker(void *arg)
{
  sssp_job *j = (sssp_job *)arg;
  double *d = malloc(sizeof(double) * j->num_nodes);
  int64_t *p = malloc(sizeof(int64_t) * j->num_nodes);
  int64_t *q = malloc(sizeof(int64_t) * j->num_nodes);
  for (int64_t s = j->o_lo; s <= j->o_hi; s++) {
    if (j->lll)
      ora_sssp_pape_lll(j->Va, j->Ea, j->Wa, j->num_nodes, s, j->first_thru, p, d, q);
    else
      ora_sssp_pape(j->Va, j->Ea, j->Wa, j->num_nodes, s, j->first_thru, p, d, q);
    for (int64_t v = 0; v < j->num_nodes; v++) {
      j->dist[v * j->num_zones + s] = d[v];
      j->pred[v * j->num_zones + s] = p[v];
    }
  }
  free(d); free(p); free(q);
  return NULL;
}

void ora_sssp_all(const int64_t *Va, const int64_t *Ea, const double *Wa,
                  int64_t num_nodes, int64_t num_zones, int64_t first_thru_node,
                  int lll, double *dist, int64_t *pred, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > num_zones) nthreads = (int)num_zones;
  if (nthreads < 1) nthreads = 1;
  sssp_job *jobs = calloc(nthreads, sizeof(sssp_job));
  pthread_t *th = calloc(nthreads, sizeof(pthread_t));
  int64_t per = (num_zones + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = (sssp_job){Va, Ea, Wa, num_nodes, num_zones, first_thru_node,
                         1 + t * per, t * per + per, lll, dist, pred};
    if (jobs[t].o_hi > num_zones) jobs[t].o_hi = num_zones;
    pthread_create(&th[t], NULL, sssp_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++)
    pthread_join(th[t], NULL);
  free(jobs); free(th);
}

/* ---- link functions (tap_functions.c) ----------------------------------- */

/* Power function used by BPR and the objective.  Mode 0 (default) is libm
 * pow, as in the reference.  Mode 1 returns pow(x, 4) correctly rounded,
 * computed independently of the product's code: Veltkamp/Dekker exact
 * products (no fma; this file is built with -ffp-contract=off) give
 * x^2 = h1 + l1 and h1^2 = h2 + l2 exactly, and x^4 = h2 + l2 + 2 h1 l1 + l1^2
 * is rounded once.  glibc's pow(x, 4) is within 0.52 ulp but not always
 * correctly rounded, so the two modes differ by one ulp on a small fraction
 * of arguments; mode 1 lets tests compare the GPU engine bit for bit. */
static int g_pow_mode = 0;

void ora_set_pow_mode(int mode) { g_pow_mode = mode; }

static void exact_mul(double a, double b, double *hi, double *lo)
{
  const double split = 134217729.0;            /* 2^27 + 1 */
  const double p = a * b;
  const double ca = split * a, ah = ca - (ca - a), al = a - ah;
  const double cb = split * b, bh = cb - (cb - b), bl = b - bh;
  *hi = p;
  *lo = ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}

static double pow4_rounded_once(double x)
{
  double h1, l1, h2, l2;
  if (!(fabs(x) < 1e75) || fabs(x) < 1e-70)     /* outside the exact-split range */
    return pow(x, 4.0);
  exact_mul(x, x, &h1, &l1);
  exact_mul(h1, h1, &h2, &l2);
  const double tail = l2 + (2.0 * h1) * l1 + l1 * l1;
  return h2 + tail;
}

double ora_pow4_exact_check(double x) { return pow4_rounded_once(x); }

static double ora_pow(double x, double p)
{
  if (g_pow_mode == 1 && p == 4.0)
    return pow4_rounded_once(x);
  return pow(x, p);
}

/* BPR cost t0*(1 + B*(Q/c)^power), tap_functions.c:55-65 */
double ora_bpr(double t0, double cap, double B, double power, double Q)
{
  return t0 * (1.0 + B * ora_pow(Q / cap, power));
}

/* integral of BPR from 0 to Q in the reference's algebraic form,
   tap_functions.c:112-126 (0 when power < 0) */
double ora_integral(double t0, double cap, double B, double power, double Q)
{
  if (!(power >= 0))
    return 0.0;
  return Q * t0 * (1 + ((B / ora_pow(cap, power)) / (power + 1)) * ora_pow(Q, power));
}

/* sum_k t_k(x_k + lambda d_k) d_k in link order, tap_functions.c:83-97 */
double ora_dirderiv(const ora_net *net, const double *cap, const double *x,
                    const double *d, double lambda)
{
  double s = 0;
  for (int64_t k = 0; k < net->num_links; k++) {
    const double q = x[k] + lambda * d[k];
    s += ora_bpr(net->fftime[k], cap[k], net->B[k], net->power[k], q) * d[k];
  }
  return s;
}

/* Beckmann objective, tap_functions.c:143-152 */
double ora_objective(const ora_net *net, const double *cap, const double *x)
{
  double s = 0;
  for (int64_t k = 0; k < net->num_links; k++)
    s += ora_integral(net->fftime[k], cap[k], net->B[k], net->power[k], x[k]);
  return s;
}

/* sum x*c (VHT when c = BPR(x)), tap_functions.c:408-417 */
double ora_total_cost(int64_t E, const double *x, const double *c)
{
  double s = 0;
  for (int64_t k = 0; k < E; k++)
    s += x[k] * c[k];
  return s;
}

/* Step size, restating tap_functions.c:469-507: 0 if g(0) >= 0, 1 if
   g(1) <= 0, otherwise halve lambda from 1/2 and return the first lambda
   with g(lambda) <= 0 (at most 1000 halvings after the first). */
double ora_line_search(const ora_net *net, const double *cap, const double *x,
                       const double *d)
{
  if (ora_dirderiv(net, cap, x, d, 0) >= 0)
    return 0;
  if (ora_dirderiv(net, cap, x, d, 1) <= 0)
    return 1;
  double left = 0, right = 1, lambda = 0.5;
  if (ora_dirderiv(net, cap, x, d, lambda) <= 0)
    left = lambda;
  else
    right = lambda;
  lambda = 0.5 * (left + right);
  for (int n = 0; left == 0 && n < 1000; n++) {
    if (ora_dirderiv(net, cap, x, d, lambda) <= 0)
      left = lambda;
    else
      right = lambda;
    lambda = 0.5 * (left + right);
  }
  return left;
}

void ora_update_costs(const ora_net *net, const double *cap, const double *x, double *c)
{
  for (int64_t k = 0; k < net->num_links; k++)
    c[k] = ora_bpr(net->fftime[k], cap[k], net->B[k], net->power[k], x[k]);
}

/* tap_functions.c:377-386 (called with num_zones+1: all origins 1..Z) */
double ora_demand_sum(const ora_demand *dem)
{
  double s = 0;
  for (int64_t o = 1; o <= dem->num_zones; o++)
    for (int64_t i = dem->start[o - 1]; i < dem->start[o]; i++)
      s += dem->demand[i];
  return s;
}

/* ---- AON (fw_assign) ----------------------------------------------------- */

typedef struct {
  const ora_net *net;
  const ora_demand *dem;
  const int64_t *Va, *Ea;
  const double *Wa;
  int64_t o_lo, o_hi;
  double *y;           /* private partial volumes */
  int64_t bad_orig, bad_dest;
  const int64_t *rv_off, *rv_link;   /* in-links by head (tie mode 1), else NULL */
} aon_job;

/* 0: each node's predecessor is the link Pape's queue order leaves in
   prevlink (the reference's choice, default); 1: on the walks, a node's
   predecessor is re-chosen among its tight in-links u->v (u expandable,
   u != v, fl(dist[u] + w) == dist[v], dist[u] < dist[v]) as the one with
   the smallest dist[u], then the smallest u -- the GPU engine's rule, which
   differs from Pape's only where several in-links are tight (exact ties).
   A node with no such in-link keeps Pape's link. */
static int g_tie_mode = 0;

void ora_set_tie_mode(int mode) { g_tie_mode = mode; }

static int64_t pred_by_rule(const aon_job *j, const double *dist, int64_t v, int64_t o,
                            int64_t pape)
{
  const double dv = dist[v];
  int64_t best = -1, best_u = 0;
  double best_du = dv;
  for (int64_t i = j->rv_off[v]; i < j->rv_off[v + 1]; i++) {
    const int64_t k = j->rv_link[i];
    const int64_t u = j->net->init[k];
    if (u == v || (u < j->net->first_thru_node && u != o))
      continue;
    const double du = dist[u];
    if (du == DBL_MAX || !(du < dv) || du + j->Wa[k] != dv)
      continue;
    if (best < 0 || du < best_du || (du == best_du && u < best_u)) {
      best = k;
      best_du = du;
      best_u = u;
    }
  }
  return best >= 0 ? best : pape;
}

/* one block of origins: Pape-LLL tree, then walk each positive-demand path
   from the destination back to the origin, restating
   tap_frankwolfe.c:131-163 / tap_frankwolfe_pthread.c:270-307 */
static void *aon_worker(void *arg)
{
  aon_job *j = (aon_job *)arg;
  const int64_t n = j->net->num_nodes + 1;
  double *dist = malloc(sizeof(double) * n);
  int64_t *pred = malloc(sizeof(int64_t) * n);
  int64_t *q = malloc(sizeof(int64_t) * n);
  j->bad_orig = 0;
  for (int64_t o = j->o_lo; o <= j->o_hi && !j->bad_orig; o++) {
    ora_sssp_pape_lll(j->Va, j->Ea, j->Wa, n, o, j->net->first_thru_node, pred, dist, q);
    for (int64_t i = j->dem->start[o - 1]; i < j->dem->start[o]; i++) {
      const int64_t dst = j->dem->dest[i];
      const double f = j->dem->demand[i];
      if (dst == o || f <= 0)
        continue;
      for (int64_t v = dst; v != o;) {
        int64_t k = pred[v];
        if (k >= 0 && j->rv_off)
          k = pred_by_rule(j, dist, v, o, k);
        if (k < 0) {
          j->bad_orig = o;
          j->bad_dest = dst;
          break;
        }
        j->y[k] += f;
        v = j->net->init[k];
      }
      if (j->bad_orig)
        break;
    }
  }
  free(dist); free(pred); free(q);
  return NULL;
}

int ora_fw_assign(const ora_net *net, const ora_demand *dem, const int64_t *Va,
                  const int64_t *Ea, const double *Wa, double *y, int nthreads,
                  int64_t *bad_orig, int64_t *bad_dest)
{
  const int64_t E = net->num_links, Z = net->num_zones;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > Z) nthreads = (int)(Z > 0 ? Z : 1);
  aon_job *jobs = calloc(nthreads, sizeof(aon_job));
  pthread_t *th = calloc(nthreads, sizeof(pthread_t));
  const int64_t per = (Z + nthreads - 1) / nthreads;   /* contiguous blocks */
  int64_t *rv_off = NULL, *rv_link = NULL;
  if (g_tie_mode == 1) {          /* in-links grouped by head, in link order */
    const int64_t n = net->num_nodes + 1;
    rv_off = calloc(n + 1, sizeof(int64_t));
    rv_link = malloc(sizeof(int64_t) * (E > 0 ? E : 1));
    for (int64_t k = 0; k < E; k++)
      rv_off[net->term[k] + 1]++;
    for (int64_t v = 0; v < n; v++)
      rv_off[v + 1] += rv_off[v];
    int64_t *fill = malloc(sizeof(int64_t) * n);
    for (int64_t v = 0; v < n; v++)
      fill[v] = rv_off[v];
    for (int64_t k = 0; k < E; k++)
      rv_link[fill[net->term[k]]++] = k;
    free(fill);
  }
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = (aon_job){net, dem, Va, Ea, Wa, 1 + t * per, t * per + per,
                        calloc(E, sizeof(double)), 0, 0, rv_off, rv_link};
    if (jobs[t].o_hi > Z) jobs[t].o_hi = Z;
    if (nthreads == 1)
      aon_worker(&jobs[t]);
    else
      pthread_create(&th[t], NULL, aon_worker, &jobs[t]);
  }
  if (nthreads > 1)
    for (int t = 0; t < nthreads; t++)
      pthread_join(th[t], NULL);
  int rc = 0;
  for (int64_t k = 0; k < E; k++)
    y[k] = 0;
  for (int t = 0; t < nthreads; t++) {
    if (jobs[t].bad_orig && !rc) {
      rc = -1;
      if (bad_orig) *bad_orig = jobs[t].bad_orig;
      if (bad_dest) *bad_dest = jobs[t].bad_dest;
    }
    for (int64_t k = 0; k < E; k++)
      y[k] += jobs[t].y[k];
    free(jobs[t].y);
  }
  free(jobs); free(th);
  free(rv_off); free(rv_link);
  return rc;
}

/* ---- Frank-Wolfe loop ---------------------------------------------------- */

/* Restates tap_frankwolfe_pthread.c:654-803 (same arithmetic as
   tap_frankwolfe.c:384-476) without signals or file output.  The warm-start
   branch initialises iter/relgap/best_lowerbound, which the reference leaves
   undefined (tap_frankwolfe_pthread.c:702-710). */
int ora_fw_solve(const ora_net *net, const ora_demand *dem, int max_iters,
                 double target_relgap, int nthreads, const double *x_init,
                 double *x_out, double *c_out, ora_iter *trace, int trace_cap,
                 int *iters_out, double *relgap_out, double *obj_out, double *vht_out)
{
  const int64_t E = net->num_links, N1 = net->num_nodes + 1;
  double *cap = malloc(sizeof(double) * E);
  int64_t *Va = malloc(sizeof(int64_t) * (N1 + 1));
  int64_t *Ea = malloc(sizeof(int64_t) * (E > 0 ? E : 1));
  double *c = malloc(sizeof(double) * (E > 0 ? E : 1));
  double *x = malloc(sizeof(double) * (E > 0 ? E : 1));
  double *y = malloc(sizeof(double) * (E > 0 ? E : 1));
  int rc = 0;

  /* capacity <= 0 -> smallest positive capacity (pthread.c:654-669) */
  double min_cap = DBL_MAX;
  for (int64_t k = 0; k < E; k++)
    if (net->capacity[k] > 0.0 && net->capacity[k] < min_cap)
      min_cap = net->capacity[k];
  for (int64_t k = 0; k < E; k++)
    cap[k] = net->capacity[k] > 0.0 ? net->capacity[k] : min_cap;

  ora_build_csr(net, Va, Ea, c);
  const double total_demand = ora_demand_sum(dem);
  int iter = 0;
  double relgap = DBL_MAX, best_lb = -DBL_MAX, obj = 0;

  if (x_init) {
    memcpy(x, x_init, sizeof(double) * E);
    ora_update_costs(net, cap, x, c);
  } else {
    rc = ora_fw_assign(net, dem, Va, Ea, c, x, nthreads, NULL, NULL);
    if (rc) goto done;
    ora_update_costs(net, cap, x, c);
  }

  while ((max_iters <= 0 || iter < max_iters) &&
         (target_relgap <= 0 || relgap > target_relgap)) {
    rc = ora_fw_assign(net, dem, Va, Ea, c, y, nthreads, NULL, NULL);
    if (rc) goto done;
    for (int64_t k = 0; k < E; k++)
      y[k] = y[k] - x[k];                       /* direction d = y - x */
    const double lambda = ora_line_search(net, cap, x, y);
    const double gap = -ora_dirderiv(net, cap, x, y, 0);
    obj = ora_objective(net, cap, x);
    const double lb = obj - gap;
    if (lb > best_lb)
      best_lb = lb;
    if (best_lb != 0)
      relgap = (obj - best_lb) / fabs(best_lb);
    const double avg_excess = -ora_total_cost(E, y, c) / total_demand;
    const double vht = ora_total_cost(E, x, c);
    for (int64_t k = 0; k < E; k++)
      x[k] += lambda * y[k];
    ora_update_costs(net, cap, x, c);
    if (trace && iter < trace_cap)
      trace[iter] = (ora_iter){lambda, gap, obj, lb, best_lb, relgap, avg_excess, vht};
    iter++;
  }

done:
  if (x_out) memcpy(x_out, x, sizeof(double) * E);
  if (c_out) memcpy(c_out, c, sizeof(double) * E);
  if (iters_out) *iters_out = iter;
  if (relgap_out) *relgap_out = relgap;
  if (obj_out) *obj_out = obj;
  if (vht_out) *vht_out = ora_total_cost(E, x, c);
  free(cap); free(Va); free(Ea); free(c); free(x); free(y);
  return rc;
}


/* ------------------------------------------------------------------------ */
/* greedy step-cost assignment                                               */
/* ------------------------------------------------------------------------ */

/* distance_from_next_step, tap_functions.c:241-266 (NUM_STEPS = 10,
   tap_functions.h:24): volume left to the next step of the cost step
   function; note the reference returns vol itself whenever the next step
   is at or above vol, which ceil() makes the usual case. */
static double ora_next_step(double cap, double vol)
{
  const double flowstepsize = cap / (double)10;
  if (vol <= 0.0)
    return flowstepsize;
  if (vol >= cap)
    return DBL_MAX;
  const double next_step = ceil(vol / flowstepsize);
  const double next_step_vol = next_step * flowstepsize;
  if (next_step_vol >= vol)
    return vol;
  return next_step_vol - vol;
}

/* tap_greedy.c: greedy_assign (:84-150) inside the main loop (:300-330):
   per iteration, for each origin a plain d'Esopo-Pape tree at the current
   costs; for each of its OD pairs in file order the flow that brings the
   first link of the path to its next step (at most the remaining demand)
   is added along the path; then BPR costs; stop when no demand is left or
   the unsatisfied total changed by less than EPS = 1e-8
   (tap_functions.h:23).  max_iters <= 0: no limit.  Returns 0, or -2 on
   bad input; *stalled = 1 when it stopped on the EPS test. */
int ora_greedy(const ora_net *net, const ora_demand *dem, int max_iters, double *x, double *c,
               int *iters_out, double *unsat_out, double *total_cost_out, int *stalled,
               double *unsat_trace, int trace_cap)
{
  const int64_t E = net->num_links, N1 = net->num_nodes + 1, Z = net->num_zones;
  if (E <= 0 || dem->num_zones > Z)
    return -2;
  double *cap = malloc(sizeof(double) * E);
  int64_t *Va = malloc(sizeof(int64_t) * (N1 + 1)), *Ea = malloc(sizeof(int64_t) * E);
  double *Wa = malloc(sizeof(double) * E);
  int64_t *prev = malloc(sizeof(int64_t) * N1), *nxt = malloc(sizeof(int64_t) * N1);
  double *dist = malloc(sizeof(double) * N1);
  const int64_t nitems = dem->start[dem->num_zones];
  double *rem = malloc(sizeof(double) * (nitems + 1));
  ora_build_csr(net, Va, Ea, Wa);
  /* capacity <= 0 -> smallest positive capacity (tap_greedy.c:270-285) */
  double min_cap = DBL_MAX;
  for (int64_t k = 0; k < E; k++)
    if (net->capacity[k] > 0.0 && net->capacity[k] < min_cap)
      min_cap = net->capacity[k];
  for (int64_t k = 0; k < E; k++)
    cap[k] = net->capacity[k] > 0.0 ? net->capacity[k] : min_cap;
  for (int64_t k = 0; k < E; k++) {
    x[k] = 0.0;
    c[k] = Wa[k];                     /* link_costs = free-flow times */
  }
  for (int64_t o = 1; o <= dem->num_zones; o++)
    for (int64_t i = dem->start[o - 1]; i < dem->start[o]; i++)
      rem[i] = (dem->dest[i] == o) ? 0.0 : dem->demand[i];
  double unsat = 0.0;
  for (int64_t i = 0; i < nitems; i++)
    unsat += rem[i];
  int iter = 0;
  *stalled = 0;
  while (unsat > 0 && (max_iters <= 0 || iter < max_iters)) {
    for (int64_t o = 1; o <= dem->num_zones; o++) {
      ora_sssp_pape(Va, Ea, c, N1, o, net->first_thru_node, prev, dist, nxt);
      for (int64_t i = dem->start[o - 1]; i < dem->start[o]; i++) {
        const int64_t d = dem->dest[i];
        const double routeflow = rem[i];
        if (o == d || routeflow <= 0)
          continue;
        double addflow = routeflow;
        int fail = 0;
        for (int64_t v = d; v != o;) {
          const int64_t k = prev[v];
          if (k < 0) { fail = 1; break; }
          const double fns = ora_next_step(cap[k], x[k]);
          if (fns < addflow)
            addflow = fns;
          v = net->init[k];
        }
        if (fail)
          continue;
        for (int64_t v = d; v != o;) {
          const int64_t k = prev[v];
          x[k] += addflow;
          v = net->init[k];
        }
        rem[i] -= addflow;
      }
    }
    ora_update_costs(net, cap, x, c);
    const double prev_unsat = unsat;
    unsat = 0.0;
    for (int64_t i = 0; i < nitems; i++)
      unsat += rem[i];
    if (unsat_trace && iter < trace_cap)
      unsat_trace[iter] = unsat;
    iter++;
    if (fabs(unsat - prev_unsat) < 1e-08) {
      *stalled = 1;
      break;
    }
  }
  *iters_out = iter;
  *unsat_out = unsat;
  *total_cost_out = ora_total_cost(E, x, c);
  free(cap); free(Va); free(Ea); free(Wa); free(prev); free(nxt); free(dist); free(rem);
  return 0;
}
