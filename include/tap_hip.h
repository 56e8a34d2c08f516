/*
 * tap_hip.h -- the reference's host API for the Frank-Wolfe path, served by
 * the MI355X engine (libtap_hip.so, built on include/tapgpu.h).
 *
 * A program written against trunk/tap (tap_types.h, parsetapfiles.h,
 * tap_functions.h, tap_frankwolfe_pthread.h, pape_cuda_host.h,
 * tap_frankwolfe_cuda.h, volume_update_host.h) links this library instead
 * and keeps its calls.  The structs below have the reference's field names,
 * types and order (trunk/tap/tap_types.h:20-103) so that they are
 * layout-compatible with a caller compiled against the reference headers.
 *
 * Behaviour that differs on purpose:
 *   - tap_frankwolfe() initialises the iteration count, relative gap and
 *     best lower bound on a warm start (the reference leaves them
 *     uninitialised, trunk/tap/tap_frankwolfe_pthread.c:702-710);
 *   - num_threads / num_cores are accepted and ignored (one GPU does the
 *     work);
 *   - fw_assign() leaves its dist / predlink work arrays untouched.
 * Failure paths keep the reference's behaviour: fw_assign() and pape_cuda()
 * print an error and exit(1) where the reference exits (no path, device
 * error); tap_frankwolfe() returns 1.
 */
#ifndef TAP_HIP_H
#define TAP_HIP_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- data model (trunk/tap/tap_types.h:20-103) ---------------------------- */

typedef struct link_data_s {
  long init_node;
  long term_node;
  double capacity;
  double length;
  double free_flow_time;
  double B;
  double power;
  double speed_limit;
  double toll;
  short link_type;
} link_data_t;

typedef struct net_data_s {
  long num_zones;
  long num_nodes;
  long first_thru_node;
  long num_links;
  link_data_t *links;
} net_data_t;

typedef struct demand_data_s {
  long dest;       /* 0 terminates an origin's array */
  double demand;
} demand_data_t;

typedef struct link_flow_s {
  long init_node;
  long term_node;
  double volume;
  double cost;
} link_flow_t;

#define MAX_CHANGE_ID_LEN 127

typedef enum net_mod_type_e {
  MOD_TYPE_INVALID_E = 0,
  MOD_TYPE_ADD_E = 1,
  MOD_TYPE_CHANGE_E = 2
} net_mod_type_e;

typedef struct net_mod_s {
  char change_id[MAX_CHANGE_ID_LEN + 1];
  net_mod_type_e modtype;
  link_data_t mod_link;
  double project_cost;
} net_mod_t;

typedef enum trip_mod_type_e {
  TRIP_MOD_TYPE_INVALID_E = 0,
  TRIP_MOD_TYPE_POINT_E = 1,
  TRIP_MOD_TYPE_UNIFORM_E = 2,
  TRIP_MOD_TYPE_ALL_E = 3
} trip_mod_type_e;

typedef struct trip_mod_s {
  char trip_change_id[MAX_CHANGE_ID_LEN + 1];
  trip_mod_type_e trip_modtype;
  long origin;
  long dest;
  double multiplier;
} trip_mod_t;

/* --- TNTP files (trunk/tap/parsetapfiles.h:26-50) -------------------------- */

int parse_net_file(FILE *fp, net_data_t *net);                             /* :62-146  */
int parse_trips_file(FILE *fp, demand_data_t **demands[], long *num_zones); /* :181-311 */
int parse_flows_file(FILE *flow_fp, link_flow_t **link_flows, long *num_links); /* :344-423 */
int parse_net_mod_file(FILE *fp, net_mod_t **mods, int *num_mods);          /* :457-516 */
int link_data_compar(const void *ent1, const void *ent2);                   /* :528-540 */
int parse_trip_mod_file(FILE *fp, trip_mod_t **mods, int *num_mods);        /* :573-638 */

/* trunk/tap/tap_functions.c:431-445 and :521-598 */
void print_flow(FILE *fp, const net_data_t *net, const double link_volumes[],
                const double link_costs[]);
int apply_trip_modification(demand_data_t *demands[], trip_mod_t *mod, long num_zones);

/* --- solver (trunk/tap/tap_frankwolfe_pthread.h:23-38) ---------------------- */

extern unsigned int num_cores;
extern unsigned int num_threads;

int tap_frankwolfe(net_data_t *net, demand_data_t **demands, int warm_start_mode,
                   FILE *flows_input_fp, int target_iterations, double target_relgap,
                   FILE *flow_fp, double *total_cost_at_end, double *objvalue_at_end,
                   int *iterations_at_end, double *relgap_at_end, int time_each_iteration);

/* --- GPU entry points of the CUDA build ------------------------------------ */

/* trunk/tap/tap_frankwolfe_cuda.h:26-29 */
void fw_assign(net_data_t *net, demand_data_t *demands[], long Va[], long Ea[], double Wa[],
               double link_volumes[], double dist[], long predlink[]);

/* trunk/tap/pape_cuda_host.h:15-19: distances / predecessors are
   [num_nodes * (1 + num_start_nodes)] with entry [v * num_start_nodes + s]
   for start node s (the reference's layout) */
void pape_cuda(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
               long start_nodes[], long num_start_nodes, long first_thru_node,
               double *distances, long *predecessors);

/* trunk/tap/volume_update_host.h:18-27 */
void link_volume_update(long num_start_nodes, long num_edges, long num_nodes,
                        double link_volumes[], long predlink[]);
void link_volume_data_setup(link_data_t *links, long num_links, long num_start_nodes,
                            demand_data_t *demands[]);
void link_volume_data_cleanup(void);

/* --- single-source shortest paths (trunk/tap/sssp_pape.h:36-64) ---------------- */

/* The reference's CPU d'Esopo-Pape family, each from one start node on the
   GPU engine: Va/Ea/Wa the CSR (num_nodes entries of Va + 1; node ids
   1-based, node 0 without links), prev[v] = link index of the last link of
   the path to v (-1: none), dist[v] = DBL_MAX where unreachable; nodes below
   first_thru_node other than the start are labelled but not expanded
   (trunk/tap/sssp_pape.c:355).  Labels are bit-identical to sssp_pape and
   sssp_pape_lll; predecessors equal theirs wherever the shortest path is
   unique and form a tree.  queue_next (the reference's work array) is
   unused.  The calls are serialised on one cached engine context; for a
   zone start node the trees of every zone are computed in one launch and
   kept for the calls with the same costs (the fw_assign loop,
   trunk/tap/tap_frankwolfe.c:131-147).  Out-of-range start nodes exit(1),
   as the reference's assert does. */
void sssp_pape(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
               long start_node, long first_thru_node, long prev[], double dist[],
               long queue_next[]);                                    /* :36-38 */
void sssp_slf(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
              long start_node, long first_thru_node, long prev[], double dist[],
              long queue_next[]);                                     /* :40-43 */
void sssp_slf_lll(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
                  long start_node, long first_thru_node, long prev[], double dist[],
                  long queue_next[]);                                 /* :45-48 */
void sssp_pape_lll(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
                   long start_node, long first_thru_node, long prev[], double dist[],
                   long queue_next[]);                                /* :50-53 */
/* trunk/tap/sssp_pape.h:55-56: the path dest .. orig along predecessor nodes */
void get_shortest_path(long pred[], long orig, long dest, long path[], long *pathlen);
/* trunk/tap/sssp_pape.h:59-64: also prevnode[v] = tail of prevlink[v] */
void sssp_prevnodefirst(long Va[], long Ea[], double Wa[], long num_nodes, long num_edges,
                        long start_node, long first_thru_node, long old_prevnode[],
                        long prevlink[], long prevnode[], double dist[], long queue_next[]);

/* The reference's greedy assignment program as a library call: the body of
   tap_greedy's main (trunk/tap/tap_greedy.c:296-335) after parsing, with the
   same stderr lines ("initial unsatifised demand", "iter = ...", the
   unchanged-demand warning, "total cost = ...") and print_flow on flow_fp.
   The reference has no library function for it.  Returns 0, or 1 on error. */
int tap_greedy_assign(net_data_t *net, demand_data_t **demands, FILE *flow_fp,
                      double *total_cost_at_end, int *iterations_at_end);

/* global of the CUDA driver (trunk/tap/tap_frankwolfe_cuda.h:23), kept for
   link compatibility; ignored */
extern int using_fermi_architecture;

#ifdef __cplusplus
}
#endif
#endif /* TAP_HIP_H */
