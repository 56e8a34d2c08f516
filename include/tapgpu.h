/*
 * tapgpu.h -- C ABI of the MI355X (gfx950) Frank-Wolfe user-equilibrium
 * engine.  Plain pointers and sizes only; every function returns a status
 * code (TAPGPU_OK == 0) and never exits the process.  Host pointers unless
 * stated otherwise; outputs are written before the call returns.
 *
 * Data model (the reference's, trunk/tap/tap_types.h:20-54): node ids are
 * 1-based, zones are nodes 1..num_zones, nodes below first_thru_node are
 * zones that are labelled but never expanded (trunk/tap/sssp_pape.c:355),
 * links are grouped by init node (the reference sorts them by (init, term))
 * so that link index == CSR edge index (trunk/tap/tap_functions.c:313-339).  OD demand is an origin-major CSR:
 * origin o's items are [od_start[o-1], od_start[o]); items with dest == o
 * or demand <= 0 are ignored (trunk/tap/tap_frankwolfe.c:146-147).
 *
 * One context = one network topology holding num_variants >= 1 independent
 * TAP instances (the batched upgrade-subset form); a plain solve is
 * num_variants == 1.  A context is used from one host thread at a time.
 */
#ifndef TAPGPU_H
#define TAPGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TAPGPU_OK = 0,
  TAPGPU_E_ARG = -1,      /* invalid argument / malformed network            */
  TAPGPU_E_HIP = -2,      /* HIP runtime error (message in tapgpu_last_error) */
  TAPGPU_E_NOPATH = -3,   /* positive demand between disconnected zones      */
  TAPGPU_E_NOMEM = -4,    /* device allocation failed                        */
  TAPGPU_E_STATE = -5,    /* call out of order (e.g. run before init)        */
  TAPGPU_E_LIMIT = -6     /* a size or sweep limit was exceeded              */
};

typedef struct tapgpu_ctx tapgpu_ctx;

/* network (one topology); per-link arrays have num_links entries */
typedef struct {
  int64_t num_zones, num_nodes, first_thru_node, num_links;
  const int64_t *init_node, *term_node;        /* 1-based, grouped by init    */
  const double *capacity, *free_flow_time, *B, *power;
} tapgpu_network;

typedef struct {
  int64_t num_zones;
  const int64_t *od_start;   /* [num_zones + 1] */
  const int64_t *od_dest;    /* 1-based zone ids */
  const double *od_demand;
} tapgpu_demand;

/* per-iteration record of a Frank-Wolfe run (one per completed iteration) */
typedef struct {
  double lambda, gap, objective, lowerbound, best_lowerbound, relgap;
  double avg_excess, vht;    /* vht: total link cost at the pre-step volumes */
} tapgpu_iter;

typedef struct {
  int32_t iterations;        /* iterations completed in total          */
  int32_t converged;         /* relgap <= target reached               */
  double relgap, objective;  /* at the last completed iteration        */
  double vht;                /* total link cost after the last step    */
} tapgpu_result;

/* --- context ------------------------------------------------------------- */

/* Replaces the per-call setup of pape_cuda (trunk/tap/pape_cuda_host.cu:51-95)
   and link_volume_data_setup (trunk/tap/volume_update_host.h:22-24): the
   graph, link parameters and demand are uploaded once and stay in HBM.
   Capacities <= 0 are replaced by the smallest positive capacity
   (trunk/tap/tap_frankwolfe_pthread.c:654-669). */
int tapgpu_create(tapgpu_ctx **ctx, const tapgpu_network *net, const tapgpu_demand *dem);

/* Batched form: num_variants instances over the union topology `net`.
   present[v*num_links + k] (NULL = all present) says whether link k exists
   in variant v; capacity_v / free_flow_time_v / B_v / power_v (each NULL =
   the base value) are [num_variants * num_links] per-variant link records
   (ADD / CHANGE modifications, trunk/tap/tap_mpi_main.c:201-240).  All
   variants share the demand. */
int tapgpu_create_batched(tapgpu_ctx **ctx, const tapgpu_network *net,
                          const tapgpu_demand *dem, int64_t num_variants,
                          const uint8_t *present, const double *capacity_v,
                          const double *free_flow_time_v, const double *B_v,
                          const double *power_v);

void tapgpu_destroy(tapgpu_ctx *ctx);

/* last error message of this context (NULL: of the calling thread's last
   failed create) */
const char *tapgpu_last_error(const tapgpu_ctx *ctx);

int64_t tapgpu_num_variants(const tapgpu_ctx *ctx);

/* --- shortest paths -------------------------------------------------------- */

/* Set the link costs Wa (variant v, link order).  Replaces the per-call Wa
   upload of pape_cuda (trunk/tap/pape_cuda_host.cu:88-91). */
int tapgpu_set_costs(tapgpu_ctx *ctx, int64_t variant, const double *costs);

/* All-origin label-correcting SSSP for variant v from the given start nodes
   (1-based), replacing pape_cuda (trunk/tap/pape_cuda_host.h:15-19).
   dist / pred (either may be NULL) receive node-major results:
   dist[node * ld + col_of(i)] for node 0..num_nodes, where col_of(i) =
   col_base + i for start node i.  With ld = num_start and col_base = 0
   this is a dense [(num_nodes+1) x num_start] matrix; pape_cuda's own
   layout is ld = num_start, col_base = 1 (column index = start node id when
   start nodes are 1..num_start; column num_start of row v is then the unused
   column 0 of row v+1, as in the reference).  The buffers must hold
   num_nodes * ld + col_base + num_start elements.  dist is DBL_MAX where
   unreachable; pred is
   the link index of the last link of the path (-1: none).  Distances are
   bit-identical to the reference's CPU d'Esopo-Pape; pred matches wherever
   the shortest path is unique, and the pred links of each start node form a
   tree rooted at it, zero-cost links included. */
int tapgpu_sssp(tapgpu_ctx *ctx, int64_t variant, const int64_t *start_nodes,
                int64_t num_start, double *dist, int64_t *pred, int64_t ld,
                int64_t col_base);

/* tapgpu_sssp without allocating device memory, for callers that cannot
   call hipMalloc (code under HIP graph capture, tight serving loops,
   harnesses that forbid allocation).  It has no counterpart in the
   reference.  The start nodes (1-based, num_start of them) and the outputs
   are device buffers of the caller: d_dist[(node - 1) * ld + i] and
   d_pred[(node - 1) * ld + i] (link index, -1: none) for node 1..num_nodes
   and start i, ld >= num_start.  Labels and links are those of
   tapgpu_sssp.  Its scratch (the start list and, on the labels-in-HBM
   kernel, 16 bytes per node for each resident workgroup, fewer workgroups
   when they do not fit) is a 64 MiB array in device memory that each GPU
   holds for the life of the process; a start list or network beyond it is
   rejected with TAPGPU_E_LIMIT.  Calls that overlap on one GPU share that
   array and race.  The context's device must be current; the call waits
   for its kernels.  It runs the per-origin kernel with labels in LDS where
   the context uses it, else the one with labels in HBM. */
int tapgpu_sssp_noalloc(tapgpu_ctx *ctx, int64_t variant, const int64_t *d_start_nodes,
                        int64_t num_start, double *d_dist, int32_t *d_pred, int64_t ld);

/* All-or-nothing loading with the current costs of variant v into
   volumes[num_links] (overwritten), replacing fw_assign
   (trunk/tap/tap_frankwolfe_cuda.h:26-29) and link_volume_update
   (trunk/tap/volume_update_host.h:18-20).  Returns TAPGPU_E_NOPATH (with
   the origin/dest in the error message) where the reference exits. */
int tapgpu_aon(tapgpu_ctx *ctx, int64_t variant, double *volumes);

/* All-or-nothing loading along caller-supplied shortest-path trees: predlink
   holds the last link of the path from each zone origin, in the node-major
   layout predlink[node * ld + col_base + (origin - 1)] (node 1..num_nodes,
   -1 = none), the layout pape_cuda returns.  Replaces link_volume_update
   (trunk/tap/volume_update_host.h:18-20, kernel
   trunk/tap/volume_update_kernels.cu:54-86).  The trees must be the
   shortest-path forests of the current costs (acyclic). */
int tapgpu_load_pred(tapgpu_ctx *ctx, int64_t variant, const int64_t *predlink, int64_t ld,
                     int64_t col_base, double *volumes);

/* --- Frank-Wolfe ----------------------------------------------------------- */

/* Initial solution for every variant: all-or-nothing at free-flow costs
   (warm_volumes == NULL) or the given volumes [num_variants * num_links]
   (warm start, trunk/tap/tap_frankwolfe_pthread.c:549-604,682-690); costs
   are then BPR(volumes).  Resets iteration counters and the trace. */
int tapgpu_fw_init(tapgpu_ctx *ctx, const double *warm_volumes);

/* Run Frank-Wolfe iterations (the loop of tap_frankwolfe_pthread.c:712-793)
   until each variant has done max_iters iterations in total (<= 0: no cap)
   or its relative gap is <= target_relgap (<= 0: no target), whichever
   comes first.  Every variant stops independently; its state is then
   frozen.  results (may be NULL) receives num_variants records. */
int tapgpu_fw_run(tapgpu_ctx *ctx, int max_iters, double target_relgap,
                  tapgpu_result *results);

/* Volumes and costs of variant v (link order of the context's network). */
int tapgpu_get_flows(tapgpu_ctx *ctx, int64_t variant, double *volumes, double *costs);

/* Beckmann objective (the sum of link cost integrals of
   links_objective_function, trunk/tap/tap_functions.c:112-152) and VHT
   (sum x * cost) of variant v at its current volumes, compensated sums in a
   fixed order.  (fw_run's per-iteration objective is the one before that
   iteration's step.) */
int tapgpu_get_totals(tapgpu_ctx *ctx, int64_t variant, double *objective, double *vht);

/* Replace variant v's link capacities and/or free-flow times (num_links
   host values each; NULL keeps the current ones), with the capacity rule of
   tapgpu_create (<= 0: the variant's smallest positive capacity).  No
   counterpart in the reference: it lets one context serve many link
   parameter sets (gradient screening) without rebuilding.  The volumes no
   longer match the links: tapgpu_fw_init must come next. */
int tapgpu_set_links(tapgpu_ctx *ctx, int64_t variant, const double *capacity,
                     const double *free_flow_time);

/* Copy up to cap per-iteration records of variant v; *count = number held. */
int tapgpu_get_trace(tapgpu_ctx *ctx, int64_t variant, tapgpu_iter *trace, int cap,
                     int *count);

/* Copy the record of iteration `index` (0-based) of variant v. */
int tapgpu_get_iter(tapgpu_ctx *ctx, int64_t variant, int index, tapgpu_iter *rec);

/* Device buffers of the current state (for zero-copy consumers such as the
   torch extension): volumes / costs are [num_variants * num_links] FP64 in
   HBM, valid until the next call on the context. */
int tapgpu_device_state(tapgpu_ctx *ctx, const double **d_volumes, const double **d_costs);

/* --- instrumentation ------------------------------------------------------- */

typedef struct {
  double sssp_ms, aon_ms, link_ms, total_ms;   /* wall time of the last fw_run */
  int64_t sssp_sweeps, aon_levels;             /* launches in the last fw_run  */
  int64_t sssp_items, aon_items;               /* (node, origin-group) visits  */
  int64_t link_passes;
  int64_t sssp_inlinks;                        /* in-links scanned by sweeps   */
  int64_t sssp_updates;                        /* labels lowered (lanes)       */
  int64_t sssp_outlinks;                       /* out-links marked             */
  int64_t aon_lanes;                           /* tree nodes loaded (lanes)    */
  /* with profiling on: summed device time of the sweep / level launches,
     measured with HIP events on the engine's stream around each launch */
  double sweep_kernel_ms, level_kernel_ms;
  int64_t sweep_launches, level_launches;
  /* one-origin-per-workgroup path: device time of its launches, launches,
     and (origin, variant) workgroups run */
  double origin_kernel_ms;
  int64_t origin_launches, origin_blocks;
  int64_t origin_expansions;   /* node expansions (label-correcting)        */
  int64_t origin_scanned;      /* out-links relaxed                         */
  int64_t origin_updates;      /* labels lowered (labels-in-HBM and lanes   */
                               /* kernels: atomic mins issued, one per      */
                               /* candidate below the label read)           */
  int64_t origin_relax_rounds;  /* relaxation rounds, summed over origins    */
  int64_t origin_load_rounds;   /* longest loading walk, summed over origins */
  int64_t origin_loaded;        /* tree links given flow                     */
  int64_t origin_walk_steps;    /* destination-to-origin walk steps          */
  double origin_relax_ms;       /* workgroup time per phase, summed over     */
  double origin_pred_ms;        /*   workgroups (device wall clock)          */
  double origin_load_ms;
  double origin_tune_ms;        /* fastest launch of the schedule tuning     */
  int64_t origin_chase;         /* chained expansions per thread and round   */
  double origin_delta;          /* bucket width of the labels-in-HBM kernel, */
                                /*   in mean free-flow link times (0: none)  */
  int64_t origin_split;         /* origins per variant co-run by the HBM     */
                                /*   kernel beside the LDS kernel            */
  int64_t origin_deferred;      /* queued nodes kept for a later bucket      */
  int64_t origin_block;         /* LDS path: threads per workgroup           */
  int64_t origin_queue;         /* frontier queue entries of the kernel used */
  /* lanes kernel (batched loading, one variant per wavefront lane) */
  int64_t lanes_launches;       /* launches                                  */
  int64_t lanes_tasks;          /* (origin, 64-variant group) workgroups     */
  int64_t lanes_pred_scans;     /* predecessor selections (lanes)            */
  int64_t lanes_fallbacks;      /* loadings redone by a per-origin kernel    */
  double lanes_delta;           /* its bucket width in mean free-flow times  */
  int64_t graph_launches;       /* iterations replayed from captured graphs  */
} tapgpu_stats;

/* Counters and wall times of the last tapgpu_fw_run (or of the calls since
   the last fw_run for sssp/aon). */
int tapgpu_get_stats(tapgpu_ctx *ctx, tapgpu_stats *stats);

/* on != 0: time every sweep / level launch with HIP events (small overhead). */
int tapgpu_set_profiling(tapgpu_ctx *ctx, int on);

/* Shortest-path / loading strategy: 0 = automatic (default: 2 where it
   fits, else 3), 1 = node-major multi-origin sweeps (any size), 2 = one
   origin per workgroup with labels in LDS (networks whose labels fit a
   workgroup's LDS: about 13k nodes, fewer than 65535 links), 3 = one origin
   per workgroup with labels in HBM (frontier bitmaps in LDS: up to about
   430k nodes).  tapgpu_get_strategy returns the one in use (1, 2 or 3). */
int tapgpu_set_strategy(tapgpu_ctx *ctx, int strategy);

/* Worklist schedule of the one-origin-per-workgroup path (strategy 2),
   rounds that scan the frontier bitmap (see tapgpu_set_origin_queue): how
   many labels a thread lowers and then expands itself, in a chain, within
   one round (0: every lowered node waits for the next round).  -1 (the
   default): tuned on the next loading by timing each candidate
   (0, 2, 4, 8, 16, 32) on the real launch.  Results do not depend on it. */
int tapgpu_set_origin_chase(tapgpu_ctx *ctx, int chase);
int tapgpu_get_origin_chase(const tapgpu_ctx *ctx);

/* Frontier queues of the one-origin-per-workgroup paths (strategies 2 and
   3) and of the lanes kernel: a round whose frontier (the nodes lowered or
   kept in the round before) has at most this many nodes takes them from an
   LDS list; a larger frontier is found by scanning its bitmap.  -1 (the
   default): the most each kernel's spare LDS holds (up to 4096), except on
   the LDS path, where the next loading picks it with the workgroup size
   (the most, 1024 or 256: a shorter queue takes less LDS, so more
   workgroups share a CU); 0: bitmap scans only; larger values are clamped
   to that most.  tapgpu_get_origin_queue returns the value of the kernel in
   use.  Results do not depend on it. */
int tapgpu_set_origin_queue(tapgpu_ctx *ctx, int entries);
int tapgpu_get_origin_queue(const tapgpu_ctx *ctx);

/* Workgroup size of the one-origin-per-workgroup LDS path (strategy 2):
   1024, 512 or 256 threads (a size holds networks of up to 16 nodes per
   thread).  -1 (the default): tuned on the next loading, with the queue
   size, by timing every allowed pair on the real launch; get returns -1
   until then.  Results do not depend on it. */
int tapgpu_set_origin_block(tapgpu_ctx *ctx, int threads);
int tapgpu_get_origin_block(const tapgpu_ctx *ctx);

/* Worklist schedule of the labels-in-HBM path (strategy 3): bucket width of
   its label-ordered rounds, as a multiple of the mean free-flow link time
   (0: no buckets, every queued node is expanded each round).  -1 (the
   default): tuned on the first loading: each candidate (0, 1, 2, 4, 8,
   16) runs one wave of workgroups, and the smallest busy time summed over
   the workgroups wins.  Results do not depend on it. */
int tapgpu_set_origin_delta(tapgpu_ctx *ctx, double mult);
double tapgpu_get_origin_delta(const tapgpu_ctx *ctx);

/* Co-scheduling on the LDS path (strategy 2): the last `origins` origins of
   every variant run in the labels-in-HBM kernel on a second queue, beside
   the LDS kernel on the same CUs (whose workgroups leave half of each CU's
   wave slots idle).  0 (the default): off; -1: tuned on the next loading
   over 0, 10, 15, 20, 25, 33 percent of the zones.  Results do not depend
   on it.  (On metro the best split gains about 2.5%.) */
int tapgpu_set_origin_split(tapgpu_ctx *ctx, int origins);
int tapgpu_get_origin_split(const tapgpu_ctx *ctx);
int tapgpu_get_strategy(const tapgpu_ctx *ctx);

/* Batched loading with the variants of one origin in the lanes of a
   wavefront (one workgroup per (origin, group of up to 64 variants), labels
   node-major in HBM, predecessors taken along the demand walks).  A loading
   of a batch uses it when at least min_active of its variants are active;
   0: never; -1 (the default): 16 for batched contexts, never for plain ones.
   A loading that meets a node whose tight in-links all come from tails with
   its own label (zero-cost links) is redone by the per-origin strategy.
   Results do not depend on it (same trees, same volumes bit for bit). */
int tapgpu_set_lanes(tapgpu_ctx *ctx, int min_active);
int tapgpu_get_lanes(const tapgpu_ctx *ctx);

/* Worklist schedule of the lanes kernel: bucket width of its label-ordered
   rounds (on the smallest label over a node's lanes), as a multiple of the
   mean free-flow link time (0: no buckets).  -1 (the default): tuned on the
   first loading: each candidate (0, 1, 2, 4, 8) runs one wave of
   workgroups, and the smallest busy time summed over them wins.  Results
   do not depend on it. */
int tapgpu_set_lanes_delta(tapgpu_ctx *ctx, double mult);
double tapgpu_get_lanes_delta(const tapgpu_ctx *ctx);

/* Launch form of tapgpu_fw_run's iterations (no counterpart in the
   reference, whose steps are host loops).  mode = 1 (the default): once the
   loading's schedule is tuned, each iteration (opening kernel, loading,
   both line-search windows, step, readback) is a replay of a captured HIP
   graph -- on the per-origin LDS and HBM paths and on the lanes kernel;
   elsewhere the link passes are.  On the per-origin paths iterations
   shorter than 2 ms go 8 to a replay, so the launch and the
   synchronisation are paid once for several.  mode = 2..8: that many
   iterations per replay on those paths.  mode = 0: the same launches issued
   one by one.  Profiling (tapgpu_set_profiling) issues them one by one too,
   so the events around the loading are recorded.  Results do not depend on
   it. */
int tapgpu_set_graphs(tapgpu_ctx *ctx, int mode);
int tapgpu_get_graphs(tapgpu_ctx *ctx, int *mode);

/* Form of the link passes (the line search and the step; no counterpart in
   the reference).  0 (the default): networks of up to 4,096 links run all of
   an iteration's link work in one launch, one workgroup per variant (every
   line-search window included); larger ones run multi-block passes (two
   windows queued per iteration, further ones by a host loop).  1 / 2: always
   the first / the second form.  window: step sizes 2^-j a window decides on
   (1..16; 0: 16, the default) -- smaller windows exercise the later-window
   paths on ordinary steps.  The step chosen never depends on either (the
   first 2^-j with g <= 0); the sums only change order. */
int tapgpu_set_link_passes(tapgpu_ctx *ctx, int mode, int window);

/* --- greedy step-cost assignment ------------------------------------------ */

typedef struct {
  int32_t iterations;       /* assignment iterations done                  */
  int32_t stalled;          /* stopped because unsatisfied demand changed   */
                            /*   by less than 1e-8                          */
  double unsatisfied;       /* demand left unassigned                       */
  double total_cost;        /* sum x * c at the end                         */
} tapgpu_greedy_result;

/* The reference's greedy assignment program (trunk/tap/tap_greedy.c:84-150
   greedy_assign, :300-330 its loop) on variant v, from zero volumes and
   free-flow costs: per iteration one all-origin shortest-path launch, then
   the OD pairs in input order each add the flow that brings the first link
   of their path to the next cost step (trunk/tap/tap_functions.c:241-266),
   then BPR costs.  Stops when no demand is left, when the unsatisfied total
   changes by less than 1e-8, or after max_iters (<= 0: no limit).
   volumes / costs [num_links] (may be NULL) receive the result; the
   variant's costs are left at the final costs (tapgpu_fw_init resets the FW
   state).  unsat_trace / ms_trace [trace_cap]: unsatisfied demand and wall
   time of each iteration. */
int tapgpu_greedy(tapgpu_ctx *ctx, int64_t variant, int max_iters, double *volumes,
                  double *costs, double *unsat_trace, double *ms_trace, int trace_cap,
                  tapgpu_greedy_result *result);

/* --- equilibrium gradients ------------------------------------------------ */

typedef struct {
  int64_t cycles;           /* active cycles (origin, tight non-tree link)   */
  int64_t nonzeros;         /* links summed over all cycles                   */
  int32_t useful_rounds;    /* rounds of the usefulness fixpoint              */
  int32_t cg_iters;         /* conjugate-gradient iterations                  */
  int32_t cg_status;        /* 1 converged, 2 breakdown, 0 iteration limit    */
  int32_t reserved;
  double cg_relres;         /* preconditioned residual / initial residual    */
  double setup_ms, cg_ms;   /* wall time: shortest paths + cycle matrix, CG  */
} tapgpu_grad_info;

/* Gradients at the variant's current volumes (after tapgpu_fw_run, or
   tapgpu_fw_init with warm volumes), which should be a user equilibrium.
   The reference has no counterpart; the operator boundary it serves is the
   Python extension (trunk/python_extension/pape_module.c:301-473), whose
   successor carries a torch.autograd.Function over this call.
     dobj_dt0, dobj_dcap  d(Beckmann objective)/d(free-flow time, capacity)
                          by the envelope theorem;
     dvht_dcap, dvht_dt0  d(sum_a x_a t_a)/d(capacity, free-flow time) by
                          implicit differentiation over the active paths:
                          links with reduced cost <= eps * label of their head
                          in some origin's shortest-path tree, on the way to
                          a destination with demand; a conjugate-gradient
                          solve to relative residual cg_tol, at most max_cg
                          iterations.  Link curvatures dt/dx below 1e-8 of
                          the largest count as zero: detours of constant cost
                          keep their (undetermined) flow split.
   Any output pointer may be NULL; each has num_links entries (0 for links
   absent from the variant).  Uses the node-major label buffers (allocated on
   first use) and temporary device memory; the FW state is unchanged. */
int tapgpu_gradient(tapgpu_ctx *ctx, int64_t variant, double eps, double cg_tol, int max_cg,
                    double *dobj_dt0, double *dobj_dcap, double *dvht_dcap, double *dvht_dt0,
                    tapgpu_grad_info *info);

#ifdef __cplusplus
}
#endif
#endif /* TAPGPU_H */
