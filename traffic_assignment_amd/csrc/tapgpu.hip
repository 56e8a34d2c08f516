// tapgpu.hip -- gfx950 Frank-Wolfe user-equilibrium engine (C ABI in
// include/tapgpu.h).
//
// Replaces the reference's per-iteration work (trunk/tap/tap_frankwolfe_*.c,
// trunk/tap/pape_*.cu, trunk/tap/volume_update_*.cu) with HBM-resident state
// and four kernel families, all FP64 and none of them MFMA-shaped:
//
//  * all-origin shortest paths: node-major labels dist[v][col] with one
//    origin per wavefront lane (64 origins per "group"); pull relaxations
//    over an explicit worklist of (node, group) items; a node's out-neighbours
//    are queued for the next sweep when any lane improved, deduplicated by a
//    per-(node, group) stamp and compacted with a wavefront ballot/prefix.
//    Labels are the least fixed point of fl(dist[u] + w) over paths whose
//    inner nodes are not zones (trunk/tap/sssp_pape.c:355), so they are
//    bit-identical to d'Esopo-Pape's for any relaxation order.
//  * all-or-nothing loading: bottom-up accumulation on the predecessor trees
//    (children counted in a dense pass, then ready nodes push their subtree
//    demand to the parent level by level); per-group partial link volumes are
//    summed in a fixed order, so volumes are deterministic (and exact for
//    integer demand, where the reference's path walk is exact too).
//  * link passes: one fused pass evaluates g(0) = sum c*d, the Beckmann
//    objective, VHT and g(2^-j) for a window of j (the only step sizes the
//    reference's bisection ever tests, trunk/tap/tap_functions.c:469-507)
//    with compensated block sums; a decision kernel per variant picks lambda
//    and updates gap / relative gap; an update pass does x += lambda*d and
//    BPR.
//  * everything is per "variant": a batch of TAPs over one union topology
//    (uneven link sets via a presence mask) runs in the same launches.
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off (no FMA
// contraction: every + and * is one IEEE rounding, as on the reference host).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tapgpu.h"
#include "tapmath.h"

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;                 // 4 wavefronts
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kJ = 16;                      // step sizes 2^-j evaluated per link pass
constexpr int kQ = 3 + kJ;                  // g0, objective, vht, g(2^-j)...
constexpr int kMaxHalvings = 1001;          // lambda = 2^-1 ... 2^-1001
constexpr int kLinkBlocks = 120;            // partial-sum blocks per variant

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ unsigned long long ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// ---------------------------------------------------------------------------
// kernel arguments
// ---------------------------------------------------------------------------

struct Graph {
  int N, E, ft0;                  // ft0 = first_thru_node - 1 (0-based)
  const int* fw_off;              // [N+1]
  const int* fw_dst;              // [E]  head of link k (CSR edge k == link k)
  const int* rv_off;              // [N+1] in-links of v
  const int* rv_src;              // [E]  tail, reverse order
  const int* rv_eid;              // [E]  link index, reverse order
  const int* tail;                // [E]  tail of link k
};

struct Labels {
  double* dist;                   // [N * P]   (flow during loading)
  int* pred;                      // [N * P]
  int* pend;                      // [N * P]   children left / -1 done
  long long P;                    // row stride (columns allocated)
  int Galloc;                     // group stride of stamp / items
  const int* col_src;             // [P] origin node of each column, -1 = padding
  const int* grp_var;             // [Galloc] variant of each group
  const unsigned char* var_on;    // [V] or null (all on)
  const double* rvw;              // [V * E] link costs in reverse-CSR order
};

__device__ __forceinline__ bool group_on(const Labels& L, int g) {
  return L.var_on == nullptr || L.var_on[L.grp_var[g]] != 0;
}

// Frontier bookkeeping: stamp[v * Galloc + g] holds the tag of the round in
// which item (v, g) must be processed; flags[r] != 0 iff round r has work.
// Writers only store (no atomics); a round's kernel finds its items with a
// dense scan compacted per block in LDS.  An item whose stamp is advanced to
// the next round before the scan reads it is simply processed one round
// later, which both worklists tolerate (labels only decrease; loading lanes
// carry the round from which they may run).
struct Frontier {
  int* stamp;     // [N * Galloc]
  int* flags;     // [rounds]
  unsigned long long* counters;   // [4] items, in-links scanned, lanes improved, out-links
};

// per-block scan of slots [blockIdx.x*kBlock, +kBlock) of the (v, g) space
// of groups [g0, g0+Gn); returns the number of items compacted into s_items
__device__ __forceinline__ int scan_slots(const Frontier& F, int N, int Galloc, int g0, int Gn,
                                          int tag, unsigned* s_items, int* s_n) {
  const long long slot = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long total = (long long)N * Gn;
  bool act = false;
  unsigned item = 0;
  if (slot < total) {
    const int v = (int)(slot / Gn);
    const int g = g0 + (int)(slot - (long long)v * Gn);
    item = (unsigned)v * (unsigned)Galloc + (unsigned)g;
    act = F.stamp[item] == tag;
  }
  const unsigned long long m = ballot(act);
  const int lane = lane_id();
  int base = 0;
  if (lane == 0 && m) base = atomicAdd(s_n, (int)__popcll(m));
  base = __shfl(base, 0);
  if (act) s_items[base + (int)__popcll(m & ((1ull << lane) - 1ull))] = item;
  __syncthreads();
  return *s_n;
}

// Mark the out-neighbours of v for group g in the next round (wave-wide).
__device__ __forceinline__ void mark_out(const Graph& G, const Frontier& F, int v, int g, int Galloc,
                                         int tag_next) {
  const int beg = G.fw_off[v], end = G.fw_off[v + 1];
  for (int e = beg + lane_id(); e < end; e += kWave)
    F.stamp[(size_t)G.fw_dst[e] * Galloc + g] = tag_next;
}

__device__ __forceinline__ void block_flush(const Frontier& F, int* s_any, int* flag_next,
                                            unsigned long long c0, unsigned long long c1,
                                            unsigned long long c2, unsigned long long c3,
                                            unsigned long long* s_cnt) {
  // c0..c3 are per-wave totals held by lane 0
  if (lane_id() == 0) {
    atomicAdd(&s_cnt[0], c0); atomicAdd(&s_cnt[1], c1);
    atomicAdd(&s_cnt[2], c2); atomicAdd(&s_cnt[3], c3);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (*s_any) *flag_next = 1;
    if (s_cnt[0]) {
      atomicAdd(&F.counters[0], s_cnt[0]); atomicAdd(&F.counters[1], s_cnt[1]);
      atomicAdd(&F.counters[2], s_cnt[2]); atomicAdd(&F.counters[3], s_cnt[3]);
    }
  }
}

// ---------------------------------------------------------------------------
// shortest paths
// ---------------------------------------------------------------------------

// dist = DBL_MAX, pred = -1 over the rows of groups [g0, g0+G).
__global__ __launch_bounds__(kBlock) void k_fill_labels(Graph G, Labels L, int g0, int Gn) {
  const long long waves = (long long)gridDim.x * kWavesPerBlock;
  const long long items = (long long)G.N * Gn;
  const int lane = lane_id();
  for (long long it = (long long)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; it < items;
       it += waves) {
    const int v = (int)(it / Gn);
    const int g = g0 + (int)(it - (long long)v * Gn);
    if (!group_on(L, g)) continue;
    const size_t idx = (size_t)v * L.P + (size_t)g * kWave + lane;
    L.dist[idx] = DBL_MAX;
    L.pred[idx] = -1;
  }
}

// Origins get label 0; their out-neighbours form round 0.
__global__ __launch_bounds__(kBlock) void k_sssp_seed(Graph G, Labels L, Frontier F, int g0,
                                                       int Gn, int tag0) {
  const int wave = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (wave >= Gn) return;
  const int g = g0 + wave;
  if (!group_on(L, g)) return;
  const int col = g * kWave + lane_id();
  const int src = L.col_src[col];
  bool any = false;
  if (src >= 0) {
    L.dist[(size_t)src * L.P + col] = 0.0;
    for (int e = G.fw_off[src]; e < G.fw_off[src + 1]; ++e) {
      F.stamp[(size_t)G.fw_dst[e] * L.Galloc + g] = tag0;
      any = true;
    }
  }
  if (ballot(any) && lane_id() == 0) F.flags[0] = 1;
}

// Round r: every marked (v, g) pulls min over in-links of dist[u] + w for its
// 64 origins (zones other than the lane's origin are not expanded).
__global__ __launch_bounds__(kBlock) void k_sssp_sweep(Graph G, Labels L, Frontier F, int g0,
                                                        int Gn, int r, int tag, int tag_next) {
  __shared__ unsigned s_items[kBlock];
  __shared__ int s_n, s_any;
  __shared__ unsigned long long s_cnt[4];
  if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) { s_n = 0; s_any = 0; }
  __syncthreads();
  const int n = scan_slots(F, G.N, L.Galloc, g0, Gn, tag, s_items, &s_n);
  const int lane = lane_id();
  unsigned long long c_items = 0, c_scan = 0, c_upd = 0, c_out = 0;
  for (int it = threadIdx.x / kWave; it < n; it += kWavesPerBlock) {
    const unsigned item = s_items[it];
    const int v = uniform((int)(item / (unsigned)L.Galloc));
    const int g = uniform((int)(item - (unsigned)v * (unsigned)L.Galloc));
    const int col = g * kWave + lane;
    const int src = L.col_src[col];
    const double* __restrict__ w = L.rvw + (size_t)L.grp_var[g] * G.E;
    const size_t me = (size_t)v * L.P + col;
    const double cur = L.dist[me];
    double best = cur, arg_du = INFINITY;
    int arg = -1;
    const int kb = G.rv_off[v], ke = G.rv_off[v + 1];
    for (int k = kb; k < ke; ++k) {
      const int u = G.rv_src[k];
      if (u == v || (u < G.ft0 && u != src)) continue;   // zone: not a through node
      const double du = L.dist[(size_t)u * L.P + col];
      if (du < DBL_MAX) {
        const double cand = du + w[k];
        // equal-cost candidates: the tail with the smaller label wins (the
        // one a label-ordered search expands first)
        if (cand < best || (cand == best && du < arg_du)) {
          best = cand;
          arg = G.rv_eid[k];
          arg_du = du;
        }
      }
    }
    const bool improved = best < cur;
    if (improved) {
      L.dist[me] = best;
      L.pred[me] = arg;
    } else if (arg >= 0 && arg_du < cur && L.pred[me] != arg) {
      L.pred[me] = arg;     // same label, tail with a strictly smaller label
    }
    const unsigned long long m = ballot(improved);
    c_items += 1;
    c_scan += (unsigned long long)(ke - kb);
    c_upd += (unsigned long long)__popcll(m);
    if (ballot(improved && (v >= G.ft0 || v == src))) {
      mark_out(G, F, v, g, L.Galloc, tag_next);
      c_out += (unsigned long long)(G.fw_off[v + 1] - G.fw_off[v]);
      if (lane == 0) s_any = 1;
    }
  }
  block_flush(F, &s_any, &F.flags[r + 1], c_items, c_scan, c_upd, c_out, s_cnt);
}

// ---------------------------------------------------------------------------
// all-or-nothing loading on the predecessor trees
// ---------------------------------------------------------------------------

struct Loading {
  const double* odc;     // [Z * Pv] demand origin(col) -> zone row
  long long Pv;          // columns per variant
  int Z;
  double* ypart;         // [Galloc * E] per-group link volumes
  int* err;              // [3] flag, origin, dest (0-based)
};

// Lane state in pend[]: >= 0 children still to report; -1 done or not in the
// tree; -(L+2) ready from round L on (its subtree flow is complete once all
// launches before round L have finished).

// Dense pass: children count, initial node flow (= demand to this zone),
// no-path check, leaves marked for round 0.
__global__ __launch_bounds__(kBlock) void k_aon_init(Graph G, Labels L, Loading A, Frontier F,
                                                      int g0, int Gn, int tag0) {
  const long long waves = (long long)gridDim.x * kWavesPerBlock;
  const long long items = (long long)G.N * Gn;
  const int lane = lane_id();
  bool any = false;
  for (long long it = (long long)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; it < items;
       it += waves) {
    const int v = uniform((int)(it / Gn));
    const int g = uniform(g0 + (int)(it - (long long)v * Gn));
    if (!group_on(L, g)) continue;
    const int col = g * kWave + lane;
    const int src = L.col_src[col];
    const size_t me = (size_t)v * L.P + col;
    const int p = L.pred[me];
    int kids = 0;
    for (int e = G.fw_off[v]; e < G.fw_off[v + 1]; ++e) {
      const int x = G.fw_dst[e];
      kids += (L.pred[(size_t)x * L.P + col] == e) ? 1 : 0;
    }
    double f = 0.0;
    if (v < A.Z && src >= 0 && v != src) {
      const double d = A.odc[(size_t)v * A.Pv + (col % A.Pv)];
      if (d > 0.0) {
        f = d;
        if (p < 0 && atomicCAS(&A.err[0], 0, 1) == 0) {
          A.err[1] = src;
          A.err[2] = v;
        }
      }
    }
    L.dist[me] = f;
    L.pend[me] = (p < 0) ? -1 : (kids > 0 ? kids : -2);
    if (ballot(p >= 0 && kids == 0)) {
      if (lane == 0) F.stamp[(size_t)v * L.Galloc + g] = tag0;
      any = true;
    }
  }
  if (ballot(any) && lane == 0) F.flags[0] = 1;
}

// Round r: ready lanes of each marked (v, g) add their subtree demand to the
// in-link on their path, pass it to the tail node and release it.
__global__ __launch_bounds__(kBlock) void k_aon_level(Graph G, Labels L, Loading A, Frontier F,
                                                       int g0, int Gn, int r, int tag,
                                                       int tag_next) {
  __shared__ unsigned s_items[kBlock];
  __shared__ int s_n, s_any;
  __shared__ unsigned long long s_cnt[4];
  if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) { s_n = 0; s_any = 0; }
  __syncthreads();
  const int n = scan_slots(F, G.N, L.Galloc, g0, Gn, tag, s_items, &s_n);
  const int lane = lane_id();
  unsigned long long c_items = 0, c_lanes = 0;
  for (int it = threadIdx.x / kWave; it < n; it += kWavesPerBlock) {
    const unsigned item = s_items[it];
    const int v = uniform((int)(item / (unsigned)L.Galloc));
    const int g = uniform((int)(item - (unsigned)v * (unsigned)L.Galloc));
    const int col = g * kWave + lane;
    const int src = L.col_src[col];
    const size_t me = (size_t)v * L.P + col;
    const int pe = L.pend[me];
    const int p = L.pred[me];
    // lanes released for this round or an earlier one (a lane released in
    // this round carries -(r+3) and waits for the next launch)
    const bool ready = (pe <= -2) && (-(pe + 2) <= r) && (p >= 0);
    const double f = ready ? L.dist[me] : 0.0;
    bool release = false;   // this lane's parent became ready
    if (ready) {
      const int parent = G.tail[p];
      L.pend[me] = -1;
      if (parent != src) {
        const size_t pi = (size_t)parent * L.P + col;
        if (f != 0.0) unsafeAtomicAdd(&L.dist[pi], f);
        if (atomicSub(&L.pend[pi], 1) == 1) {
          L.pend[pi] = -(r + 3);
          release = true;
        }
      }
    }
    c_items += 1;
    c_lanes += (unsigned long long)__popcll(ballot(ready));
    // per in-link: wave sum of the flow routed through it (fixed order)
    double* __restrict__ yp = A.ypart + (size_t)g * G.E;
    for (int k = G.rv_off[v]; k < G.rv_off[v + 1]; ++k) {
      const int eid = G.rv_eid[k];
      const bool mine = ready && p == eid;
      if (!ballot(mine)) continue;
      double s = mine ? f : 0.0;
      for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
      const bool rel = ballot(mine && release) != 0ull;
      if (lane == 0) {
        if (s != 0.0) yp[eid] += s;
        if (rel) {
          F.stamp[(size_t)G.rv_src[k] * L.Galloc + g] = tag_next;
          s_any = 1;
        }
      }
    }
  }
  block_flush(F, &s_any, &F.flags[r + 1], c_items, c_lanes, 0, 0, s_cnt);
}

// After the level rounds have drained every lane of a tree is done (-1).  A
// lane still waiting for children (>= 0) or released but never run (<= -2)
// lies on a predecessor cycle or hangs from one: flag it (err code 4).
__global__ __launch_bounds__(kBlock) void k_aon_check(Graph G, Labels L, Loading A, int g0, int Gn) {
  const long long n = (long long)G.N * Gn * kWave;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const long long item = i / kWave;
    const int v = (int)(item / Gn);
    const int g = g0 + (int)(item - (long long)v * Gn);
    if (!group_on(L, g)) continue;
    const int col = g * kWave + (int)(i % kWave);
    if (L.col_src[col] < 0) continue;
    if (L.pend[(size_t)v * L.P + col] != -1 && atomicCAS(&A.err[0], 0, 4) == 0) {
      A.err[1] = L.col_src[col];
      A.err[2] = v;
    }
  }
}

// y[var][e] = sum over the variant's groups of ypart (fixed order); clears
// ypart for the next loading.
__global__ __launch_bounds__(kBlock) void k_aon_reduce(int E, int Gv, int v0, int Vn,
                                                        const unsigned char* var_on,
                                                        double* ypart, double* y) {
  const long long n = (long long)E * Vn;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int var = v0 + (int)(i / E);
    const int e = (int)(i - (long long)(var - v0) * E);
    if (var_on && !var_on[var]) continue;
    double s = 0.0;
    for (int j = 0; j < Gv; ++j) {
      double* p = ypart + ((size_t)var * Gv + j) * E + e;
      s += *p;
      *p = 0.0;
    }
    y[(size_t)var * E + e] = s;
  }
}

// ---------------------------------------------------------------------------
// one origin per workgroup, labels in LDS (networks up to ~13k nodes)
// ---------------------------------------------------------------------------
//
// For a metro-sized network one origin's labels (8 B/node) fit in a CU's
// 160 KB LDS, so a 1024-thread workgroup runs the whole origin: rounds of
// label-correcting relaxation (64-bit LDS atomicMin on the label bits --
// non-negative doubles order like their bit patterns; the nodes lowered in a
// round form the next round's frontier, deduplicated by a bitmap and listed
// in an LDS queue so that each frontier node gets a thread of its own), then
// predecessor selection, then loading: each destination walks its tree path
// adding its demand to LDS node flows (FP64 LDS atomics), and every tree
// link with flow adds it to the HBM link volumes (FP64 atomics).  One launch
// covers every (variant, origin).
//
// Predecessor rule: among in-links u->v with fl(dist[u] + w) == dist[v]
// (u expandable) and dist[u] < dist[v], take the tail with the smallest
// label (the one a label-ordered search such as d'Esopo-Pape-LLL normally
// expands first), then the smallest tail id, then the first such link in
// link order -- the choice of the node-major sweeps too, and independent of
// thread timing.  A node whose tight tails all carry its own label
// (zero-cost links, or costs below half an ulp of the label) has none; such
// nodes join the tree level by level (attach_level_ties): each pass gives
// every one of them the smallest-id tight tail that was in the tree before
// the pass.  Along the chosen links the label decreases, or stays equal
// while the level decreases, so the tree is acyclic.  Where the shortest
// path is unique this is the reference's predecessor.

// Workgroup sizes of the LDS kernel; a size serves networks of up to
// kLdsPerThread nodes per thread (the per-thread register arrays of the
// predecessor phase).  1024 by default; the schedule tuner may pick a
// smaller one where it fits.
constexpr int kLdsBlock = 1024;
constexpr int kLdsBlocks[] = {1024, 512, 256};
constexpr int kDefaultChase = 4;               // before tuning / shortest-path-only calls
constexpr int kLdsPerThread = 16;
constexpr int kLdsMaxNodes = kLdsPerThread * kLdsBlock;
#ifndef TAPGPU_LDS_QUEUE_MAX
#define TAPGPU_LDS_QUEUE_MAX 4096
#endif
constexpr int kLdsQueueMax = TAPGPU_LDS_QUEUE_MAX;   // frontier queue entries per round
#ifndef TAPGPU_LDS_GROUP
#define TAPGPU_LDS_GROUP 6
#endif
constexpr int kGroup = TAPGPU_LDS_GROUP;             // lanes per expanded node (one link each)
constexpr unsigned short kNoPred = 0xFFFF;

struct OriginArgs {
  int N, E, ft0, Z;
  const int *fw_off, *fw_dst, *rv_off, *rv_src, *rv_eid, *tail;
  const double* fww;                 // [V * E] forward costs (INF: absent)
  const double* rvw;                 // [V * E] reverse-order costs
  const long long* od_start;         // [Z + 1]
  const int* od_dest;                // 0-based zone
  const double* od_val;
  const unsigned char* var_on;       // [V] or null
  int chase;                         // expansions a thread chains per round
  int qcap;                          // LDS frontier queue entries per round (0: bitmap scan only)
  int hqcap;                         // labels-in-HBM kernel: queue entries per round (0: bitmap scan only)
  double delta;                      // labels-in-HBM kernel: bucket width (<= 0: none)
  int z0, zn;                        // labels-in-HBM kernel: origins [z0, z0 + zn) of each variant
  double* y;                         // [V * E], zeroed by the caller
  int* err;                          // [3]
  unsigned long long* counters;      // [16] expansions, links scanned, updates,
                                     //     relax rounds, longest loading walk,
                                     //     tree links loaded, walk steps, -,
                                     //     wall-clock ticks of relaxation,
                                     //     predecessors, loading
  // block x handles origin src_list[x] (shortest-path-only mode: start
  // nodes; LDS loadings: origins longest first)
  const int* src_list;
  int* work;                         // [Z] relaxation rounds per origin, recorded when non-null
  // labels-in-HBM loadings: items taken from a counter (null: block-strided),
  // origin-major in `order` (origins by decreasing rounds) when non-null
  int* next;
  const int* order;
  int vn;                            // variants of the launch (with order)
  double* out_dist;                  // [(N) * ld] node-major, row v = node v+1
  int* out_pred;
  long long ld;
  // labels-in-HBM kernel: per-block scratch rows of N entries
  unsigned long long* h_dist;        // labels, then flows
  int* h_pred;                       // tree link
  int* h_ptail;                      // its tail
};

__device__ __forceinline__ double as_d(unsigned long long b) { return __longlong_as_double((long long)b); }

// wavefront-wide sum / max (every lane of the wavefront active), so that
// per-thread counters reach a workgroup total with one LDS atomic per wave
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
  for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ __forceinline__ int wave_max_i32(int x) {
  for (int o = kWave / 2; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
  return x;
}
__device__ __forceinline__ unsigned long long as_b(double d) { return (unsigned long long)__double_as_longlong(d); }

// LDS layout (bytes): labels 8N (flows while loading) | three bitmaps
// 3*8W | offs/pred 2A | ptail 2A | two frontier queues 2*2Q, with
// W = ceil(N/64), A = N+1 rounded up to 4, Q = queue entries.
size_t origin_lds_bytes(int N, int qcap) {
  const size_t W = (size_t)(N + 63) / 64, A = (size_t)((N + 1 + 3) & ~3);
  return 8ull * N + 24ull * W + 4ull * A + 4ull * (size_t)qcap;
}

// Second step of the predecessor rule: the nodes flagged in `pending` have
// no tight tail with a smaller label.  Each pass gives every pending node
// the smallest-id tight tail with its own label that was in the tree before
// the pass (not pending), and clears the flags after a barrier, so a pass's
// result does not depend on thread timing.  Every pending node hangs from a
// tree node through such links (its label came along them), so each pass
// attaches at least one node; false if a pass attaches none.  All threads
// of the workgroup call it; label(v) reads a label, set_pred(v, link, tail)
// records a choice.
template <int kThreads, class Label, class SetPred>
__device__ bool attach_level_ties(const OriginArgs& A, const double* __restrict__ rw, int src,
                                  unsigned long long* pending, unsigned long long* added,
                                  int* s_left, int* s_added, Label label, SetPred set_pred) {
  const int N = A.N, W = (N + 63) >> 6, tid = threadIdx.x;
  for (int w = tid; w < W; w += kThreads) added[w] = 0;
  for (;;) {
    if (tid == 0) { *s_left = 0; *s_added = 0; }
    __syncthreads();
    for (int v = tid; v < N; v += kThreads) {
      if (!((pending[v >> 6] >> (v & 63)) & 1ull)) continue;
      const unsigned long long dv = label(v);
      int bu = -1, be = -1;
      for (int k = A.rv_off[v]; k < A.rv_off[v + 1]; ++k) {
        const int u = A.rv_src[k];
        if (u == v || (u < A.ft0 && u != src)) continue;
        if ((pending[u >> 6] >> (u & 63)) & 1ull) continue;       // not in the tree yet
        const unsigned long long du = label(u);
        if (du != dv || as_b(as_d(du) + rw[k]) != dv) continue;
        if (bu < 0 || u < bu) { bu = u; be = A.rv_eid[k]; }
      }
      if (be >= 0) {
        set_pred(v, be, bu);
        atomicOr(&added[v >> 6], 1ull << (v & 63));
        atomicAdd(s_added, 1);
      } else {
        atomicAdd(s_left, 1);
      }
    }
    __syncthreads();
    const int left = *s_left, got = *s_added;
    for (int w = tid; w < W; w += kThreads) { pending[w] &= ~added[w]; added[w] = 0; }
    if (left == 0 || got == 0) return left == 0;
    __syncthreads();
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_origin_lds(OriginArgs A, int load) {
  extern __shared__ unsigned long long sm[];
  const int N = A.N;
  const int W = (N + 63) >> 6;
  const int Al = (N + 1 + 3) & ~3;
  unsigned long long* dist = sm;                                   // [N] (flow later)
  unsigned long long* bmA = dist + N;                              // 3 x [W]
  unsigned long long* bmB = bmA + W;
  unsigned long long* bmC = bmB + W;
  unsigned short* offs = (unsigned short*)(bmC + W);               // [N+1], then pred [N]
  unsigned short* pred = offs;
  unsigned short* ptail = offs + Al;                               // [N]
  unsigned short* queue0 = ptail + Al;                             // 2 x [qcap]
  unsigned short* queue1 = queue0 + A.qcap;
  auto qw_of = [&](int rr) { return (rr & 1) ? queue1 : queue0; };   // queue written in round rr
  __shared__ int s_qn[3];
  __shared__ unsigned long long s_cnt[8];

  const int var = blockIdx.y;
  if (A.var_on && !A.var_on[var]) return;
  const int src = A.src_list ? A.src_list[blockIdx.x] : (int)blockIdx.x;
  const int tid = threadIdx.x;
  const double* __restrict__ fw = A.fww + (size_t)var * A.E;
  const double* __restrict__ rw = A.rvw + (size_t)var * A.E;
  const unsigned long long kMax = as_b(DBL_MAX);

  for (int v = tid; v < N; v += BLOCK) dist[v] = kMax;
  for (int v = tid; v <= N; v += BLOCK) offs[v] = (unsigned short)A.fw_off[v];
  for (int w = tid; w < 3 * W; w += BLOCK) bmA[w] = 0;
  if (tid < 3) s_qn[tid] = tid == 0 ? 1 : 0;
  if (tid < 8) s_cnt[tid] = 0;
  __syncthreads();
  if (tid == 0) {
    dist[src] = 0ull;                     // +0.0
    bmA[src >> 6] = 1ull << (src & 63);
    if (A.qcap > 0) queue0[0] = (unsigned short)src;
  }
  __syncthreads();

  unsigned long long n_exp = 0, n_scan = 0, n_upd = 0;
  const long long clk0 = wall_clock64();
  // ---- label-correcting rounds, one barrier each.  Round r expands the
  //      nodes lowered in round r-1.  Normally they are listed in a queue
  //      (count s_qn[(r-1)%3], buffer (r-1)&1) and each gets a group of
  //      kGroup lanes, one out-link per lane: the group's loads are
  //      consecutive, and a round costs one link's chain of latencies.
  //      When more nodes were lowered than the queue holds, the round scans
  //      their bitmap `cur` in 16-bit slots instead, a lane expanding each
  //      set bit's node (and chaining up to A.chase further expansions of
  //      nodes it lowers itself).  A node lowered in round r is listed
  //      once: the lane whose atomicOr sets its bit in `nxt` appends it to
  //      queue r&1, one LDS atomicAdd per wave for all of the wave's
  //      appends.  Three rotating bitmaps: the third one is cleared while
  //      the first is read.  Labels are the same least fixed point for any
  //      such order. ----
  unsigned long long *cur = bmA, *nxt = bmB, *clr = bmC;
  const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  int r = 1;
  // relaxes u -> v at cost w and, if this lane lowered v and `want`, lists
  // it unless listed already; returns whether this lane lowered v
  auto relax = [&](double du, int v, double w, bool want) {
    const unsigned long long cb = as_b(du + w);
    // no pre-read of dist[v]: one LDS round trip, and measured faster
    const bool low = cb < atomicMin(&dist[v], cb);
    bool first = false;
    if (low) {
      ++n_upd;
      if (want && v >= A.ft0) {
        const unsigned long long bit = 1ull << (v & 63);
        first = !(atomicOr(&nxt[v >> 6], bit) & bit);
      }
    }
    const unsigned long long bal = __ballot(first);
    if (bal) {
      const int leader = __ffsll((long long)bal) - 1;
      int b0 = 0;
      if (lane == leader) b0 = atomicAdd(&s_qn[r % 3], __popcll(bal));
      b0 = __builtin_amdgcn_readlane(b0, leader);
      if (first) {
        const int pos = b0 + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
        if (pos < A.qcap) qw_of(r)[pos] = (unsigned short)v;
      }
    }
    return low;
  };
  for (;; ++r) {
    const unsigned short* qr = (r & 1) ? queue0 : queue1;
    const int j0 = tid / kGroup;
    const int u0 = j0 < A.qcap ? qr[j0] : 0;      // read beside the count
    const int qn = s_qn[(r - 1) % 3];
    if (qn <= A.qcap) {
      for (int idx = tid; idx / kGroup < qn; idx += BLOCK) {
        const int u = idx == tid ? u0 : qr[idx / kGroup];
        const int g = idx % kGroup;
        const double du = as_d(dist[u]);
        const int eb = offs[u], ee = offs[u + 1];
        if (g == 0) { ++n_exp; n_scan += (unsigned long long)(ee - eb); }
        for (int e = eb + g; e < ee; e += kGroup) relax(du, A.fw_dst[e], fw[e], true);
      }
    } else {
      for (int slot = tid; slot < 4 * W; slot += BLOCK) {
        const int w = slot >> 2, q = slot & 3;
        unsigned bits = (unsigned)((cur[w] >> (16 * q)) & 0xFFFFull);
        while (bits) {
          const int b = __ffs(bits) - 1;
          bits &= bits - 1;
          int u = (w << 6) + 16 * q + b;
          for (int hops = 0;; ++hops) {
            const double du = as_d(dist[u]);
            const int eb = offs[u], ee = offs[u + 1];
            ++n_exp;
            n_scan += (unsigned long long)(ee - eb);
            int next = -1;
            for (int e = eb; e < ee; ++e) {
              const int v = A.fw_dst[e];
              // the first node this lane lowers is chained, not listed
              const bool chain = next < 0 && hops < A.chase;
              if (relax(du, v, fw[e], !chain) && chain && v >= A.ft0) next = v;
            }
            if (next < 0) break;
            u = next;
          }
        }
      }
    }
    for (int w = tid; w < W; w += BLOCK) clr[w] = 0;
    if (tid == 0) s_qn[(r + 1) % 3] = 0;
    __syncthreads();
    if (!s_qn[r % 3]) break;
    if (r > N + 8) {            // only possible with negative costs
      if (tid == 0 && atomicCAS(&A.err[0], 0, 2) == 0) { A.err[1] = src; A.err[2] = 0; }
      return;
    }
    unsigned long long* t = cur; cur = nxt; nxt = clr; clr = t;
  }
  const int relax_rounds = r;
  const long long clk1 = wall_clock64();
  if (A.work && tid == 0 && blockIdx.y == 0) A.work[src] = relax_rounds;

  // ---- predecessor links (label, then tail id; see above).  The in-link
  //      offsets are staged in LDS (the out-link offsets are dead); each
  //      node's tails, costs and link ids are loaded together, so a node
  //      costs one L2 round trip.  Results stay in registers until every
  //      thread is done reading offsets (pred shares their space). ----
  __shared__ int s_level_ties;
  for (int v = tid; v <= N; v += BLOCK) offs[v] = (unsigned short)A.rv_off[v];
  for (int w = tid; w < W; w += BLOCK) bmB[w] = 0;     // nodes left for attach_level_ties
  if (tid == 0) s_level_ties = 0;
  __syncthreads();
  unsigned short my_tail[kLdsPerThread], my_pred[kLdsPerThread];
  {
    int j = 0;
    for (int v = tid; v < N; v += BLOCK, ++j) {
      unsigned short p = kNoPred, pt = 0;
      const unsigned long long dv = dist[v];
      if (v != src && dv != kMax) {
        // starting from (dv, 0) only tails with a smaller label can win
        unsigned best = 0;
        unsigned long long best_du = dv;
        const int kb = offs[v], ke = offs[v + 1];
        for (int k0 = kb; k0 < ke; k0 += 8) {
          int hu[8], he[8];
          double hw[8];
#pragma unroll
          for (int i = 0; i < 8; ++i)
            if (k0 + i < ke) { hu[i] = A.rv_src[k0 + i]; hw[i] = rw[k0 + i]; he[i] = A.rv_eid[k0 + i]; }
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            if (k0 + i >= ke) break;
            const int u = hu[i];
            if (u == v || (u < A.ft0 && u != src)) continue;
            const unsigned long long du = dist[u];
            if (du == kMax) continue;
            if (as_b(as_d(du) + hw[i]) != dv) continue;
            const unsigned key = (unsigned)u;
            if (du < best_du || (du == best_du && key < best)) {
              best = key;
              best_du = du;
              p = (unsigned short)he[i];
              pt = (unsigned short)u;
            }
          }
        }
      }
      my_tail[j] = pt;
      my_pred[j] = p;
      // reached, not the origin, and no tight tail with a smaller label:
      // every tight tail has this label
      const bool level_tie = p == kNoPred && v != src && dv != kMax;
      if (level_tie) {
        atomicOr(&bmB[v >> 6], 1ull << (v & 63));
        s_level_ties = 1;
      }
      if (A.out_dist) {
        const size_t o = (size_t)v * A.ld + blockIdx.x;
        A.out_dist[o] = as_d(dv);
        if (!level_tie) A.out_pred[o] = (p == kNoPred) ? -1 : (int)p;
      }
    }
  }
  __syncthreads();        // every offs[] read is done
  const long long clk2 = wall_clock64();
  {
    int j = 0;
    for (int v = tid; v < N; v += BLOCK, ++j) { pred[v] = my_pred[j]; ptail[v] = my_tail[j]; }
  }
  if (s_level_ties) {
    __shared__ int s_left, s_added;
    const bool ok = attach_level_ties<BLOCK>(
        A, rw, src, bmB, bmC, &s_left, &s_added, [&](int v) { return dist[v]; },
        [&](int v, int e, int u) {
          pred[v] = (unsigned short)e;
          ptail[v] = (unsigned short)u;
          if (A.out_dist) A.out_pred[(size_t)v * A.ld + blockIdx.x] = e;
        });
    if (!ok && tid == 0 && atomicCAS(&A.err[0], 0, 3) == 0) { A.err[1] = src; A.err[2] = 0; }
  }
  if (!load) {
    if (A.counters) {
      n_exp = wave_sum_u64(n_exp); n_scan = wave_sum_u64(n_scan); n_upd = wave_sum_u64(n_upd);
      if (lane == 0) {
        atomicAdd(&s_cnt[0], n_exp); atomicAdd(&s_cnt[1], n_scan); atomicAdd(&s_cnt[2], n_upd);
      }
      __syncthreads();
      if (tid == 0) {
        atomicAdd(&A.counters[0], s_cnt[0]); atomicAdd(&A.counters[1], s_cnt[1]);
        atomicAdd(&A.counters[2], s_cnt[2]); atomicAdd(&A.counters[3], (unsigned long long)relax_rounds);
      }
    }
    return;
  }

  // ---- all-or-nothing loading: every destination with demand walks its
  //      tree path to the origin, adding its demand to each node's flow
  //      (the flow on the node's tree in-link); no rounds, no barriers.
  //      Integer demands make every partial sum exact, so the order of the
  //      LDS atomics does not matter. ----
  double* flow = (double*)dist;
  for (int v = tid; v < N; v += BLOCK) flow[v] = 0.0;
  __shared__ unsigned long long s_steps;
  __shared__ int s_maxlen, s_loaded;
  if (tid == 0) { s_steps = 0; s_maxlen = 0; s_loaded = 0; }
  __syncthreads();
  unsigned long long steps = 0;
  int maxlen = 0;
  for (long long i = A.od_start[src] + tid; i < A.od_start[src + 1]; i += BLOCK) {
    const int d = A.od_dest[i];
    const double f = A.od_val[i];
    if (d == src || !(f > 0.0)) continue;
    if (pred[d] == kNoPred) {
      if (atomicCAS(&A.err[0], 0, 1) == 0) { A.err[1] = src; A.err[2] = d; }
      continue;
    }
    int v = d, len = 0;
    while (v != src) {
      if (len > N) {        // a predecessor cycle
        if (atomicCAS(&A.err[0], 0, 3) == 0) { A.err[1] = src; A.err[2] = 0; }
        break;
      }
      atomicAdd(&flow[v], f);
      v = ptail[v];
      ++len;
    }
    steps += (unsigned long long)len;
    maxlen = max(maxlen, len);
  }
  __syncthreads();
  double* __restrict__ yv = A.y + (size_t)var * A.E;
  int loaded = 0;
  for (int v = tid; v < N; v += BLOCK) {
    const double f = flow[v];
    if (f != 0.0 && pred[v] != kNoPred) {
      unsafeAtomicAdd(&yv[pred[v]], f);
      ++loaded;
    }
  }
  if (A.counters) {
    n_exp = wave_sum_u64(n_exp); n_scan = wave_sum_u64(n_scan); n_upd = wave_sum_u64(n_upd);
    steps = wave_sum_u64(steps);
    maxlen = wave_max_i32(maxlen);
    const unsigned long long ld = wave_sum_u64((unsigned long long)loaded);
    if (lane == 0) {
      atomicAdd(&s_cnt[0], n_exp); atomicAdd(&s_cnt[1], n_scan); atomicAdd(&s_cnt[2], n_upd);
      atomicAdd(&s_steps, steps);
      atomicMax(&s_maxlen, maxlen);
      atomicAdd(&s_loaded, (int)ld);
    }
  }
  __syncthreads();
  if (tid == 0 && A.counters) {
    atomicAdd(&A.counters[0], s_cnt[0]); atomicAdd(&A.counters[1], s_cnt[1]);
    atomicAdd(&A.counters[2], s_cnt[2]); atomicAdd(&A.counters[3], (unsigned long long)relax_rounds);
    atomicAdd(&A.counters[4], (unsigned long long)s_maxlen);
    atomicAdd(&A.counters[5], (unsigned long long)s_loaded);
    atomicAdd(&A.counters[6], s_steps);
    const long long clk3 = wall_clock64();
    atomicAdd(&A.counters[8], (unsigned long long)(clk1 - clk0));
    atomicAdd(&A.counters[9], (unsigned long long)(clk2 - clk1));
    atomicAdd(&A.counters[10], (unsigned long long)(clk3 - clk2));
  }
}


// ---------------------------------------------------------------------------
// one origin per workgroup, labels in HBM (networks of any size)
// ---------------------------------------------------------------------------
//
// The same three phases as k_origin_lds for networks whose labels do not fit
// LDS: only the three frontier bitmaps (3 x 8 x ceil(N/64) bytes) live in
// LDS; labels, predecessors and flows are a per-workgroup scratch
// row in HBM (L2 / infinity cache resident while the workgroup runs).  A
// persistent grid loops over (origin, variant) items.  A row is private to
// its workgroup, whose waves share one CU, so every scratch access is a
// workgroup-scope atomic: loads skip the CU's L1 and read the XCD's L2,
// where the atomics execute, and the barrier closing a round orders them.
// (Device-scope accesses would have to be coherent across the XCDs' L2s and
// cost an L2 writeback and invalidation per round.)
// Relaxation within a round may read a label another thread is lowering;
// the lowered node is queued for the next round, so the fixed point is the
// same as with strict rounds.

#ifndef TAPGPU_HBM_BLOCK
#define TAPGPU_HBM_BLOCK 512
#endif
constexpr int kHbmBlock = TAPGPU_HBM_BLOCK;    // 512: 2 workgroups per CU, each with room for queues
#ifndef TAPGPU_HBM_PER_CU
#define TAPGPU_HBM_PER_CU (1024 / TAPGPU_HBM_BLOCK)
#endif
constexpr int kHbmPerCu = TAPGPU_HBM_PER_CU;    // workgroups sharing a CU's LDS (registers: <= 128 VGPRs)
constexpr int kHbmQueueMax = 4096;              // frontier queue entries per round

__global__ void k_zero_i32(int* p) { *p = 0; }

// LDS of the labels-in-HBM kernel: three bitmaps 3*8W | two queues of
// (label 8 B, node 4 B) entries 2*12Q.
size_t origin_hbm_lds_bytes(int N, int hqcap) {
  return 24ull * (size_t)((N + 63) / 64) + 24ull * (size_t)hqcap;
}

// workgroup-scope loads of a workgroup's private scratch row
__device__ __forceinline__ unsigned long long ldg_u64(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned ldg_u32(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int ldg_i32(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ __launch_bounds__(kHbmBlock, kHbmPerCu * kHbmBlock / kWave / 4) void k_origin_hbm(OriginArgs A, int load, int items) {
  extern __shared__ unsigned long long sm[];
  const int N = A.N;
  const int W = (N + 63) >> 6;
  unsigned long long* bmA = sm;
  unsigned long long* bmB = bmA + W;
  unsigned long long* bmC = bmB + W;
  unsigned long long* qlab0 = bmC + W;                  // 2 x [hqcap] labels when listed
  unsigned long long* qlab1 = qlab0 + A.hqcap;
  unsigned* qnode0 = (unsigned*)(qlab1 + A.hqcap);      // 2 x [hqcap] nodes
  unsigned* qnode1 = qnode0 + A.hqcap;
  const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  __shared__ int s_qn[3], s_exp[3];
  __shared__ unsigned long long s_minq[3];
  __shared__ unsigned long long s_cnt[8];
  __shared__ int s_maxlen;
  __shared__ int s_level_ties, s_left, s_added;
  const int tid = threadIdx.x;
  const unsigned long long kMax = as_b(DBL_MAX);
  unsigned long long* dist = A.h_dist + (size_t)blockIdx.x * N;
  int* pred = A.h_pred + (size_t)blockIdx.x * N;
  int* ptail = A.h_ptail + (size_t)blockIdx.x * N;
  double* flow = (double*)dist;
  if (tid < 8) s_cnt[tid] = 0;
  if (tid == 0) s_maxlen = 0;
  unsigned long long n_exp = 0, n_scan = 0, n_upd = 0, n_rounds = 0, n_steps = 0, n_loaded = 0;
  unsigned long long n_def = 0;
  unsigned long long c_relax = 0, c_pred = 0, c_load = 0;

  __shared__ int s_item;
  for (int k = 0;; ++k) {
    // the next item: from the launch's counter (the first free workgroup
    // takes it), or block-strided
    if (tid == 0) s_item = A.next ? atomicAdd(A.next, 1) : (int)blockIdx.x + k * (int)gridDim.x;
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= items) break;
    int var, src;
    if (A.src_list) {
      var = 0;
      src = A.src_list[item];
    } else if (A.order) {                     // longest origins first, all variants of each
      const int rank = item / A.vn;
      var = item - rank * A.vn;
      src = A.order[rank];
    } else {
      var = item / A.zn;
      src = A.z0 + item - var * A.zn;
    }
    if (A.var_on && !A.var_on[var]) continue;
    const double* __restrict__ fw = A.fww + (size_t)var * A.E;
    const double* __restrict__ rw = A.rvw + (size_t)var * A.E;
    const long long clk0 = wall_clock64();
    for (int v = tid; v < N; v += kHbmBlock) {
      __hip_atomic_store(&dist[v], kMax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    for (int w = tid; w < 3 * W; w += kHbmBlock) bmA[w] = 0;
    if (tid < 3) { s_qn[tid] = tid == 0 ? 1 : 0; s_exp[tid] = 0; s_minq[tid] = ~0ull; }
    __syncthreads();
    if (tid == 0) {
      __hip_atomic_store(&dist[src], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      bmA[src >> 6] = 1ull << (src & 63);
      if (A.hqcap > 0) { qnode0[0] = (unsigned)src; qlab0[0] = 0ull; }
    }
    __syncthreads();

    // ---- label-correcting rounds, bucketed: a round expands the queued
    //      nodes whose label is below T and keeps the others queued; a
    //      round that expands nothing moves T to the smallest queued label
    //      plus delta (delta <= 0: no buckets, every queued node expands).
    //      The nodes queued for round r (lowered or kept in round r-1) are
    //      listed in an LDS queue with their label when listed (count
    //      s_qn[(r-1)%3], buffer (r-1)&1), one node per thread, so keeping a
    //      node needs no HBM read; the label only decides the bucket (the
    //      current label, never higher, is read to expand).  When more were
    //      queued than the queue holds, the round scans their bitmap `cur`
    //      instead.  A node is listed once per round: the lane whose
    //      atomicOr sets its bit in `nxt` appends it, one LDS atomicAdd per
    //      wave for all of the wave's appends. ----
    unsigned long long *cur = bmA, *nxt = bmB, *clr = bmC;
    const bool buckets = A.delta > 0.0;
    double T = buckets ? A.delta : INFINITY;
    int r = 1;
    // sets v's bit in nxt; the lane that set it appends (v, lab)
    auto list = [&](bool want, int v, unsigned long long lab) {
      bool first = false;
      if (want) {
        const unsigned long long bit = 1ull << (v & 63);
        first = !(atomicOr(&nxt[v >> 6], bit) & bit);
      }
      const unsigned long long bal = __ballot(first);
      if (bal) {
        const int leader = __ffsll((long long)bal) - 1;
        int b0 = 0;
        if (lane == leader) b0 = atomicAdd(&s_qn[r % 3], __popcll(bal));
        b0 = __builtin_amdgcn_readlane(b0, leader);
        if (first) {
          const int pos = b0 + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
          if (pos < A.hqcap) {
            ((r & 1) ? qlab1 : qlab0)[pos] = lab;
            ((r & 1) ? qnode1 : qnode0)[pos] = (unsigned)v;
          }
        }
      }
    };
    // u (listed with label lab) either stays queued (lab >= T) or expands;
    // the smallest label kept queued by this lane in the round is min_kept
    unsigned long long min_kept = ~0ull;
    auto visit = [&](int u, unsigned long long lab) {
      const int sl = r % 3;
      if (!(as_d(lab) < T)) lab = ldg_u64(&dist[u]);   // the listed label may be stale
      if (!(as_d(lab) < T)) {                 // later bucket: stays queued
        min_kept = min(min_kept, lab);
        ++n_def;
        list(true, u, lab);
        return;
      }
      s_exp[sl] = 1;
      const double du = as_d(ldg_u64(&dist[u]));
      const int eb = A.fw_off[u], ee = A.fw_off[u + 1];
      ++n_exp;
      n_scan += (unsigned long long)(ee - eb);
      for (int e0 = eb; e0 < ee; e0 += 8) {
        int hv[8];
        double hw[8];
        unsigned long long hd[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (e0 + k < ee) { hv[k] = A.fw_dst[e0 + k]; hw[k] = fw[e0 + k]; }
        // the heads' labels are read first: most links lower nothing, and an
        // L2 atomic costs more than a read
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (e0 + k < ee) hd[k] = ldg_u64(&dist[hv[k]]);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (e0 + k >= ee) break;
          const int v = hv[k];
          const unsigned long long cb = as_b(du + hw[k]);
          bool low = false;
          if (cb < hd[k]) {
            // no wait for the old value: a head whose label this lowers is
            // listed (a node listed twice is deduplicated by the bitmap, and
            // its key -- cb, at or above the label -- is re-read before a
            // bucket decision), so the atomic need not return
            (void)__hip_atomic_fetch_min(&dist[v], cb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            ++n_upd;
            low = v >= A.ft0;
          }
          list(low, v, cb);
        }
      }
    };
    for (;; ++r) {
      const int sl = r % 3;
      const int qn = s_qn[(r - 1) % 3];
      if (qn <= A.hqcap) {
        const unsigned* qn_r = (r & 1) ? qnode0 : qnode1;
        const unsigned long long* ql_r = (r & 1) ? qlab0 : qlab1;
        for (int i = tid; i < qn; i += kHbmBlock) visit((int)qn_r[i], ql_r[i]);
      } else {
        for (int slot = tid; slot < 4 * W; slot += kHbmBlock) {
          const int w = slot >> 2, q = slot & 3;
          unsigned bits = (unsigned)((cur[w] >> (16 * q)) & 0xFFFFull);
          while (bits) {
            const int b = __ffs(bits) - 1;
            bits &= bits - 1;
            const int u = (w << 6) + 16 * q + b;
            visit(u, ldg_u64(&dist[u]));
          }
        }
      }
      // one LDS atomic per wavefront for the round's smallest kept label
      for (int o = kWave / 2; o > 0; o >>= 1) min_kept = min(min_kept, __shfl_xor(min_kept, o));
      if (lane == 0 && min_kept != ~0ull) atomicMin(&s_minq[sl], min_kept);
      min_kept = ~0ull;
      for (int w = tid; w < W; w += kHbmBlock) clr[w] = 0;
      if (tid == 0) {
        const int sn = (r + 1) % 3;
        s_qn[sn] = 0;
        s_exp[sn] = 0;
        s_minq[sn] = ~0ull;
      }
      __syncthreads();
      if (!s_qn[sl]) break;
      if (buckets && !s_exp[sl]) T = as_d(s_minq[sl]) + A.delta;
      if (r > 64 * N + 64) {
        if (tid == 0 && atomicCAS(&A.err[0], 0, 2) == 0) { A.err[1] = src; A.err[2] = 0; }
        break;
      }
      unsigned long long* t = cur; cur = nxt; nxt = clr; clr = t;
    }
    n_rounds += (unsigned long long)r;
    if (A.work && tid == 0 && var == 0) A.work[src] = r;
    const long long clk1 = wall_clock64();

    // ---- predecessor links: smallest tail label below the node's, then
    //      tail id; nodes with none join by attach_level_ties.  Tails,
    //      costs and link ids of up to 4 in-links are loaded together, then
    //      the tails' labels. ----
    for (int w = tid; w < W; w += kHbmBlock) bmB[w] = 0;     // nodes left for attach_level_ties
    if (tid == 0) s_level_ties = 0;
    __syncthreads();
    for (int v = tid; v < N; v += kHbmBlock) {
      int p = -1, pt = -1;
      const unsigned long long dv = ldg_u64(&dist[v]);
      if (v != src && dv != kMax) {
        unsigned long long best_du = dv;      // only tails with a smaller label can win
        const int kb = A.rv_off[v], ke = A.rv_off[v + 1];
        for (int k0 = kb; k0 < ke; k0 += 4) {
          int hu[4], he[4];
          double hw[4];
          unsigned long long hd[4];
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (k0 + i < ke) { hu[i] = A.rv_src[k0 + i]; hw[i] = rw[k0 + i]; he[i] = A.rv_eid[k0 + i]; }
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (k0 + i < ke) hd[i] = ldg_u64(&dist[hu[i]]);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (k0 + i >= ke) break;
            const int u = hu[i];
            if (u == v || (u < A.ft0 && u != src)) continue;
            const unsigned long long du = hd[i];
            if (du == kMax) continue;
            if (as_b(as_d(du) + hw[i]) != dv) continue;
            bool take = du < best_du;
            if (!take && du == best_du && pt >= 0) take = u < pt;
            if (take) {
              best_du = du;
              p = he[i];
              pt = u;
            }
          }
        }
      }
      __hip_atomic_store(&pred[v], p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_store(&ptail[v], pt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      // reached, not the origin, and no tight tail with a smaller label
      const bool level_tie = p < 0 && v != src && dv != kMax;
      if (level_tie) {
        atomicOr(&bmB[v >> 6], 1ull << (v & 63));
        s_level_ties = 1;
      }
      if (A.out_dist) {
        const size_t o = (size_t)v * A.ld + item;
        A.out_dist[o] = as_d(dv);
        if (!level_tie) A.out_pred[o] = p;
      }
    }
    __syncthreads();
    if (s_level_ties) {
      const bool ok = attach_level_ties<kHbmBlock>(
          A, rw, src, bmB, bmC, &s_left, &s_added, [&](int v) { return ldg_u64(&dist[v]); },
          [&](int v, int e, int u) {
            __hip_atomic_store(&pred[v], e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_store(&ptail[v], u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (A.out_dist) A.out_pred[(size_t)v * A.ld + item] = e;
          });
      if (!ok && tid == 0 && atomicCAS(&A.err[0], 0, 3) == 0) { A.err[1] = src; A.err[2] = 0; }
      __syncthreads();
    }
    // every label read is done: the row now holds flows
    const long long clk2 = wall_clock64();
    c_relax += (unsigned long long)(clk1 - clk0);
    c_pred += (unsigned long long)(clk2 - clk1);
    if (!load) continue;

    // ---- loading: destination walks, then tree links with flow ----
    for (int v = tid; v < N; v += kHbmBlock)
      __hip_atomic_store(&dist[v], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);   // +0.0
    __syncthreads();
    int maxlen = 0;
    for (long long i = A.od_start[src] + tid; i < A.od_start[src + 1]; i += kHbmBlock) {
      const int d = A.od_dest[i];
      const double f = A.od_val[i];
      if (d == src || !(f > 0.0)) continue;
      if (ldg_i32(&pred[d]) < 0) {
        if (atomicCAS(&A.err[0], 0, 1) == 0) { A.err[1] = src; A.err[2] = d; }
        continue;
      }
      int v = d, len = 0;
      while (v != src) {
        if (len > N || v < 0) {
          if (atomicCAS(&A.err[0], 0, 3) == 0) { A.err[1] = src; A.err[2] = 0; }
          break;
        }
        __hip_atomic_fetch_add(&flow[v], f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        v = ldg_i32(&ptail[v]);
        ++len;
      }
      n_steps += (unsigned long long)len;
      maxlen = max(maxlen, len);
    }
    if (maxlen) atomicMax(&s_maxlen, maxlen);
    __syncthreads();
    double* __restrict__ yv = A.y + (size_t)var * A.E;
    for (int v = tid; v < N; v += kHbmBlock) {
      const double f = as_d(ldg_u64(&dist[v]));
      if (f != 0.0) {
        const int p = ldg_i32(&pred[v]);
        if (p >= 0) {
          unsafeAtomicAdd(&yv[p], f);
          ++n_loaded;
        }
      }
    }
    __syncthreads();          // the row is reused by the next item
    c_load += (unsigned long long)(wall_clock64() - clk2);
  }
  if (A.counters) {
    atomicAdd(&s_cnt[0], n_exp); atomicAdd(&s_cnt[1], n_scan); atomicAdd(&s_cnt[2], n_upd);
    atomicAdd(&s_cnt[5], n_loaded); atomicAdd(&s_cnt[6], n_steps); atomicAdd(&s_cnt[7], n_def);
    __syncthreads();
    if (tid == 0) {
      atomicAdd(&A.counters[7], s_cnt[7]);
      atomicAdd(&A.counters[0], s_cnt[0]); atomicAdd(&A.counters[1], s_cnt[1]);
      atomicAdd(&A.counters[2], s_cnt[2]); atomicAdd(&A.counters[3], n_rounds);
      atomicAdd(&A.counters[4], (unsigned long long)s_maxlen);
      atomicAdd(&A.counters[5], s_cnt[5]); atomicAdd(&A.counters[6], s_cnt[6]);
      atomicAdd(&A.counters[8], c_relax); atomicAdd(&A.counters[9], c_pred);
      atomicAdd(&A.counters[10], c_load);
    }
  }
}


// ---------------------------------------------------------------------------
// link passes
// ---------------------------------------------------------------------------

struct Links {
  int E, V;
  const double *t0, *cap, *B, *pw;   // [V * E]
  const unsigned char* present;      // [V * E]
  const int* rvpos;                  // [E] position of link k in reverse order
  double *x, *y, *c;                 // [V * E]
  double* rvw;                       // [V * E]  costs, reverse-CSR order, INF: absent
  double* fww;                       // [V * E]  costs, link order, INF: absent
};

struct VarState {
  double best_lb, relgap, lambda, objective, total_demand, vht;
  int iter, active, stepping, need_more, jnext, found;
};

constexpr int kMaxGraphIters = 8;    // iterations per whole-iteration replay, at most

// Arguments of a tapgpu_fw_run call that the decision kernel reads from
// device memory, so a captured iteration graph serves every call.
struct FwParams {
  tapgpu_iter* trace;
  int trace_cap;
  int max_iters;
  double target;
  int jwin;                 // step sizes a line-search window decides on (<= kJ)
};

// mode 0: c = t0 (free-flow, initial Wa); mode 1: c = BPR(x); mode 2: c given
__global__ __launch_bounds__(kBlock) void k_costs(Links K, int mode, const unsigned char* var_on,
                                                  int v0, int Vn) {
  const long long n = (long long)K.E * Vn;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock) {
    const int var = v0 + (int)(i / K.E);
    const int k = (int)(i - (long long)(var - v0) * K.E);
    if (var_on && !var_on[var]) continue;
    const size_t j = (size_t)var * K.E + k;
    double c;
    if (mode == 0) c = K.t0[j];
    else if (mode == 1) c = tap_bpr(K.t0[j], K.cap[j], K.B[j], K.pw[j], K.x[j]);
    else c = K.c[j];
    K.c[j] = c;
    const double wv = K.present[j] ? c : INFINITY;
    K.rvw[(size_t)var * K.E + K.rvpos[k]] = wv;
    K.fww[j] = wv;
  }
}

// All kQ compensated sums of a block at once, in a fixed order: a butterfly
// inside each wavefront (lane 0's result is kept), then thread 0 merges the
// wavefronts in order.  One barrier; the result is in thread 0's v[].
__device__ void block_reduce_acc(tap_acc* v, tap_acc* sh /* [kBlock / kWave][kQ] */) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    for (int off = kWave / 2; off > 0; off >>= 1) {
      tap_acc o;
      o.s = __shfl_xor(v[q].s, off);
      o.c = __shfl_xor(v[q].c, off);
      tap_acc_merge(&v[q], o);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) sh[w * kQ + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
      tap_acc a = sh[q];
      for (int j = 1; j < kBlock / kWave; ++j) tap_acc_merge(&a, sh[j * kQ + q]);
      v[q] = a;
    }
  }
}

// Partial sums per (variant, block): g0 = sum c*d, objective, vht and
// g(2^-j) for j in [jbase, jbase + kJ) (j = 0 is lambda = 1).  A block is
// one wavefront (kLinkThreads): each lane sums several links, then one
// shuffle reduction per accumulator, with no LDS stage.
constexpr int kLinkThreads = kWave;

// acc[3 + j] += BPR(x + 2^-(jbase+j) d) * d for the kJ step sizes of a
// window (lam0 = 2^-jbase).  Halving is exact for every step size used
// (2^-1016 is still a normal double), so lam equals ldexp(1, -(jbase + j)).
// Unrolled, so the sums stay in registers.
__device__ __forceinline__ void step_sums(tap_acc* acc, double lam0, double x, double d, double t0,
                                          double cap, double B, double pw) {
  double lam = lam0;
#pragma unroll
  for (int j = 0; j < kJ; ++j) {
    const double q = x + lam * d;
    tap_acc_add(&acc[3 + j], tap_bpr(t0, cap, B, pw, q) * d);
    lam *= 0.5;
  }
}

__global__ __launch_bounds__(kLinkThreads) void k_link_partials(Links K, const VarState* st,
                                                                tap_acc* part, int only_g) {
  const int var = blockIdx.y;
  const VarState s = st[var];
  if (!s.stepping || (only_g && !s.need_more)) return;
  const int jbase = only_g ? s.jnext : 0;
  tap_acc acc[kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) acc[q] = tap_acc{0.0, 0.0};
  const double lam0 = ldexp(1.0, -jbase);
  const size_t off = (size_t)var * K.E;
  for (int k = blockIdx.x * kLinkThreads + threadIdx.x; k < K.E; k += gridDim.x * kLinkThreads) {
    const size_t i = off + k;
    if (!K.present[i]) continue;
    const double x = K.x[i];
    const double d = K.y[i] - x;
    const double t0 = K.t0[i], cap = K.cap[i], B = K.B[i], pw = K.pw[i];
    if (!only_g) {
      const double c = K.c[i];
      tap_acc_add(&acc[0], c * d);
      tap_acc_add(&acc[1], tap_integral(t0, cap, B, pw, x));
      tap_acc_add(&acc[2], x * c);
    }
    step_sums(acc, lam0, x, d, t0, cap, B, pw);
  }
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    for (int o = kWave / 2; o > 0; o >>= 1) {
      tap_acc t;
      t.s = __shfl_xor(acc[q].s, o);
      t.c = __shfl_xor(acc[q].c, o);
      tap_acc_merge(&acc[q], t);
    }
  }
  if (threadIdx.x == 0) {
    tap_acc* out = part + ((size_t)var * gridDim.x + blockIdx.x) * kQ;
#pragma unroll
    for (int q = 0; q < kQ; ++q) out[q] = acc[q];
  }
}

// Per variant: combine partials, pick lambda (first 2^-j with g <= 0, the
// reference's rule), then gap, objective, lower bound, relative gap, average
// excess cost, trace record, stop test.
// The decision of one window for one variant from its merged sums acc[kQ]
// (thread 0 of the caller): on the first window the iteration's scalars
// and trace record, then lambda = the first 2^-j with g <= 0, or a request
// for the next window (need_more, jnext).
__device__ __forceinline__ void decide_window(VarState& s, const tap_acc* acc, int only_g,
                                              const FwParams* prm, int var) {
  tapgpu_iter* const trace = prm->trace;
  const int trace_cap = prm->trace_cap, max_iters = prm->max_iters;
  const double target = prm->target;
  const int jwin = prm->jwin;
  double gj[kJ];
#pragma unroll
  for (int j = 0; j < kJ; ++j) gj[j] = tap_acc_value(acc[3 + j]);
  const int jbase = only_g ? s.jnext : 0;
  if (!only_g) {
    const double g0 = tap_acc_value(acc[0]);
    s.objective = tap_acc_value(acc[1]);
    s.vht = tap_acc_value(acc[2]);
    s.found = 0;
    s.need_more = 0;
    if (g0 >= 0) {
      s.lambda = 0.0; s.found = 1;
    } else if (gj[0] <= 0) {
      s.lambda = 1.0; s.found = 1;
    }
    // record the iteration's scalars (lambda filled below)
    const double gap = -g0;
    const double lb = s.objective - gap;
    if (lb > s.best_lb) s.best_lb = lb;
    if (s.best_lb != 0) s.relgap = (s.objective - s.best_lb) / fabs(s.best_lb);
    if (s.iter < trace_cap) {
      tapgpu_iter t;
      t.lambda = 0.0; t.gap = gap; t.objective = s.objective; t.lowerbound = lb;
      t.best_lowerbound = s.best_lb; t.relgap = s.relgap;
      t.avg_excess = -g0 / s.total_demand; t.vht = s.vht;
      trace[(size_t)var * trace_cap + s.iter] = t;
    }
  }
  if (!s.found) {
    for (int j = std::max(jbase, 1); j < jbase + jwin && j <= kMaxHalvings; ++j) {
      if (gj[j - jbase] <= 0) {
        s.lambda = ldexp(1.0, -j);
        s.found = 1;
        break;
      }
    }
    if (!s.found) {
      if (jbase + jwin > kMaxHalvings) {
        s.lambda = 0.0;            // bisection cap reached: lambda_left == 0
        s.found = 1;
      } else {
        s.need_more = 1;
        s.jnext = jbase + jwin;
      }
    }
  }
  if (s.found) {
    s.need_more = 0;
    if (s.iter < trace_cap) trace[(size_t)var * trace_cap + s.iter].lambda = s.lambda;
    s.iter += 1;
    s.active = (max_iters <= 0 || s.iter < max_iters) && (target <= 0 || s.relgap > target);
  }
}

__global__ __launch_bounds__(kBlock) void k_link_decide(VarState* st, const tap_acc* part,
                                                        int nparts, int only_g,
                                                        const FwParams* prm) {
  const int var = blockIdx.x;
  VarState s = st[var];
  if (!s.stepping || (only_g && !s.need_more)) return;
  tap_acc acc[kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) acc[q] = tap_acc{0.0, 0.0};
  for (int b = threadIdx.x; b < nparts; b += kBlock) {
    const tap_acc* in = part + ((size_t)var * nparts + b) * kQ;
#pragma unroll
    for (int q = 0; q < kQ; ++q) tap_acc_merge(&acc[q], in[q]);
  }
  __shared__ tap_acc sh[kBlock / kWave * kQ];
  block_reduce_acc(acc, sh);
  if (threadIdx.x != 0) return;
  decide_window(s, acc, only_g, prm, var);
  st[var] = s;
}

// x += lambda * (y - x); c = BPR(x); reverse-order costs for the next sweep.
__global__ __launch_bounds__(kBlock) void k_link_step(Links K, VarState* st) {
  const int var = blockIdx.y;
  if (!st[var].stepping || st[var].need_more) return;     // no step until lambda is known
  const double lambda = st[var].lambda;
  const size_t off = (size_t)var * K.E;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < K.E; k += gridDim.x * kBlock) {
    const size_t i = off + k;
    const double d = K.y[i] - K.x[i];
    const double x = K.x[i] + lambda * d;
    K.x[i] = x;
    const double c = tap_bpr(K.t0[i], K.cap[i], K.B[i], K.pw[i], x);
    K.c[i] = c;
    const double wv = K.present[i] ? c : INFINITY;
    K.rvw[off + K.rvpos[k]] = wv;
    K.fww[i] = wv;
  }
}

// All link passes of an iteration for a small network, one workgroup per
// variant and one launch: the window sums (block reduction), the decision,
// further windows while the step is undecided, the step (x, costs, both
// cost mirrors) and the flags.  A workgroup owns its variant's links, so
// nothing waits on other workgroups.  Used where links <= kSmallLinks;
// the sums are the same compensated sums as the multi-block passes.
constexpr int kSmallLinks = 4096;

__global__ __launch_bounds__(kBlock) void k_link_small(Links K, VarState* st, unsigned char* var_on,
                                                       const FwParams* prm) {
  const int var = blockIdx.x;
  VarState s = st[var];
  if (!s.stepping) return;
  __shared__ tap_acc sh[kBlock / kWave * kQ];
  __shared__ int s_more, s_jnext;
  __shared__ double s_lambda;
  const size_t off = (size_t)var * K.E;
  int jbase = 0;                     // first window, then the ones thread 0's decisions ask for
  for (int only_g = 0;; only_g = 1) {
    tap_acc acc[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) acc[q] = tap_acc{0.0, 0.0};
    const double lam0 = ldexp(1.0, -jbase);
    for (int k = threadIdx.x; k < K.E; k += kBlock) {
      const size_t i = off + k;
      if (!K.present[i]) continue;
      const double x = K.x[i];
      const double d = K.y[i] - x;
      const double t0 = K.t0[i], cap = K.cap[i], B = K.B[i], pw = K.pw[i];
      if (!only_g) {
        const double c = K.c[i];
        tap_acc_add(&acc[0], c * d);
        tap_acc_add(&acc[1], tap_integral(t0, cap, B, pw, x));
        tap_acc_add(&acc[2], x * c);
      }
      step_sums(acc, lam0, x, d, t0, cap, B, pw);
    }
    block_reduce_acc(acc, sh);
    if (threadIdx.x == 0) {
      decide_window(s, acc, only_g, prm, var);
      s_more = s.need_more;
      s_jnext = s.jnext;
      s_lambda = s.lambda;
    }
    __syncthreads();
    if (!s_more) break;
    jbase = s_jnext;
    __syncthreads();                   // (s_more and s_jnext are rewritten next window)
  }
  const double lambda = s_lambda;
  for (int k = threadIdx.x; k < K.E; k += kBlock) {
    const size_t i = off + k;
    const double d = K.y[i] - K.x[i];
    const double x = K.x[i] + lambda * d;
    K.x[i] = x;
    const double c = tap_bpr(K.t0[i], K.cap[i], K.B[i], K.pw[i], x);
    K.c[i] = c;
    const double wv = K.present[i] ? c : INFINITY;
    K.rvw[off + K.rvpos[k]] = wv;
    K.fww[i] = wv;
  }
  if (threadIdx.x == 0) {
    s.stepping = 0;
    st[var] = s;
    var_on[var] = (unsigned char)s.active;
  }
}

// Per variant at the current volumes: out[2v] = VHT = sum x * c, out[2v+1]
// = the Beckmann objective (sum of the link cost integrals), compensated
// sums in a fixed order.
__global__ __launch_bounds__(kBlock) void k_totals(Links K, double* out) {
  const int var = blockIdx.x;
  tap_acc acc[kQ];
  for (int q = 0; q < kQ; ++q) acc[q] = tap_acc{0.0, 0.0};
  const size_t off = (size_t)var * K.E;
  for (int k = threadIdx.x; k < K.E; k += kBlock) {
    const size_t i = off + k;
    if (!K.present[i]) continue;
    tap_acc_add(&acc[0], K.x[i] * K.c[i]);
    tap_acc_add(&acc[1], tap_integral(K.t0[i], K.cap[i], K.B[i], K.pw[i], K.x[i]));
  }
  __shared__ tap_acc sh[kBlock / kWave * kQ];
  block_reduce_acc(acc, sh);
  if (threadIdx.x == 0) {
    out[2 * var] = tap_acc_value(acc[0]);
    out[2 * var + 1] = tap_acc_value(acc[1]);
  }
}

__global__ void k_finish_step(VarState* st, int V, unsigned char* var_on) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  if (st[v].stepping && st[v].need_more) return;          // still searching lambda
  st[v].stepping = 0;
  var_on[v] = (unsigned char)st[v].active;
}

// Opens an FW iteration in one launch: the step flags, and the loading's
// volumes cleared; the error flag and counters too unless `keep` (later
// iterations of a multi-iteration graph, which report the first error and
// the sum of the counters).
__global__ __launch_bounds__(kBlock) void k_iter_begin(VarState* st, int V, const unsigned char* var_on,
                                                       int* err, unsigned long long* counters,
                                                       double* y, long long ny, int keep) {
  const long long i0 = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long v = i0; v < V; v += stride) {
    st[v].stepping = var_on[v] ? 1 : 0;
    st[v].need_more = 0;
  }
  if (!keep) {
    if (i0 < 4) err[i0] = 0;
    if (i0 < 16) counters[i0] = 0ull;
  }
  for (long long i = i0; i < ny; i += stride) y[i] = 0.0;
}

__global__ void k_begin_step(VarState* st, int V, const unsigned char* var_on) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  st[v].stepping = var_on[v] ? 1 : 0;
  st[v].need_more = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct tapgpu_ctx {
  std::string err;
  hipStream_t stream = nullptr;
  int device = 0;
  // sizes
  int N = 0, E = 0, Z = 0, ft0 = 0, V = 1, Gv = 0, Galloc = 0;
  long long Pv = 0, P = 0;
  double total_demand = 0;
  // graph
  int *fw_off = nullptr, *fw_dst = nullptr, *rv_off = nullptr, *rv_src = nullptr,
      *rv_eid = nullptr, *tail = nullptr, *rvpos = nullptr;
  // link records per variant
  double *t0 = nullptr, *cap = nullptr, *B = nullptr, *pw = nullptr;
  unsigned char* present = nullptr;
  double *x = nullptr, *y = nullptr, *c = nullptr, *rvw = nullptr, *fww = nullptr;
  // origin-major demand (per-origin workgroups)
  long long* od_start = nullptr;
  int* od_dest = nullptr;
  double* od_val = nullptr;
  // strategy: 0 auto, 1 node-major sweeps, 2 one origin per workgroup (LDS)
  int strategy = 0;
  int chase = -1;                     // LDS path: chained expansions (-1: tune)
  int lds_block = -1;                 // LDS path: workgroup size (-1: tune)
  // LDS loadings dispatch origins longest first: the first loading records
  // each origin's relaxation rounds, later ones launch in that order
  int* d_origin_work = nullptr;          // [Z] rounds per origin
  int* d_hbm_next = nullptr;          // item counter of a labels-in-HBM loading
  int* d_origin_order = nullptr;         // [Z] origins by decreasing rounds
  bool origin_recorded = false, origin_order_ready = false;
  std::vector<int> h_origin_order;       // its host copy (the upload's source)
  bool qcap_auto = true;              // LDS path: queue size picked with the workgroup size
  int split = 0;                      // LDS path: origins per variant run by the HBM kernel
                                      // alongside it (-1: tune, 0: none)
  hipStream_t stream2 = nullptr;      // second queue for that co-running kernel
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool hbm_ok = false;                // labels-in-HBM path: bitmaps fit LDS
  size_t hbm_lds = 0;
  int hbm_grid = 0;                   // persistent workgroups
  double delta_mult = -1;             // bucket width / mean free-flow time (-1: tune)
  double mean_t0 = 1;                 // mean free-flow time of variant 0's links
  unsigned long long* h_dist = nullptr;
  int *h_pred = nullptr, *h_ptail = nullptr;
  int clock_khz = 0;                  // device wall clock
  bool labels_ready = false;          // node-major buffers allocated
  std::vector<double> h_odc;          // demand by (zone row, column), host copy
  std::vector<long long> h_od_start;  // origin-major demand, host copy (input order)
  std::vector<int> h_od_dest;
  std::vector<double> h_od_val;
  std::vector<int> h_tail;            // tail node of each link (0-based)
  std::vector<int> h_head;            // head node of each link (0-based)
  bool lds_ok = false;
  size_t lds_bytes = 0;
  int qcap = 0;                       // LDS frontier queue entries in use (per buffer)
  int qcap_max = 0;                   // entries the LDS allocation holds
  int hqcap = 0;                      // labels-in-HBM kernel: queue entries in use (per buffer)
  int hqcap_max = 0;                  // entries its LDS allocation holds
  // labels
  double* dist = nullptr;
  int *pred = nullptr, *pend = nullptr, *col_src = nullptr, *grp_var = nullptr;
  int *stamp = nullptr, *flags = nullptr;
  unsigned long long* counters = nullptr;
  int flags_cap = 0;
  int tag = 0;
  std::vector<int> zone_order;   // column -> zone (landmark order)
  // loading
  double *odc = nullptr, *ypart = nullptr;
  int* err_dev = nullptr;
  // FW state
  unsigned char* var_on = nullptr;
  VarState* st = nullptr;
  tap_acc* part = nullptr;
  double* vht = nullptr;              // [2V] VHT, objective at the current volumes
  FwParams* prm = nullptr;            // the fw_run call's limits and trace
  // captured iteration graphs (replayed while profiling is off):
  // g_iter = iteration begin + per-origin loading + link passes + readback,
  // g_link = link passes + readback (for the lanes path)
  bool graphs = true;
  int link_mode = 0;                  // link passes: 0 by size, 1 one workgroup per variant, 2 multi-block
  int line_window = kJ;               // step sizes per line-search window
  int g_link_mode = 0;                // the mode g_link was captured with
  int graph_iters = 0;                // iterations per whole-iteration replay (0: from the iteration time)
  double last_iter_ms = 0;            // wall time of the last iteration (per iteration of a replay)
  hipGraphExec_t g_iter = nullptr, g_link = nullptr;
  long long g_iter_key[6] = {-1, -1, -1, -1, -1, -1};
  int64_t graph_launches = 0;
  tapgpu_iter* trace = nullptr;
  int trace_cap = 0;
  bool initialised = false;
  // pinned host scratch
  int* h_int = nullptr;
  // pinned readback of one FW iteration: error flag, origin-kernel
  // counters, variant states and flags (one synchronisation per iteration)
  int* pin_err = nullptr;                   // [4]
  unsigned long long* pin_cnt = nullptr;    // [16]
  VarState* pin_st = nullptr;               // [V]
  unsigned char* pin_on = nullptr;          // [V]
  void* pin_block = nullptr;
  // loading whose counters and error flag are read at the next
  // synchronisation instead of right after the launch
  bool defer_origin = false;
  bool origin_pending = false;
  int pending_Vn = 0;
  // the volumes, error flag and counters of the next loading were cleared
  // by k_iter_begin (the loading skips its own clears)
  bool pre_cleared = false;
  // error flag, counters, variant states and flags: one device block laid
  // out like pin_block, read back with one copy
  unsigned char* iter_dev = nullptr;
  size_t iter_bytes = 0;
  // lanes kernel (batched loading, variants in wavefront lanes)
  int lanes_min = 0;                  // active variants needed to use it (0: never)
  bool lanes_ok = false;              // its LDS (bitmaps + queues) fits
  int lanes_qcap = 0;                 // LDS frontier queue entries per round
  int lanes_qcap_max = 0;             // entries its LDS allocation holds
  double lanes_delta_mult = -1;       // bucket width / mean free-flow time (-1: tune)
  size_t lanes_lds = 0;
  int lanes_grid = 0;                 // persistent workgroups (rows allocated)
  int lanes_grid_max = 0;             // resident workgroups on this device
  unsigned long long *l_dist = nullptr, *l_memo = nullptr;   // [grid][N][64]
  double *l_W = nullptr, *l_Y = nullptr;                     // [groups][E][64]
  int* l_lane_var = nullptr;          // [groups * 64]
  std::vector<int> h_lane_var;        // its host copy (as uploaded)
  unsigned l_stamp = 1;               // stamp of the next task's memo entries
  unsigned* l_stamp_dev = nullptr;    // its device copy, read by the kernel
  int* l_next = nullptr;              // task counter of a lanes launch
  bool capturing = false;             // launches are being recorded into a graph
  // stats
  tapgpu_stats stats{};
  bool profiling = false;
  hipEvent_t ev[2 * 8] = {};
  std::vector<void*> allocs;

  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
};

namespace {

thread_local std::string g_create_error;   // message of this thread's last failed create

#define HIPCHK(ctx, call)                                                               \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return (ctx)->fail(TAPGPU_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_));  \
  } while (0)

template <class T>
int dalloc(tapgpu_ctx* ctx, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return ctx->fail(TAPGPU_E_NOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                     hipGetErrorString(e));
  }
  ctx->allocs.push_back(*p);
  return TAPGPU_OK;
}

template <class T>
int upload(tapgpu_ctx* ctx, T* d, const T* h, size_t n) {
  if (n == 0) return TAPGPU_OK;
  HIPCHK(ctx, hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return TAPGPU_OK;
}

int grid_for(long long waves) {
  long long b = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (int)b;
}

constexpr int kSweepBlocks = 2048;
constexpr int kCheckEvery = 8;

Graph make_graph(tapgpu_ctx* c) {
  return Graph{c->N, c->E, c->ft0, c->fw_off, c->fw_dst, c->rv_off, c->rv_src, c->rv_eid, c->tail};
}

Labels make_labels(tapgpu_ctx* c, const int* col_src, const int* grp_var,
                   const unsigned char* var_on) {
  return Labels{c->dist, c->pred, c->pend, c->P, c->Galloc, col_src, grp_var, var_on, c->rvw};
}

Links make_links(tapgpu_ctx* c) {
  return Links{c->E, c->V, c->t0, c->cap, c->B, c->pw, c->present, c->rvpos,
               c->x, c->y, c->c, c->rvw, c->fww};
}

// Round loop shared by the sweeps and the loading levels: launch
// kCheckEvery rounds, then read whether the next round has work.
template <class Launch>
int run_rounds(tapgpu_ctx* c, Launch launch, int64_t* rounds_out, double* kernel_ms,
               int64_t* launches) {
  int r = 0;
  for (;;) {
    if (r + kCheckEvery + 2 >= c->flags_cap)
      return c->fail(TAPGPU_E_LIMIT, "worklist did not drain in %d rounds", r);
    for (int k = 0; k < kCheckEvery; ++k, ++r) {
      if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[2 * k], c->stream));
      launch(r, c->tag + r, c->tag + r + 1);
      if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[2 * k + 1], c->stream));
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->h_int, c->flags + r - kCheckEvery + 1, sizeof(int) * kCheckEvery,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->profiling) {
      // rounds without work still launch; count every launch that was timed
      for (int k = 0; k < kCheckEvery; ++k) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]));
        *kernel_ms += ms;
        *launches += 1;
      }
    }
    if (c->h_int[kCheckEvery - 1] == 0) {
      if (rounds_out) {
        int last = r - kCheckEvery;
        for (int k = 0; k < kCheckEvery; ++k)
          if (c->h_int[k]) last = r - kCheckEvery + 1 + k;
        *rounds_out += last + 1;
      }
      break;
    }
  }
  c->tag += r + 2;
  return TAPGPU_OK;
}

// tags are compared for equality with stamp[]; restart them long before
// 32-bit overflow
int maybe_reset_tags(tapgpu_ctx* c) {
  if (c->tag < 2000000000) return TAPGPU_OK;
  HIPCHK(c, hipMemsetAsync(c->stamp, 0xFF, sizeof(int) * (size_t)c->N * c->Galloc, c->stream));
  c->tag = 0;
  return TAPGPU_OK;
}

int scan_blocks(tapgpu_ctx* c, int Gn) {
  return (int)(((long long)c->N * Gn + kBlock - 1) / kBlock);
}

int read_counters(tapgpu_ctx* c, int64_t* a, int64_t* b, int64_t* d, int64_t* e) {
  unsigned long long h[4];
  HIPCHK(c, hipMemcpyAsync(h, c->counters, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (a) *a += (int64_t)h[0];
  if (b) *b += (int64_t)h[1];
  if (d) *d += (int64_t)h[2];
  if (e) *e += (int64_t)h[3];
  return TAPGPU_OK;
}

// Node-major buffers of the multi-origin sweep path, allocated on first use
// (the one-origin-per-workgroup path never needs them).
int ensure_labels(tapgpu_ctx* c) {
  if (c->labels_ready) return TAPGPU_OK;
  const size_t NP = (size_t)c->N * c->P;
  int rc;
  if ((rc = dalloc(c, &c->dist, NP))) return rc;
  if ((rc = dalloc(c, &c->pred, NP))) return rc;
  if ((rc = dalloc(c, &c->pend, NP))) return rc;
  if ((rc = dalloc(c, &c->stamp, (size_t)c->N * c->Galloc))) return rc;
  if ((rc = dalloc(c, &c->odc, c->h_odc.size()))) return rc;
  if ((rc = dalloc(c, &c->ypart, (size_t)c->Galloc * c->E))) return rc;
  if ((rc = upload(c, c->odc, c->h_odc.data(), c->h_odc.size()))) return rc;
  HIPCHK(c, hipMemsetAsync(c->stamp, 0xFF, sizeof(int) * (size_t)c->N * c->Galloc, c->stream));
  HIPCHK(c, hipMemsetAsync(c->ypart, 0, sizeof(double) * (size_t)c->Galloc * c->E, c->stream));
  c->labels_ready = true;
  return TAPGPU_OK;
}

// Shortest paths for groups [g0, g0+Gn).
int sssp_groups(tapgpu_ctx* c, int g0, int Gn, const int* col_src, const int* grp_var,
                const unsigned char* var_on) {
  int rc = ensure_labels(c);
  if (rc) return rc;
  rc = maybe_reset_tags(c);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  Graph G = make_graph(c);
  Labels L = make_labels(c, col_src, grp_var, var_on);
  Frontier F{c->stamp, c->flags, c->counters};
  HIPCHK(c, hipMemsetAsync(c->flags, 0, sizeof(int) * c->flags_cap, c->stream));
  HIPCHK(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long) * 4, c->stream));
  hipLaunchKernelGGL(k_fill_labels, dim3(grid_for((long long)c->N * Gn)), dim3(kBlock), 0,
                     c->stream, G, L, g0, Gn);
  hipLaunchKernelGGL(k_sssp_seed, dim3((Gn + kWavesPerBlock - 1) / kWavesPerBlock), dim3(kBlock), 0,
                     c->stream, G, L, F, g0, Gn, c->tag);
  const int nb = scan_blocks(c, Gn);
  rc = run_rounds(
      c,
      [&](int r, int tag, int tag_next) {
        hipLaunchKernelGGL(k_sssp_sweep, dim3(nb), dim3(kBlock), 0, c->stream, G, L, F, g0, Gn, r,
                           tag, tag_next);
      },
      &c->stats.sssp_sweeps, &c->stats.sweep_kernel_ms, &c->stats.sweep_launches);
  if (rc) return rc;
  rc = read_counters(c, &c->stats.sssp_items, &c->stats.sssp_inlinks, &c->stats.sssp_updates,
                     &c->stats.sssp_outlinks);
  c->stats.sssp_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

// Loading for groups [g0, g0+Gn) of variants [v0, v0+Vn) into y.  With
// check_tree the predecessor links came from the caller: lanes that never
// drained (a cycle) are reported as TAPGPU_E_ARG.
int aon_groups(tapgpu_ctx* c, int g0, int Gn, int v0, int Vn, const unsigned char* var_on,
               bool check_tree = false) {
  int rc = ensure_labels(c);
  if (rc) return rc;
  rc = maybe_reset_tags(c);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  Graph G = make_graph(c);
  Labels L = make_labels(c, c->col_src, c->grp_var, var_on);
  Loading A{c->odc, c->Pv, c->Z, c->ypart, c->err_dev};
  Frontier F{c->stamp, c->flags, c->counters};
  HIPCHK(c, hipMemsetAsync(c->flags, 0, sizeof(int) * c->flags_cap, c->stream));
  HIPCHK(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long) * 4, c->stream));
  HIPCHK(c, hipMemsetAsync(c->err_dev, 0, sizeof(int) * 3, c->stream));
  hipLaunchKernelGGL(k_aon_init, dim3(grid_for((long long)c->N * Gn)), dim3(kBlock), 0, c->stream,
                     G, L, A, F, g0, Gn, c->tag);
  const int nb = scan_blocks(c, Gn);
  rc = run_rounds(
      c,
      [&](int r, int tag, int tag_next) {
        hipLaunchKernelGGL(k_aon_level, dim3(nb), dim3(kBlock), 0, c->stream, G, L, A, F, g0, Gn,
                           r, tag, tag_next);
      },
      &c->stats.aon_levels, &c->stats.level_kernel_ms, &c->stats.level_launches);
  if (rc) return rc;
  if ((rc = read_counters(c, &c->stats.aon_items, &c->stats.aon_lanes, nullptr, nullptr))) return rc;
  if (check_tree) {
    hipLaunchKernelGGL(k_aon_check, dim3(grid_for((long long)c->N * Gn)), dim3(kBlock), 0,
                       c->stream, G, L, A, g0, Gn);
    HIPCHK(c, hipGetLastError());
  }
  int herr[3];
  HIPCHK(c, hipMemcpyAsync(herr, c->err_dev, sizeof herr, hipMemcpyDeviceToHost, c->stream));
  hipLaunchKernelGGL(k_aon_reduce, dim3(grid_for((long long)c->E * Vn / kWave + 1)), dim3(kBlock), 0,
                     c->stream, c->E, c->Gv, v0, Vn, var_on, c->ypart, c->y);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->stats.aon_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (herr[0] == 4)
    return c->fail(TAPGPU_E_ARG, "predecessor links do not form a tree (origin %d, node %d)",
                   herr[1] + 1, herr[2] + 1);
  if (herr[0])
    return c->fail(TAPGPU_E_NOPATH, "no path from %d to %d", herr[1] + 1, herr[2] + 1);
  return TAPGPU_OK;
}

bool use_lds(tapgpu_ctx* c) {
  return c->lds_ok && (c->strategy == 0 || c->strategy == 2);
}

// auto: labels in LDS where they fit, else labels in HBM
bool use_hbm(tapgpu_ctx* c) {
  return c->hbm_ok && (c->strategy == 3 || (c->strategy == 0 && !c->lds_ok));
}

OriginArgs make_origin_args(tapgpu_ctx* c, const unsigned char* var_on) {
  OriginArgs A{};
  A.N = c->N; A.E = c->E; A.ft0 = c->ft0; A.Z = c->Z;
  A.fw_off = c->fw_off; A.fw_dst = c->fw_dst; A.rv_off = c->rv_off; A.rv_src = c->rv_src;
  A.rv_eid = c->rv_eid; A.tail = c->tail; A.fww = c->fww; A.rvw = c->rvw;
  A.od_start = c->od_start; A.od_dest = c->od_dest; A.od_val = c->od_val;
  A.var_on = var_on; A.y = c->y; A.err = c->err_dev; A.counters = c->counters;
  A.chase = c->chase >= 0 ? c->chase : kDefaultChase;
  A.qcap = c->qcap;
  A.hqcap = c->hqcap;
  A.h_dist = c->h_dist; A.h_pred = c->h_pred; A.h_ptail = c->h_ptail;
  A.delta = c->delta_mult > 0 ? c->delta_mult * c->mean_t0 : 0.0;
  A.z0 = 0;
  A.zn = c->Z;
  return A;
}

// Workgroup size the LDS kernel runs with (before tuning: kLdsBlock).
int lds_threads(const tapgpu_ctx* c) { return c->lds_block > 0 ? c->lds_block : kLdsBlock; }

// One launch of the LDS kernel at the context's workgroup size, with the
// LDS its queue size needs (a smaller queue leaves room for more
// workgroups per CU).
void launch_lds(tapgpu_ctx* c, dim3 grid, hipStream_t st, const OriginArgs& A, int load) {
  const size_t sh = origin_lds_bytes(c->N, A.qcap);
  switch (lds_threads(c)) {
    case 256: hipLaunchKernelGGL(k_origin_lds<256>, grid, dim3(256), sh, st, A, load); break;
    case 512: hipLaunchKernelGGL(k_origin_lds<512>, grid, dim3(512), sh, st, A, load); break;
    default: hipLaunchKernelGGL(k_origin_lds<1024>, grid, dim3(1024), sh, st, A, load); break;
  }
}

// Shortest paths + loading for variants [v0, v0+Vn) with one workgroup per
// (origin, variant); y of those variants is overwritten.
//
// Worklist schedule tuning: the first loading on a context with chase < 0
// times one full launch for each candidate chain length and keeps the
// fastest (the result is the same for every candidate; the last launch
// below, with the chosen value, is the one whose volumes are kept).
constexpr int kChaseCandidates[] = {0, 2, 4, 8, 16, 32};

int ensure_hbm(tapgpu_ctx* c);

// Per-origin round counts of the recording loading (zeroed: origins a
// loading skips keep 0 and go last).
int ensure_origin_work(tapgpu_ctx* c) {
  int rc = dalloc(c, &c->d_origin_work, (size_t)c->Z);
  if (rc) return rc;
  HIPCHK(c, hipMemsetAsync(c->d_origin_work, 0, sizeof(int) * (size_t)c->Z, c->stream));
  return TAPGPU_OK;
}

// One loading launch of the LDS path.  With split > 0 the last `split`
// origins of every variant go to the labels-in-HBM kernel on a second queue,
// running on the same CUs at the same time: the LDS kernel's workgroup takes
// most of a CU's LDS but only half of its wave slots, and the HBM kernel's
// workgroups (a few KB of LDS each) fill the rest.  Both add integer flows
// to y with atomics, so the split changes no result.
int launch_origin(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on, float* ms) {
  if (!c->pre_cleared)
    HIPCHK(c, hipMemsetAsync(c->y + (size_t)v0 * c->E, 0, sizeof(double) * (size_t)Vn * c->E, c->stream));
  if (c->Z > 0) {
    OriginArgs A = make_origin_args(c, var_on);
    A.fww += (size_t)v0 * c->E; A.rvw += (size_t)v0 * c->E; A.y += (size_t)v0 * c->E;
    if (var_on) A.var_on = var_on + v0;
    const int zh = (c->hbm_ok && c->split > 0) ? std::min(c->split, c->Z - 1) : 0;
    const int zl = c->Z - zh;
    if (zh > 0) {
      int rc = ensure_hbm(c);
      if (rc) return rc;
      A.h_dist = c->h_dist; A.h_pred = c->h_pred; A.h_ptail = c->h_ptail;
    }
    if (zh == 0) {
      if (c->origin_order_ready) {
        A.src_list = c->d_origin_order;
      } else if (!c->capturing && c->Z > 1) {
        if (!c->d_origin_work) {
          int rc = ensure_origin_work(c);
          if (rc) return rc;
        }
        A.work = c->d_origin_work;
        c->origin_recorded = true;
      }
    }
    if (ms) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    if (zh > 0) {
      HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
      HIPCHK(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
      OriginArgs B = A;
      B.z0 = zl;
      B.zn = zh;
      B.src_list = nullptr;
      B.work = nullptr;
      const int items = zh * Vn;
      hipLaunchKernelGGL(k_origin_hbm, dim3(std::min(c->hbm_grid, items)), dim3(kHbmBlock),
                         c->hbm_lds, c->stream2, B, 1, items);
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipEventRecord(c->ev_join, c->stream2));
    }
    launch_lds(c, dim3(zl, Vn), c->stream, A, 1);
    HIPCHK(c, hipGetLastError());
    if (zh > 0) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    if (ms) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    if (ms) {
      HIPCHK(c, hipEventSynchronize(c->ev[1]));
      HIPCHK(c, hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
    }
  }
  return TAPGPU_OK;
}

// Share of origins for the co-running HBM kernel: each candidate (as a
// fraction of Z) is timed on the full launch; the fastest is kept.
constexpr int kSplitCandidates[] = {0, 10, 15, 20, 25, 33};   // percent of Z

int tune_split(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on) {
  if (!c->hbm_ok || c->Z < 8) {
    c->split = 0;
    return TAPGPU_OK;
  }
  if (!c->ev[0])
    for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
  float best_ms = 0.f;
  int best = 0;
  for (int pct : kSplitCandidates) {
    c->split = (int)((long long)c->Z * pct / 100);
    float ms = 0.f;
    int rc = launch_origin(c, v0, Vn, var_on, &ms);
    if (rc) return rc;
    int herr = 0;
    HIPCHK(c, hipMemcpyAsync(&herr, c->err_dev, sizeof herr, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (herr) break;
    if (best_ms == 0.f || ms < best_ms) { best_ms = ms; best = c->split; }
  }
  c->split = best;
  c->stats.origin_tune_ms += best_ms;
  return TAPGPU_OK;
}

int tune_chase(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on) {
  if (!c->ev[0])
    for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
  float best_ms = 0.f;
  int best = kDefaultChase;
  for (int k : kChaseCandidates) {
    c->chase = k;
    float ms = 0.f;
    int rc = launch_origin(c, v0, Vn, var_on, &ms);
    if (rc) return rc;
    int herr = 0;
    HIPCHK(c, hipMemcpyAsync(&herr, c->err_dev, sizeof herr, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (herr) break;                    // reported by the caller's launch
    if (best_ms == 0.f || ms < best_ms) { best_ms = ms; best = k; }
  }
  c->chase = best;
  c->stats.origin_tune_ms += best_ms;
  return TAPGPU_OK;
}

// Workgroup size and queue size of the LDS kernel: every allowed size
// (kLdsPerThread nodes per thread at most) with, unless the queue size was
// set, the full queue and two shorter ones (less LDS per workgroup, so more
// workgroups per CU), each timed on one full launch; the fastest pair is
// kept.  Results do not depend on them.
int tune_lds_shape(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on) {
  if (!c->ev[0])
    for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
  std::vector<int> qs{c->qcap};
  if (c->qcap_auto) {
    qs = {c->qcap_max};
    for (int q : {1024, 256})
      if (q < c->qcap_max) qs.push_back(q);
  }
  float best_ms = 0.f;
  int best_b = kLdsBlock, best_q = c->qcap;
  for (int b : kLdsBlocks) {
    if (c->N > kLdsPerThread * b) continue;
    for (int q : qs) {
      c->lds_block = b;
      c->qcap = q;
      float ms = 0.f;
      int rc = launch_origin(c, v0, Vn, var_on, &ms);
      if (rc) return rc;
      int herr = 0;
      HIPCHK(c, hipMemcpyAsync(&herr, c->err_dev, sizeof herr, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      if (herr) {                       // reported by the caller's launch
        c->lds_block = best_b;
        c->qcap = best_q;
        return TAPGPU_OK;
      }
      if (best_ms == 0.f || ms < best_ms) { best_ms = ms; best_b = b; best_q = q; }
    }
  }
  c->lds_block = best_b;
  c->qcap = best_q;
  c->stats.origin_tune_ms += best_ms;
  return TAPGPU_OK;
}

int absorb_origin(tapgpu_ctx* c, const int* herr, const unsigned long long* hc, int Vn,
                  std::chrono::steady_clock::time_point t0);

// Reads the error flag and counters of an origin-kernel launch into stats:
// now, or (defer_origin) from the pinned copy of the iteration block that
// the caller makes, absorbed by absorb_pending after its synchronisation.
int finish_origin(tapgpu_ctx* c, int Vn, std::chrono::steady_clock::time_point t0) {
  if (c->defer_origin) {            // read back with the iteration's block copy
    c->origin_pending = true;
    c->pending_Vn = Vn;
    return TAPGPU_OK;
  }
  int herr[3];
  unsigned long long hc[16];
  HIPCHK(c, hipMemcpyAsync(herr, c->err_dev, sizeof herr, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hc, c->counters, sizeof hc, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return absorb_origin(c, herr, hc, Vn, t0);
}

// The deferred readback of finish_origin, after the stream has synchronised.
int absorb_pending(tapgpu_ctx* c) {
  if (!c->origin_pending) return TAPGPU_OK;
  c->origin_pending = false;
  return absorb_origin(c, c->pin_err, c->pin_cnt, c->pending_Vn, std::chrono::steady_clock::now());
}

int absorb_origin(tapgpu_ctx* c, const int* herr, const unsigned long long* hc, int Vn,
                  std::chrono::steady_clock::time_point t0) {
  c->stats.origin_expansions += (int64_t)hc[0];
  c->stats.origin_scanned += (int64_t)hc[1];
  c->stats.origin_updates += (int64_t)hc[2];
  c->stats.origin_relax_rounds += (int64_t)hc[3];
  c->stats.origin_load_rounds += (int64_t)hc[4];
  c->stats.origin_loaded += (int64_t)hc[5];
  c->stats.origin_walk_steps += (int64_t)hc[6];
  c->stats.origin_deferred += (int64_t)hc[7];
  if (c->clock_khz > 0) {
    c->stats.origin_relax_ms += (double)hc[8] / c->clock_khz;
    c->stats.origin_pred_ms += (double)hc[9] / c->clock_khz;
    c->stats.origin_load_ms += (double)hc[10] / c->clock_khz;
  }
  if (c->profiling && c->Z > 0) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->stats.origin_kernel_ms += ms;
    c->stats.origin_launches += 1;
  }
  c->stats.origin_blocks += (int64_t)c->Z * Vn;
  c->stats.origin_chase = c->chase;
  c->stats.origin_block = use_lds(c) ? lds_threads(c) : 0;
  c->stats.origin_queue = use_lds(c) ? c->qcap : c->hqcap;
  c->stats.origin_split = c->split;
  c->stats.aon_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (herr[0] == 1)
    return c->fail(TAPGPU_E_NOPATH, "no path from %d to %d", herr[1] + 1, herr[2] + 1);
  if (herr[0] == 2)
    return c->fail(TAPGPU_E_LIMIT, "label-correcting rounds exceeded from origin %d (negative costs?)",
                   herr[1] + 1);
  if (herr[0] == 3)
    return c->fail(TAPGPU_E_LIMIT, "predecessor cycle in the tree of origin %d", herr[1] + 1);
  return TAPGPU_OK;
}

// After a loading that recorded each origin's relaxation rounds: the
// origins sorted by decreasing rounds (then id), the dispatch order of later
// LDS loadings.  The hardware starts workgroups in block order as slots
// free up, so the longest searches start first and the short ones fill the
// tail.  Results do not depend on the order.
int origin_order_update(tapgpu_ctx* c) {
  if (!c->origin_recorded || c->origin_order_ready || c->capturing) return TAPGPU_OK;
  std::vector<int> work(c->Z), order(c->Z);
  HIPCHK(c, hipMemcpyAsync(work.data(), c->d_origin_work, sizeof(int) * c->Z, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int z = 0; z < c->Z; ++z) order[z] = z;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
  int rc;
  if (!c->d_origin_order && (rc = dalloc(c, &c->d_origin_order, (size_t)c->Z))) return rc;
  if ((rc = upload(c, c->d_origin_order, order.data(), (size_t)c->Z))) return rc;
  c->h_origin_order = std::move(order);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->origin_order_ready = true;
  return TAPGPU_OK;
}

int aon_lds(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on) {
  auto t0 = std::chrono::steady_clock::now();
  if (int rc = origin_order_update(c)) return rc;
  if (!c->pre_cleared) HIPCHK(c, hipMemsetAsync(c->err_dev, 0, sizeof(int) * 3, c->stream));
  if (c->lds_block < 0 && c->Z > 0) {
    // workgroup and queue size first, then the chain length, on the LDS
    // kernel alone
    const int keep = c->split;
    c->split = 0;
    int rc = tune_lds_shape(c, v0, Vn, var_on);
    c->split = keep;
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->err_dev, 0, sizeof(int) * 3, c->stream));
  }
  if (c->chase < 0 && c->Z > 0) {
    const int keep = c->split;
    c->split = 0;
    int rc = tune_chase(c, v0, Vn, var_on);
    c->split = keep;
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->err_dev, 0, sizeof(int) * 3, c->stream));
  }
  if (c->split < 0 && c->Z > 0) {
    int rc = tune_split(c, v0, Vn, var_on);
    if (rc) return rc;
    HIPCHK(c, hipMemsetAsync(c->err_dev, 0, sizeof(int) * 3, c->stream));
  }
  if (!c->pre_cleared)
    HIPCHK(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long) * 16, c->stream));
  if (c->Z > 0) {
    if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    int rc = launch_origin(c, v0, Vn, var_on, nullptr);
    if (rc) return rc;
    if (c->profiling) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  }
  return finish_origin(c, Vn, t0);
}

// Per-workgroup scratch rows of the labels-in-HBM kernel (allocated on
// first use: hbm_grid rows of N labels, predecessors and tails).
int ensure_hbm(tapgpu_ctx* c) {
  if (c->h_dist) return TAPGPU_OK;
  const size_t n = (size_t)c->hbm_grid * c->N;
  int rc;
  if ((rc = dalloc(c, &c->d_hbm_next, 1))) return rc;
  if ((rc = dalloc(c, &c->h_dist, n))) return rc;
  if ((rc = dalloc(c, &c->h_pred, n))) return rc;
  if ((rc = dalloc(c, &c->h_ptail, n))) return rc;
  return TAPGPU_OK;
}

// Shortest paths + loading of variants [v0, v0+Vn) with the labels-in-HBM
// kernel (persistent grid over the Z * Vn (origin, variant) items).
//
// Bucket-width tuning: the first loading on a context with delta_mult < 0
// runs one wave of workgroups (the first hbm_grid items) for each candidate
// width, in units of the mean free-flow link time, and keeps the one with
// the smallest busy time summed over the workgroups; the full launch
// follows.  Results do not depend on the width.
constexpr double kDeltaCandidates[] = {0.0, 1.0, 2.0, 4.0, 8.0, 16.0};

int launch_hbm(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on, int items_cap,
               float* ms) {
  if (!c->pre_cleared)
    HIPCHK(c, hipMemsetAsync(c->y + (size_t)v0 * c->E, 0, sizeof(double) * (size_t)Vn * c->E, c->stream));
  if (c->Z == 0) return TAPGPU_OK;
  OriginArgs A = make_origin_args(c, var_on);
  A.fww += (size_t)v0 * c->E; A.rvw += (size_t)v0 * c->E; A.y += (size_t)v0 * c->E;
  if (var_on) A.var_on = var_on + v0;
  const int items = std::min(c->Z * Vn, items_cap);
  const int G = std::min(c->hbm_grid, items);
  if (int rc = ensure_hbm(c)) return rc;        // (the item counter; no allocation once made)
  if (c->origin_order_ready) {
    A.order = c->d_origin_order;
    A.vn = Vn;
  } else if (!c->capturing && c->Z > 1) {
    if (!c->d_origin_work) {
      int rc = ensure_origin_work(c);
      if (rc) return rc;
    }
    A.work = c->d_origin_work;
    c->origin_recorded = true;
  }
  A.next = c->d_hbm_next;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, c->stream, c->d_hbm_next);
  if (ms) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  hipLaunchKernelGGL(k_origin_hbm, dim3(G), dim3(kHbmBlock), c->hbm_lds, c->stream, A, 1, items);
  HIPCHK(c, hipGetLastError());
  if (ms) {
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[1]));
    HIPCHK(c, hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
  }
  return TAPGPU_OK;
}

// Busy time of a tuning launch: the workgroups' phase clocks (relaxation,
// predecessors, loading), summed over its tasks.  At full occupancy the
// smallest sum is the highest throughput; one wave's wall time would be set
// by its slowest task alone.
int tuning_work(tapgpu_ctx* c, double* work, int* herr) {
  unsigned long long hc[16];
  HIPCHK(c, hipMemcpyAsync(hc, c->counters, sizeof hc, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(herr, c->err_dev, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *work = (double)hc[8] + (double)hc[9] + (double)hc[10];
  return TAPGPU_OK;
}

int tune_delta(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on) {
  if (!c->ev[0])
    for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
  float best_ms = 0.f;
  double best = 4.0, best_work = 0.0;
  for (double m : kDeltaCandidates) {
    c->delta_mult = m;
    HIPCHK(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long) * 16, c->stream));
    float ms = 0.f;
    int rc = launch_hbm(c, v0, Vn, var_on, c->hbm_grid, &ms);
    if (rc) return rc;
    int herr = 0;
    double work = 0.0;
    if ((rc = tuning_work(c, &work, &herr))) return rc;
    if (herr) break;
    if (best_work == 0.0 || work < best_work) { best_work = work; best_ms = ms; best = m; }
  }
  c->delta_mult = best;
  c->stats.origin_tune_ms += best_ms;
  HIPCHK(c, hipMemsetAsync(c->err_dev, 0, sizeof(int) * 3, c->stream));
  HIPCHK(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long) * 16, c->stream));
  return TAPGPU_OK;
}

int aon_hbm(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on) {
  auto t0 = std::chrono::steady_clock::now();
  int rc = ensure_hbm(c);
  if (rc) return rc;
  if ((rc = origin_order_update(c))) return rc;
  if (!c->pre_cleared) HIPCHK(c, hipMemsetAsync(c->err_dev, 0, sizeof(int) * 3, c->stream));
  if (c->delta_mult < 0 && c->Z > 0 && (rc = tune_delta(c, v0, Vn, var_on))) return rc;
  if (!c->pre_cleared)
    HIPCHK(c, hipMemsetAsync(c->counters, 0, sizeof(unsigned long long) * 16, c->stream));
  if (c->profiling && c->Z > 0) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  if ((rc = launch_hbm(c, v0, Vn, var_on, INT32_MAX, nullptr))) return rc;
  if (c->profiling && c->Z > 0) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  c->stats.origin_delta = c->delta_mult;
  return finish_origin(c, Vn, t0);
}

#include "taplanes.inc"

// The lanes kernel loads a batch when at least lanes_min of its variants
// are active (lanes_min 0: never).
bool use_lanes(tapgpu_ctx* c, int v0, int Vn, const unsigned char* h_on) {
  if (c->lanes_min <= 0 || !c->lanes_ok) return false;
  int n = 0;
  for (int v = v0; v < v0 + Vn; ++v) n += (!h_on || h_on[v]) ? 1 : 0;
  return n >= c->lanes_min;
}

// All-or-nothing loading of variants [v0, v0 + Vn) (groups of those
// variants) with the current costs, by the configured strategy.  h_on is
// the host copy of var_on (null: every variant of the range).
int load_per_origin(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on) {
  if (use_lds(c)) return aon_lds(c, v0, Vn, var_on);
  if (use_hbm(c)) return aon_hbm(c, v0, Vn, var_on);
  int rc = sssp_groups(c, v0 * c->Gv, Vn * c->Gv, c->col_src, c->grp_var, var_on);
  if (!rc) rc = aon_groups(c, v0 * c->Gv, Vn * c->Gv, v0, Vn, var_on);
  return rc;
}

// A loading redone by the per-origin strategy after the lanes kernel met a
// level tie: from cleared volumes, error flag and counters.
int redo_per_origin(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on) {
  const bool pc = c->pre_cleared;
  c->pre_cleared = false;
  const int rc = load_per_origin(c, v0, Vn, var_on);
  c->pre_cleared = pc;
  return rc;
}

int load_variants(tapgpu_ctx* c, int v0, int Vn, const unsigned char* var_on,
                  const unsigned char* h_on) {
  if (use_lanes(c, v0, Vn, h_on)) {
    bool fallback = false;
    int rc = aon_lanes(c, v0, Vn, var_on, h_on, &fallback);
    if (rc || !fallback) return rc;
    return redo_per_origin(c, v0, Vn, var_on);
  }
  return load_per_origin(c, v0, Vn, var_on);
}

int launch_costs(tapgpu_ctx* c, int mode, const unsigned char* var_on, int v0, int Vn) {
  hipLaunchKernelGGL(k_costs, dim3(grid_for((long long)c->E * Vn / kWave + 1)), dim3(kBlock), 0,
                     c->stream, make_links(c), mode, var_on, v0, Vn);
  HIPCHK(c, hipGetLastError());
  return TAPGPU_OK;
}

// Free-flow shortest-path distances from one node (binary-heap Dijkstra on
// the host; zones other than the start are labelled but not expanded).
std::vector<double> host_dijkstra(int N, int ft0, const std::vector<int>& off,
                                  const std::vector<int>& dst, const std::vector<double>& w,
                                  int s) {
  std::vector<double> d(N, INFINITY);
  std::vector<std::pair<double, int>> heap;
  d[s] = 0.0;
  heap.push_back({0.0, s});
  auto cmp = [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
    return a.first > b.first;
  };
  while (!heap.empty()) {
    std::pop_heap(heap.begin(), heap.end(), cmp);
    const auto [du, u] = heap.back();
    heap.pop_back();
    if (du > d[u] || (u < ft0 && u != s)) continue;
    for (int e = off[u]; e < off[u + 1]; ++e) {
      const double nd = du + w[e];
      if (nd < d[dst[e]]) {
        d[dst[e]] = nd;
        heap.push_back({nd, dst[e]});
        std::push_heap(heap.begin(), heap.end(), cmp);
      }
    }
  }
  return d;
}

// Column order of the zones: origins that are close in the network share a
// 64-lane group, so their wavefronts reach each node in nearby rounds.  The
// "coordinates" of a zone are its free-flow distances from two far-apart
// landmark nodes; zones are sorted along the Morton curve of those.  Results
// do not depend on the order (each column is an independent search and the
// group partial volumes are exact for integer demand).
std::vector<int> landmark_zone_order(int N, int Z, int ft0, const std::vector<int>& off,
                                     const std::vector<int>& dst, const std::vector<double>& w) {
  std::vector<int> order(Z);
  for (int z = 0; z < Z; ++z) order[z] = z;
  if (Z <= kWave) return order;
  auto farthest = [&](const std::vector<double>& d) {
    int best = 0;
    double bd = -1;
    for (int v = 0; v < N; ++v)
      if (std::isfinite(d[v]) && d[v] > bd) { bd = d[v]; best = v; }
    return best;
  };
  const int l1 = farthest(host_dijkstra(N, ft0, off, dst, w, 0));
  const std::vector<double> d1 = host_dijkstra(N, ft0, off, dst, w, l1);
  const int l2 = farthest(d1);
  const std::vector<double> d2 = host_dijkstra(N, ft0, off, dst, w, l2);
  double m1 = 0, m2 = 0;
  for (int z = 0; z < Z; ++z) {
    if (std::isfinite(d1[z])) m1 = std::max(m1, d1[z]);
    if (std::isfinite(d2[z])) m2 = std::max(m2, d2[z]);
  }
  auto q = [](double x, double m) -> uint32_t {
    if (!std::isfinite(x) || m <= 0) return 0xFFFFu;
    return (uint32_t)std::min(65535.0, std::floor(x / m * 65535.0));
  };
  auto spread = [](uint32_t x) -> uint64_t {
    uint64_t r = 0;
    for (int b = 0; b < 16; ++b) r |= (uint64_t)((x >> b) & 1u) << (2 * b);
    return r;
  };
  std::vector<uint64_t> key(Z);
  for (int z = 0; z < Z; ++z) key[z] = spread(q(d1[z], m1)) | (spread(q(d2[z], m2)) << 1);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  return order;
}

int build(tapgpu_ctx* c, const tapgpu_network* net, const tapgpu_demand* dem, int64_t V,
          const uint8_t* present, const double* cap_v, const double* t0_v, const double* B_v,
          const double* pw_v) {
  if (!net || !dem) return c->fail(TAPGPU_E_ARG, "null network or demand");
  const int64_t N = net->num_nodes, E = net->num_links, Z = net->num_zones;
  if (N < 1 || E < 0 || Z < 0 || Z > N || V < 1)
    return c->fail(TAPGPU_E_ARG, "bad sizes N=%lld E=%lld Z=%lld V=%lld", (long long)N,
                   (long long)E, (long long)Z, (long long)V);
  if (N >= (1LL << 30) || E >= (1LL << 30))
    return c->fail(TAPGPU_E_LIMIT, "network too large for 32-bit indices");
  if (dem->num_zones != Z)
    return c->fail(TAPGPU_E_ARG, "demand has %lld zones, network %lld", (long long)dem->num_zones,
                   (long long)Z);
  c->N = (int)N; c->E = (int)E; c->Z = (int)Z; c->V = (int)V;
  int64_t ft = net->first_thru_node;
  if (ft < 1) ft = 1;
  if (ft > N + 1) ft = N + 1;
  c->ft0 = (int)(ft - 1);
  c->Gv = (int)((std::max<int64_t>(Z, 1) + kWave - 1) / kWave);
  c->Pv = (long long)c->Gv * kWave;
  c->Galloc = c->Gv * c->V;
  c->P = c->Pv * c->V;
  if ((unsigned long long)N * (unsigned long long)c->Galloc >= (1ull << 32))
    return c->fail(TAPGPU_E_LIMIT, "nodes x origin groups exceeds 2^32");

  // --- graph (host build) ---
  std::vector<int> fw_off(N + 1, 0), fw_dst(E), rv_off(N + 1, 0), rv_src(E), rv_eid(E), tail(E),
      rvpos(E);
  for (int64_t k = 0; k < E; ++k) {
    const int64_t a = net->init_node[k], b = net->term_node[k];
    if (a < 1 || a > N || b < 1 || b > N)
      return c->fail(TAPGPU_E_ARG, "link %lld has node out of range (%lld, %lld)", (long long)k,
                     (long long)a, (long long)b);
    if (k > 0 && a < net->init_node[k - 1])
      return c->fail(TAPGPU_E_ARG, "links not grouped by init node at %lld", (long long)k);
    fw_off[a]++;  // count into slot a (0-based a-1 -> prefix below)
    rv_off[b]++;
    fw_dst[k] = (int)(b - 1);
    tail[k] = (int)(a - 1);
  }
  for (int64_t v = 0; v < N; ++v) {
    fw_off[v + 1] += fw_off[v];
    rv_off[v + 1] += rv_off[v];
  }
  {
    std::vector<int> fill(rv_off.begin(), rv_off.end() - 1);
    for (int64_t k = 0; k < E; ++k) {   // stable: in-links of v in link order
      const int b = fw_dst[k];
      const int pos = fill[b]++;
      rv_src[pos] = tail[k];
      rv_eid[pos] = (int)k;
      rvpos[k] = pos;
    }
  }
  // --- origin columns in landmark order ---
  {
    std::vector<double> w0(E);
    for (int64_t k = 0; k < E; ++k)
      w0[k] = (present && !present[k]) ? INFINITY : (t0_v ? t0_v[k] : net->free_flow_time[k]);
    c->zone_order = landmark_zone_order(c->N, c->Z, c->ft0, fw_off, fw_dst, w0);
  }
  std::vector<int> col_of_zone(Z);
  for (int64_t j = 0; j < Z; ++j) col_of_zone[c->zone_order[j]] = (int)j;
  // --- demand: dense per zone row in column order ---
  std::vector<double> odc((size_t)std::max<int64_t>(Z, 1) * c->Pv, 0.0);
  double tot = 0;
  for (int64_t o = 1; o <= Z; ++o) {
    for (int64_t i = dem->od_start[o - 1]; i < dem->od_start[o]; ++i) {
      const int64_t d = dem->od_dest[i];
      const double f = dem->od_demand[i];
      tot += f;
      if (d == o || !(f > 0)) continue;
      if (d < 1 || d > Z)
        return c->fail(TAPGPU_E_ARG, "demand destination %lld is not a zone", (long long)d);
      odc[(size_t)(d - 1) * c->Pv + col_of_zone[o - 1]] += f;
    }
  }
  c->total_demand = tot;
  std::vector<long long> h_od_start(Z + 1, 0);
  std::vector<int> h_od_dest;
  std::vector<double> h_od_val;
  for (int64_t o = 1; o <= Z; ++o) {
    for (int64_t i = dem->od_start[o - 1]; i < dem->od_start[o]; ++i) {
      const int64_t d = dem->od_dest[i];
      const double f = dem->od_demand[i];
      if (d == o || !(f > 0)) continue;
      h_od_dest.push_back((int)(d - 1));
      h_od_val.push_back(f);
    }
    h_od_start[o] = (long long)h_od_dest.size();
  }
  // one-origin-per-workgroup path: labels must fit LDS, link ids and CSR
  // offsets 16-bit
  {
    int dev = 0, lds_max = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      lds_max = 0;
    c->lds_bytes = origin_lds_bytes(c->N, 0);
    if (hipDeviceGetAttribute(&c->clock_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
      c->clock_khz = 0;
    c->lds_ok = c->lds_bytes + 128 <= (size_t)lds_max && E < 65535 && N <= kLdsMaxNodes;
    if (c->lds_ok) {
      // the rest of the LDS, up to kLdsQueueMax entries, holds the frontier queues
      const size_t spare = (size_t)lds_max - 128 - c->lds_bytes;
      c->qcap_max = (int)std::min<size_t>(kLdsQueueMax, spare / 4) & ~63;
      c->qcap = c->qcap_max;
      c->lds_bytes = origin_lds_bytes(c->N, c->qcap_max);
    }
    c->hbm_lds = origin_hbm_lds_bytes(N, 0);
    c->hbm_ok = c->hbm_lds + 256 <= (size_t)lds_max;
    if (c->hbm_ok) {
      // queues from the LDS share of one of the kHbmPerCu workgroups on a CU
      int lds_cu = 0;
      if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
              hipSuccess || lds_cu <= 0)
        lds_cu = lds_max;
      const size_t share = std::min<size_t>((size_t)lds_cu / kHbmPerCu, (size_t)lds_max) - 256;
      const size_t spare = share > c->hbm_lds ? share - c->hbm_lds : 0;
      c->hqcap_max = (int)std::min<size_t>(kHbmQueueMax, spare / 24) & ~63;
      c->hqcap = c->hqcap_max;
      c->hbm_lds = origin_hbm_lds_bytes(N, c->hqcap_max);
    }
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) ncu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_origin_hbm, kHbmBlock, c->hbm_lds) !=
        hipSuccess || per_cu < 1)
      per_cu = 1;
    c->hbm_grid = std::max(1, ncu * per_cu);
    // lanes kernel: three bitmaps, then queues from the LDS share of the
    // kLanesPerCu workgroups on a CU
    {
      int lds_cu = 0;
      if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
              hipSuccess || lds_cu <= 0)
        lds_cu = lds_max;
      const size_t base = lanes_lds_bytes(c->N, 0);
      c->lanes_ok = base + 256 <= (size_t)lds_max;
      if (c->lanes_ok) {
        const size_t share = std::min<size_t>((size_t)lds_cu / kLanesPerCu, (size_t)lds_max) - 256;
        const size_t spare = share > base ? share - base : 0;
        c->lanes_qcap_max = (int)std::min<size_t>(kLanesQueueMax, spare / 24) & ~63;
        c->lanes_qcap = c->lanes_qcap_max;
        c->lanes_lds = lanes_lds_bytes(c->N, c->lanes_qcap_max);
        int pc = 0;
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_origin_lanes),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lanes_lds) !=
                hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, k_origin_lanes, kLanesBlock,
                                                          c->lanes_lds) != hipSuccess ||
            pc < 1)
          pc = 1;
        c->lanes_grid_max = std::max(1, ncu * pc);
      }
      c->lanes_min = V > 1 ? kLanesMinDefault : 0;
    }
    double sum = 0;
    int cnt = 0;
    for (int64_t k = 0; k < E; ++k) {
      const double t = (t0_v ? t0_v[k] : net->free_flow_time[k]);
      const bool on = !present || present[k];
      if (on && t > 0 && std::isfinite(t)) { sum += t; ++cnt; }
    }
    c->mean_t0 = cnt ? sum / cnt : 1.0;
  }

  // --- link records per variant (capacity fix-up per variant) ---
  const size_t VE = (size_t)V * E;
  std::vector<double> hcap(VE), ht0(VE), hB(VE), hpw(VE);
  std::vector<unsigned char> hpres(VE, 1);
  for (int64_t v = 0; v < V; ++v) {
    double minpos = DBL_MAX;
    for (int64_t k = 0; k < E; ++k) {
      const size_t j = (size_t)v * E + k;
      hcap[j] = cap_v ? cap_v[j] : net->capacity[k];
      ht0[j] = t0_v ? t0_v[j] : net->free_flow_time[k];
      hB[j] = B_v ? B_v[j] : net->B[k];
      hpw[j] = pw_v ? pw_v[j] : net->power[k];
      if (present) hpres[j] = present[j] ? 1 : 0;
      if (hpres[j] && hcap[j] > 0.0 && hcap[j] < minpos) minpos = hcap[j];
      // the kernels order labels by their bit patterns, which holds for
      // non-negative costs only: BPR costs are non-negative when t0 and B are
      if (hpres[j] && !(ht0[j] >= 0.0 && hB[j] >= 0.0))
        return c->fail(TAPGPU_E_ARG, "link %lld of variant %lld: negative or NaN free-flow time or B",
                       (long long)k, (long long)v);
    }
    for (int64_t k = 0; k < E; ++k) {
      const size_t j = (size_t)v * E + k;
      if (!(hcap[j] > 0.0)) hcap[j] = minpos;
    }
  }
  std::vector<int> col_src(c->P, -1), grp_var(c->Galloc);
  for (int v = 0; v < c->V; ++v) {
    for (int64_t j = 0; j < Z; ++j) col_src[(size_t)v * c->Pv + j] = c->zone_order[j];
    for (int g = 0; g < c->Gv; ++g) grp_var[v * c->Gv + g] = v;
  }

  // --- device ---
  HIPCHK(c, hipGetDevice(&c->device));
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  HIPCHK(c, hipHostMalloc((void**)&c->h_int, 64 * sizeof(int), hipHostMallocDefault));
  {
    const size_t bytes = 16 + 128 + sizeof(VarState) * (size_t)V + (size_t)V;   // as iter_dev
    HIPCHK(c, hipHostMalloc(&c->pin_block, bytes, hipHostMallocDefault));
    unsigned char* b = (unsigned char*)c->pin_block;
    c->pin_err = (int*)b;
    c->pin_cnt = (unsigned long long*)(b + 16);
    c->pin_st = (VarState*)(b + 144);
    c->pin_on = b + 144 + sizeof(VarState) * (size_t)V;
  }
  int rc;
#define ALLOC(ptr, n) if ((rc = dalloc(c, &c->ptr, (n)))) return rc
  ALLOC(fw_off, N + 1); ALLOC(fw_dst, E); ALLOC(rv_off, N + 1); ALLOC(rv_src, E);
  ALLOC(rv_eid, E); ALLOC(tail, E); ALLOC(rvpos, E);
  ALLOC(t0, VE); ALLOC(cap, VE); ALLOC(B, VE); ALLOC(pw, VE); ALLOC(present, VE);
  ALLOC(x, VE); ALLOC(y, VE); ALLOC(c, VE); ALLOC(rvw, VE); ALLOC(fww, VE);
  ALLOC(col_src, c->P); ALLOC(grp_var, c->Galloc);
  c->flags_cap = (int)std::min<int64_t>(std::max<int64_t>(4 * N + 64, 1024), 1 << 24);
  ALLOC(flags, c->flags_cap);
  {
    c->iter_bytes = 16 + 128 + sizeof(VarState) * (size_t)V + (size_t)V;
    if ((rc = dalloc(c, &c->iter_dev, c->iter_bytes))) return rc;
    c->err_dev = (int*)c->iter_dev;
    c->counters = (unsigned long long*)(c->iter_dev + 16);
    c->st = (VarState*)(c->iter_dev + 144);
    c->var_on = c->iter_dev + 144 + sizeof(VarState) * (size_t)V;
  }


  ALLOC(part, (size_t)V * kLinkBlocks * kQ);
  ALLOC(prm, 1);
#undef ALLOC
  if ((rc = upload(c, c->fw_off, fw_off.data(), N + 1))) return rc;
  if ((rc = upload(c, c->fw_dst, fw_dst.data(), E))) return rc;
  if ((rc = upload(c, c->rv_off, rv_off.data(), N + 1))) return rc;
  if ((rc = upload(c, c->rv_src, rv_src.data(), E))) return rc;
  if ((rc = upload(c, c->rv_eid, rv_eid.data(), E))) return rc;
  if ((rc = upload(c, c->tail, tail.data(), E))) return rc;
  if ((rc = upload(c, c->rvpos, rvpos.data(), E))) return rc;
  if ((rc = upload(c, c->t0, ht0.data(), VE))) return rc;
  if ((rc = upload(c, c->cap, hcap.data(), VE))) return rc;
  if ((rc = upload(c, c->B, hB.data(), VE))) return rc;
  if ((rc = upload(c, c->pw, hpw.data(), VE))) return rc;
  if ((rc = upload(c, c->present, hpres.data(), VE))) return rc;
  if ((rc = upload(c, c->col_src, col_src.data(), c->P))) return rc;
  if ((rc = upload(c, c->grp_var, grp_var.data(), c->Galloc))) return rc;
  c->h_odc = std::move(odc);
  if ((rc = dalloc(c, &c->od_start, Z + 1))) return rc;
  if ((rc = dalloc(c, &c->od_dest, h_od_dest.size()))) return rc;
  if ((rc = dalloc(c, &c->od_val, h_od_val.size()))) return rc;
  if ((rc = upload(c, c->od_start, h_od_start.data(), Z + 1))) return rc;
  if ((rc = upload(c, c->od_dest, h_od_dest.data(), h_od_dest.size()))) return rc;
  if ((rc = upload(c, c->od_val, h_od_val.data(), h_od_val.size()))) return rc;
  if (c->lds_ok) {
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_origin_lds<1024>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_bytes));
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_origin_lds<512>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_bytes));
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(k_origin_lds<256>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->lds_bytes));
  }

  c->h_od_start = std::move(h_od_start);
  c->h_od_dest = std::move(h_od_dest);
  c->h_od_val = std::move(h_od_val);
  c->h_tail = std::move(tail);
  c->h_head = fw_dst;

  HIPCHK(c, hipMemsetAsync(c->x, 0, sizeof(double) * VE, c->stream));
  HIPCHK(c, hipMemsetAsync(c->y, 0, sizeof(double) * VE, c->stream));
  HIPCHK(c, hipMemsetAsync(c->var_on, 1, V, c->stream));
  // Wa = free-flow time (trunk/tap/tap_functions.c:331)
  if ((rc = launch_costs(c, 0, nullptr, 0, c->V))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TAPGPU_OK;
}

int check_variant(tapgpu_ctx* c, int64_t v) {
  if (v < 0 || v >= c->V) return c->fail(TAPGPU_E_ARG, "variant %lld out of range", (long long)v);
  return TAPGPU_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int tapgpu_create_batched(tapgpu_ctx** out, const tapgpu_network* net, const tapgpu_demand* dem,
                          int64_t num_variants, const uint8_t* present, const double* capacity_v,
                          const double* free_flow_time_v, const double* B_v,
                          const double* power_v) {
  if (!out) return TAPGPU_E_ARG;
  *out = nullptr;
  tapgpu_ctx* c = new (std::nothrow) tapgpu_ctx();
  if (!c) return TAPGPU_E_NOMEM;
  const int rc = build(c, net, dem, num_variants, present, capacity_v, free_flow_time_v, B_v,
                       power_v);
  if (rc) {
    g_create_error = c->err;
    tapgpu_destroy(c);
    return rc;
  }
  *out = c;
  return TAPGPU_OK;
}

int tapgpu_create(tapgpu_ctx** out, const tapgpu_network* net, const tapgpu_demand* dem) {
  return tapgpu_create_batched(out, net, dem, 1, nullptr, nullptr, nullptr, nullptr, nullptr);
}

void tapgpu_destroy(tapgpu_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->g_iter) (void)hipGraphExecDestroy(c->g_iter);
  if (c->g_link) (void)hipGraphExecDestroy(c->g_link);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->trace) (void)hipFree(c->trace);
  if (c->h_int) (void)hipHostFree(c->h_int);
  if (c->pin_block) (void)hipHostFree(c->pin_block);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  delete c;
}

const char* tapgpu_last_error(const tapgpu_ctx* c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

int64_t tapgpu_num_variants(const tapgpu_ctx* c) { return c ? c->V : 0; }

int tapgpu_set_costs(tapgpu_ctx* c, int64_t variant, const double* costs) {
  if (!c || !costs) return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  for (int k = 0; k < c->E; ++k)
    if (!(costs[k] >= 0.0))
      return c->fail(TAPGPU_E_ARG, "link %d: negative or NaN cost", k);
  if ((rc = upload(c, c->c + (size_t)variant * c->E, costs, c->E))) return rc;
  if ((rc = launch_costs(c, 2, nullptr, (int)variant, 1))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TAPGPU_OK;
}

int tapgpu_sssp(tapgpu_ctx* c, int64_t variant, const int64_t* start_nodes, int64_t num_start,
                double* dist, int64_t* pred, int64_t ld, int64_t col_base) {
  if (!c || (num_start > 0 && !start_nodes) || num_start < 0 || col_base < 0)
    return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  // row v starts at v * ld; with pape_cuda's layout (ld = num_start,
  // col_base = 1) the last column of row v is the unused column 0 of v + 1
  if (ld < num_start || (col_base > 0 && ld < num_start + col_base - 1))
    return c->fail(TAPGPU_E_ARG, "ld %lld too small for %lld columns", (long long)ld,
                   (long long)num_start);
  for (int64_t i = 0; i < num_start; ++i)
    if (start_nodes[i] < 1 || start_nodes[i] > c->N)
      return c->fail(TAPGPU_E_ARG, "start node %lld out of range", (long long)start_nodes[i]);
  const int64_t N1 = (int64_t)c->N + 1;
  // node 0 does not exist: row 0 stays unreachable
  for (int64_t i = 0; i < num_start; ++i) {
    if (dist) dist[col_base + i] = DBL_MAX;
    if (pred) pred[col_base + i] = -1;
  }
  if (use_lds(c) || use_hbm(c)) {
    // one workgroup per start node (labels in LDS), or the persistent
    // labels-in-HBM grid; chunks of up to 1024 starts
    const bool hbm = !use_lds(c);
    if (hbm && (rc = ensure_hbm(c))) return rc;
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(num_start, 1), 1024);
    int* d_src = nullptr;
    double* d_dist = nullptr;
    int* d_pred = nullptr;
    if ((rc = dalloc(c, &d_src, chunk))) return rc;
    if ((rc = dalloc(c, &d_dist, (size_t)c->N * chunk))) return rc;
    if ((rc = dalloc(c, &d_pred, (size_t)c->N * chunk))) return rc;
    std::vector<int> hs(chunk);
    std::vector<double> hd((size_t)c->N * chunk);
    std::vector<int> hp((size_t)c->N * chunk);
    for (int64_t s0 = 0; !rc && s0 < num_start; s0 += chunk) {
      const int64_t ns = std::min<int64_t>(chunk, num_start - s0);
      for (int64_t j = 0; j < ns; ++j) hs[j] = (int)(start_nodes[s0 + j] - 1);
      if ((rc = upload(c, d_src, hs.data(), ns))) break;
      // a failed launch must not hand back stale links: predecessors start
      // at -1 and the error flag at 0 for every chunk
      if (hipMemsetAsync(d_pred, 0xFF, sizeof(int) * (size_t)c->N * chunk, c->stream) != hipSuccess ||
          hipMemsetAsync(c->err_dev, 0, sizeof(int) * 3, c->stream) != hipSuccess) {
        rc = c->fail(TAPGPU_E_HIP, "clearing the shortest-path outputs failed");
        break;
      }
      OriginArgs A = make_origin_args(c, nullptr);
      A.fww += (size_t)variant * c->E; A.rvw += (size_t)variant * c->E;
      A.src_list = d_src; A.out_dist = d_dist; A.out_pred = d_pred; A.ld = chunk;
      if (hbm)
        hipLaunchKernelGGL(k_origin_hbm, dim3((unsigned)std::min<int64_t>(ns, c->hbm_grid)),
                           dim3(kHbmBlock), c->hbm_lds, c->stream, A, 0, (int)ns);
      else
        launch_lds(c, dim3((unsigned)ns, 1), c->stream, A, 0);
      int herr[3] = {0, 0, 0};
      const hipError_t e0 = hipGetLastError();
      const hipError_t e1 = hipMemcpyAsync(hd.data(), d_dist, sizeof(double) * c->N * chunk,
                                           hipMemcpyDeviceToHost, c->stream);
      const hipError_t e2 = hipMemcpyAsync(hp.data(), d_pred, sizeof(int) * c->N * chunk,
                                           hipMemcpyDeviceToHost, c->stream);
      const hipError_t e4 = hipMemcpyAsync(herr, c->err_dev, sizeof herr, hipMemcpyDeviceToHost,
                                           c->stream);
      const hipError_t e3 = hipStreamSynchronize(c->stream);
      if (e0 != hipSuccess || e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess ||
          e4 != hipSuccess) {
        const hipError_t e = e0 != hipSuccess ? e0 : e1 != hipSuccess ? e1 : e2 != hipSuccess ? e2
                             : e4 != hipSuccess ? e4 : e3;
        rc = c->fail(TAPGPU_E_HIP, "shortest-path launch failed: %s", hipGetErrorString(e));
        break;
      }
      if (herr[0] == 2) {
        rc = c->fail(TAPGPU_E_LIMIT, "label-correcting rounds exceeded from origin %d (negative costs?)",
                     herr[1] + 1);
        break;
      }
      if (herr[0] == 3) {
        rc = c->fail(TAPGPU_E_LIMIT, "predecessor cycle in the tree of origin %d", herr[1] + 1);
        break;
      }
      if (herr[0] != 0) {
        rc = c->fail(TAPGPU_E_LIMIT, "shortest-path kernel reported error %d at origin %d", herr[0],
                     herr[1] + 1);
        break;
      }
      for (int64_t v = 1; v < N1; ++v)
        for (int64_t j = 0; j < ns; ++j) {
          const size_t o = (size_t)v * ld + col_base + s0 + j;
          if (dist) dist[o] = hd[(size_t)(v - 1) * chunk + j];
          if (pred) pred[o] = hp[(size_t)(v - 1) * chunk + j];
        }
    }
    for (void* q : {(void*)d_src, (void*)d_dist, (void*)d_pred}) {
      c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), q), c->allocs.end());
      (void)hipFree(q);
    }
    return rc;
  }
  // chunks of the variant's columns; groups [0, Gv) of a private mapping
  int* d_col = nullptr;
  int* d_gv = nullptr;
  if ((rc = dalloc(c, &d_col, c->Pv))) return rc;
  if ((rc = dalloc(c, &d_gv, c->Gv))) return rc;
  std::vector<int> hgv(c->Gv, (int)variant), hcol(c->Pv);
  std::vector<double> hd((size_t)c->N * c->Pv);
  std::vector<int> hp((size_t)c->N * c->Pv);
  rc = upload(c, d_gv, hgv.data(), c->Gv);
  for (int64_t s0 = 0; !rc && s0 < num_start; s0 += c->Pv) {
    const int64_t ns = std::min<int64_t>(c->Pv, num_start - s0);
    for (int64_t j = 0; j < c->Pv; ++j) hcol[j] = j < ns ? (int)(start_nodes[s0 + j] - 1) : -1;
    if ((rc = upload(c, d_col, hcol.data(), c->Pv))) break;
    const int Gn = (int)((ns + kWave - 1) / kWave);
    if ((rc = sssp_groups(c, 0, Gn, d_col, d_gv, nullptr))) break;
    // rows are P apart on the device; copy the first Pv columns of each row
    const hipError_t e1 = hipMemcpy2DAsync(hd.data(), c->Pv * sizeof(double), c->dist,
                                           c->P * sizeof(double), c->Pv * sizeof(double), c->N,
                                           hipMemcpyDeviceToHost, c->stream);
    const hipError_t e2 = hipMemcpy2DAsync(hp.data(), c->Pv * sizeof(int), c->pred,
                                           c->P * sizeof(int), c->Pv * sizeof(int), c->N,
                                           hipMemcpyDeviceToHost, c->stream);
    const hipError_t e3 = hipStreamSynchronize(c->stream);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
      rc = c->fail(TAPGPU_E_HIP, "copy of labels failed");
      break;
    }
    for (int64_t v = 1; v < N1; ++v) {
      const size_t row = (size_t)(v - 1) * c->Pv;
      for (int64_t j = 0; j < ns; ++j) {
        const size_t o = (size_t)v * ld + col_base + s0 + j;
        if (dist) dist[o] = hd[row + j];
        if (pred) pred[o] = hp[row + j];
      }
    }
  }
  // private buffers are freed here (not kept in allocs)
  c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)d_col), c->allocs.end());
  c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)d_gv), c->allocs.end());
  (void)hipFree(d_col);
  (void)hipFree(d_gv);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// shortest paths without allocation (tapgpu_sssp_noalloc)
// ---------------------------------------------------------------------------
//
// The per-origin kernels' scratch (the labels-in-HBM kernel's label,
// predecessor and tail rows, and the 0-based start list of either kernel)
// lives in this file-scope device array instead of buffers allocated by the
// context.  Each GPU holds its own copy for the life of the process; host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (kept per device in g_scratch_ptr).  Every call on a
// GPU shares that GPU's copy: calls that overlap on one GPU race.

constexpr size_t kNoAllocScratchBytes = 64ull << 20;       // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device `dev` (which is current).
int scratch_address(tapgpu_ctx* c, int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return c->fail(TAPGPU_E_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return c->fail(TAPGPU_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TAPGPU_OK;
}

// 1-based start nodes -> 0-based list; clears the error flag first; a start
// node outside 1..N is flagged (code 5) and replaced by node 0.  One block.
__global__ __launch_bounds__(kBlock) void k_noalloc_starts(const int64_t* starts, int64_t n, int N,
                                                           int* out, int* err) {
  if (threadIdx.x < 3) err[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = threadIdx.x; i < n; i += kBlock) {
    const int64_t s = starts[i];
    if (s < 1 || s > N) {
      if (atomicCAS(&err[0], 0, 5) == 0) err[1] = (int)i;
      out[i] = 0;
    } else {
      out[i] = (int)(s - 1);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_i32(int* p, long long n, int v) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    p[i] = v;
}

}  // namespace

extern "C" {

int tapgpu_sssp_noalloc(tapgpu_ctx* c, int64_t variant, const int64_t* d_start_nodes,
                        int64_t num_start, double* d_dist, int32_t* d_pred, int64_t ld) {
  if (!c || num_start < 0 || (num_start > 0 && (!d_start_nodes || !d_dist || !d_pred)))
    return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  if (num_start == 0) return TAPGPU_OK;
  if (ld < num_start)
    return c->fail(TAPGPU_E_ARG, "ld %lld too small for %lld columns", (long long)ld,
                   (long long)num_start);
  if (num_start > INT32_MAX) return c->fail(TAPGPU_E_LIMIT, "too many start nodes");
  int dev = -1;
  HIPCHK(c, hipGetDevice(&dev));
  if (dev != c->device)
    return c->fail(TAPGPU_E_STATE, "the context belongs to device %d, device %d is current",
                   c->device, dev);
  const bool lds = use_lds(c);
  if (!lds && !c->hbm_ok)
    return c->fail(TAPGPU_E_LIMIT, "the per-origin kernels do not fit this network");
  // scratch: the start list, then (labels-in-HBM kernel) rows of 16 B per node
  const size_t list_bytes = ((size_t)num_start * sizeof(int) + 255) & ~(size_t)255;
  const size_t row_bytes = 16ull * (size_t)c->N;
  if (list_bytes + (lds ? 0 : row_bytes) > kNoAllocScratchBytes)
    return c->fail(TAPGPU_E_LIMIT,
                   "%lld start nodes on %d nodes exceed the %zu-byte scratch of the no-allocation "
                   "entry point", (long long)num_start, c->N, kNoAllocScratchBytes);
  void* base = nullptr;
  if ((rc = scratch_address(c, dev, &base))) return rc;
  unsigned char* p = (unsigned char*)base;
  int* d_src = (int*)p;
  OriginArgs A = make_origin_args(c, nullptr);
  A.fww += (size_t)variant * c->E;
  A.rvw += (size_t)variant * c->E;
  A.src_list = d_src;
  A.out_dist = d_dist;
  A.out_pred = d_pred;
  A.ld = ld;
  int rows = 0;
  if (!lds) {
    rows = (int)std::min<size_t>((size_t)c->hbm_grid, (kNoAllocScratchBytes - list_bytes) / row_bytes);
    const size_t rN = (size_t)rows * c->N;
    A.h_dist = (unsigned long long*)(p + list_bytes);
    A.h_pred = (int*)(p + list_bytes + 8 * rN);
    A.h_ptail = (int*)(p + list_bytes + 12 * rN);
  }
  hipLaunchKernelGGL(k_noalloc_starts, dim3(1), dim3(kBlock), 0, c->stream, d_start_nodes, num_start,
                     c->N, d_src, c->err_dev);
  // predecessors start at -1, so a failed search hands back no stale link
  const long long npred = (long long)c->N * ld;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(npred / kWave + 1)), dim3(kBlock), 0, c->stream,
                     d_pred, npred, -1);
  if (lds)
    launch_lds(c, dim3((unsigned)num_start, 1), c->stream, A, 0);
  else
    hipLaunchKernelGGL(k_origin_hbm, dim3((unsigned)std::min<int64_t>(num_start, rows)),
                       dim3(kHbmBlock), c->hbm_lds, c->stream, A, 0, (int)num_start);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->h_int, c->err_dev, sizeof(int) * 3, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int* herr = c->h_int;
  if (herr[0] == 5)
    return c->fail(TAPGPU_E_ARG, "start node %d of the list is out of range", herr[1]);
  if (herr[0] == 2)
    return c->fail(TAPGPU_E_LIMIT, "label-correcting rounds exceeded from origin %d (negative costs?)",
                   herr[1] + 1);
  if (herr[0] == 3)
    return c->fail(TAPGPU_E_LIMIT, "predecessor cycle in the tree of origin %d", herr[1] + 1);
  if (herr[0] != 0)
    return c->fail(TAPGPU_E_LIMIT, "shortest-path kernel reported error %d at origin %d", herr[0],
                   herr[1] + 1);
  return TAPGPU_OK;
}

int tapgpu_aon(tapgpu_ctx* c, int64_t variant, double* volumes) {
  if (!c) return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  const int v = (int)variant;
  std::vector<unsigned char> on(c->V, 0);
  on[v] = 1;
  unsigned char* d_on = nullptr;
  if ((rc = dalloc(c, &d_on, c->V))) return rc;
  rc = upload(c, d_on, on.data(), c->V);
  if (!rc) rc = load_variants(c, v, 1, d_on, on.data());
  if (!rc && volumes) {
    const hipError_t e = hipMemcpy(volumes, c->y + (size_t)v * c->E, sizeof(double) * c->E,
                                   hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = c->fail(TAPGPU_E_HIP, "copy of volumes failed");
  }
  c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)d_on), c->allocs.end());
  (void)hipFree(d_on);
  return rc;
}

static int ensure_trace(tapgpu_ctx* c, int need) {
  if (need <= c->trace_cap) return TAPGPU_OK;
  int cap = std::max(need, 1024);
  tapgpu_iter* t = nullptr;
  hipError_t e = hipMalloc((void**)&t, sizeof(tapgpu_iter) * (size_t)cap * c->V);
  if (e != hipSuccess) return c->fail(TAPGPU_E_NOMEM, "trace allocation failed");
  HIPCHK(c, hipMemsetAsync(t, 0, sizeof(tapgpu_iter) * (size_t)cap * c->V, c->stream));
  if (c->trace) {
    for (int v = 0; v < c->V; ++v)
      HIPCHK(c, hipMemcpyAsync(t + (size_t)v * cap, c->trace + (size_t)v * c->trace_cap,
                               sizeof(tapgpu_iter) * c->trace_cap, hipMemcpyDeviceToDevice,
                               c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->trace);
  }
  c->trace = t;
  c->trace_cap = cap;
  return TAPGPU_OK;
}

int tapgpu_fw_init(tapgpu_ctx* c, const double* warm_volumes) {
  if (!c) return TAPGPU_E_ARG;
  int rc = ensure_trace(c, 1024);
  if (rc) return rc;
  std::vector<VarState> hs(c->V);
  for (auto& s : hs) {
    s = VarState{};
    s.best_lb = -DBL_MAX;
    s.relgap = DBL_MAX;
    s.total_demand = c->total_demand;
    s.active = 1;
  }
  if ((rc = upload(c, c->st, hs.data(), c->V))) return rc;
  HIPCHK(c, hipMemsetAsync(c->var_on, 1, c->V, c->stream));
  if (warm_volumes) {
    if ((rc = upload(c, c->x, warm_volumes, (size_t)c->V * c->E))) return rc;
  } else {
    if ((rc = launch_costs(c, 0, nullptr, 0, c->V))) return rc;   // Wa = t0
    if ((rc = load_variants(c, 0, c->V, nullptr, nullptr))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->x, c->y, sizeof(double) * (size_t)c->V * c->E,
                             hipMemcpyDeviceToDevice, c->stream));
  }
  if ((rc = launch_costs(c, 1, nullptr, 0, c->V))) return rc;      // c = BPR(x)
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->initialised = true;
  return TAPGPU_OK;
}

__global__ void k_set_params(FwParams* p, FwParams v) { *p = v; }

// (a kernel argument, not a copy from host memory that the next call
// overwrites)
static int upload_params(tapgpu_ctx* c, int max_iters, double target) {
  hipLaunchKernelGGL(k_set_params, dim3(1), dim3(1), 0, c->stream, c->prm,
                     FwParams{c->trace, c->trace_cap, max_iters, target, c->line_window});
  HIPCHK(c, hipGetLastError());
  return TAPGPU_OK;
}

// The link passes of an iteration: two line-search windows (the second does
// nothing for variants the first decided), the step, the flags, and the
// copy of the iteration block (error flag, counters, states, flags) to
// pinned host memory.
static int issue_link_passes(tapgpu_ctx* c, const Links& K, dim3 lgrid, bool readback = true) {
  const bool small = c->link_mode == 1 || (c->link_mode == 0 && c->E <= kSmallLinks);
  if (small) {                        // one launch, one workgroup per variant
    hipLaunchKernelGGL(k_link_small, dim3(c->V), dim3(kBlock), 0, c->stream, K, c->st, c->var_on,
                       c->prm);
    HIPCHK(c, hipGetLastError());
    if (readback)
      HIPCHK(c, hipMemcpyAsync(c->pin_block, c->iter_dev, c->iter_bytes, hipMemcpyDeviceToHost,
                               c->stream));
    return TAPGPU_OK;
  }
  hipLaunchKernelGGL(k_link_partials, lgrid, dim3(kLinkThreads), 0, c->stream, K, c->st, c->part, 0);
  hipLaunchKernelGGL(k_link_decide, dim3(c->V), dim3(kBlock), 0, c->stream, c->st, c->part,
                     kLinkBlocks, 0, c->prm);
  hipLaunchKernelGGL(k_link_partials, lgrid, dim3(kLinkThreads), 0, c->stream, K, c->st, c->part, 1);
  hipLaunchKernelGGL(k_link_decide, dim3(c->V), dim3(kBlock), 0, c->stream, c->st, c->part,
                     kLinkBlocks, 1, c->prm);
  hipLaunchKernelGGL(k_link_step, lgrid, dim3(kBlock), 0, c->stream, K, c->st);
  hipLaunchKernelGGL(k_finish_step, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->st,
                     c->V, c->var_on);
  HIPCHK(c, hipGetLastError());
  if (readback)
    HIPCHK(c, hipMemcpyAsync(c->pin_block, c->iter_dev, c->iter_bytes, hipMemcpyDeviceToHost,
                             c->stream));
  return TAPGPU_OK;
}

// Records what `issue` puts on the context's stream into an executable
// graph (replacing *exec).  Nothing in a captured sequence may allocate,
// synchronise, record events or clear memory with a memset: the captured
// kernels clear what they need themselves.
extern "C++" template <class F>
static int capture(tapgpu_ctx* c, hipGraphExec_t* exec, F&& issue) {
  if (*exec) {
    HIPCHK(c, hipGraphExecDestroy(*exec));
    *exec = nullptr;
  }
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;               // (launch paths then skip host-side set-up)
  const int rc = issue();
  c->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess || !g) return c->fail(TAPGPU_E_HIP, "stream capture failed: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    *exec = nullptr;
    return c->fail(TAPGPU_E_HIP, "graph instantiation failed: %s", hipGetErrorString(ei));
  }
  return TAPGPU_OK;
}

// The whole iteration of a tuned per-origin schedule as one graph:
// k_iter_begin, the loading kernel, the link passes and the readback.  It is
// captured again when the schedule that its launches bake in changes.
// With m > 1 the graph holds m iterations back to back and one readback: a
// variant that stops (or converges) in one of them is skipped by the rest
// (its flag is off), and a step that needs a third line-search window is
// retried unchanged by the following iterations and finished by the host
// after the replay, so the results are those of m single replays.  Used
// where an iteration is short enough for the launch and synchronisation
// to matter.
static int iter_graph(tapgpu_ctx* c, const Links& K, dim3 lgrid, int m) {
  const bool lds = use_lds(c);
  if (!lds) {
    int rc = ensure_hbm(c);
    if (rc) return rc;
  }
  long long dbits;
  std::memcpy(&dbits, &c->delta_mult, sizeof dbits);
  const long long key[6] = {(lds ? 1 : 2) + 4 * m + (c->origin_order_ready ? 64 : 0) + 128 * c->link_mode,
                            c->chase, c->qcap, c->hqcap, dbits, lds_threads(c)};
  if (c->g_iter && std::equal(key, key + 6, c->g_iter_key)) return TAPGPU_OK;
  int rc = capture(c, &c->g_iter, [&]() -> int {
    for (int it = 0; it < m; ++it) {
      const long long ny = (long long)c->V * c->E;
      hipLaunchKernelGGL(k_iter_begin, dim3(grid_for(ny / kWave + 1)), dim3(kBlock), 0, c->stream,
                         c->st, c->V, c->var_on, c->err_dev, c->counters, c->y, ny, it > 0 ? 1 : 0);
      c->defer_origin = true;
      c->pre_cleared = true;
      int r = lds ? aon_lds(c, 0, c->V, c->var_on) : aon_hbm(c, 0, c->V, c->var_on);
      c->defer_origin = false;
      c->pre_cleared = false;
      c->origin_pending = false;
      if (r) return r;
      if ((r = issue_link_passes(c, K, lgrid, it == m - 1))) return r;
    }
    return TAPGPU_OK;
  });
  if (rc) return rc;
  std::copy(key, key + 6, c->g_iter_key);
  return TAPGPU_OK;
}

// The whole iteration of a batch on the lanes kernel as one graph:
// k_iter_begin (flags, error flag, counters), pack, the lanes kernel, unpack,
// the gate, the link passes and the readback.  The lane map and the memo
// stamps are device data, set before each replay; the graph is captured
// again when the number of groups or the kernel's schedule changes.
static int lanes_graph(tapgpu_ctx* c, const Links& K, dim3 lgrid, int G) {
  long long dbits;
  std::memcpy(&dbits, &c->lanes_delta_mult, sizeof dbits);
  const long long key[6] = {3 + (c->origin_order_ready ? 64 : 0) + 128 * c->link_mode, G, c->lanes_qcap,
                            dbits, (long long)c->lanes_lds, c->lanes_grid};
  if (c->g_iter && std::equal(key, key + 6, c->g_iter_key)) return TAPGPU_OK;
  int rc = capture(c, &c->g_iter, [&]() {
    hipLaunchKernelGGL(k_iter_begin, dim3(1), dim3(kBlock), 0, c->stream, c->st, c->V, c->var_on,
                       c->err_dev, c->counters, c->y, 0ll, 0);
    int r = launch_lanes(c, G, (long long)c->Z * G, nullptr);
    if (r) return r;
    hipLaunchKernelGGL(k_lanes_gate, dim3(1), dim3(kBlock), 0, c->stream, c->err_dev, c->st, c->V);
    return issue_link_passes(c, K, lgrid);
  });
  if (rc) return rc;
  std::copy(key, key + 6, c->g_iter_key);
  return TAPGPU_OK;
}

static int link_graph(tapgpu_ctx* c, const Links& K, dim3 lgrid) {
  if (c->g_link && c->g_link_mode == c->link_mode) return TAPGPU_OK;
  c->g_link_mode = c->link_mode;
  return capture(c, &c->g_link, [&]() { return issue_link_passes(c, K, lgrid); });
}

int tapgpu_set_link_passes(tapgpu_ctx* c, int mode, int window) {
  if (!c || mode < 0 || mode > 2 || window < 0 || window > kJ) return TAPGPU_E_ARG;
  c->link_mode = mode;
  c->line_window = window > 0 ? window : kJ;
  return TAPGPU_OK;
}

int tapgpu_set_graphs(tapgpu_ctx* c, int mode) {
  if (!c || mode < 0 || mode > kMaxGraphIters) return TAPGPU_E_ARG;
  c->graphs = mode != 0;
  c->graph_iters = mode > 1 ? mode : 0;
  return TAPGPU_OK;
}

int tapgpu_get_graphs(tapgpu_ctx* c, int* mode) {
  if (!c || !mode) return TAPGPU_E_ARG;
  *mode = !c->graphs ? 0 : (c->graph_iters > 1 ? c->graph_iters : 1);
  return TAPGPU_OK;
}

int tapgpu_fw_run(tapgpu_ctx* c, int max_iters, double target_relgap, tapgpu_result* results) {
  if (!c) return TAPGPU_E_ARG;
  if (!c->initialised) return c->fail(TAPGPU_E_STATE, "tapgpu_fw_init was not called");
  int rc = ensure_trace(c, max_iters > 0 ? max_iters : 1024);
  if (rc) return rc;
  c->stats = tapgpu_stats{};
  auto t_start = std::chrono::steady_clock::now();
  Links K = make_links(c);
  const dim3 lgrid(kLinkBlocks, c->V);
  std::vector<unsigned char> hon(c->V);
  std::vector<VarState> hs(c->V);
  // variants already past the stop rule do not run (the loop test comes first)
  HIPCHK(c, hipMemcpy(hs.data(), c->st, sizeof(VarState) * c->V, hipMemcpyDeviceToHost));
  for (int v = 0; v < c->V; ++v) {
    hs[v].active = (max_iters <= 0 || hs[v].iter < max_iters) &&
                   (target_relgap <= 0 || hs[v].relgap > target_relgap);
    hon[v] = (unsigned char)hs[v].active;
  }
  if ((rc = upload(c, c->st, hs.data(), c->V))) return rc;
  if ((rc = upload(c, c->var_on, hon.data(), c->V))) return rc;
  // One synchronisation per iteration (or per replay of several): the
  // loading's error flag and counters, the variants' states and flags are
  // read back together at the end.  On networks above kSmallLinks links a
  // second line-search window pass is always queued and does nothing for
  // variants whose lambda the first window decided; a variant that needs
  // more windows skips the step and is finished by the host loop below
  // (rare: lambda below 2^-31).  k_link_small runs every window itself.
  if (!c->ev[2])
    for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
  if ((rc = upload_params(c, max_iters, target_relgap))) return rc;
  for (;;) {
    bool any = false;
    for (int v = 0; v < c->V; ++v) any = any || hon[v];
    if (!any) break;
    if (c->trace_cap < 1 << 30) {
      int maxit = 0;
      for (int v = 0; v < c->V; ++v) maxit = std::max(maxit, hs[v].iter);
      if (maxit + kMaxGraphIters >= c->trace_cap) {
        if ((rc = ensure_trace(c, std::max(2 * c->trace_cap, maxit + kMaxGraphIters + 1)))) return rc;
        if ((rc = upload_params(c, max_iters, target_relgap))) return rc;
      }
    }
    auto t_it = std::chrono::steady_clock::now();
    if ((rc = origin_order_update(c))) return rc;      // (before any capture)
    // with the schedule already tuned, one kernel opens the iteration and
    // clears what the loading would clear
    const bool tuned = use_lds(c) ? (c->chase >= 0 && c->split >= 0 && c->lds_block > 0)
                                  : (use_hbm(c) && c->delta_mult >= 0);
    const bool graphs = c->graphs && !c->profiling;
    const bool lanes_now = use_lanes(c, 0, c->V, hon.data());
    float aon = 0.f;
    // how the iteration runs: a captured graph of all of it (per-origin
    // kernel or lanes kernel), or the loading issued launch by launch and
    // the link passes (captured or not) after it
    int mode = 0;
    int G = 0;
    if (graphs && lanes_now && c->lanes_delta_mult >= 0 && c->l_dist && c->Z > 0) {
      if ((rc = lanes_map(c, 0, c->V, hon.data(), &G))) return rc;
      if (G > 0) mode = 2;
    } else if (graphs && tuned && !lanes_now && (use_lds(c) ? c->split == 0 : use_hbm(c))) {
      mode = 1;
    }
    int m = 1;                          // iterations in this replay
    if (mode == 1) {
      // short iterations go several to a replay: one host synchronisation
      // per replay instead of per iteration (on metro, ~0.9 ms iterations,
      // 8 per replay beat 2 by 0.9%; the iterations a replay runs past
      // convergence are gated off and cost little)
      m = c->graph_iters > 0 ? c->graph_iters
          : (c->last_iter_ms > 0 && c->last_iter_ms < 2.0) ? kMaxGraphIters : 1;
      // the loading's error flag and counters come back in the iteration block
      if ((rc = iter_graph(c, K, lgrid, m))) return rc;
      HIPCHK(c, hipGraphLaunch(c->g_iter, c->stream));
      c->graph_launches += 1;
      c->origin_pending = true;
      c->pending_Vn = c->V;
      c->stats.origin_blocks += (int64_t)(m - 1) * c->Z * c->V;
      c->stats.link_passes += m - 1;
      if (!use_lds(c)) c->stats.origin_delta = c->delta_mult;
    } else if (mode == 2) {
      const long long tasks = (long long)c->Z * G;
      if ((rc = lanes_reserve(c, tasks))) return rc;
      if ((rc = lanes_graph(c, K, lgrid, G))) return rc;
      HIPCHK(c, hipGraphLaunch(c->g_iter, c->stream));
      c->graph_launches += 1;
      HIPCHK(c, hipStreamSynchronize(c->stream));
      bool fallback = false;
      if ((rc = absorb_lanes(c, c->pin_err, c->pin_cnt, tasks, c->V, std::chrono::steady_clock::now(),
                             &fallback)))
        return rc;
      if (fallback) {
        // the gate kept every variant from stepping: open the step again,
        // load with a per-origin kernel and run the link passes
        hipLaunchKernelGGL(k_begin_step, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->st,
                           c->V, c->var_on);
        if ((rc = redo_per_origin(c, 0, c->V, c->var_on))) return rc;
        if ((rc = issue_link_passes(c, K, lgrid))) return rc;
      }
    } else {
      if (tuned) {
        const long long ny = (long long)c->V * c->E;
        hipLaunchKernelGGL(k_iter_begin, dim3(grid_for(ny / kWave + 1)), dim3(kBlock), 0, c->stream,
                           c->st, c->V, c->var_on, c->err_dev, c->counters, c->y, ny, 0);
      } else {
        hipLaunchKernelGGL(k_begin_step, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->st,
                           c->V, c->var_on);
      }
      HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
      c->defer_origin = true;
      c->pre_cleared = tuned;
      rc = load_variants(c, 0, c->V, c->var_on, hon.data());
      c->defer_origin = false;
      c->pre_cleared = false;
      if (rc) return rc;
      HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
      if (graphs) {
        if ((rc = link_graph(c, K, lgrid))) return rc;
        HIPCHK(c, hipGraphLaunch(c->g_link, c->stream));
        c->graph_launches += 1;
      } else if ((rc = issue_link_passes(c, K, lgrid))) {
        return rc;
      }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = absorb_pending(c))) return rc;
    c->stats.link_passes += 1;
    auto need_more = [&]() {
      bool more = false;
      for (int v = 0; v < c->V; ++v) more = more || (c->pin_st[v].stepping && c->pin_st[v].need_more);
      return more;
    };
    bool more = need_more();
    if (more) c->stats.link_passes += 1;              // the queued second window was used
    while (more) {
      hipLaunchKernelGGL(k_link_partials, lgrid, dim3(kLinkThreads), 0, c->stream, K, c->st, c->part, 1);
      hipLaunchKernelGGL(k_link_decide, dim3(c->V), dim3(kBlock), 0, c->stream, c->st, c->part,
                         kLinkBlocks, 1, c->prm);
      hipLaunchKernelGGL(k_link_step, lgrid, dim3(kBlock), 0, c->stream, K, c->st);
      hipLaunchKernelGGL(k_finish_step, dim3((c->V + 255) / 256), dim3(256), 0, c->stream, c->st,
                         c->V, c->var_on);
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipMemcpyAsync(c->pin_st, c->st, sizeof(VarState) * c->V, hipMemcpyDeviceToHost,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(c->pin_on, c->var_on, c->V, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      c->stats.link_passes += 1;
      more = need_more();
    }
    std::copy(c->pin_st, c->pin_st + c->V, hs.begin());
    std::copy(c->pin_on, c->pin_on + c->V, hon.begin());
    const double it_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_it).count();
    c->last_iter_ms = it_ms / m;
    // (a captured iteration is timed as a whole, in link_ms)
    if (mode == 0) HIPCHK(c, hipEventElapsedTime(&aon, c->ev[2], c->ev[3]));
    c->stats.aon_ms += aon;
    c->stats.link_ms += std::max(0.0, it_ms - aon);
  }
  c->stats.graph_launches = c->graph_launches;
  c->graph_launches = 0;
  c->stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (results) {
    // VHT after the last step: sum x*c per variant on the device (fixed
    // order), read back with the variant states
    if (!c->vht) {
      int rc2 = dalloc(c, &c->vht, 2 * (size_t)c->V);
      if (rc2) return rc2;
    }
    hipLaunchKernelGGL(k_totals, dim3(c->V), dim3(kBlock), 0, c->stream, K, c->vht);
    HIPCHK(c, hipGetLastError());
    std::vector<double> tot(2 * (size_t)c->V);
    HIPCHK(c, hipMemcpyAsync(hs.data(), c->st, sizeof(VarState) * c->V, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(tot.data(), c->vht, sizeof(double) * 2 * c->V, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int v = 0; v < c->V; ++v) {
      tapgpu_result& r = results[v];
      r.iterations = hs[v].iter;
      r.relgap = hs[v].relgap;
      r.objective = hs[v].objective;
      r.converged = (target_relgap > 0 && hs[v].relgap <= target_relgap) ? 1 : 0;
      r.vht = tot[2 * (size_t)v];
    }
  }
  return TAPGPU_OK;
}

int tapgpu_get_flows(tapgpu_ctx* c, int64_t variant, double* volumes, double* costs) {
  if (!c) return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  if (volumes)
    HIPCHK(c, hipMemcpy(volumes, c->x + (size_t)variant * c->E, sizeof(double) * c->E,
                        hipMemcpyDeviceToHost));
  if (costs)
    HIPCHK(c, hipMemcpy(costs, c->c + (size_t)variant * c->E, sizeof(double) * c->E,
                        hipMemcpyDeviceToHost));
  return TAPGPU_OK;
}

int tapgpu_get_totals(tapgpu_ctx* c, int64_t variant, double* objective, double* vht) {
  if (!c) return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  if (!c->initialised) return c->fail(TAPGPU_E_STATE, "tapgpu_fw_init was not called");
  if (!c->vht && (rc = dalloc(c, &c->vht, 2 * (size_t)c->V))) return rc;
  hipLaunchKernelGGL(k_totals, dim3(c->V), dim3(kBlock), 0, c->stream, make_links(c), c->vht);
  HIPCHK(c, hipGetLastError());
  double t[2];
  HIPCHK(c, hipMemcpyAsync(t, c->vht + 2 * variant, sizeof t, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (vht) *vht = t[0];
  if (objective) *objective = t[1];
  return TAPGPU_OK;
}

int tapgpu_set_links(tapgpu_ctx* c, int64_t variant, const double* capacity,
                     const double* free_flow_time) {
  if (!c) return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  if (free_flow_time)       // (costs must stay non-negative; checked before any change)
    for (int k = 0; k < c->E; ++k)
      if (!(free_flow_time[k] >= 0.0))
        return c->fail(TAPGPU_E_ARG, "link %d: negative or NaN free-flow time", k);
  const size_t off = (size_t)variant * c->E;
  if (capacity) {
    std::vector<unsigned char> pres(c->E);
    HIPCHK(c, hipMemcpy(pres.data(), c->present + off, c->E, hipMemcpyDeviceToHost));
    std::vector<double> cap(capacity, capacity + c->E);
    double minpos = DBL_MAX;
    for (int64_t k = 0; k < c->E; ++k)
      if (pres[k] && cap[k] > 0.0 && cap[k] < minpos) minpos = cap[k];
    for (auto& v : cap)
      if (!(v > 0.0)) v = minpos;
    HIPCHK(c, hipMemcpy(c->cap + off, cap.data(), sizeof(double) * c->E, hipMemcpyHostToDevice));
  }
  if (free_flow_time)
    HIPCHK(c, hipMemcpy(c->t0 + off, free_flow_time, sizeof(double) * c->E, hipMemcpyHostToDevice));
  // the volumes no longer belong to these links: fw_init comes next
  c->initialised = false;
  return TAPGPU_OK;
}

int tapgpu_get_trace(tapgpu_ctx* c, int64_t variant, tapgpu_iter* trace, int cap, int* count) {
  if (!c) return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  VarState s;
  HIPCHK(c, hipMemcpy(&s, c->st + variant, sizeof s, hipMemcpyDeviceToHost));
  const int held = std::min(s.iter, c->trace_cap);
  if (count) *count = s.iter;
  const int n = std::min(held, cap);
  if (trace && n > 0)
    HIPCHK(c, hipMemcpy(trace, c->trace + (size_t)variant * c->trace_cap, sizeof(tapgpu_iter) * n,
                        hipMemcpyDeviceToHost));
  return TAPGPU_OK;
}

int tapgpu_load_pred(tapgpu_ctx* c, int64_t variant, const int64_t* predlink, int64_t ld,
                     int64_t col_base, double* volumes) {
  if (!c || !predlink) return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  if ((rc = ensure_labels(c))) return rc;
  const int v = (int)variant;
  // predecessor links of zone origins into the variant's columns
  std::vector<int> hp((size_t)c->N * c->Pv, -1);
  for (int64_t node = 1; node <= c->N; ++node)
    for (int64_t j = 0; j < c->Z; ++j) {
      const int64_t e = predlink[(size_t)node * ld + col_base + c->zone_order[j]];
      if (e < -1 || e >= c->E)
        return c->fail(TAPGPU_E_ARG, "predecessor link %lld out of range", (long long)e);
      if (e >= 0 && c->h_head[e] != node - 1)
        return c->fail(TAPGPU_E_ARG, "predecessor link %lld of node %lld (origin %d) ends at node %d",
                       (long long)e, (long long)node, c->zone_order[j] + 1, c->h_head[e] + 1);
      hp[(size_t)(node - 1) * c->Pv + j] = (int)e;
    }
  HIPCHK(c, hipMemcpy2DAsync(c->pred + (size_t)v * c->Pv, c->P * sizeof(int), hp.data(),
                             c->Pv * sizeof(int), c->Pv * sizeof(int), c->N,
                             hipMemcpyHostToDevice, c->stream));
  std::vector<unsigned char> on(c->V, 0);
  on[v] = 1;
  unsigned char* d_on = nullptr;
  if ((rc = dalloc(c, &d_on, c->V))) return rc;
  rc = upload(c, d_on, on.data(), c->V);
  if (!rc) rc = aon_groups(c, v * c->Gv, c->Gv, v, 1, d_on, true);
  if (!rc && volumes) {
    const hipError_t e = hipMemcpy(volumes, c->y + (size_t)v * c->E, sizeof(double) * c->E,
                                   hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = c->fail(TAPGPU_E_HIP, "copy of volumes failed");
  }
  c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)d_on), c->allocs.end());
  (void)hipFree(d_on);
  return rc;
}

int tapgpu_get_iter(tapgpu_ctx* c, int64_t variant, int index, tapgpu_iter* rec) {
  if (!c || !rec) return TAPGPU_E_ARG;
  int rc = check_variant(c, variant);
  if (rc) return rc;
  if (index < 0 || index >= c->trace_cap) return c->fail(TAPGPU_E_ARG, "iteration %d not held", index);
  HIPCHK(c, hipMemcpy(rec, c->trace + (size_t)variant * c->trace_cap + index, sizeof(tapgpu_iter),
                      hipMemcpyDeviceToHost));
  return TAPGPU_OK;
}

int tapgpu_device_state(tapgpu_ctx* c, const double** d_volumes, const double** d_costs) {
  if (!c) return TAPGPU_E_ARG;
  if (d_volumes) *d_volumes = c->x;
  if (d_costs) *d_costs = c->c;
  return TAPGPU_OK;
}

int tapgpu_get_stats(tapgpu_ctx* c, tapgpu_stats* s) {
  if (!c || !s) return TAPGPU_E_ARG;
  *s = c->stats;
  return TAPGPU_OK;
}

int tapgpu_set_strategy(tapgpu_ctx* c, int strategy) {
  if (!c || strategy < 0 || strategy > 3) return TAPGPU_E_ARG;
  if (strategy == 2 && !c->lds_ok)
    return c->fail(TAPGPU_E_ARG, "network does not fit the one-origin-per-workgroup path");
  if (strategy == 3 && !c->hbm_ok)
    return c->fail(TAPGPU_E_ARG, "frontier bitmaps of this network do not fit LDS");
  c->strategy = strategy;
  return TAPGPU_OK;
}

int tapgpu_set_origin_chase(tapgpu_ctx* c, int chase) {
  if (!c || chase < -1 || chase > 1024) return TAPGPU_E_ARG;
  c->chase = chase;
  return TAPGPU_OK;
}

int tapgpu_set_origin_delta(tapgpu_ctx* c, double mult) {
  if (!c || !(mult >= -1.0) || !std::isfinite(mult)) return TAPGPU_E_ARG;
  c->delta_mult = mult < 0 ? -1.0 : mult;
  return TAPGPU_OK;
}

double tapgpu_get_origin_delta(const tapgpu_ctx* c) { return c ? c->delta_mult : -2.0; }

int tapgpu_set_origin_split(tapgpu_ctx* c, int origins) {
  if (!c || origins < -1) return TAPGPU_E_ARG;
  c->split = origins;
  return TAPGPU_OK;
}

int tapgpu_get_origin_split(const tapgpu_ctx* c) { return c ? c->split : TAPGPU_E_ARG; }

int tapgpu_set_origin_block(tapgpu_ctx* c, int threads) {
  if (!c) return TAPGPU_E_ARG;
  if (threads == -1) {
    c->lds_block = -1;
    return TAPGPU_OK;
  }
  bool known = false;
  for (int b : kLdsBlocks) known = known || b == threads;
  if (!known) return c->fail(TAPGPU_E_ARG, "workgroup size %d: 1024, 512, 256 or -1", threads);
  if (c->N > kLdsPerThread * threads)
    return c->fail(TAPGPU_E_ARG, "%d threads per workgroup hold at most %d nodes", threads,
                   kLdsPerThread * threads);
  c->lds_block = threads;
  return TAPGPU_OK;
}

int tapgpu_get_origin_block(const tapgpu_ctx* c) {
  if (!c) return TAPGPU_E_ARG;
  return c->lds_block;
}

int tapgpu_get_origin_chase(const tapgpu_ctx* c) {
  if (!c) return TAPGPU_E_ARG;
  return c->chase;
}

int tapgpu_set_origin_queue(tapgpu_ctx* c, int entries) {
  if (!c || entries < -1) return TAPGPU_E_ARG;
  c->qcap = entries < 0 ? c->qcap_max : std::min(entries, c->qcap_max);
  c->qcap_auto = entries < 0;
  if (entries < 0) c->lds_block = -1;     // picked again with the workgroup size
  c->hqcap = entries < 0 ? c->hqcap_max : std::min(entries, c->hqcap_max);
  c->lanes_qcap = entries < 0 ? c->lanes_qcap_max : std::min(entries, c->lanes_qcap_max);
  return TAPGPU_OK;
}

int tapgpu_get_origin_queue(const tapgpu_ctx* c) {
  if (!c) return TAPGPU_E_ARG;
  tapgpu_ctx* m = const_cast<tapgpu_ctx*>(c);
  return use_hbm(m) ? c->hqcap : c->qcap;
}

int tapgpu_get_strategy(const tapgpu_ctx* c) {
  if (!c) return TAPGPU_E_ARG;
  tapgpu_ctx* m = const_cast<tapgpu_ctx*>(c);
  return use_lds(m) ? 2 : (use_hbm(m) ? 3 : 1);
}

int tapgpu_set_lanes(tapgpu_ctx* c, int min_active) {
  if (!c || min_active < -1) return TAPGPU_E_ARG;
  if (min_active != 0 && !c->lanes_ok)
    return c->fail(TAPGPU_E_ARG, "frontier bitmaps of this network do not fit LDS");
  c->lanes_min = min_active < 0 ? (c->V > 1 ? kLanesMinDefault : 0) : min_active;
  return TAPGPU_OK;
}

int tapgpu_get_lanes(const tapgpu_ctx* c) { return c ? c->lanes_min : TAPGPU_E_ARG; }

int tapgpu_set_lanes_delta(tapgpu_ctx* c, double mult) {
  if (!c || !(mult >= -1.0) || !std::isfinite(mult)) return TAPGPU_E_ARG;
  c->lanes_delta_mult = mult < 0 ? -1.0 : mult;
  return TAPGPU_OK;
}

double tapgpu_get_lanes_delta(const tapgpu_ctx* c) { return c ? c->lanes_delta_mult : -2.0; }

int tapgpu_set_profiling(tapgpu_ctx* c, int on) {
  if (!c) return TAPGPU_E_ARG;
  if (on && !c->ev[0])
    for (auto& e : c->ev) HIPCHK(c, hipEventCreate(&e));
  c->profiling = on != 0;
  return TAPGPU_OK;
}

}  // extern "C"

#include "tapgrad.inc"
#include "tapgreedy.inc"
