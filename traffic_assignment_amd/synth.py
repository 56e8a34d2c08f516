"""Seeded synthetic road networks, OD demand and upgrade variants.

Every network is produced in the data model the reference's TNTP readers
build (trunk/tap/tap_types.h:20-54): links carry {init, term, capacity,
length, free-flow time, B, power, speed, toll, type}, node ids are 1-based,
zones are nodes 1..Z and links are sorted by (init, term), which is the order
`adjlist_to_packed_arrays` (trunk/tap/tap_functions.c:313-339) relies on so
that link index == CSR edge index.

Workloads (all float64, all seeded, every free parameter recorded in
`Network.params`):

* ``grid5``      24 nodes / 76 links, jittered 5x5 grid minus a corner,
                 all nodes are zones (FIRST THRU NODE = 1), integer uniform
                 demand U{0..1000} with 40% of off-diagonal pairs zeroed.
* ``metro``      jittered 100x100 grid + 1,000 centroid zones numbered first
                 (FIRST THRU NODE = 1001), 2 bidirectional connectors each,
                 gravity/Poisson demand with the free-flow zone-to-zone
                 shortest-path time as impedance.
* ``regional``   jittered 300x300 grid + 2,000 bidirectional arterial
                 shortcuts (5-20 km) + 4,000 zones with 3 connectors each;
                 gravity/Poisson trip counts with the free-flow zone-to-zone
                 shortest-path time as impedance, scaled so that 30% of the
                 16M OD pairs are nonzero; the assigned demand is counts x
                 2^-6 (peak-hour share; a power of two, so every partial sum
                 of demand stays exact in FP64; at reduced scale the power
                 of two that keeps the load per area).
* ``upgrades``   32 candidate upgrades on a metro network and a list of
                 budgeted subsets (the batched-variant workload).

`scale` shrinks the grid/zone counts for tests (same generator, same rules).
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

# link classes: freeway / arterial / local
CLASS_PROB = (0.05, 0.25, 0.70)
CLASS_SPEED_KMH = (100.0, 70.0, 50.0)
CLASS_CAP_PER_LANE = (4000.0, 1800.0, 800.0)
CLASS_LANES = (3, 2, 1)
BPR_B = 0.15
BPR_POWER = 4.0
CONNECTOR_T0 = 0.5  # minutes
CONNECTOR_CAP = 1e5


@dataclass
class Network:
    num_zones: int
    num_nodes: int
    first_thru_node: int
    init: np.ndarray          # int64 [E], 1-based
    term: np.ndarray          # int64 [E], 1-based
    capacity: np.ndarray      # float64 [E]
    length: np.ndarray
    fftime: np.ndarray        # free-flow time t0 (minutes)
    B: np.ndarray
    power: np.ndarray
    speed: np.ndarray
    toll: np.ndarray
    ltype: np.ndarray         # int16 [E]
    xy: np.ndarray | None = None   # node coordinates (km), row i = node i+1
    params: dict = field(default_factory=dict)

    @property
    def num_links(self) -> int:
        return int(self.init.shape[0])

    def sorted(self) -> "Network":
        order = np.lexsort((self.term, self.init))
        kw = {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}
        for name in ("init", "term", "capacity", "length", "fftime", "B", "power",
                     "speed", "toll", "ltype"):
            kw[name] = kw[name][order]
        return Network(**kw)

    def copy(self) -> "Network":
        kw = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            kw[f.name] = v.copy() if isinstance(v, np.ndarray) else (dict(v) if isinstance(v, dict) else v)
        return Network(**kw)


@dataclass
class Demand:
    """OD demand in origin-major CSR form (origins 1..Z, dests ascending)."""
    num_zones: int
    start: np.ndarray    # int64 [Z+1]; origin o's entries at [start[o-1], start[o])
    dest: np.ndarray     # int64 [nnz] 1-based zone ids
    demand: np.ndarray   # float64 [nnz]
    params: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.dest.shape[0])

    def total(self) -> float:
        return float(self.demand.sum())

    @staticmethod
    def from_dense(od: np.ndarray, keep_zero: bool = False, params=None) -> "Demand":
        Z = od.shape[0]
        mask = np.ones_like(od, dtype=bool) if keep_zero else (od != 0)
        np.fill_diagonal(mask, False)
        counts = mask.sum(axis=1)
        start = np.zeros(Z + 1, dtype=np.int64)
        np.cumsum(counts, out=start[1:])
        oi, dj = np.nonzero(mask)
        return Demand(Z, start, (dj + 1).astype(np.int64), od[oi, dj].astype(np.float64),
                      params or {})

    def to_dense(self) -> np.ndarray:
        Z = self.num_zones
        od = np.zeros((Z, Z), dtype=np.float64)
        orig = np.repeat(np.arange(Z), np.diff(self.start))
        od[orig, self.dest - 1] = self.demand
        return od


# ---------------------------------------------------------------------------
# building blocks
# ---------------------------------------------------------------------------

def _jittered_grid(rng, rows, cols, spacing, jitter):
    r, c = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    x = (c + rng.uniform(-jitter, jitter, size=(rows, cols))) * spacing
    y = (r + rng.uniform(-jitter, jitter, size=(rows, cols))) * spacing
    return np.stack([x.ravel(), y.ravel()], axis=1)


def _grid_edges(rows, cols, alive=None):
    """Undirected 4-neighbour edges (a, b) with a < b, 0-based grid indices."""
    idx = np.arange(rows * cols).reshape(rows, cols)
    h = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    v = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    e = np.concatenate([h, v])
    if alive is not None:
        e = e[alive[e[:, 0]] & alive[e[:, 1]]]
    return e


def _classed_links(rng, a, b, xy):
    """Bidirectional road links of a random class between node ids a<->b."""
    n = a.shape[0]
    cls = rng.choice(3, size=n, p=CLASS_PROB)
    length = np.hypot(*(xy[a - 1] - xy[b - 1]).T)
    speed = np.asarray(CLASS_SPEED_KMH)[cls]
    cap = np.asarray(CLASS_CAP_PER_LANE)[cls] * np.asarray(CLASS_LANES)[cls]
    t0 = length / speed * 60.0
    rec = dict(init=np.concatenate([a, b]), term=np.concatenate([b, a]),
               capacity=np.tile(cap, 2), length=np.tile(length, 2),
               fftime=np.tile(t0, 2), speed=np.tile(speed, 2),
               ltype=np.tile(cls + 1, 2).astype(np.int16))
    return rec


def _assemble(num_zones, num_nodes, first_thru, parts, xy, params):
    cat = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    E = cat["init"].shape[0]
    net = Network(num_zones=num_zones, num_nodes=num_nodes, first_thru_node=first_thru,
                  init=cat["init"].astype(np.int64), term=cat["term"].astype(np.int64),
                  capacity=cat["capacity"].astype(np.float64),
                  length=cat["length"].astype(np.float64),
                  fftime=cat["fftime"].astype(np.float64),
                  B=np.full(E, BPR_B), power=np.full(E, BPR_POWER),
                  speed=cat["speed"].astype(np.float64), toll=np.zeros(E),
                  ltype=cat["ltype"].astype(np.int16), xy=xy, params=params)
    return net.sorted()


def _nearest_grid(xy_grid, rows, cols, spacing, pts, k):
    """k nearest (jittered) grid nodes for each point; 0-based grid indices."""
    out = np.empty((pts.shape[0], k), dtype=np.int64)
    rad = 3
    for i, (px, py) in enumerate(pts):
        c0 = int(round(px / spacing)); r0 = int(round(py / spacing))
        rr = np.arange(max(0, r0 - rad), min(rows, r0 + rad + 1))
        cc = np.arange(max(0, c0 - rad), min(cols, c0 + rad + 1))
        cand = (rr[:, None] * cols + cc[None, :]).ravel()
        d = np.hypot(xy_grid[cand, 0] - px, xy_grid[cand, 1] - py)
        out[i] = cand[np.argsort(d, kind="stable")[:k]]
    return out


def _connectors(zone_ids, grid_node_ids, xy):
    """Bidirectional centroid connectors (t0 = 0.5 min, c = 1e5)."""
    z = np.repeat(zone_ids, grid_node_ids.shape[1])
    g = grid_node_ids.ravel()
    length = np.hypot(*(xy[z - 1] - xy[g - 1]).T)
    n = z.shape[0]
    return dict(init=np.concatenate([z, g]), term=np.concatenate([g, z]),
                capacity=np.full(2 * n, CONNECTOR_CAP), length=np.tile(length, 2),
                fftime=np.full(2 * n, CONNECTOR_T0), speed=np.zeros(2 * n),
                ltype=np.zeros(2 * n, dtype=np.int16))


def _zone_times_native(net: Network) -> np.ndarray:
    """Zone-to-zone free-flow times on host threads (libtapsynth.so, built
    by traffic_assignment_amd.build): the same rule as freeflow_zone_times
    (zones other than the source are not expanded), for networks where a
    single-threaded library Dijkstra over every zone is too slow."""
    import ctypes as C
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtapsynth.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built (run traffic_assignment_amd.build)")
    L = C.CDLL(path)
    L.tap_zone_times.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_int64, C.c_int64, C.c_void_p, C.c_int]
    Z = net.num_zones
    out = np.empty((Z, Z), dtype=np.float64)
    a = np.ascontiguousarray(net.init, dtype=np.int64)
    b = np.ascontiguousarray(net.term, dtype=np.int64)
    w = np.ascontiguousarray(net.fftime, dtype=np.float64)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    rc = L.tap_zone_times(net.num_nodes, net.num_links, a.ctypes.data, b.ctypes.data,
                          w.ctypes.data, Z, net.first_thru_node, out.ctypes.data, threads)
    if rc != 0:
        raise RuntimeError("tap_zone_times failed")
    return out


def freeflow_zone_times(net: Network) -> np.ndarray:
    """Zone-to-zone free-flow shortest-path times, zones not traversable.

    Each zone is split into a source copy (its out-links) and a sink copy
    (its in-links) so no path can pass through a centroid, which is the
    first-thru-node rule of trunk/tap/sssp_pape.c:355.
    """
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import dijkstra
    Z, N = net.num_zones, net.num_nodes
    u = net.init - 1
    v = net.term - 1
    v2 = np.where(v < net.first_thru_node - 1, N + v, v)
    g = csr_matrix((net.fftime, (u, v2)), shape=(N + Z, N + Z))
    d = dijkstra(g, directed=True, indices=np.arange(Z))
    return d[:, N:N + Z]


def _gravity(rng, Z, t0, total, beta=0.1, nonzero=None):
    """Poisson OD counts with mean k * P_i * A_j * exp(-beta * t0_ij), P and
    A ~ LogNormal(5, 1).  k makes the expected total `total`, or (nonzero
    given) the expected share of nonzero off-diagonal pairs `nonzero`
    (bisection on k, in log space)."""
    P = rng.lognormal(5.0, 1.0, size=Z)
    A = rng.lognormal(5.0, 1.0, size=Z)
    w = P[:, None] * A[None, :] * np.exp(-beta * t0)
    np.fill_diagonal(w, 0.0)
    w[~np.isfinite(w)] = 0.0
    if nonzero is None:
        k = total / w.sum()
    else:
        ww = w[~np.eye(Z, dtype=bool)]
        lo, hi = -40.0, 10.0                  # log k
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            if np.mean(-np.expm1(-np.exp(mid) * ww)) < nonzero:
                lo = mid
            else:
                hi = mid
        k = float(np.exp(0.5 * (lo + hi)))
    od = rng.poisson(k * w).astype(np.float64)
    np.fill_diagonal(od, 0.0)
    return od, k


# ---------------------------------------------------------------------------
# the workloads
# ---------------------------------------------------------------------------

def grid5(seed: int = 1):
    """24 nodes, 76 links: jittered 5x5 grid minus its last corner node."""
    rng = np.random.default_rng(seed)
    rows = cols = 5
    xy_all = _jittered_grid(rng, rows, cols, 1.0, 0.2)
    alive = np.ones(rows * cols, dtype=bool)
    alive[rows * cols - 1] = False
    ids = np.cumsum(alive)                  # 1-based id of each alive grid point
    e = _grid_edges(rows, cols, alive)
    a, b = ids[e[:, 0]], ids[e[:, 1]]
    init = np.concatenate([a, b]); term = np.concatenate([b, a])
    E = init.shape[0]
    xy = xy_all[alive]
    length = np.hypot(*(xy[init - 1] - xy[term - 1]).T)
    part = dict(init=init, term=term,
                capacity=rng.uniform(4000.0, 26000.0, size=E),
                length=length, fftime=rng.uniform(2.0, 10.0, size=E),
                speed=np.zeros(E), ltype=np.ones(E, dtype=np.int16))
    N = int(alive.sum())
    net = _assemble(N, N, 1, [part], xy,
                    dict(workload="grid5", seed=seed, rows=rows, cols=cols, jitter=0.2,
                         removed="corner (5,5)", t0="U[2,10]", capacity="U[4000,26000]"))
    od = rng.integers(0, 1001, size=(N, N)).astype(np.float64)
    np.fill_diagonal(od, 0.0)
    off = np.flatnonzero(~np.eye(N, dtype=bool).ravel())
    nzero = int(round(0.4 * off.size))
    zero_idx = rng.choice(off, size=nzero, replace=False)
    od.ravel()[zero_idx] = 0.0
    dem = Demand.from_dense(od, keep_zero=True,
                            params=dict(rule="U{0..1000}, 40% of off-diagonal zeroed"))
    return net, dem


def _grid_city(seed, rows, cols, zones, conn_per_zone, shortcuts, spacing, jitter,
               total_demand, impedance, workload, nonzero=None, demand_factor=1.0):
    rng = np.random.default_rng(seed)
    G = rows * cols
    xy_grid = _jittered_grid(rng, rows, cols, spacing, jitter)
    ext = np.array([(cols - 1) * spacing, (rows - 1) * spacing])
    xy_zone = rng.uniform(0.0, 1.0, size=(zones, 2)) * ext
    xy = np.concatenate([xy_zone, xy_grid])
    gid = lambda gidx: gidx + zones + 1          # grid index -> node id
    e = _grid_edges(rows, cols)
    parts = [_classed_links(rng, gid(e[:, 0]), gid(e[:, 1]), xy)]
    if shortcuts:
        have = set()
        sa, sb = [], []
        while len(sa) < shortcuts:
            a = int(rng.integers(G))
            L = rng.uniform(5.0, 20.0)
            th = rng.uniform(0.0, 2 * np.pi)
            px, py = xy_grid[a] + L * np.array([np.cos(th), np.sin(th)])
            if not (0 <= px <= ext[0] and 0 <= py <= ext[1]):
                continue
            b = int(_nearest_grid(xy_grid, rows, cols, spacing, np.array([[px, py]]), 1)[0, 0])
            key = (min(a, b), max(a, b))
            if a == b or key in have:
                continue
            have.add(key); sa.append(a); sb.append(b)
        sa = gid(np.array(sa)); sb = gid(np.array(sb))
        length = np.hypot(*(xy[sa - 1] - xy[sb - 1]).T)
        cap = CLASS_CAP_PER_LANE[1] * CLASS_LANES[1]
        parts.append(dict(init=np.concatenate([sa, sb]), term=np.concatenate([sb, sa]),
                          capacity=np.full(2 * len(sa), cap),
                          length=np.tile(length, 2),
                          fftime=np.tile(length / CLASS_SPEED_KMH[1] * 60.0, 2),
                          speed=np.full(2 * len(sa), CLASS_SPEED_KMH[1]),
                          ltype=np.full(2 * len(sa), 2, dtype=np.int16)))
    near = _nearest_grid(xy_grid, rows, cols, spacing, xy_zone, conn_per_zone)
    parts.append(_connectors(np.arange(1, zones + 1), gid(near), xy))
    params = dict(workload=workload, seed=seed, rows=rows, cols=cols, zones=zones,
                  connectors_per_zone=conn_per_zone, shortcuts=shortcuts,
                  spacing_km=spacing, jitter=jitter, class_prob=CLASS_PROB,
                  class_speed_kmh=CLASS_SPEED_KMH, class_cap_per_lane=CLASS_CAP_PER_LANE,
                  class_lanes=CLASS_LANES, connector_t0=CONNECTOR_T0,
                  connector_cap=CONNECTOR_CAP, shortcuts_bidirectional=True)
    net = _assemble(zones, zones + G, zones + 1, parts, xy, params)
    if impedance == "sp":
        t0 = freeflow_zone_times(net)
    elif impedance == "sp-native":
        t0 = _zone_times_native(net)
    else:  # straight-line proxy at 50 km/h plus the two connectors
        d = np.hypot(xy_zone[:, None, 0] - xy_zone[None, :, 0],
                     xy_zone[:, None, 1] - xy_zone[None, :, 1])
        t0 = d / 50.0 * 60.0 + 2 * CONNECTOR_T0
    od, k = _gravity(rng, zones, t0, total_demand, nonzero=nonzero)
    counts_total = float(od.sum())
    nz = int(np.count_nonzero(od))
    if demand_factor is None:
        # a power of two near target / counts: every partial sum of demand
        # stays exact in FP64
        demand_factor = 2.0 ** round(np.log2(total_demand / max(counts_total, 1.0)))
    if demand_factor != 1.0:
        od = od * demand_factor
    dem = Demand.from_dense(od, params=dict(model="gravity Poisson", beta=0.1,
                                            P_A="LogNormal(5,1)",
                                            impedance=("free-flow shortest path"
                                                       if impedance.startswith("sp")
                                                       else impedance),
                                            target_total=total_demand,
                                            target_nonzero_share=nonzero, scale_k=k,
                                            count_total=counts_total,
                                            nonzero_pairs=nz,
                                            nonzero_share=nz / max(zones * (zones - 1), 1),
                                            demand_per_count=demand_factor,
                                            total=float(od.sum())))
    return net, dem


def metro(seed: int = 2, scale: float = 1.0, total_demand: float | None = None):
    """Jittered 100x100 grid + 1,000 zones (x scale^2 in grid area and zones)."""
    rows = cols = max(4, int(round(100 * scale)))
    zones = max(4, int(round(1000 * scale * scale)))
    if total_demand is None:
        total_demand = 400000.0 * scale * scale
    return _grid_city(seed, rows, cols, zones, 2, 0, 0.5, 0.25, total_demand, "sp",
                      "metro" if scale == 1.0 else f"metro@{scale:g}")


REGIONAL_NONZERO = 0.30          # share of nonzero OD pairs (Poisson trip counts)
REGIONAL_LOAD = 1.8e6            # assigned demand at full scale (x scale^2)


def regional(seed: int = 3, scale: float = 1.0):
    """Jittered 300x300 grid + 2,000 shortcuts + 4,000 zones (x scale^2).

    Demand: gravity Poisson trip counts with the free-flow shortest-path
    impedance, k chosen so that 30% of the off-diagonal OD pairs are
    expected nonzero (about 114M counted trips at full scale).  The assigned
    demand is counts x 2^m, the power of two nearest 1.8M x scale^2 / counts
    (2^-6 at full scale: about 1.79M), so the network carries the load of a
    congested peak hour rather than many times its capacity, at every
    scale."""
    rows = cols = max(6, int(round(300 * scale)))
    zones = max(4, int(round(4000 * scale * scale)))
    short = int(round(2000 * scale * scale))
    return _grid_city(seed, rows, cols, zones, 3, short, 0.5, 0.25, REGIONAL_LOAD * scale * scale,
                      "sp-native", "regional" if scale == 1.0 else f"regional@{scale:g}",
                      nonzero=REGIONAL_NONZERO, demand_factor=None)


# ---------------------------------------------------------------------------
# upgrade variants (network modifications, as in tap_mpi_main.c:201-240)
# ---------------------------------------------------------------------------

@dataclass
class NetMod:
    change_id: str
    modtype: str                 # "ADD" or "CHANGE"
    link: dict                   # full link record (init, term, capacity, ...)
    project_cost: float


def upgrades(net: Network, seed: int = 4, n_change: int = 20, n_add: int = 12):
    """32 candidate upgrades: capacity x1.5 on freeway/arterial roads (both
    directions, one CHANGE id) and new bidirectional arterials of 1-5 km
    between grid nodes 2-4 grid hops apart (one ADD id, two rows)."""
    rng = np.random.default_rng(seed)
    rows, cols, Z = net.params["rows"], net.params["cols"], net.num_zones
    mods = []
    key = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(net.init, net.term))}
    cand = np.flatnonzero((net.ltype == 1) | (net.ltype == 2))
    cand = cand[net.init[cand] < net.term[cand]]
    pick = rng.choice(cand, size=n_change, replace=False)
    for j, k in enumerate(pick):
        cost = float(rng.uniform(1.0, 10.0))
        for kk in (k, key[(int(net.term[k]), int(net.init[k]))]):
            rec = {f: (int(getattr(net, f)[kk]) if f in ("init", "term", "ltype")
                       else float(getattr(net, f)[kk]))
                   for f in ("init", "term", "capacity", "length", "fftime", "B",
                             "power", "speed", "toll", "ltype")}
            rec["capacity"] = rec["capacity"] * 1.5
            mods.append(NetMod(f"U{j:02d}", "CHANGE", rec, cost))
    added = set()
    j = n_change
    while j < n_change + n_add:
        r, c = int(rng.integers(rows)), int(rng.integers(cols))
        hops = int(rng.integers(2, 5))
        dr = int(rng.integers(0, hops + 1)); dc = hops - dr
        r2 = r + dr * (1 if rng.random() < 0.5 else -1)
        c2 = c + dc * (1 if rng.random() < 0.5 else -1)
        if not (0 <= r2 < rows and 0 <= c2 < cols):
            continue
        a = Z + 1 + r * cols + c; b = Z + 1 + r2 * cols + c2
        if (a, b) in key or (min(a, b), max(a, b)) in added:
            continue
        added.add((min(a, b), max(a, b)))
        length = float(rng.uniform(1.0, 5.0))
        cost = float(rng.uniform(1.0, 10.0))
        for (u, v) in ((a, b), (b, a)):
            rec = dict(init=u, term=v, capacity=CLASS_CAP_PER_LANE[1] * CLASS_LANES[1],
                       length=length, fftime=length / CLASS_SPEED_KMH[1] * 60.0,
                       B=BPR_B, power=BPR_POWER, speed=CLASS_SPEED_KMH[1], toll=0.0,
                       ltype=2)
            mods.append(NetMod(f"U{j:02d}", "ADD", rec, cost))
        j += 1
    return mods


def change_ids(mods):
    seen = []
    for m in mods:
        if m.change_id not in seen:
            seen.append(m.change_id)
    return seen


def budget_subsets(mods, n_variants: int = 512, budget: float = 40.0, seed: int = 5):
    """Random subsets of change ids whose summed project cost is <= budget."""
    rng = np.random.default_rng(seed)
    ids = change_ids(mods)
    cost = {m.change_id: m.project_cost for m in mods}
    out = []
    while len(out) < n_variants:
        order = rng.permutation(len(ids))
        target = int(rng.integers(1, len(ids) + 1))
        chosen, tot = [], 0.0
        for i in order:
            c = cost[ids[i]]
            if tot + c <= budget:
                chosen.append(ids[i]); tot += c
            if len(chosen) >= target:
                break
        out.append(sorted(chosen))
    return out


def apply_mods(net: Network, mods, chosen) -> Network:
    """Apply the rows of the chosen change ids: ADD appends a link record,
    CHANGE overwrites the whole record of the (init, term) link (error if
    absent); then re-sort by (init, term) (tap_mpi_main.c:201-240, :753)."""
    n = net.copy()
    fields = ("init", "term", "capacity", "length", "fftime", "B", "power", "speed",
              "toll", "ltype")
    chosen = set(chosen)
    add = {f: [] for f in fields}
    index = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(n.init, n.term))}
    for m in mods:
        if m.change_id not in chosen:
            continue
        if m.modtype == "ADD":
            for f in fields:
                add[f].append(m.link[f])
        else:
            k = index.get((m.link["init"], m.link["term"]))
            if k is None:
                raise KeyError(f"link not found for change {m.change_id}")
            for f in fields:
                getattr(n, f)[k] = m.link[f]
    if add["init"]:
        for f in fields:
            cur = getattr(n, f)
            setattr(n, f, np.concatenate([cur, np.asarray(add[f], dtype=cur.dtype)]))
    return n.sorted()


# ---------------------------------------------------------------------------
# named configurations
# ---------------------------------------------------------------------------

def make(name: str, seed: int | None = None, scale: float = 1.0):
    """(net, demand) for a named workload: grid5 | metro | regional."""
    if name == "grid5":
        return grid5(1 if seed is None else seed)
    if name == "metro":
        return metro(2 if seed is None else seed, scale=scale)
    if name == "regional":
        return regional(3 if seed is None else seed, scale=scale)
    raise ValueError(name)
