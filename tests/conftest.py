"""Shared fixtures.  `-m gpu` tests need an MI355X and the in-tree
libtapgpu.so; everything else runs on the CPU against the oracle."""
from __future__ import annotations

import base64
import functools
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


@functools.lru_cache(maxsize=None)
def network(name: str, scale: float = 1.0):
    from traffic_assignment_amd import synth
    return synth.make(name, scale=scale)


@functools.lru_cache(maxsize=None)
def golden(name: str):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def b64_array(s: str, dtype) -> np.ndarray:
    return np.frombuffer(base64.b64decode(s), dtype="<f8" if dtype == np.float64 else "<i8").copy()


def hexarr(rows):
    return np.array([[float.fromhex(v) for v in r] for r in rows])


def sha(a: np.ndarray) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).astype(
        "<f8" if a.dtype.kind == "f" else "<i8").tobytes()).hexdigest()


def rmat(tag: str):
    """(Network, 1-based) from a reference SSSP input graph fixture."""
    from traffic_assignment_amd import synth
    z = np.load(os.path.join(GOLDEN, f"rmat_{tag}.npz"))
    fr, to, cost = z["frm"], z["to"], z["cost"]
    n = int(max(fr.max(), to.max()))
    order = np.lexsort((to, fr))
    E = fr.size
    net = synth.Network(num_zones=1, num_nodes=n, first_thru_node=1, init=fr[order].astype(np.int64),
                        term=to[order].astype(np.int64), capacity=np.ones(E), length=np.ones(E),
                        fftime=cost[order].astype(np.float64), B=np.zeros(E), power=np.ones(E),
                        speed=np.zeros(E), toll=np.zeros(E), ltype=np.ones(E, np.int16))
    return net


def empty_demand(Z):
    from traffic_assignment_amd import synth
    return synth.Demand(Z, np.zeros(Z + 1, np.int64), np.zeros(0, np.int64), np.zeros(0))


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    oracle.lib().ora_set_pow_mode(0)
    return oracle


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def engine_mod():
    """The product's Python front end; where a GPU is present a missing
    library fails instead of skipping."""
    if not gpu_available():
        pytest.skip("no GPU")
    from traffic_assignment_amd import engine
    engine.lib()
    return engine


def pred_tie_check(net, costs, dist, pred, o_pred, origins):
    """Number of (node, origin) where pred differs from the oracle's and the
    difference is NOT explained by an exact tie; and the number of ties.
    dist/pred are [(N+1), len(origins)] arrays (row = node id)."""
    ft = net.first_thru_node
    term_sorted = np.argsort(net.term, kind="stable")
    starts = np.searchsorted(net.term[term_sorted], np.arange(net.num_nodes + 2))
    bad = ties = 0
    diff = np.argwhere(pred != o_pred)
    for v, j in diff:
        if v == 0:
            continue
        src = origins[j]
        ks = term_sorted[starts[v]:starts[v + 1]]
        cand = [k for k in ks if (net.init[k] >= ft or net.init[k] == src)
                and dist[net.init[k], j] < 1.7e308
                and dist[net.init[k], j] + costs[k] == dist[v, j]]
        if pred[v, j] in cand and o_pred[v, j] in cand:
            ties += 1
        else:
            bad += 1
    return bad, ties


def random_network(seed, cost_kind):
    """A seeded random network with the awkward cases a road file can hold:
    parallel links, self-loops, zero-cost links (zero-cost cycles with
    cost_kind "ints"), non-through zones, start nodes that are not zones,
    costs over twelve orders of magnitude ("wide").  Through nodes are
    joined by a two-way ring and every non-through node by a two-way link to
    one through node, so every zone reaches every other.  "pints": costs
    1, 2 or 3 (thousands of exact ties, every link costs something)."""
    from traffic_assignment_amd import synth
    rng = np.random.default_rng(seed)
    N = int(rng.integers(40, 2500))
    Z = int(rng.integers(2, min(40, N // 2)))
    ft = int([1, Z // 2 + 1, Z + 1][seed % 3])
    E0 = int(N * rng.uniform(1.5, 4.0))
    init = [rng.integers(1, N + 1, E0)]
    term = [rng.integers(1, N + 1, E0)]
    thru = np.arange(ft, N + 1)
    init += [thru, np.roll(thru, -1)]
    term += [np.roll(thru, -1), thru]
    if ft > 1:
        z = np.arange(1, ft)
        t = rng.integers(ft, N + 1, z.size)
        init += [z, t]
        term += [t, z]
    dup = rng.integers(0, E0, E0 // 20)                  # parallel links
    init += [init[0][dup]]
    term += [term[0][dup]]
    loops = rng.integers(1, N + 1, 5)                    # self-loops
    init += [loops]
    term += [loops]
    init = np.concatenate(init).astype(np.int64)
    term = np.concatenate(term).astype(np.int64)
    order = np.lexsort((term, init))
    init, term = init[order], term[order]
    E = init.size
    if cost_kind == "ints":
        cost = rng.integers(0, 4, E).astype(np.float64)
    elif cost_kind == "pints":           # exact ties, but no zero-cost links
        cost = rng.integers(1, 4, E).astype(np.float64)
    elif cost_kind == "wide":
        cost = 10.0 ** rng.uniform(-6, 6, E)
    else:
        cost = rng.uniform(0.01, 5.0, E)
        if cost_kind == "cont":
            cost[rng.random(E) < 0.15] = 0.0
    net = synth.Network(num_zones=Z, num_nodes=N, first_thru_node=ft, init=init, term=term,
                        capacity=rng.uniform(20.0, 200.0, E), length=np.ones(E), fftime=cost,
                        B=np.full(E, 0.15), power=np.full(E, 4.0), speed=np.zeros(E),
                        toll=np.zeros(E), ltype=np.ones(E, np.int16))
    return net, rng


def random_demand(net, rng):
    """Integer demand from every zone to a random subset of the others."""
    from traffic_assignment_amd import synth
    Z = net.num_zones
    dest = [np.sort(rng.choice(np.setdiff1d(np.arange(1, Z + 1), [o]),
                               size=int(rng.integers(0, Z)), replace=False)) for o in range(1, Z + 1)]
    start = np.r_[0, np.cumsum([d.size for d in dest])].astype(np.int64)
    allr = np.concatenate(dest).astype(np.int64)
    return synth.Demand(Z, start, allr, rng.integers(0, 30, allr.size).astype(np.float64))


def pred_reaches_origin(net, dist, pred, origins):
    """Every reachable node's predecessor chain ends at its origin (the
    predecessor links form a tree, zero-cost cycles included)."""
    n1 = net.num_nodes + 1
    for j, s in enumerate(origins):
        p = pred[:, j]
        anc = np.arange(n1)
        has = p >= 0
        anc[has] = net.init[p[has]]
        for _ in range(int(np.ceil(np.log2(n1))) + 1):
            anc = anc[anc]
        reach = dist[:, j] < 1.7e308
        reach[0] = False
        assert np.all(anc[reach] == s), (s, np.nonzero(anc[reach] != s)[0][:5])
