"""Frank-Wolfe iterations replayed from captured HIP graphs against the same
launches issued one by one: every replay (one per iteration) must give the
eager run's step sizes, relative gaps, volumes and costs bit for bit.

Each run goes in chunks of a few iterations with the program's own eager
work between the replays (volume readback, an extra loading, synchronising
copies), then to relgap 1e-4 (at most 400 iterations); both runs are compared after every chunk and
over every iteration's trace record.  Cases: the per-origin LDS and
labels-in-HBM kernels (whole iteration captured; also 4 and 8 iterations
per replay, where a chunk's iteration cap falls inside a replay), the node
sweeps and the
LDS kernel with HBM co-run origins (link passes captured), the lanes kernel
on one variant and on a 12-variant batch (whole iteration captured), and
the lanes kernel on a network with zero-cost links, whose level ties stop
the captured step and send every loading to a per-origin kernel."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import network, random_demand, random_network

pytestmark = pytest.mark.gpu
CASES = ["lds", "hbm", "sweeps", "lds-split", "lanes", "lanes-batch", "lanes-ties", "lds-x4", "hbm-x8"]


def _engine(engine_mod, case):
    from traffic_assignment_amd import synth, variants
    if case == "lanes-batch":
        base, dem = synth.metro(scale=0.3)
        mods = synth.upgrades(base, n_change=6, n_add=4)
        subsets = synth.budget_subsets(mods, n_variants=12, budget=25.0)
        eng = variants.make_engine(base, dem, mods, subsets)[0]
        eng.set_lanes(4)
        return eng
    if case == "lanes-ties":
        net, rng = random_network(13, "ints")
        eng = engine_mod.TapEngine(net, random_demand(net, rng))
        eng.set_lanes(1)
        return eng
    net, dem = network("metro", 0.3)
    eng = engine_mod.TapEngine(net, dem)
    case = case.split("-x")[0]
    if case == "lanes":
        eng.set_lanes(1)
    elif case == "lds-split":
        eng.set_strategy("lds")
        eng.set_origin_split(50)
    else:
        eng.set_strategy(case)
    return eng


def _run(eng, graphs, chunks):
    eng.set_graphs(graphs)
    assert eng.graphs == graphs
    V = eng.V
    eng.fw_init()
    out, launches, fallbacks = [], [], 0
    for k in chunks:
        res = eng.fw_run(max_iters=k) if k else eng.fw_run(max_iters=400, target_relgap=1e-4)
        st = eng.stats()
        launches.append(st["graph_launches"])
        fallbacks += st["lanes_fallbacks"]
        out.append(([r["iterations"] for r in res], [eng.flows(v) for v in range(V)]))
        if graphs:
            eng.aon()                       # eager loading between replays
    traces = [eng.trace(v) for v in range(V)]
    return out, traces, launches, fallbacks


@pytest.mark.parametrize("case", CASES)
def test_graph_replays_equal_eager_launches(engine_mod, case):
    print(f"case graphs {case}", flush=True)
    chunks = (3, 7, 11, 19, 0)
    with _engine(engine_mod, case) as eng:
        a, ta, la, fa = _run(eng, False, chunks)
    mode = int(case.split("-x")[1]) if "-x" in case else 1
    with _engine(engine_mod, case) as eng:
        b, tb, lb, fb = _run(eng, mode, chunks)
    assert all(n == 0 for n in la)
    assert all(n > 0 for n in lb), lb
    assert fa == fb and (fb > 0) == (case == "lanes-ties"), (fa, fb)
    for (ia, fa), (ib, fb) in zip(a, b):
        assert ia == ib
        for (xa, ca), (xb, cb) in zip(fa, fb):
            assert np.array_equal(xa, xb) and np.array_equal(ca, cb)
    for x, y in zip(ta, tb):
        for k in ("lambda_", "relgap", "objective"):
            assert np.array_equal(x[k], y[k]), k
    if case != "lanes-ties":
        assert all(np.all(t["relgap"][-1] <= 1e-4) for t in tb)


def test_graph_recaptured_when_schedule_changes(engine_mod):
    """The captured iteration bakes in the loading's schedule: changing the
    chain length, the queue size or the strategy between runs must be
    picked up (same volumes as the eager run, captured again)."""
    print("case graphs recapture", flush=True)
    net, dem = network("metro", 0.3)
    res = {}
    for graphs in (False, True):
        with engine_mod.TapEngine(net, dem) as eng:
            eng.set_graphs(graphs)
            eng.set_strategy("lds")
            eng.fw_init()
            eng.fw_run(max_iters=4)
            eng.set_origin_chase(8)
            eng.fw_run(max_iters=8)
            eng.set_origin_queue(16)
            eng.fw_run(max_iters=12)
            eng.set_strategy("hbm")
            eng.set_origin_delta(2.0)
            eng.fw_run(max_iters=16)
            eng.set_origin_queue(-1)
            eng.set_origin_delta(0.0)
            eng.fw_run(max_iters=20)
            res[graphs] = (eng.flows(), eng.trace()["lambda_"], eng.stats()["graph_launches"])
    assert res[True][2] >= 1 and res[False][2] == 0   # (short iterations go several to a replay)
    assert np.array_equal(res[True][0][0], res[False][0][0])
    assert np.array_equal(res[True][0][1], res[False][0][1])
    assert np.array_equal(res[True][1], res[False][1])
