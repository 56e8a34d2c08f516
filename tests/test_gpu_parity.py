"""Parity of the gfx950 engine (through the C ABI) with the oracle and the
reference's golden vectors.  Bar: distances bit-exact, predecessor links
equal wherever the shortest path is unique, loads exact (integer demand),
the step-size sequence identical, volumes and costs bit-identical to the
oracle with the correctly-rounded power (and within 1e-9 relative of the
reference's glibc-pow run), relative gap / objective / VHT within 1e-9."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import (b64_array, empty_demand, golden, hexarr, network, pred_reaches_origin,
                      pred_tie_check, random_demand, random_network, rmat)

pytestmark = pytest.mark.gpu
TOL = 1e-9
STRATS = ["sweeps", "lds", "hbm"]
# loading paths: the three strategies plus the lanes kernel (variants in
# wavefront lanes; here one variant, lane 0 of its group)
LOAD_STRATS = STRATS + ["lanes"]


def make_engine(engine_mod, net, dem, strategy, **kw):
    eng = engine_mod.TapEngine(net, dem, **kw)
    if strategy == "lanes":
        eng.set_lanes(1)
        return eng
    if strategy == "lds" and eng.strategy != "lds":
        eng.close()
        pytest.skip("network does not fit the LDS path")
    eng.set_strategy(strategy)
    return eng


def rel(a, b):
    a = np.asarray(a, dtype=float); b = np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.size else 0.0


def _sssp_case(engine_mod, oracle_lib, net, dem, costs, origins, strategy="sweeps"):
    with make_engine(engine_mod, net, dem, strategy) as eng:
        if costs is not net.fftime:
            eng.set_costs(costs)
        d, p = eng.sssp(origins)
    od = np.empty_like(d); op = np.empty_like(p)
    for j, s in enumerate(origins):
        od[:, j], op[:, j] = oracle_lib.sssp(net, costs, int(s))
    od[0] = np.finfo(float).max; op[0] = -1
    assert np.array_equal(d.view(np.int64), od.view(np.int64))
    bad, ties = pred_tie_check(net, costs, d, p, op, origins)
    assert bad == 0, (bad, ties)
    pred_reaches_origin(net, d, p, origins)
    return ties


@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("name,scale", [("grid5", 1.0), ("metro", 0.3), ("regional", 0.25)])
def test_sssp_synthetic(engine_mod, oracle_lib, name, scale, strategy):
    print(f"case sssp {name}@{scale} {strategy}", flush=True)
    net, dem = network(name, scale)
    origins = np.arange(1, net.num_zones + 1)
    _sssp_case(engine_mod, oracle_lib, net, dem, net.fftime, origins, strategy)
    rng = np.random.default_rng(5)
    costs = net.fftime * rng.uniform(1.0, 4.0, net.num_links)
    _sssp_case(engine_mod, oracle_lib, net, dem, costs, origins, strategy)


@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("tag", ["martin", "okuyama", "chicagosketch_flows",
                                 "chicagoregional_flows"])
def test_sssp_reference_graphs(engine_mod, oracle_lib, tag, strategy):
    """The reference's own SSSP inputs (zero-cost links included)."""
    print(f"case sssp rmat {tag} {strategy}", flush=True)
    net = rmat(tag)
    n = net.num_nodes
    origins = np.unique(np.r_[1, n, np.random.default_rng(9).integers(1, n + 1, 90)])
    ties = _sssp_case(engine_mod, oracle_lib, net, empty_demand(1), net.fftime, origins, strategy)
    print("equal-cost ties:", ties)


@pytest.mark.parametrize("strategy", STRATS)
@pytest.mark.parametrize("seed,cost_kind", [(1, "cont"), (2, "ints"), (3, "wide"), (4, "ints"),
                                            (5, "cont"), (6, "wide")])
def test_sssp_random_networks(engine_mod, oracle_lib, seed, cost_kind, strategy):
    """Random networks (see random_network): labels bit-exact, predecessor
    links equal to the oracle's except at exact ties, and a tree."""
    print(f"case sssp random seed={seed} {cost_kind} {strategy}", flush=True)
    net, rng = random_network(seed, cost_kind)
    origins = np.unique(np.r_[np.arange(1, net.num_zones + 1),
                              rng.integers(1, net.num_nodes + 1, 10)])
    with make_engine(engine_mod, net, empty_demand(net.num_zones), strategy) as eng:
        d, p = eng.sssp(origins)
    od = np.empty_like(d); op = np.empty_like(p)
    for j, s in enumerate(origins):
        od[:, j], op[:, j] = oracle_lib.sssp(net, net.fftime, int(s))
    od[0] = np.finfo(float).max; op[0] = -1
    assert np.array_equal(d.view(np.int64), od.view(np.int64))
    bad, ties = pred_tie_check(net, net.fftime, d, p, op, origins)
    assert bad == 0, (bad, ties)
    pred_reaches_origin(net, d, p, origins)
    print("equal-cost ties:", ties)


def _shortest_path_loading_identities(oracle_lib, net, dem, y):
    """Exact checks that hold for every shortest-path loading (integers
    throughout): total cost sum y*c equals sum over OD of demand x label,
    and flow balances at every node."""
    Z = net.num_zones
    orig = np.repeat(np.arange(1, Z + 1), np.diff(dem.start))
    keep = (dem.dest != orig) & (dem.demand > 0)
    tot = 0.0
    for o in range(1, Z + 1):
        m = keep & (orig == o)
        if m.any():
            d, _ = oracle_lib.sssp(net, net.fftime, o)
            tot += float(np.sum(dem.demand[m] * d[dem.dest[m]]))
    assert float(np.sum(y * net.fftime)) == tot
    bal = np.zeros(net.num_nodes + 1)
    np.add.at(bal, net.term, y)
    np.add.at(bal, net.init, -y)
    want = np.zeros(net.num_nodes + 1)
    np.add.at(want, dem.dest[keep], dem.demand[keep])
    np.add.at(want, orig[keep], -dem.demand[keep])
    assert np.array_equal(bal, want)


@pytest.mark.parametrize("strategy", LOAD_STRATS)
@pytest.mark.parametrize("seed,cost_kind", [(7, "cont"), (8, "wide"), (9, "cont"), (12, "wide"),
                                            (13, "ints"), (14, "pints"), (15, "pints")])
def test_aon_random_networks(engine_mod, oracle_lib, seed, cost_kind, strategy):
    """Integer demand between random zone pairs on random networks: volumes
    exactly the oracle's."""
    print(f"case aon random seed={seed} {cost_kind} {strategy}", flush=True)
    net, rng = random_network(seed, cost_kind)
    dem = random_demand(net, rng)
    with make_engine(engine_mod, net, dem, strategy) as eng:
        y = eng.aon()
        st = eng.stats()
    if cost_kind == "ints":
        # small-integer costs, zero-cost cycles: thousands of exact ties,
        # which Pape's queue order and the engine's rules break differently,
        # so the trees (and volumes) may differ from the oracle's; the
        # identities of any shortest-path loading hold.  The per-origin
        # kernels and the lanes kernel (which hands level ties to them)
        # share one rule: bit-identical volumes.
        _shortest_path_loading_identities(oracle_lib, net, dem, y)
        if strategy != "sweeps":
            with make_engine(engine_mod, net, dem, "hbm") as ref:
                assert np.array_equal(y, ref.aon())
    elif cost_kind == "pints":
        # costs 1, 2, 3: thousands of exact ties, no zero-cost links.  The
        # per-origin and lanes kernels pick by the engine's rule, which the
        # oracle's tie mode 1 restates: bit-identical volumes.  The node
        # sweeps pick during their rounds; the identities hold.
        _shortest_path_loading_identities(oracle_lib, net, dem, y)
        if strategy != "sweeps":
            lib = oracle_lib.lib()
            lib.ora_set_tie_mode(1)
            try:
                o1 = oracle_lib.fw_assign(net, dem, net.fftime, nthreads=8)
            finally:
                lib.ora_set_tie_mode(0)
            assert np.array_equal(y, o1)
    else:
        assert np.array_equal(y, oracle_lib.fw_assign(net, dem, net.fftime, nthreads=8))
    if strategy == "lanes":
        # costs over twelve orders of magnitude but none zero: the lanes
        # kernel itself loads; zero-cost links make level ties, which a
        # per-origin kernel redoes
        print("lanes launches", st["lanes_launches"], "fallbacks", st["lanes_fallbacks"])
        assert st["lanes_launches"] + st["lanes_fallbacks"] == 1
        if cost_kind in ("wide", "pints"):
            assert st["lanes_fallbacks"] == 0


@pytest.mark.parametrize("strategy", LOAD_STRATS)
@pytest.mark.parametrize("seed", [10, 11])
def test_fw_random_networks(engine_mod, oracle_lib, seed, strategy):
    """30 FW iterations on random networks (BPR, power 4, random
    capacities, positive free-flow times): step sizes, volumes and costs
    bit-identical to the oracle with the correctly rounded power."""
    print(f"case fw random seed={seed} {strategy}", flush=True)
    net, rng = random_network(seed, "pos")
    dem = random_demand(net, rng)
    with make_engine(engine_mod, net, dem, strategy) as eng:
        eng.fw_init()
        eng.fw_run(max_iters=30)
        x, c = eng.flows()
        tr = eng.trace()
    oracle_lib.lib().ora_set_pow_mode(1)
    try:
        o = oracle_lib.fw_solve(net, dem, max_iters=30, nthreads=8)
    finally:
        oracle_lib.lib().ora_set_pow_mode(0)
    assert np.array_equal(tr["lambda_"], o["trace"]["lambda_"])
    assert np.array_equal(x, o["x"]) and np.array_equal(c, o["c"])


def test_sssp_pape_cuda_layout(engine_mod, oracle_lib):
    """Output layout of pape_cuda: [v * Z + s] with s the 1-based start node."""
    import ctypes as C
    net, dem = network("grid5")
    Z = net.num_zones
    with engine_mod.TapEngine(net, dem) as eng:
        s = np.arange(1, Z + 1, dtype=np.int64)
        d = np.full((net.num_nodes + 1) * (Z + 1), -7.0)
        p = np.full((net.num_nodes + 1) * (Z + 1), -7, dtype=np.int64)
        rc = engine_mod.lib().tapgpu_sssp(eng._h, 0, s.ctypes.data, Z, d.ctypes.data,
                                          p.ctypes.data, Z, 1)
        assert rc == 0
    for o in range(1, Z + 1):
        od, op = oracle_lib.sssp(net, net.fftime, o)
        for v in range(1, net.num_nodes + 1):
            assert d[v * Z + o] == od[v] and p[v * Z + o] == op[v]


def test_sssp_unreachable_and_bad_args(engine_mod):
    from traffic_assignment_amd import synth
    # 1 -> 2, 3 isolated
    net = synth.Network(num_zones=1, num_nodes=3, first_thru_node=1, init=np.array([1]),
                        term=np.array([2]), capacity=np.ones(1), length=np.ones(1),
                        fftime=np.array([1.5]), B=np.zeros(1), power=np.ones(1),
                        speed=np.zeros(1), toll=np.zeros(1), ltype=np.ones(1, np.int16))
    with engine_mod.TapEngine(net, empty_demand(1)) as eng:
        d, p = eng.sssp([1, 3])
        assert d[2, 0] == 1.5 and p[2, 0] == 0
        assert d[3, 0] == np.finfo(float).max and p[3, 0] == -1
        assert d[3, 1] == 0.0 and d[1, 1] == np.finfo(float).max
        with pytest.raises(engine_mod.TapGpuError):
            eng.sssp([4])


@pytest.mark.parametrize("strategy", LOAD_STRATS)
@pytest.mark.parametrize("name,scale", [("grid5", 1.0), ("metro", 0.3), ("metro", 1.0),
                                        ("regional", 0.25)])
def test_aon_exact(engine_mod, oracle_lib, name, scale, strategy):
    print(f"case aon {name}@{scale} {strategy}", flush=True)
    net, dem = network(name, scale)
    rng = np.random.default_rng(11)
    costs = net.fftime * rng.uniform(1.0, 2.0, net.num_links)
    with make_engine(engine_mod, net, dem, strategy) as eng:
        y0 = eng.aon()
        eng.set_costs(costs)
        y1 = eng.aon()
    assert np.array_equal(y0, oracle_lib.fw_assign(net, dem, net.fftime, nthreads=8))
    assert np.array_equal(y1, oracle_lib.fw_assign(net, dem, costs, nthreads=8))


@pytest.mark.parametrize("strategy", LOAD_STRATS)
def test_aon_no_path_is_an_error(engine_mod, strategy):
    from traffic_assignment_amd import synth
    net = synth.Network(num_zones=2, num_nodes=2, first_thru_node=1, init=np.array([1]),
                        term=np.array([1]), capacity=np.ones(1), length=np.ones(1),
                        fftime=np.ones(1), B=np.zeros(1), power=np.ones(1), speed=np.zeros(1),
                        toll=np.zeros(1), ltype=np.ones(1, np.int16))
    dem = synth.Demand(2, np.array([0, 1, 1]), np.array([2]), np.array([5.0]))
    with make_engine(engine_mod, net, dem, strategy) as eng:
        with pytest.raises(engine_mod.NoPathError, match="no path from 1 to 2"):
            eng.aon()


@pytest.mark.parametrize("strategy", LOAD_STRATS)
@pytest.mark.parametrize("name,scale,fix", [("grid5", 1.0, "grid5_fw.json"),
                                            ("metro", 0.3, "metro_s03_fw.json"),
                                            ("metro", 1.0, "metro_fw.json")])
def test_fw_against_reference_golden(engine_mod, oracle_lib, name, scale, fix, strategy):
    print(f"case fw {name}@{scale} {strategy}", flush=True)
    g = golden(fix)
    net, dem = network(name, scale)
    it = g["iterations"]
    with make_engine(engine_mod, net, dem, strategy) as eng:
        eng.fw_init()
        r = eng.fw_run(max_iters=it)[0]
        x, c = eng.flows()
        tr = eng.trace()
    assert r["iterations"] == it
    ref = hexarr(g["trace"])
    fields = engine_mod.ITER_FIELDS
    mine = np.stack([tr[f] for f in fields], axis=1)
    assert np.array_equal(mine[:, 0], ref[:, 0])                  # step sizes
    for j in range(1, len(fields)):
        assert rel(mine[:, j], ref[:, j]) <= TOL, fields[j]
    assert rel(r["vht"], float.fromhex(g["final_vht"])) <= TOL
    if "x_b64" in g:
        assert rel(x, b64_array(g["x_b64"], np.float64)) <= TOL
        assert rel(c, b64_array(g["c_b64"], np.float64)) <= TOL
    # bit-identical to the oracle when both round pow(x, 4) correctly
    oracle_lib.lib().ora_set_pow_mode(1)
    try:
        o = oracle_lib.fw_solve(net, dem, max_iters=it, nthreads=8)
    finally:
        oracle_lib.lib().ora_set_pow_mode(0)
    assert np.array_equal(tr["lambda_"], o["trace"]["lambda_"])
    assert np.array_equal(x, o["x"]) and np.array_equal(c, o["c"])


@pytest.mark.parametrize("links,window", [(1, 0), (2, 0), (1, 2), (2, 2)])
def test_fw_line_search_beyond_first_window(engine_mod, oracle_lib, links, window):
    """Past iteration ~7,600 on grid5 the line search needs steps below
    2^-15, beyond the first window of 16 halvings: in the one-launch link
    kernel (links=1) every window runs in the kernel; in the multi-block
    passes (links=2) the second window is queued with every iteration and
    later ones are run by the host.  With windows of 2 step sizes
    (window=2) nearly every iteration needs later windows, on the host loop
    too, inside multi-iteration graph replays.  8,450
    iterations (ten such steps) bit-identical to the oracle; in a batched
    engine whose second variant has other capacities, the variants need the
    extra windows at different iterations and each still equals its plain
    solve."""
    import dataclasses
    from traffic_assignment_amd import synth
    print(f"case fw deep line search grid5, link passes {links}", flush=True)
    it = 8450
    net, dem = network("grid5")
    with engine_mod.TapEngine(net, dem) as eng:
        eng.set_link_passes(links, window)
        eng.fw_init()
        r = eng.fw_run(max_iters=it)[0]
        x, c = eng.flows()
        tr = eng.trace()
    assert r["iterations"] == it
    assert np.count_nonzero(tr["lambda_"] < 2.0 ** -15) >= 5
    oracle_lib.lib().ora_set_pow_mode(1)
    try:
        o = oracle_lib.fw_solve(net, dem, max_iters=it, nthreads=8)
    finally:
        oracle_lib.lib().ora_set_pow_mode(0)
    assert np.array_equal(tr["lambda_"], o["trace"]["lambda_"])
    assert np.array_equal(x, o["x"]) and np.array_equal(c, o["c"])
    # two variants: the base capacities and 0.8x on every link
    E = net.num_links
    cap2 = np.stack([net.capacity, 0.8 * net.capacity])
    rec = {k: np.stack([getattr(net, k)] * 2) for k in ("fftime", "B", "power")}
    with engine_mod.TapEngine(net, dem, num_variants=2, present=np.ones((2, E), np.uint8),
                              capacity=cap2, **rec) as be:
        be.set_link_passes(links, window)
        be.fw_init()
        res = be.fw_run(max_iters=it)
        bx = [be.flows(v)[0] for v in range(2)]
        blam = [be.trace(v)["lambda_"] for v in range(2)]
    assert [q["iterations"] for q in res] == [it, it]
    assert np.array_equal(blam[0], tr["lambda_"]) and np.array_equal(bx[0], x)
    net2 = dataclasses.replace(net, capacity=0.8 * net.capacity)
    with engine_mod.TapEngine(net2, dem) as pe:
        pe.set_link_passes(links, window)
        pe.fw_init()
        pe.fw_run(max_iters=it)
        assert np.array_equal(blam[1], pe.trace()["lambda_"])
        assert np.array_equal(bx[1], pe.flows()[0])
    deep = [set(np.nonzero(l < 2.0 ** -15)[0]) for l in blam]
    assert deep[1] and deep[0] != deep[1]


@pytest.mark.parametrize("strategy", LOAD_STRATS)
def test_fw_relgap_target_and_resume(engine_mod, oracle_lib, strategy):
    """Stop rule (relgap > target is the loop test) and running in two
    calls gives the same result as one call."""
    net, dem = network("metro", 0.3)
    with make_engine(engine_mod, net, dem, strategy) as eng:
        eng.fw_init()
        r1 = eng.fw_run(target_relgap=1e-4)[0]
        x1, _ = eng.flows()
        eng.fw_init()
        eng.fw_run(max_iters=5)
        r2 = eng.fw_run(target_relgap=1e-4)[0]
        x2, _ = eng.flows()
        tr = eng.trace()
    assert r1["converged"] and r1["relgap"] <= 1e-4
    assert np.all(tr["relgap"][:-1] > 1e-4)
    assert r1["iterations"] == r2["iterations"] and np.array_equal(x1, x2)
    o = oracle_lib.fw_solve(net, dem, target_relgap=1e-4, nthreads=8)
    assert o["iterations"] == r1["iterations"]


@pytest.mark.parametrize("strategy", ["hbm", "sweeps"])
def test_fw_regional_first_iterations(engine_mod, oracle_lib, strategy):
    """The regional network (94k nodes, 4,000 zones; too large for LDS
    labels): two FW iterations bit-identical to the oracle (correctly
    rounded x^4 in both)."""
    print(f"case fw regional {strategy}", flush=True)
    net, dem = network("regional")
    with make_engine(engine_mod, net, dem, strategy) as eng:
        eng.fw_init()
        eng.fw_run(max_iters=2)
        x, c = eng.flows()
        tr = eng.trace()
    o = _regional_oracle(oracle_lib)
    assert np.array_equal(tr["lambda_"], o["trace"]["lambda_"])
    assert np.array_equal(x, o["x"]) and np.array_equal(c, o["c"])


_REGIONAL = {}


def _regional_oracle(oracle_lib):
    if not _REGIONAL:
        net, dem = network("regional")
        oracle_lib.lib().ora_set_pow_mode(1)
        try:
            _REGIONAL["o"] = oracle_lib.fw_solve(net, dem, max_iters=2, nthreads=16)
        finally:
            oracle_lib.lib().ora_set_pow_mode(0)
    return _REGIONAL["o"]


def test_schedule_knobs_do_not_change_results(engine_mod):
    """Frontier queue size, chained expansions and workgroup size (LDS
    kernel) and bucket widths (HBM kernel) are schedules only: volumes
    bit-identical for every setting; bad values rejected.  Queue 0 runs
    every round as a bitmap scan; queue 16 mixes queued rounds with bitmap
    rounds.  A workgroup size too small for the network is refused."""
    print("case schedule knobs", flush=True)
    net, dem = network("metro", 0.3)
    out = []
    with engine_mod.TapEngine(net, dem) as eng:
        eng.set_strategy("lds")
        qmax = eng.origin_queue
        assert qmax >= 64
        for q, ch in ((-1, 0), (0, 0), (0, 1), (0, 4), (0, 32), (16, 0), (16, 4)):
            eng.set_origin_queue(q)
            eng.set_origin_chase(ch)
            out.append(eng.aon())
        assert eng.origin_queue == 16
        with pytest.raises(engine_mod.TapGpuError):
            eng.set_origin_chase(-2)
        with pytest.raises(engine_mod.TapGpuError):
            eng.set_origin_queue(-2)
        eng.set_origin_queue(10 ** 6)         # clamped to what the LDS holds
        assert eng.origin_queue == qmax
        eng.set_origin_queue(-1)
        eng.set_origin_chase(0)
        for b, q in ((1024, 256), (512, -1), (512, 1024), (256, 0), (256, -1)):
            eng.set_origin_queue(q)           # (-1 also asks for the size to be tuned)
            eng.set_origin_block(b)
            out.append(eng.aon())
            assert eng.stats()["origin_block"] == b
        eng.set_origin_block(-1)              # tuned with the queue size on the next loading
        out.append(eng.aon())
        assert eng.origin_block in (1024, 512, 256)
        for bad in (300, 2048, 0):
            with pytest.raises(engine_mod.TapGpuError):
                eng.set_origin_block(bad)
        for sp in (50, 200):                  # co-run share for the HBM kernel
            eng.set_origin_split(sp)
            out.append(eng.aon())
        eng.set_origin_split(0)
        eng.set_strategy("hbm")
        for d in (0.0, 0.5, 4.0, 64.0):
            eng.set_origin_delta(d)
            out.append(eng.aon())
        for q in (0, 16):                     # bitmap scans only; mixed rounds
            eng.set_origin_queue(q)
            out.append(eng.aon())
        assert eng.origin_queue == 16
        eng.set_origin_queue(-1)
        assert eng.origin_queue >= 64
        with pytest.raises(engine_mod.TapGpuError):
            eng.set_origin_delta(float("nan"))
        eng.set_origin_delta(-1)          # tune again on the next loading
        out.append(eng.aon())
        assert eng.origin_delta >= 0
        eng.set_lanes(1)                  # lanes kernel: queue sizes and bucket widths
        for q, d in ((-1, 0.0), (0, 0.0), (16, 0.0), (-1, 1.0), (-1, 4.0), (0, 2.0), (16, 2.0),
                     (-1, -1.0)):
            eng.set_origin_queue(q)
            eng.set_lanes_delta(d)
            out.append(eng.aon())
        assert eng.stats()["lanes_launches"] == 8
        assert eng.lanes_delta >= 0               # tuned
        with pytest.raises(engine_mod.TapGpuError):
            eng.set_lanes_delta(float("inf"))
        eng.set_origin_queue(-1)
    for y in out[1:]:
        assert np.array_equal(y, out[0])
    big, bdem = network("metro")                  # 11,000 nodes: 1024 threads only
    with engine_mod.TapEngine(big, bdem) as eng:
        with pytest.raises(engine_mod.TapGpuError, match="at most 8192 nodes"):
            eng.set_origin_block(512)
        eng.set_origin_block(1024)
        eng.set_origin_block(-1)
        eng.aon()
        assert eng.origin_block == 1024


def test_fw_warm_start_continues(engine_mod):
    net, dem = network("metro", 0.3)
    with engine_mod.TapEngine(net, dem) as eng:
        eng.fw_init()
        eng.fw_run(max_iters=10)
        x10, _ = eng.flows()
        eng.fw_run(max_iters=12)
        x12, _ = eng.flows()
        eng.fw_init(warm_volumes=x10)
        eng.fw_run(max_iters=2)
        xw, _ = eng.flows()
    assert np.array_equal(x12, xw)


def test_zero_demand_and_isolated(engine_mod):
    net, _ = network("metro", 0.3)
    dem = empty_demand(net.num_zones)
    with engine_mod.TapEngine(net, dem) as eng:
        assert np.all(eng.aon() == 0)
        eng.fw_init()
        r = eng.fw_run(max_iters=3)[0]
        x, c = eng.flows()
    assert np.all(x == 0) and np.array_equal(c, net.fftime)
    assert r["iterations"] == 3


def test_python_pape_module(engine_mod, oracle_lib):
    """The reference's Python operator interface: predecessor nodes."""
    from traffic_assignment_amd.python_extension import pape
    net, _ = network("metro", 0.3)
    graph = {}
    for u, v, c in zip(net.init, net.term, net.fftime):
        graph.setdefault(int(u), {})[int(v)] = float(c)
    pape.build_adjlist(graph, net.num_nodes + 1, net.first_thru_node)
    rng = np.random.default_rng(4)
    w = net.fftime * rng.uniform(1, 2, net.num_links)
    pape.update_weights(list(w))
    for s in (1, 17, net.num_zones):
        preds = pape.pape(s)
        _, op = oracle_lib.sssp(net, w, s)
        ref = [-1 if k < 0 else int(net.init[k]) for k in op]
        assert preds == ref
    with pytest.raises(IndexError):
        pape.update_weights([1.0])
    with pytest.raises(IndexError):
        pape.pape(net.num_nodes + 1)
