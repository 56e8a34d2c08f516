"""Batched upgrade-subset solves (one launch per Frank-Wolfe step for every
variant) against plain solves of each variant network and the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from traffic_assignment_amd import synth, variants

pytestmark = pytest.mark.gpu


def _small():
    base, dem = synth.metro(scale=0.3)
    mods = synth.upgrades(base, n_change=6, n_add=4)
    subsets = synth.budget_subsets(mods, n_variants=12, budget=25.0)
    return base, dem, mods, subsets


@pytest.mark.parametrize("strategy", ["sweeps", "lds", "hbm", "lanes"])
def test_batched_equals_plain_and_oracle(engine_mod, oracle_lib, strategy):
    """Every loading path; "lanes" puts the 12 variants in the lanes of one
    group (the per-origin kernel takes over once fewer than 4 are active)."""
    print(f"case batched {strategy}", flush=True)
    base, dem, mods, subsets = _small()
    eng, u, present = variants.make_engine(base, dem, mods, subsets)
    with eng:
        if strategy == "lanes":
            eng.set_lanes(4)
        else:
            eng.set_strategy(strategy)
            eng.set_lanes(0)
        eng.fw_init()
        res = eng.fw_run(target_relgap=1e-4)
        out = [(eng.flows(v), eng.trace(v)) for v in range(len(subsets))]
        st = eng.stats()
    assert all(r["converged"] for r in res)
    if strategy == "lanes":
        assert st["lanes_launches"] > 0 and st["lanes_fallbacks"] == 0
    for v in (0, 3, 7, 11):
        net = variants.variant_network(base, mods, subsets[v])
        m = present[v].astype(bool)
        with engine_mod.TapEngine(net, dem) as pe:
            if strategy != "lanes":
                pe.set_strategy(strategy)
            pe.fw_init()
            pr = pe.fw_run(target_relgap=1e-4)[0]
            px, pc = pe.flows()
            ptr = pe.trace()
        (bx, bc), btr = out[v]
        assert pr["iterations"] == res[v]["iterations"]
        assert np.array_equal(btr["lambda_"], ptr["lambda_"])
        assert np.array_equal(bx[m], px) and np.all(bx[~m] == 0)
        assert np.array_equal(bc[m], pc)
        assert np.allclose(btr["relgap"], ptr["relgap"], rtol=1e-12, atol=0)
        if v in (0, 7):
            oracle_lib.lib().ora_set_pow_mode(1)
            try:
                o = oracle_lib.fw_solve(net, dem, target_relgap=1e-4, nthreads=8)
            finally:
                oracle_lib.lib().ora_set_pow_mode(0)
            assert o["iterations"] == pr["iterations"]
            assert np.array_equal(o["trace"]["lambda_"], ptr["lambda_"])
            assert np.array_equal(o["x"], px)


def test_lanes_exact_ties_many_variants(engine_mod, oracle_lib):
    """70 variants (two lane groups) of a random network whose free-flow
    times are 1, 2 or 3 per variant: the free-flow loading meets thousands
    of exact ties in every lane.  The lanes kernel (each lane walks its own
    tree by the tie rule) equals the per-origin kernel variant by variant
    over fw_init and 3 iterations, and sampled variants equal the oracle
    with the engine's tie rule (tie mode 1), bit for bit."""
    from conftest import random_demand, random_network
    print("case lanes exact ties, 70 variants", flush=True)
    net, rng = random_network(31, "pints")
    dem = random_demand(net, rng)
    V = 70
    t0 = rng.integers(1, 4, (V, net.num_links)).astype(np.float64)
    out = {}
    for path in ("lanes", "hbm"):
        with engine_mod.TapEngine(net, dem, num_variants=V, fftime=t0) as eng:
            if path == "lanes":
                eng.set_lanes(1)
            else:
                eng.set_lanes(0)
                eng.set_strategy("hbm")
            eng.fw_init()
            eng.fw_run(max_iters=3)
            st = eng.stats()
            out[path] = [eng.flows(v) for v in range(V)]
        if path == "lanes":
            assert st["lanes_launches"] > 0 and st["lanes_fallbacks"] == 0
    for v in range(V):
        assert np.array_equal(out["lanes"][v][0], out["hbm"][v][0]), v
        assert np.array_equal(out["lanes"][v][1], out["hbm"][v][1]), v
    lib = oracle_lib.lib()
    lib.ora_set_pow_mode(1)
    lib.ora_set_tie_mode(1)
    try:
        for v in (0, 41, 69):
            import dataclasses
            nv = dataclasses.replace(net, fftime=t0[v].copy())
            o = oracle_lib.fw_solve(nv, dem, max_iters=3, nthreads=8)
            assert np.array_equal(o["x"], out["lanes"][v][0]), v
            assert np.array_equal(o["c"], out["lanes"][v][1]), v
    finally:
        lib.ora_set_tie_mode(0)
        lib.ora_set_pow_mode(0)


def test_batched_variants_stop_independently(engine_mod):
    base, dem, mods, subsets = _small()
    eng, u, present = variants.make_engine(base, dem, mods, subsets)
    with eng:
        eng.fw_init()
        res = eng.fw_run(max_iters=50, target_relgap=1e-4)
        iters = [r["iterations"] for r in res]
        x_before = [eng.flows(v)[0] for v in range(len(subsets))]
        res2 = eng.fw_run(max_iters=50, target_relgap=1e-4)
        x_after = [eng.flows(v)[0] for v in range(len(subsets))]
    # converged variants are frozen: a second run changes nothing
    assert all(r["converged"] for r in res)
    assert [r["iterations"] for r in res2] == iters
    assert all(np.array_equal(a, b) for a, b in zip(x_before, x_after))
    # an iteration cap stops each variant at the cap; continuing resumes it
    eng2, _, _ = variants.make_engine(base, dem, mods, subsets)
    with eng2:
        eng2.fw_init()
        capped = eng2.fw_run(max_iters=5)
        assert all(r["iterations"] == 5 for r in capped)
        done = eng2.fw_run(target_relgap=1e-4)
        x_resumed = [eng2.flows(v)[0] for v in range(len(subsets))]
    assert [r["iterations"] for r in done] == iters
    assert all(np.array_equal(a, b) for a, b in zip(x_before, x_resumed))


@pytest.mark.parametrize("strategy", ["lds", "lanes"])
def test_batched_warm_start_from_base_equilibrium(engine_mod, oracle_lib, strategy):
    """Every variant warm-started from the base network's equilibrium (added
    links start empty): each variant's λ sequence and volumes equal its plain
    solve from the same volumes and the oracle's, and the warm start needs
    fewer iterations than the cold one."""
    print(f"case batched warm start {strategy}", flush=True)
    base, dem, mods, subsets = _small()
    with engine_mod.TapEngine(base, dem) as be:
        be.fw_init()
        be.fw_run(target_relgap=1e-4)
        xb, _ = be.flows()
    eng, u, present = variants.make_engine(base, dem, mods, subsets)
    index = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(u.init, u.term))}
    xu = np.zeros(u.num_links)
    xu[[index[(int(a), int(b))] for a, b in zip(base.init, base.term)]] = xb
    with eng:
        if strategy == "lanes":
            eng.set_lanes(4)
        else:
            eng.set_strategy(strategy)
            eng.set_lanes(0)
        eng.fw_init()
        cold = eng.fw_run(target_relgap=1e-4)
        eng.fw_init(np.tile(xu, (len(subsets), 1)))
        res = eng.fw_run(target_relgap=1e-4)
        out = [(eng.flows(v)[0], eng.trace(v)["lambda_"]) for v in range(len(subsets))]
    assert all(r["converged"] for r in res)
    assert sum(r["iterations"] for r in res) < sum(r["iterations"] for r in cold)
    for v in (0, 5, 11):
        net = variants.variant_network(base, mods, subsets[v])
        m = present[v].astype(bool)
        with engine_mod.TapEngine(net, dem) as pe:
            pe.fw_init(xu[m])
            pr = pe.fw_run(target_relgap=1e-4)[0]
            px, _ = pe.flows()
            plam = pe.trace()["lambda_"]
        bx, blam = out[v]
        assert pr["iterations"] == res[v]["iterations"], v
        assert np.array_equal(blam, plam) and np.array_equal(bx[m], px), v
        if v == 0:
            oracle_lib.lib().ora_set_pow_mode(1)
            try:
                o = oracle_lib.fw_solve(net, dem, target_relgap=1e-4, nthreads=8, x_init=xu[m])
            finally:
                oracle_lib.lib().ora_set_pow_mode(0)
            assert o["iterations"] == pr["iterations"]
            assert np.array_equal(o["trace"]["lambda_"], plam) and np.array_equal(o["x"], px)
