"""The stated workloads at their stated sizes, under -m gpu.

* metro (11,000 nodes / 43,600 links / 1,000 zones) solved to relgap 1e-4:
  every step size, volume and cost bit-identical to the oracle (correctly
  rounded x^4 in both), on the per-origin LDS path and on the lanes path.
* the batched workload: 512 budgeted upgrade subsets of metro (32 candidate
  upgrades) in one context, to relgap 1e-4: every variant converges; sampled
  variants equal their plain solves and the oracle bit for bit.
* regional (94,000 nodes / 386,800 links / 4,000 zones) solved to relgap 1e-4
  on the labels-in-HBM path, bit-identical to the oracle with the engine's
  tie rule; against Pape's tie choice, the same step sizes and a few links
  apart.
* equilibrium gradients on the regional network: at a quarter of its size
  against the CPU path-space gradient and central differences of re-solved
  equilibria on 64 random links (step 1e-4 c, 1e-4 relative); at full size
  the closed forms, Euler's identity sum t0 dZ/dt0 = Z and determinism.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import network

pytestmark = pytest.mark.gpu
TARGET = 1e-4


def _oracle_solve(oracle_lib, net, dem, **kw):
    oracle_lib.lib().ora_set_pow_mode(1)
    try:
        return oracle_lib.fw_solve(net, dem, nthreads=16, **kw)
    finally:
        oracle_lib.lib().ora_set_pow_mode(0)


_METRO = {}


@pytest.mark.parametrize("path", ["lds", "lanes"])
def test_metro_solve_to_target_bit_identical(engine_mod, oracle_lib, path):
    print(f"case metro full solve {path}", flush=True)
    net, dem = network("metro")
    with engine_mod.TapEngine(net, dem) as eng:
        if path == "lanes":
            eng.set_lanes(1)
        else:
            eng.set_strategy("lds")
        eng.fw_init()
        r = eng.fw_run(target_relgap=TARGET)[0]
        x, c = eng.flows()
        tr = eng.trace()
        st = eng.stats()
    if "o" not in _METRO:
        _METRO["o"] = _oracle_solve(oracle_lib, net, dem, target_relgap=TARGET)
    o = _METRO["o"]
    print("iterations", r["iterations"], "oracle", o["iterations"], flush=True)
    assert r["converged"] and r["iterations"] == o["iterations"] > 100
    assert np.array_equal(tr["lambda_"], o["trace"]["lambda_"])
    assert np.array_equal(x, o["x"]) and np.array_equal(c, o["c"])
    assert abs(tr["relgap"][-1] - o["relgap"]) <= 1e-9 * o["relgap"]
    if path == "lanes":
        assert st["lanes_launches"] > 0 and st["lanes_fallbacks"] == 0


def test_batched_workload_512_variants(engine_mod, oracle_lib):
    from traffic_assignment_amd import variants as V
    print("case batched 512 variants of metro", flush=True)
    base, dem, mods, subsets = V.workload(n_variants=512)
    assert len(subsets) == 512 and len({m.change_id for m in mods}) == 32
    eng, u, present = V.make_engine(base, dem, mods, subsets)
    with eng:
        eng.fw_init()
        done = 0
        while True:
            done += 50
            res = eng.fw_run(max_iters=done, target_relgap=TARGET)
            running = sum(1 for q in res if not q["converged"] and q["iterations"] >= done)
            print(f"{done} iterations, {running} variants running", flush=True)
            if running == 0:
                break
        sample = (0, 137, 300, 511)
        out = {v: (eng.flows(v), eng.trace(v)) for v in sample}
        st = eng.stats()
    assert all(q["converged"] for q in res)
    iters = [q["iterations"] for q in res]
    print("iterations min/median/max", min(iters), int(np.median(iters)), max(iters), flush=True)
    for v in sample:
        net = V.variant_network(base, mods, subsets[v])
        m = present[v].astype(bool)
        with engine_mod.TapEngine(net, dem) as pe:
            pe.fw_init()
            pr = pe.fw_run(target_relgap=TARGET)[0]
            px, pc = pe.flows()
            ptr = pe.trace()
        (bx, bc), btr = out[v]
        assert pr["iterations"] == res[v]["iterations"], v
        assert np.array_equal(btr["lambda_"], ptr["lambda_"]), v
        assert np.array_equal(bx[m], px) and np.all(bx[~m] == 0), v
        assert np.array_equal(bc[m], pc), v
        o = _oracle_solve(oracle_lib, net, dem, target_relgap=TARGET)
        assert o["iterations"] == pr["iterations"], v
        assert np.array_equal(o["trace"]["lambda_"], ptr["lambda_"]), v
        assert np.array_equal(o["x"], px) and np.array_equal(o["c"], pc), v


def test_regional_solve_to_target_bit_identical(engine_mod, oracle_lib):
    """The regional network (94,000 nodes / 386,800 links / 4,000 zones,
    gravity demand on free-flow times) solved to relgap 1e-4 on the
    labels-in-HBM path.

    Some later iterations meet exact ties (two shortest paths of equal
    cost), where Pape's queue order and the engine's predecessor rule load
    different, equally short paths.  Against the oracle with the engine's
    rule (tie mode 1): every step size, volume and cost bit-identical.
    Against the oracle with Pape's choice (the reference's): the same
    iterations and step sizes, the relative gap within 1e-9, and volumes
    that differ on a few links only (12 of 386,800 on this network)."""
    print("case regional full solve", flush=True)
    net, dem = network("regional")
    with engine_mod.TapEngine(net, dem) as eng:
        eng.set_strategy("hbm")
        eng.fw_init()
        r = eng.fw_run(target_relgap=TARGET)[0]
        x, c = eng.flows()
        tr = eng.trace()
    print("iterations", r["iterations"], flush=True)
    lib = oracle_lib.lib()
    lib.ora_set_tie_mode(1)
    try:
        o1 = _oracle_solve(oracle_lib, net, dem, target_relgap=TARGET)
    finally:
        lib.ora_set_tie_mode(0)
    assert r["converged"] and r["iterations"] == o1["iterations"] > 10
    assert np.array_equal(tr["lambda_"], o1["trace"]["lambda_"])
    assert np.array_equal(x, o1["x"]) and np.array_equal(c, o1["c"])
    assert abs(tr["relgap"][-1] - o1["relgap"]) <= 1e-9 * o1["relgap"]
    o0 = _oracle_solve(oracle_lib, net, dem, target_relgap=TARGET)
    assert o0["iterations"] == r["iterations"]
    assert np.array_equal(tr["lambda_"], o0["trace"]["lambda_"])
    assert abs(tr["relgap"][-1] - o0["relgap"]) <= 1e-9 * o0["relgap"]
    differ = int(np.sum(x != o0["x"]))
    print("links differing from the Pape-order solve:", differ, flush=True)
    assert differ <= 1e-4 * net.num_links


def test_regional_quarter_gradient_against_reference_and_fd(engine_mod):
    """Regional at a quarter of its size (5,875 nodes, 23,950 links, 250
    zones, gravity demand on free-flow paths, v/c up to ~1.6): the GPU
    gradients at the precise equilibrium against the CPU path-space
    gradient (1e-7 relative) and against central differences of re-solved
    equilibria on 64 random loaded links.

    Step 1e-4 c: each re-solve (path-based, relative gap 1e-14) fixes VHT to
    about 2e-13 relative, so a central difference carries up to
    4e-13 VHT / (2 h) of noise (h = 1e-4 c: ~3e-6 on an 800-veh/h link,
    above the smaller gradients here); d VHT/d c is held to 1e-4 relative
    plus that noise, d Beckmann/d c to 1e-4 relative plus 1e-5 of the
    largest.  Step 2e-3 c (twenty times less noise, truncation ~1e-5
    measured): 1e-4 relative on every link above 1e-2 of the largest, 1e-4
    of the largest absolute below."""
    from oracle import gradient as G
    print("case gradient regional@0.25", flush=True)
    net, dem = network("regional", 0.25)
    eq = G.PathEquilibrium(net, dem)
    eq.solve(1e-14, 5000)
    x, _ = eq.flows()
    ref = G.reference_gradient(net, x, eq.paths())
    vht = eq.vht()
    eq.close()
    with engine_mod.TapEngine(net, dem) as eng:
        eng.fw_init(x)
        g = eng.gradient(eps=1e-9, cg_tol=1e-13)
    assert g["info"]["cg_status"] == 1 and g["info"]["cycles"] > 0
    for k in ("dobj_dt0", "dobj_dcap"):
        np.testing.assert_allclose(g[k], ref[k], rtol=1e-12, atol=1e-300)
    for k in ("dvht_dcap", "dvht_dt0"):
        scale = np.abs(ref[k]).max()
        np.testing.assert_allclose(g[k], ref[k], rtol=1e-7, atol=1e-9 * scale)
    rng = np.random.default_rng(11)
    links = rng.choice(np.nonzero(x > 0)[0], 64, replace=False)
    fo, fv = G.fd_gradient(net, dem, links, rel_step=1e-4)
    a = g["dobj_dcap"][links]
    np.testing.assert_allclose(fo, a, rtol=1e-4, atol=1e-5 * np.abs(a).max())
    a = g["dvht_dcap"][links]
    h = 1e-4 * net.capacity[links]
    noise = 4e-13 * abs(vht) / (2 * h)
    assert np.all(np.abs(fv - a) <= 1e-4 * np.abs(a) + noise), np.max(np.abs(fv - a) / noise)
    _, fv2 = G.fd_gradient(net, dem, links, rel_step=2e-3)
    scale = np.abs(a).max()
    big = np.abs(a) >= 1e-2 * scale
    assert big.sum() >= 5
    np.testing.assert_allclose(fv2[big], a[big], rtol=1e-4)
    np.testing.assert_allclose(fv2, a, rtol=1e-4, atol=1e-4 * scale)


def test_regional_full_gradient_identities(engine_mod):
    """The full regional network (4,000 zones, 4.8M OD pairs with demand)
    after a Frank-Wolfe solve to relgap 1e-5: the envelope terms are the
    closed forms at the solved volumes, Euler's identity
    sum_a t0_a dZ/dt0_a = Z (the Beckmann objective is homogeneous of degree
    one in t0), the VHT gradients are finite, and two calls give the same
    bits (conjugate gradients capped at 2,000 iterations: the implicit terms
    are then partial, but deterministic)."""
    from oracle import gradient as G
    import oracle
    print("case gradient regional full", flush=True)
    net, dem = network("regional")
    with engine_mod.TapEngine(net, dem) as eng:
        eng.fw_init()
        done = 0
        while True:
            done += 100
            r = eng.fw_run(max_iters=done, target_relgap=1e-5)[0]
            print(f"{r['iterations']} iterations, relgap {r['relgap']:.3e}", flush=True)
            if r["converged"] or done >= 3000:
                break
        x, _ = eng.flows()
        a = eng.gradient(eps=max(1e-9, 20 * r["relgap"]), cg_tol=1e-10, max_cg=2000)
        b = eng.gradient(eps=max(1e-9, 20 * r["relgap"]), cg_tol=1e-10, max_cg=2000)
    assert r["converged"]
    print("gradient", a["info"], flush=True)
    P = G.link_partials(net.fftime, oracle.capacity_fixed(net), net.B, net.power, x)
    for k in ("dobj_dt0", "dobj_dcap"):
        np.testing.assert_allclose(a[k], P[k], rtol=1e-12, atol=1e-300)
    Z = oracle.objective(net, x)
    euler = float(np.sum(net.fftime * a["dobj_dt0"]))
    assert abs(euler - Z) <= 1e-10 * abs(Z), (euler, Z)
    for k in ("dobj_dt0", "dobj_dcap", "dvht_dcap", "dvht_dt0"):
        assert np.all(np.isfinite(a[k])) and np.array_equal(a[k], b[k]), k
    assert a["info"]["cycles"] > 0
