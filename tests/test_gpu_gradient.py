"""Equilibrium gradients on the GPU (tapgpu_gradient and the autograd
Function) against the CPU path-space reference (oracle/gradient.py) and
central finite differences of re-solved equilibria.

The GPU builds the active-path span from fundamental cycles of its own
shortest-path trees and solves it by conjugate gradients; the reference
enumerates paths from a path-based equilibrium and solves densely.  Both are
fed the same equilibrium volumes (solved to rounding level by the path-based
solver and loaded into the engine as warm-start volumes).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

from conftest import network

pytestmark = pytest.mark.gpu

CASES = [("grid5", 1.0), ("metro", 0.3)]


def _precise(name, sc):
    from oracle import gradient as G
    net, dem = network(name, sc)
    eq = G.PathEquilibrium(net, dem)
    eq.solve(1e-14, 5000)
    x, _ = eq.flows()
    ref = G.reference_gradient(net, x, eq.paths())
    eq.close()
    return net, dem, x, ref


def _gpu_grad(engine_mod, net, dem, x, **kw):
    with engine_mod.TapEngine(net, dem) as eng:
        eng.fw_init(x)
        return eng.gradient(**kw)


@pytest.mark.parametrize("name,sc", CASES, ids=["grid5", "metro0.3"])
def test_gradient_matches_path_reference(engine_mod, name, sc):
    print(f"case gradient-vs-reference {name}", flush=True)
    net, dem, x, ref = _precise(name, sc)
    g = _gpu_grad(engine_mod, net, dem, x, eps=1e-9, cg_tol=1e-13)
    info = g["info"]
    assert info["cg_status"] == 1 and info["cycles"] > 0
    # envelope terms: closed forms, same arithmetic up to pow
    for k in ("dobj_dt0", "dobj_dcap"):
        np.testing.assert_allclose(g[k], ref[k], rtol=1e-12, atol=1e-300)
    # implicit terms: different constructions of the same span
    for k in ("dvht_dcap", "dvht_dt0"):
        scale = np.abs(ref[k]).max()
        np.testing.assert_allclose(g[k], ref[k], rtol=1e-7, atol=1e-9 * scale)


def test_gradient_matches_finite_differences(engine_mod):
    """Gradient check at metro scale 0.3: 64 random loaded links, central
    differences of re-solved equilibria.  Step 1e-4 * c: 1e-4 relative above
    the re-solve noise (~1e-7..1e-6 absolute, bounded by 1e-5 of the largest
    gradient in the sample); step 1e-3 * c (ten times less noise, ~4e-8
    absolute, truncation < 1e-5): 1e-4 relative on every link above 1e-2 of
    the largest, 1e-6 of the largest absolute below."""
    from oracle import gradient as G
    print("case gradient-vs-fd metro0.3", flush=True)
    net, dem, x, ref = _precise("metro", 0.3)
    g = _gpu_grad(engine_mod, net, dem, x, eps=1e-9, cg_tol=1e-13)
    rng = np.random.default_rng(7)
    links = rng.choice(np.nonzero(x > 0)[0], 64, replace=False)
    fo, fv = G.fd_gradient(net, dem, links, rel_step=1e-4)
    for a, f in ((g["dobj_dcap"][links], fo), (g["dvht_dcap"][links], fv)):
        scale = np.abs(a).max()
        assert (np.abs(a) >= 1e-2 * scale).sum() >= 3
        np.testing.assert_allclose(f, a, rtol=1e-4, atol=1e-5 * scale)
    a = g["dvht_dcap"][links]
    scale = np.abs(a).max()
    _, fv3 = G.fd_gradient(net, dem, links, rel_step=1e-3)
    big = np.abs(a) >= 1e-2 * scale
    np.testing.assert_allclose(fv3[big], a[big], rtol=1e-4)
    np.testing.assert_allclose(fv3, a, rtol=1e-4, atol=1e-6 * scale)


def test_gradient_deterministic_and_state_preserving(engine_mod):
    print("case gradient-determinism", flush=True)
    net, dem, x, _ = _precise("metro", 0.3)
    with engine_mod.TapEngine(net, dem) as eng:
        eng.fw_init(x)
        x0, c0 = eng.flows()
        a = eng.gradient(eps=1e-9)
        b = eng.gradient(eps=1e-9)
        x1, c1 = eng.flows()
        # the FW state continues unchanged after a gradient
        r = eng.fw_run(3)
    for k in ("dobj_dt0", "dobj_dcap", "dvht_dcap", "dvht_dt0"):
        assert np.array_equal(a[k], b[k])
    assert np.array_equal(x0, x1) and np.array_equal(c0, c1)
    assert r[0]["iterations"] == 3


@pytest.mark.parametrize("strategy", ["sweeps", "lds"])
def test_gradient_after_fw_solve(engine_mod, strategy):
    """After an ordinary FW solve: the envelope terms are the closed forms at
    the solved volumes; at relgap ~1e-6 the VHT gradient is within a few
    percent of the precise one on the heavily used links (the equilibrium
    itself is only that accurate)."""
    from oracle import gradient as G
    import oracle
    print(f"case gradient-after-fw {strategy}", flush=True)
    net, dem, xs, ref = _precise("metro", 0.3)
    with engine_mod.TapEngine(net, dem) as eng:
        eng.set_strategy(strategy)
        eng.fw_init()
        res = eng.fw_run(3000, 1e-6)[0]
        x, _ = eng.flows()
        g = eng.gradient(eps=max(1e-9, 20 * res["relgap"]))
    P = G.link_partials(net.fftime, oracle.capacity_fixed(net), net.B, net.power, x)
    np.testing.assert_allclose(g["dobj_dcap"], P["dobj_dcap"], rtol=1e-12, atol=1e-300)
    scale = np.abs(ref["dvht_dcap"]).max()
    big = np.abs(ref["dvht_dcap"]) > 0.05 * scale
    rel = np.abs(g["dvht_dcap"][big] - ref["dvht_dcap"][big]) / np.abs(ref["dvht_dcap"][big])
    assert np.median(rel) < 0.05


def test_gradient_without_alternatives(engine_mod):
    """One path per OD pair (a tree network): no cycles, dVHT/dc is the
    partial derivative x dt/dc."""
    from traffic_assignment_amd import synth
    import oracle
    from oracle import gradient as G
    # a path 1-2-3-4 with both directions, all nodes zones
    init = np.array([1, 2, 2, 3, 3, 4]); term = np.array([2, 1, 3, 2, 4, 3])
    E = init.size
    net = synth.Network(num_zones=4, num_nodes=4, first_thru_node=1, init=init, term=term,
                        capacity=np.full(E, 100.0), length=np.ones(E),
                        fftime=np.array([1.0, 1.0, 2.0, 2.0, 1.5, 1.5]), B=np.full(E, 0.15),
                        power=np.full(E, 4.0), speed=np.zeros(E), toll=np.zeros(E),
                        ltype=np.ones(E, np.int16))
    start = np.array([0, 3, 6, 9, 12]); dest = np.array([2, 3, 4, 1, 3, 4, 1, 2, 4, 1, 2, 3])
    dem = synth.Demand(4, start, dest, np.arange(1.0, 13.0) * 10)
    with engine_mod.TapEngine(net, dem) as eng:
        eng.fw_init()
        eng.fw_run(2)
        x, _ = eng.flows()
        g = eng.gradient()
    assert g["info"]["cycles"] == 0
    P = G.link_partials(net.fftime, oracle.capacity_fixed(net), net.B, net.power, x)
    np.testing.assert_allclose(g["dvht_dcap"], x * P["dt_dcap"], rtol=1e-13)


def test_gradient_batched_variant_equals_plain(engine_mod):
    from traffic_assignment_amd import synth, variants
    print("case gradient-batched", flush=True)
    base, dem = synth.metro(scale=0.3)
    mods = synth.upgrades(base, n_change=6, n_add=4)
    subsets = synth.budget_subsets(mods, n_variants=4, budget=25.0)
    eng, u, present = variants.make_engine(base, dem, mods, subsets)
    v = 2
    with eng:
        eng.fw_init()
        eng.fw_run(200)
        bx, _ = eng.flows(v)
        bg = eng.gradient(v, eps=1e-6)
    m = present[v].astype(bool)
    net = variants.variant_network(base, mods, subsets[v])
    with engine_mod.TapEngine(net, dem) as pe:
        pe.fw_init(bx[m])
        pg = pe.gradient(eps=1e-6)
    for k in ("dobj_dt0", "dobj_dcap", "dvht_dcap", "dvht_dt0"):
        np.testing.assert_allclose(bg[k][m], pg[k], rtol=1e-10,
                                   atol=1e-12 * np.abs(pg[k]).max())
        assert np.all(bg[k][~m] == 0)


def test_gradient_errors(engine_mod):
    net, dem = network("grid5")
    with engine_mod.TapEngine(net, dem) as eng:
        with pytest.raises(engine_mod.TapGpuError):
            eng.gradient()                       # before fw_init
        eng.fw_init()
        with pytest.raises(engine_mod.TapGpuError):
            eng.gradient(eps=-1.0)
        with pytest.raises(engine_mod.TapGpuError):
            eng.gradient(variant=3)


def test_autograd_function(engine_mod):
    """backward == grad_obj * d obj + grad_vht * d vht from the kernels, and
    a directional derivative of VHT against re-solved equilibria."""
    from traffic_assignment_amd.python_extension.autograd import equilibrium
    print("case autograd", flush=True)
    net, dem, xs, ref = _precise("grid5", 1.0)
    cap = torch.tensor(net.capacity, dtype=torch.float64, device="cuda", requires_grad=True)
    t0 = torch.tensor(net.fftime, dtype=torch.float64, device="cuda", requires_grad=True)
    obj, vht, x = equilibrium(net, dem, cap, t0, max_iters=0, warm_volumes=xs, eps=1e-9)
    assert obj.dtype == torch.float64 and obj.device.type == "cuda"
    np.testing.assert_allclose(x.cpu().numpy(), xs, rtol=0, atol=0)
    (2.0 * obj + 3.0 * vht).backward()
    want_cap = 2.0 * ref["dobj_dcap"] + 3.0 * ref["dvht_dcap"]
    want_t0 = 2.0 * ref["dobj_dt0"] + 3.0 * ref["dvht_dt0"]
    np.testing.assert_allclose(cap.grad.cpu().numpy(), want_cap, rtol=1e-7,
                               atol=1e-9 * np.abs(want_cap).max())
    np.testing.assert_allclose(t0.grad.cpu().numpy(), want_t0, rtol=1e-7,
                               atol=1e-9 * np.abs(want_t0).max())
    assert not x.requires_grad
    # the forward values are the objective and VHT at those volumes
    from oracle import gradient as G
    import oracle
    Zs = G.beckmann_terms(net.fftime, oracle.capacity_fixed(net), net.B, net.power, xs).sum()
    assert abs(obj.item() - Zs) <= 1e-12 * Zs
    # a solve from scratch: forward runs FW, backward still works
    cap2 = cap.detach().clone().requires_grad_(True)
    obj2, vht2, _ = equilibrium(net, dem, cap2, t0.detach(), target_relgap=1e-5)
    vht2.backward()
    assert torch.isfinite(cap2.grad).all() and abs(obj2.item() - Zs) <= 1e-4 * Zs


def test_autograd_reuses_engine(engine_mod):
    """Calls on one network and demand share one engine (capacities and
    free-flow times replaced per call): every call's volumes equal a freshly
    built engine's solve bit for bit, objective and VHT the oracle's sums,
    the gradient the fresh engine's to 1e-8; a demand changed in place is
    seen (the engine is rebuilt)."""
    from traffic_assignment_amd.python_extension import autograd as A
    import oracle
    print("case autograd engine reuse", flush=True)
    net, dem = network("metro", 0.3)
    rng = np.random.default_rng(5)
    A._CACHE.close()
    b0 = A._CACHE.builds
    for k in range(3):
        capk = net.capacity * rng.uniform(0.8, 1.2, net.num_links)
        t0k = net.fftime * rng.uniform(0.9, 1.1, net.num_links)
        cap = torch.tensor(capk, dtype=torch.float64, device="cuda", requires_grad=True)
        t0 = torch.tensor(t0k, dtype=torch.float64, device="cuda", requires_grad=True)
        obj, vht, x = A.equilibrium(net, dem, cap, t0, target_relgap=1e-4, eps=1e-3)
        vht.backward()
        assert A._CACHE.builds == b0 + 1
        with engine_mod.TapEngine(net, dem, num_variants=1, capacity=capk[None, :],
                                  fftime=t0k[None, :]) as eng:
            eng.fw_init()
            eng.fw_run(target_relgap=1e-4)
            xf, cf = eng.flows()
            g = eng.gradient(eps=1e-3)
        assert np.array_equal(x.cpu().numpy(), xf)
        assert abs(vht.item() - float(np.dot(xf, cf))) <= 1e-12 * vht.item()
        netk = dataclasses.replace(net, capacity=capk, fftime=t0k)
        Z = oracle.objective(netk, xf)
        assert abs(obj.item() - Z) <= 1e-12 * abs(Z)
        # the zone order of a context (and with it the order of the
        # cycle sums in conjugate gradients) comes from the free-flow times
        # it was built with, so the two gradients agree to the solve's
        # tolerance, not bit for bit
        ga = cap.grad.cpu().numpy()
        np.testing.assert_allclose(ga, g["dvht_dcap"], rtol=1e-8,
                                   atol=1e-10 * np.abs(g["dvht_dcap"]).max())
    dem.demand[0] += 1.0
    try:
        cap = torch.tensor(net.capacity, dtype=torch.float64, device="cuda")
        t0 = torch.tensor(net.fftime, dtype=torch.float64, device="cuda")
        _, _, x = A.equilibrium(net, dem, cap, t0, target_relgap=1e-4)
        assert A._CACHE.builds == b0 + 2
        with engine_mod.TapEngine(net, dem) as eng:
            eng.fw_init()
            eng.fw_run(target_relgap=1e-4)
            assert np.array_equal(x.cpu().numpy(), eng.flows()[0])
    finally:
        dem.demand[0] -= 1.0
        A._CACHE.close()
