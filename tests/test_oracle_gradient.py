"""Pins of the CPU gradient reference (oracle/gradient.py, oracle/equilibrium.c).

The reference has no gradient code, so these checks stand on mathematics:
the path-based equilibrium must satisfy flow conservation and Wardrop's
condition to rounding level and agree with Frank-Wolfe; the reference
gradients must match central finite differences of re-solved equilibria
(step 1e-4 * theta, as the gradient configuration prescribes) to 1e-4
relative wherever the difference quotient rises above the re-solve noise.
"""
import numpy as np
import pytest

import oracle
from oracle import gradient as G
from conftest import network


@pytest.fixture(scope="module", params=[("grid5", 1.0), ("metro", 0.3)], ids=["grid5", "metro0.3"])
def solved(request):
    name, sc = request.param
    net, dem = network(name, sc)
    eq = G.PathEquilibrium(net, dem)
    sweeps, rg = eq.solve(1e-14, 5000)
    x, t = eq.flows()
    paths = eq.paths()
    yield dict(name=name, net=net, dem=dem, eq=eq, x=x, t=t, paths=paths, sweeps=sweeps, relgap=rg)
    eq.close()


def test_path_equilibrium_is_wardrop(solved):
    net, dem, x, t, P = solved["net"], solved["dem"], solved["x"], solved["t"], solved["paths"]
    assert solved["relgap"] <= 1e-13
    # volumes are the sum of path flows; path flows of an OD pair sum to its demand
    xs = np.zeros(net.num_links)
    for p in range(P["flow"].size):
        xs[P["links"][P["off"][p]:P["off"][p + 1]]] += P["flow"][p]
    np.testing.assert_allclose(xs, x, rtol=1e-12, atol=1e-9)
    tot = np.bincount(P["od"], weights=P["flow"], minlength=P["od_demand"].size)
    np.testing.assert_allclose(tot, P["od_demand"], rtol=1e-12)
    # every used path is a shortest path (Wardrop), checked against the oracle's trees
    dist = {}
    for o in np.unique(P["od_origin"]):
        dist[int(o)] = oracle.sssp(net, t, int(o))[0]
    excess, total = 0.0, 0.0
    for p in np.nonzero(P["flow"] > 0)[0]:
        k = P["od"][p]
        c = t[P["links"][P["off"][p]:P["off"][p + 1]]].sum()
        u = dist[int(P["od_origin"][k])][int(P["od_dest"][k])]
        assert c >= u * (1 - 1e-14)
        if P["flow"][p] > 1e-6:
            assert c <= u * (1 + 1e-9)
        excess += P["flow"][p] * (c - u)
        total += P["flow"][p] * u
    assert excess <= 1e-13 * total


def test_path_equilibrium_matches_frank_wolfe(solved):
    """FW converges to the same (unique) link volumes, slowly."""
    net, dem, x = solved["net"], solved["dem"], solved["x"]
    r = oracle.fw_solve(net, dem, 400, 0.0, 4)
    assert np.abs(r["x"] - x).max() <= 0.01 * np.abs(x).max()
    eq_obj = solved["eq"].objective()
    assert eq_obj <= r["objective"] * (1 + 1e-12)   # the minimum of Beckmann


def test_envelope_closed_forms(solved):
    """dZ/dt0 and dZ/dc at fixed volumes against central differences of the
    Beckmann integrand, and Euler's identity sum t0 dZ/dt0 = Z."""
    net, x = solved["net"], solved["x"]
    cap = oracle.capacity_fixed(net)
    P = G.link_partials(net.fftime, cap, net.B, net.power, x)
    Z = G.beckmann_terms(net.fftime, cap, net.B, net.power, x)
    np.testing.assert_allclose(np.sum(net.fftime * P["dobj_dt0"]), Z.sum(), rtol=1e-13)
    h = 1e-3 * cap     # truncation ~ (p+1)(p+2)/6 * 1e-6 relative
    fd = (G.beckmann_terms(net.fftime, cap + h, net.B, net.power, x)
          - G.beckmann_terms(net.fftime, cap - h, net.B, net.power, x)) / (2 * h)
    big = np.abs(P["dobj_dcap"]) > 1e-3 * np.abs(P["dobj_dcap"]).max()
    np.testing.assert_allclose(fd[big], P["dobj_dcap"][big], rtol=2e-5)
    # dt/dx against the oracle's BPR
    k = np.argmax(x / cap)
    hx = 1e-6 * x[k]
    tp = oracle.bpr(net.fftime[k], cap[k], net.B[k], net.power[k], x[k] + hx)
    tm = oracle.bpr(net.fftime[k], cap[k], net.B[k], net.power[k], x[k] - hx)
    assert abs((tp - tm) / (2 * hx) - P["dt_dx"][k]) <= 1e-6 * P["dt_dx"][k]


def _fd_check(solved, which):
    net, dem, x, P = solved["net"], solved["dem"], solved["x"], solved["paths"]
    g = G.reference_gradient(net, x, P)
    rng = np.random.default_rng(11)
    cand = np.nonzero(x > 0)[0]
    links = rng.choice(cand, min(64, cand.size), replace=False)
    fo, fv = G.fd_gradient(net, dem, links, rel_step=1e-4, which=which)
    key = "cap" if which == "cap" else "t0"
    go, gv = g["dobj_d" + key][links], g["dvht_d" + key][links]
    for a, f in ((go, fo), (gv, fv)):
        scale = np.abs(a).max()
        # 1e-4 relative where the quotient stands clear of the re-solve noise
        # (about 1e-7 absolute for these networks); a noise floor elsewhere
        big = np.abs(a) >= 1e-2 * scale
        assert big.sum() >= 3
        np.testing.assert_allclose(f[big], a[big], rtol=1e-4)
        np.testing.assert_allclose(f, a, rtol=1e-4, atol=2e-6 * scale)
    return g


def test_reference_gradient_vs_fd_capacity(solved):
    g = _fd_check(solved, "cap")
    assert g["rank"] >= 1


def test_reference_gradient_vs_fd_fftime(solved):
    _fd_check(solved, "t0")


def test_fd_larger_step_all_loaded_links():
    """With step 1e-3 * c the re-solve noise (~4e-8..8e-8 absolute here) is 10x
    smaller relative to the quotient: every sampled link with a gradient
    above 1e-2 of the largest agrees to 1e-4 (truncation O(h^2) stays below
    1e-5 here), the rest within 1e-6 of the largest."""
    net, dem = network("metro", 0.3)
    eq = G.PathEquilibrium(net, dem)
    eq.solve(1e-14, 5000)
    x, _ = eq.flows()
    g = G.reference_gradient(net, x, eq.paths())
    eq.close()
    rng = np.random.default_rng(5)
    links = rng.choice(np.nonzero(x > 0)[0], 64, replace=False)
    scale = np.abs(g["dvht_dcap"]).max()
    _, fv = G.fd_gradient(net, dem, links, rel_step=1e-3)
    a = g["dvht_dcap"][links]
    scale = np.abs(a).max()
    big = np.abs(a) >= 1e-2 * scale
    assert big.sum() >= 5
    np.testing.assert_allclose(fv[big], a[big], rtol=1e-4)
    np.testing.assert_allclose(fv, a, rtol=1e-4, atol=1e-6 * scale)


def test_reference_gradient_detects_a_planted_error(solved):
    """A gradient without the flow response (z = 0: only the partial
    derivative at fixed volumes) must fail the finite-difference check on
    some link: the check has teeth."""
    net, dem, x, P = solved["net"], solved["dem"], solved["x"], solved["paths"]
    g = G.reference_gradient(net, x, P)
    partial = x * G.link_partials(net.fftime, oracle.capacity_fixed(net), net.B, net.power,
                                  x)["dt_dcap"]
    diff = np.abs(partial - g["dvht_dcap"])
    k = np.argsort(diff)[-3:]
    _, fv = G.fd_gradient(net, dem, k, rel_step=1e-4)
    assert np.any(np.abs(fv - partial[k]) > 1e-3 * np.abs(fv))
    np.testing.assert_allclose(fv, g["dvht_dcap"][k], rtol=1e-4, atol=1e-6 * np.abs(g["dvht_dcap"]).max())
