"""The CPU oracle against golden vectors written by the compiled reference
(tests/golden/make_golden.py runs trunk/tap's own functions and binary)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import b64_array, golden, hexarr, network, rmat, sha


def _digest(net, dem):
    import hashlib
    h = hashlib.sha256()
    for a in (net.init, net.term, net.capacity, net.fftime, net.B, net.power, dem.start, dem.dest,
              dem.demand):
        h.update(np.ascontiguousarray(a).astype("<f8" if a.dtype.kind == "f" else "<i8").tobytes())
    return h.hexdigest()


CASES = [("grid5", 1.0, "grid5_fw.json"), ("metro", 0.3, "metro_s03_fw.json"),
         ("metro", 1.0, "metro_fw.json")]


@pytest.mark.parametrize("name,scale,fix", CASES)
def test_generator_reproduces_golden_inputs(name, scale, fix):
    net, dem = network(name, scale)
    assert _digest(net, dem) == golden(fix)["inputs_sha256"]


@pytest.mark.parametrize("name,scale,fix", CASES)
def test_fw_trace_bit_exact(oracle_lib, name, scale, fix):
    g = golden(fix)
    net, dem = network(name, scale)
    r = oracle_lib.fw_solve(net, dem, max_iters=g["iterations"], nthreads=8)
    ref = hexarr(g["trace"])
    mine = np.stack([r["trace"][f] for f in oracle_lib.ITER_FIELDS], axis=1)
    assert mine.shape == ref.shape
    assert np.array_equal(mine.view(np.int64), ref.view(np.int64))
    assert sha(r["x"]) == g["x_sha256"] and sha(r["c"]) == g["c_sha256"]
    assert r["vht"] == float.fromhex(g["final_vht"])
    if "x_b64" in g:
        assert np.array_equal(r["x"], b64_array(g["x_b64"], np.float64))


@pytest.mark.parametrize("name,scale,fix", CASES[:2])
def test_sssp_bit_exact_both_queues(oracle_lib, name, scale, fix):
    """dist and prevlink of sssp_pape_lll and sssp_pape, at free-flow costs
    and at the costs after the golden run (non-trivial weights)."""
    g = golden(fix)
    net, dem = network(name, scale)
    final_c = b64_array(g["c_b64"], np.float64)
    for key, ent in g["sssp"].items():
        label, o, alg = key.split("/")
        costs = net.fftime if label == "freeflow" else final_c
        d, p = oracle_lib.sssp(net, costs, int(o), lll=(alg == "pape_lll"))
        assert sha(d[1:]) == ent["dist_sha256"], key
        assert sha(p[1:]) == ent["pred_sha256"], key


@pytest.mark.parametrize("tag", ["martin", "okuyama", "chicagosketch_flows",
                                 "chicagoregional_flows"])
def test_sssp_reference_graphs(oracle_lib, tag):
    g = golden("rmat_sssp.json")[tag]
    net = rmat(tag)
    for key, ent in g["sssp"].items():
        o, alg = key.split("/")
        d, p = oracle_lib.sssp(net, net.fftime, int(o), lll=(alg == "pape_lll"))
        assert sha(d[1:]) == ent["dist_sha256"], key
        assert sha(p[1:]) == ent["pred_sha256"], key


GREEDY = [("grid5", 1.0, "grid5_greedy.json"), ("metro", 0.3, "metro_s03_greedy.json")]


@pytest.mark.parametrize("name,scale,fix", GREEDY, ids=["grid5", "metro0.3"])
def test_greedy_matches_reference_binary(oracle_lib, name, scale, fix):
    """The oracle's greedy step-cost assignment against the reference's
    tap_greedy binary: same iterations, the same unsatisfied demand after
    each (as printed, %f), flows equal to the printed %.15f."""
    g = golden(fix)
    net, dem = network(name, scale)
    assert _digest(net, dem) == g["inputs_sha256"]
    o = oracle_lib.greedy(net, dem)
    assert o["iterations"] == g["iterations"] and o["stalled"] == g["stalled"]
    assert [f"{u:.6f}" for u in o["unsat_trace"]] == g["unsat_printed"]
    assert f"{o['total_cost']:.6f}" == g["total_cost_printed"]
    x = b64_array(g["x_b64"], np.float64)
    c = b64_array(g["c_b64"], np.float64)
    np.testing.assert_allclose(o["x"], x, rtol=0, atol=1e-15 * np.abs(x).max())
    np.testing.assert_allclose(o["c"], c, rtol=0, atol=5e-16 * np.abs(c).max() + 1e-15)
    assert o["unsatisfied"] == 0.0
