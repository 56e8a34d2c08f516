"""tapgpu_sssp_noalloc (the shortest-path entry point that allocates no
device memory: its scratch is a file-scope device array) against
tapgpu_sssp, through the C ABI with caller-owned device buffers: labels and
predecessor links identical, on the labels-in-LDS and the labels-in-HBM
kernels, two calls in a row with the second on fewer start nodes; sizes
beyond the scratch and bad start nodes are errors."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from conftest import network

pytestmark = pytest.mark.gpu


def _noalloc(engine_mod, eng, starts, ld=None):
    L = engine_mod.lib()
    L.tapgpu_sssp_noalloc.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                      C.c_void_p, C.c_int64]
    n = len(starts)
    ld = n if ld is None else ld
    d_s = torch.tensor(np.asarray(starts, dtype=np.int64), device="cuda")
    d_d = torch.full((eng.N * ld,), -7.0, dtype=torch.float64, device="cuda")
    d_p = torch.full((eng.N * ld,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = L.tapgpu_sssp_noalloc(eng._h, 0, d_s.data_ptr(), n, d_d.data_ptr(), d_p.data_ptr(), ld)
    torch.cuda.synchronize()
    if rc != 0:
        return rc, None, None
    d = d_d.cpu().numpy().reshape(eng.N, ld)[:, :n]
    p = d_p.cpu().numpy().reshape(eng.N, ld)[:, :n].astype(np.int64)
    return rc, d, p


@pytest.mark.parametrize("name,scale,strategy", [("metro", 0.3, "lds"), ("regional", 0.25, "hbm")])
def test_noalloc_equals_sssp(engine_mod, name, scale, strategy):
    print(f"case no-alloc sssp {name}@{scale} {strategy}", flush=True)
    net, dem = network(name, scale)
    rng = np.random.default_rng(21)
    with engine_mod.TapEngine(net, dem) as eng:
        eng.set_strategy(strategy)
        costs = net.fftime * rng.uniform(1.0, 3.0, net.num_links)
        eng.set_costs(costs)
        first = np.r_[np.arange(1, net.num_zones + 1), rng.integers(1, net.num_nodes + 1, 40)]
        second = first[[3, 0, 17, net.num_zones + 2, 5]]
        for starts, ld in ((first, len(first) + 3), (second, None)):
            rc, d, p = _noalloc(engine_mod, eng, starts, ld)
            assert rc == 0, engine_mod.lib().tapgpu_last_error(eng._h)
            rd, rp = eng.sssp(starts)
            assert np.array_equal(d.view(np.int64), rd[1:].view(np.int64))
            assert np.array_equal(p, rp[1:])
        # a start list beyond the scratch (17M entries, 68 MB as int32) is
        # rejected before anything is launched: the output buffers are never
        # touched (one element each here)
        L = engine_mod.lib()
        big = 17_000_000
        d_s = torch.ones(big, dtype=torch.int64, device="cuda")
        tiny_d = torch.zeros(1, dtype=torch.float64, device="cuda")
        tiny_p = torch.zeros(1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        rc = L.tapgpu_sssp_noalloc(eng._h, 0, d_s.data_ptr(), big, tiny_d.data_ptr(),
                                   tiny_p.data_ptr(), big)
        assert rc == -6, rc
        assert b"scratch" in L.tapgpu_last_error(eng._h)
        del d_s
        rc, _, _ = _noalloc(engine_mod, eng, [1, net.num_nodes + 1])
        assert rc == -1, rc
        # and the context still works
        rc, d, p = _noalloc(engine_mod, eng, second)
        rd, rp = eng.sssp(second)
        assert rc == 0 and np.array_equal(d.view(np.int64), rd[1:].view(np.int64))
