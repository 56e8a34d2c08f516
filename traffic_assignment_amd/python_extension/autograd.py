"""Differentiable user equilibrium: torch.autograd.Function over the engine.

The reference's Python operator boundary is its `pape` extension module
(trunk/python_extension/pape_module.c:301-473); it has no derivatives.  This
module adds the differentiable form of the Frank-Wolfe path:

    obj, vht, x = equilibrium(net, dem, capacity, fftime, target_relgap=1e-6)

solves the user equilibrium on the GPU for the given link capacities and
free-flow times (float64 tensors of num_links), and returns the Beckmann
objective and the vehicle-hours travelled sum_a x_a t_a(x_a) at the
equilibrium volumes x, as float64 tensors.  Their backward pass calls the
engine's gradient kernels (tapgpu_gradient):

    d obj / d(capacity, fftime)   envelope theorem, exact at equilibrium;
    d vht / d(capacity, fftime)   implicit differentiation over the active
                                  paths (reduced cost <= eps relative).

The gradients are those of the equilibrium the forward pass reached: at a
Frank-Wolfe relative gap g the volumes, and with them every gradient, carry
errors of the order of sqrt(g)..g; the tests pin the kernels at equilibria
solved to rounding level.  `eps` defaults to max(1e-9, 20 * relgap reached).
x is returned without a gradient (its Jacobian is not formed).

One engine (topology, demand and schedule resident on the GPU) serves
every call on the same network and demand: the arrays are compared with
the ones it was built from, and each call only replaces the link capacities
and free-flow times (tapgpu_set_links).  The objective and VHT come from the
engine (tapgpu_get_totals: the reference's link cost integral, compensated
sums).

There is no CPU fallback: the engine library and a GPU are required.
"""
from __future__ import annotations

import numpy as np
import torch

from ..engine import TapEngine


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy()


class _EngineCache:
    """The engine of the last (network, demand) seen, kept while their
    topology, BPR shapes and demand are unchanged (compared array by array,
    so arrays changed in place are seen)."""

    def __init__(self):
        self.key = None
        self.eng = None
        self.builds = 0

    @staticmethod
    def _key(net, dem):
        return ((net.num_nodes, net.num_zones, net.first_thru_node, dem.num_zones),
                tuple(np.array(a, copy=True) for a in (net.init, net.term, net.B, net.power,
                                                        dem.start, dem.dest, dem.demand)))

    def get(self, net, dem):
        key = self._key(net, dem)
        if self.eng is not None and key[0] == self.key[0] and all(
                a.shape == b.shape and np.array_equal(a, b) for a, b in zip(key[1], self.key[1])):
            return self.eng
        self.close()
        self.eng = TapEngine(net, dem)
        self.key = key
        self.builds += 1
        return self.eng

    def close(self):
        if self.eng is not None:
            self.eng.close()
        self.eng = None
        self.key = None


_CACHE = _EngineCache()


class EquilibriumFunction(torch.autograd.Function):
    """(capacity, fftime) -> (Beckmann objective, VHT, volumes) at equilibrium."""

    @staticmethod
    def forward(ctx, capacity, fftime, net, dem, target_relgap, max_iters, warm_volumes, eps,
                cg_tol):
        cap = _np(capacity)
        t0 = _np(fftime)
        if cap.shape != (net.num_links,) or t0.shape != (net.num_links,):
            raise ValueError("capacity and fftime must have num_links entries")
        if max_iters <= 0 and (warm_volumes is None or eps is None):
            raise ValueError("max_iters=0 needs warm_volumes and eps")
        eng = _CACHE.get(net, dem)
        try:
            eng.set_links(cap, t0)
            eng.fw_init(None if warm_volumes is None else np.asarray(warm_volumes, np.float64))
            if max_iters > 0:
                res = eng.fw_run(max_iters, target_relgap)[0]
            else:           # evaluate at the warm-start volumes as given
                res = dict(iterations=0, converged=False, relgap=float("nan"))
            x, _ = eng.flows()
            obj, vht = eng.totals()
            e = eps if eps is not None else max(1e-9, 20.0 * max(res["relgap"], 0.0))
            g = eng.gradient(0, eps=e, cg_tol=cg_tol)
        except BaseException:
            _CACHE.close()          # state unknown after an error: rebuild next time
            raise
        dev = capacity.device
        ctx.save_for_backward(*(torch.from_numpy(g[k]).to(dev)
                                for k in ("dobj_dcap", "dobj_dt0", "dvht_dcap", "dvht_dt0")))
        ctx.info = dict(res, **g["info"], eps=e)
        xt = torch.from_numpy(x).to(dev)
        ctx.mark_non_differentiable(xt)
        return (torch.tensor(obj, dtype=torch.float64, device=dev),
                torch.tensor(vht, dtype=torch.float64, device=dev), xt)

    @staticmethod
    def backward(ctx, g_obj, g_vht, _g_x):
        dobj_dcap, dobj_dt0, dvht_dcap, dvht_dt0 = ctx.saved_tensors
        grad_cap = g_obj * dobj_dcap + g_vht * dvht_dcap
        grad_t0 = g_obj * dobj_dt0 + g_vht * dvht_dt0
        return grad_cap, grad_t0, None, None, None, None, None, None, None


def equilibrium(net, dem, capacity, fftime, target_relgap=1e-6, max_iters=100000,
                warm_volumes=None, eps=None, cg_tol=1e-12):
    """(objective, vht, volumes) at the user equilibrium for these link
    parameters; objective and vht are differentiable w.r.t. capacity and
    fftime (float64 tensors).  FW runs from warm_volumes (or all-or-nothing
    at free flow) until target_relgap or max_iters; max_iters=0 takes
    warm_volumes as the equilibrium (eps must then be given)."""
    return EquilibriumFunction.apply(capacity, fftime, net, dem, float(target_relgap),
                                     int(max_iters), warm_volumes, eps, float(cg_tol))
