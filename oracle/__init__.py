"""ctypes front end of the CPU oracle (liboracle.so, built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py.  The product (traffic_assignment_amd) never
imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None

i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class _Net(C.Structure):
    _fields_ = [("num_zones", C.c_int64), ("num_nodes", C.c_int64),
                ("first_thru_node", C.c_int64), ("num_links", C.c_int64),
                ("init", C.c_void_p), ("term", C.c_void_p), ("capacity", C.c_void_p),
                ("fftime", C.c_void_p), ("B", C.c_void_p), ("power", C.c_void_p)]


class _Dem(C.Structure):
    _fields_ = [("num_zones", C.c_int64), ("start", C.c_void_p), ("dest", C.c_void_p),
                ("demand", C.c_void_p)]


ITER_FIELDS = ("lambda_", "gap", "objective", "lowerbound", "best_lowerbound", "relgap",
               "avg_excess", "vht")


class _Iter(C.Structure):
    _fields_ = [(f, C.c_double) for f in ITER_FIELDS]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_sssp_pape_lll.argtypes = [i64p, i64p, f64p, C.c_int64, C.c_int64, C.c_int64,
                                        i64p, f64p, i64p]
        L.ora_sssp_pape.argtypes = L.ora_sssp_pape_lll.argtypes
        L.ora_sssp_all.argtypes = [i64p, i64p, f64p, C.c_int64, C.c_int64, C.c_int64,
                                   C.c_int, f64p, i64p, C.c_int]
        L.ora_build_csr.argtypes = [C.POINTER(_Net), i64p, i64p, f64p]
        L.ora_bpr.argtypes = [C.c_double] * 5
        L.ora_bpr.restype = C.c_double
        L.ora_integral.argtypes = [C.c_double] * 5
        L.ora_integral.restype = C.c_double
        L.ora_dirderiv.argtypes = [C.POINTER(_Net), f64p, f64p, f64p, C.c_double]
        L.ora_dirderiv.restype = C.c_double
        L.ora_objective.argtypes = [C.POINTER(_Net), f64p, f64p]
        L.ora_objective.restype = C.c_double
        L.ora_line_search.argtypes = [C.POINTER(_Net), f64p, f64p, f64p]
        L.ora_line_search.restype = C.c_double
        L.ora_fw_assign.argtypes = [C.POINTER(_Net), C.POINTER(_Dem), i64p, i64p, f64p, f64p,
                                    C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ora_fw_solve.argtypes = [C.POINTER(_Net), C.POINTER(_Dem), C.c_int, C.c_double,
                                   C.c_int, C.c_void_p, f64p, f64p, C.c_void_p, C.c_int,
                                   C.POINTER(C.c_int), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


class _Keep:
    """Holds the numpy arrays a ctypes struct points into."""

    def __init__(self, net, dem=None):
        self.arrs = [np.ascontiguousarray(a, dtype=t) for a, t in (
            (net.init, np.int64), (net.term, np.int64), (net.capacity, np.float64),
            (net.fftime, np.float64), (net.B, np.float64), (net.power, np.float64))]
        a = self.arrs
        self.net = _Net(net.num_zones, net.num_nodes, net.first_thru_node, net.num_links,
                        *[x.ctypes.data for x in a])
        if dem is not None:
            self.darrs = [np.ascontiguousarray(dem.start, dtype=np.int64),
                          np.ascontiguousarray(dem.dest, dtype=np.int64),
                          np.ascontiguousarray(dem.demand, dtype=np.float64)]
            self.dem = _Dem(dem.num_zones, *[x.ctypes.data for x in self.darrs])


def csr(net):
    """(Va[N+2], Ea[E], Wa[E]) over 1-based node ids, Wa = free-flow time."""
    k = _Keep(net)
    N1 = net.num_nodes + 1
    Va = np.zeros(N1 + 1, dtype=np.int64)
    Ea = np.zeros(max(net.num_links, 1), dtype=np.int64)
    Wa = np.zeros(max(net.num_links, 1), dtype=np.float64)
    lib().ora_build_csr(C.byref(k.net), Va, Ea, Wa)
    return Va, Ea[:net.num_links], Wa[:net.num_links]


def sssp(net, costs, origin, lll=True):
    """(dist[N+1], prevlink[N+1]) from one origin (1-based ids)."""
    Va, Ea, _ = csr(net)
    N1 = net.num_nodes + 1
    d = np.zeros(N1); p = np.zeros(N1, dtype=np.int64); q = np.zeros(N1, dtype=np.int64)
    f = lib().ora_sssp_pape_lll if lll else lib().ora_sssp_pape
    f(Va, Ea, np.ascontiguousarray(costs, dtype=np.float64), N1, origin,
      net.first_thru_node, p, d, q)
    return d, p


def sssp_all(net, costs, lll=True, nthreads=8):
    """All origins 1..Z in the node-major pape_cuda layout [(N+1), Z] where
    column s-1 holds origin s (the reference's [v*Z + s] minus the offset)."""
    Va, Ea, _ = csr(net)
    N1, Z = net.num_nodes + 1, net.num_zones
    d = np.zeros(N1 * Z + Z + 1); p = np.zeros(N1 * Z + Z + 1, dtype=np.int64)
    lib().ora_sssp_all(Va, Ea, np.ascontiguousarray(costs, dtype=np.float64), N1, Z,
                       net.first_thru_node, int(lll), d, p, nthreads)
    idx = np.arange(N1)[:, None] * Z + np.arange(1, Z + 1)[None, :]
    return d[idx], p[idx]


def capacity_fixed(net):
    cap = np.array(net.capacity, dtype=np.float64)
    pos = cap[cap > 0]
    if pos.size:
        cap[cap <= 0] = pos.min()
    return cap


def bpr(t0, cap, B, power, Q):
    return lib().ora_bpr(t0, cap, B, power, Q)


def integral(t0, cap, B, power, Q):
    return lib().ora_integral(t0, cap, B, power, Q)


def line_search(net, x, d):
    k = _Keep(net)
    return lib().ora_line_search(C.byref(k.net), capacity_fixed(net),
                                 np.ascontiguousarray(x, dtype=np.float64),
                                 np.ascontiguousarray(d, dtype=np.float64))


def dirderiv(net, x, d, lam):
    k = _Keep(net)
    return lib().ora_dirderiv(C.byref(k.net), capacity_fixed(net),
                              np.ascontiguousarray(x, dtype=np.float64),
                              np.ascontiguousarray(d, dtype=np.float64), lam)


def objective(net, x):
    k = _Keep(net)
    return lib().ora_objective(C.byref(k.net), capacity_fixed(net),
                               np.ascontiguousarray(x, dtype=np.float64))


def fw_assign(net, dem, costs, nthreads=1):
    k = _Keep(net, dem)
    Va, Ea, _ = csr(net)
    y = np.zeros(max(net.num_links, 1))
    bo, bd = C.c_int64(0), C.c_int64(0)
    rc = lib().ora_fw_assign(C.byref(k.net), C.byref(k.dem), Va, Ea,
                             np.ascontiguousarray(costs, dtype=np.float64), y, nthreads,
                             C.byref(bo), C.byref(bd))
    if rc:
        raise RuntimeError(f"no path from {bo.value} to {bd.value}")
    return y[:net.num_links]


def fw_solve(net, dem, max_iters=0, target_relgap=0.0, nthreads=1, x_init=None):
    """Frank-Wolfe solve; returns dict with volumes, costs, trace and scalars."""
    k = _Keep(net, dem)
    E = net.num_links
    cap_trace = max_iters if max_iters > 0 else 100000
    trace = (_Iter * cap_trace)()
    x = np.zeros(max(E, 1)); c = np.zeros(max(E, 1))
    it, rg, obj, vht = C.c_int(0), C.c_double(0), C.c_double(0), C.c_double(0)
    xi = None
    if x_init is not None:
        xi = np.ascontiguousarray(x_init, dtype=np.float64)
    rc = lib().ora_fw_solve(C.byref(k.net), C.byref(k.dem), max_iters, target_relgap, nthreads,
                            None if xi is None else xi.ctypes.data, x, c,
                            C.cast(trace, C.c_void_p), cap_trace, C.byref(it), C.byref(rg),
                            C.byref(obj), C.byref(vht))
    if rc:
        raise RuntimeError(f"oracle fw_solve failed ({rc})")
    n = it.value
    tr = {f: np.array([getattr(trace[i], f) for i in range(min(n, cap_trace))])
          for f in ITER_FIELDS}
    return dict(x=x[:E], c=c[:E], iterations=n, relgap=rg.value, objective=obj.value,
                vht=vht.value, trace=tr)


def greedy(net, dem, max_iters=0, trace_cap=100000):
    """Greedy step-cost assignment (tap_greedy.c): dict with volumes x,
    costs c, iterations, unsatisfied demand (final and per iteration),
    total cost and whether it stopped on the unchanged-demand test."""
    L = lib()
    if not getattr(L, "_greedy_ready", False):
        L.ora_greedy.argtypes = [C.POINTER(_Net), C.POINTER(_Dem), C.c_int, f64p, f64p,
                                 C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p, C.c_int]
        L._greedy_ready = True
    k = _Keep(net, dem)
    E = max(net.num_links, 1)
    x = np.zeros(E); c = np.zeros(E)
    tr = np.zeros(trace_cap)
    it, st = C.c_int(0), C.c_int(0)
    un, tot = C.c_double(0), C.c_double(0)
    rc = L.ora_greedy(C.byref(k.net), C.byref(k.dem), int(max_iters), x, c, C.byref(it),
                      C.byref(un), C.byref(tot), C.byref(st), tr.ctypes.data, trace_cap)
    if rc:
        raise RuntimeError(f"oracle greedy failed ({rc})")
    n = it.value
    return dict(x=x[:net.num_links], c=c[:net.num_links], iterations=n, unsatisfied=un.value,
                total_cost=tot.value, stalled=bool(st.value), unsat_trace=tr[:min(n, trace_cap)])


def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "tap_frankwolfe_pthread"))
