"""Python front end of the gfx950 engine (ctypes over include/tapgpu.h).

Loads the in-tree libtapgpu.so built by traffic_assignment_amd.build; there
is no fallback: if the library or a GPU is missing, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtapgpu.so")

TAPGPU_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "no path", -4: "out of device memory",
          -5: "bad call order", -6: "limit exceeded"}


class TapGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class NoPathError(TapGpuError):
    pass


class _Network(C.Structure):
    _fields_ = [("num_zones", C.c_int64), ("num_nodes", C.c_int64),
                ("first_thru_node", C.c_int64), ("num_links", C.c_int64),
                ("init_node", C.c_void_p), ("term_node", C.c_void_p),
                ("capacity", C.c_void_p), ("free_flow_time", C.c_void_p),
                ("B", C.c_void_p), ("power", C.c_void_p)]


class _Demand(C.Structure):
    _fields_ = [("num_zones", C.c_int64), ("od_start", C.c_void_p), ("od_dest", C.c_void_p),
                ("od_demand", C.c_void_p)]


ITER_FIELDS = ("lambda_", "gap", "objective", "lowerbound", "best_lowerbound", "relgap",
               "avg_excess", "vht")


class _Iter(C.Structure):
    _fields_ = [(f, C.c_double) for f in ITER_FIELDS]


class _Result(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("relgap", C.c_double),
                ("objective", C.c_double), ("vht", C.c_double)]


class _Stats(C.Structure):
    _fields_ = [("sssp_ms", C.c_double), ("aon_ms", C.c_double), ("link_ms", C.c_double),
                ("total_ms", C.c_double), ("sssp_sweeps", C.c_int64), ("aon_levels", C.c_int64),
                ("sssp_items", C.c_int64), ("aon_items", C.c_int64), ("link_passes", C.c_int64),
                ("sssp_inlinks", C.c_int64), ("sssp_updates", C.c_int64),
                ("sssp_outlinks", C.c_int64), ("aon_lanes", C.c_int64),
                ("sweep_kernel_ms", C.c_double), ("level_kernel_ms", C.c_double),
                ("sweep_launches", C.c_int64), ("level_launches", C.c_int64),
                ("origin_kernel_ms", C.c_double), ("origin_launches", C.c_int64),
                ("origin_blocks", C.c_int64), ("origin_expansions", C.c_int64),
                ("origin_scanned", C.c_int64), ("origin_updates", C.c_int64),
                ("origin_relax_rounds", C.c_int64), ("origin_load_rounds", C.c_int64),
                ("origin_loaded", C.c_int64), ("origin_walk_steps", C.c_int64),
                ("origin_relax_ms", C.c_double), ("origin_pred_ms", C.c_double),
                ("origin_load_ms", C.c_double), ("origin_tune_ms", C.c_double),
                ("origin_chase", C.c_int64), ("origin_delta", C.c_double),
                ("origin_split", C.c_int64), ("origin_deferred", C.c_int64),
                ("origin_block", C.c_int64), ("origin_queue", C.c_int64),
                ("lanes_launches", C.c_int64), ("lanes_tasks", C.c_int64),
                ("lanes_pred_scans", C.c_int64), ("lanes_fallbacks", C.c_int64),
                ("lanes_delta", C.c_double), ("graph_launches", C.c_int64)]


class _GreedyResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("stalled", C.c_int32), ("unsatisfied", C.c_double),
                ("total_cost", C.c_double)]


class _GradInfo(C.Structure):
    _fields_ = [("cycles", C.c_int64), ("nonzeros", C.c_int64), ("useful_rounds", C.c_int32),
                ("cg_iters", C.c_int32), ("cg_status", C.c_int32), ("reserved", C.c_int32),
                ("cg_relres", C.c_double), ("setup_ms", C.c_double), ("cg_ms", C.c_double)]


_lib = None
vp = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (run traffic_assignment_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.tapgpu_create.argtypes = [C.POINTER(vp), C.POINTER(_Network), C.POINTER(_Demand)]
        L.tapgpu_create_batched.argtypes = [C.POINTER(vp), C.POINTER(_Network), C.POINTER(_Demand),
                                            C.c_int64, vp, vp, vp, vp, vp]
        L.tapgpu_destroy.argtypes = [vp]
        L.tapgpu_destroy.restype = None
        L.tapgpu_last_error.argtypes = [vp]
        L.tapgpu_last_error.restype = C.c_char_p
        L.tapgpu_num_variants.argtypes = [vp]
        L.tapgpu_num_variants.restype = C.c_int64
        L.tapgpu_set_costs.argtypes = [vp, C.c_int64, vp]
        L.tapgpu_sssp.argtypes = [vp, C.c_int64, vp, C.c_int64, vp, vp, C.c_int64, C.c_int64]
        L.tapgpu_aon.argtypes = [vp, C.c_int64, vp]
        L.tapgpu_load_pred.argtypes = [vp, C.c_int64, vp, C.c_int64, C.c_int64, vp]
        L.tapgpu_fw_init.argtypes = [vp, vp]
        L.tapgpu_fw_run.argtypes = [vp, C.c_int, C.c_double, vp]
        L.tapgpu_get_flows.argtypes = [vp, C.c_int64, vp, vp]
        L.tapgpu_get_trace.argtypes = [vp, C.c_int64, vp, C.c_int, C.POINTER(C.c_int)]
        L.tapgpu_device_state.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
        L.tapgpu_get_stats.argtypes = [vp, C.POINTER(_Stats)]
        L.tapgpu_set_profiling.argtypes = [vp, C.c_int]
        L.tapgpu_set_strategy.argtypes = [vp, C.c_int]
        L.tapgpu_get_strategy.argtypes = [vp]
        L.tapgpu_set_origin_chase.argtypes = [vp, C.c_int]
        L.tapgpu_get_origin_chase.argtypes = [vp]
        L.tapgpu_set_origin_queue.argtypes = [vp, C.c_int]
        L.tapgpu_get_origin_queue.argtypes = [vp]
        L.tapgpu_set_origin_delta.argtypes = [vp, C.c_double]
        L.tapgpu_set_origin_split.argtypes = [vp, C.c_int]
        L.tapgpu_set_origin_block.argtypes = [vp, C.c_int]
        L.tapgpu_get_origin_block.argtypes = [vp]
        L.tapgpu_set_lanes.argtypes = [vp, C.c_int]
        L.tapgpu_get_lanes.argtypes = [vp]
        L.tapgpu_set_lanes_delta.argtypes = [vp, C.c_double]
        L.tapgpu_get_lanes_delta.argtypes = [vp]
        L.tapgpu_get_lanes_delta.restype = C.c_double
        L.tapgpu_get_totals.argtypes = [vp, C.c_int64, vp, vp]
        L.tapgpu_set_links.argtypes = [vp, C.c_int64, vp, vp]
        L.tapgpu_set_graphs.argtypes = [vp, C.c_int]
        L.tapgpu_set_link_passes.argtypes = [vp, C.c_int, C.c_int]
        L.tapgpu_get_graphs.argtypes = [vp, C.POINTER(C.c_int)]
        L.tapgpu_greedy.argtypes = [vp, C.c_int64, C.c_int, vp, vp, vp, vp, C.c_int,
                                    C.POINTER(_GreedyResult)]
        L.tapgpu_get_origin_split.argtypes = [vp]
        L.tapgpu_get_origin_delta.argtypes = [vp]
        L.tapgpu_get_origin_delta.restype = C.c_double
        L.tapgpu_gradient.argtypes = [vp, C.c_int64, C.c_double, C.c_double, C.c_int,
                                      vp, vp, vp, vp, C.POINTER(_GradInfo)]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class TapEngine:
    """One network topology, num_variants TAP instances, state in HBM.

    net / dem follow traffic_assignment_amd.synth (1-based ids, links sorted
    by (init, term)).  For the batched form pass per-variant arrays shaped
    [num_variants, num_links] (present: bool mask over the union links).
    """

    def __init__(self, net, dem, num_variants=1, present=None, capacity=None,
                 fftime=None, B=None, power=None):
        L = lib()
        self._keep = [np.ascontiguousarray(a, dtype=t) for a, t in (
            (net.init, np.int64), (net.term, np.int64), (net.capacity, np.float64),
            (net.fftime, np.float64), (net.B, np.float64), (net.power, np.float64),
            (dem.start, np.int64), (dem.dest, np.int64), (dem.demand, np.float64))]
        k = self._keep
        self._net = _Network(net.num_zones, net.num_nodes, net.first_thru_node, net.num_links,
                             *[a.ctypes.data for a in k[:6]])
        self._dem = _Demand(dem.num_zones, *[a.ctypes.data for a in k[6:]])
        self.E = int(net.num_links)
        self.N = int(net.num_nodes)
        self.Z = int(net.num_zones)
        self.V = int(num_variants)
        h = vp()
        if num_variants == 1 and present is None and capacity is None and fftime is None \
                and B is None and power is None:
            rc = L.tapgpu_create(C.byref(h), C.byref(self._net), C.byref(self._dem))
        else:
            extra = []
            for a, t in ((present, np.uint8), (capacity, np.float64), (fftime, np.float64),
                         (B, np.float64), (power, np.float64)):
                if a is None:
                    extra.append(None)
                else:
                    arr = np.ascontiguousarray(a, dtype=t).reshape(-1)
                    if arr.size != self.V * self.E:
                        raise ValueError("per-variant arrays must be [num_variants, num_links]")
                    extra.append(arr)
            self._keep += [a for a in extra if a is not None]
            rc = L.tapgpu_create_batched(C.byref(h), C.byref(self._net), C.byref(self._dem),
                                         self.V, *[_p(a) for a in extra])
        if rc != TAPGPU_OK:
            raise TapGpuError(rc, L.tapgpu_last_error(None).decode())
        self._h = h

    # -- plumbing --
    def _check(self, rc):
        if rc != TAPGPU_OK:
            msg = lib().tapgpu_last_error(self._h).decode()
            raise (NoPathError if rc == -3 else TapGpuError)(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            lib().tapgpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- shortest paths --
    def set_costs(self, costs, variant=0):
        c = np.ascontiguousarray(costs, dtype=np.float64)
        assert c.size == self.E
        self._check(lib().tapgpu_set_costs(self._h, variant, _p(c)))

    def sssp(self, start_nodes, variant=0):
        """(dist, pred) as [(N+1), len(start_nodes)] arrays (row = node id)."""
        s = np.ascontiguousarray(start_nodes, dtype=np.int64)
        n = s.size
        d = np.empty((self.N + 1) * max(n, 1), dtype=np.float64)
        p = np.empty((self.N + 1) * max(n, 1), dtype=np.int64)
        self._check(lib().tapgpu_sssp(self._h, variant, _p(s), n, _p(d), _p(p), n, 0))
        return d[:(self.N + 1) * n].reshape(self.N + 1, n), p[:(self.N + 1) * n].reshape(self.N + 1, n)

    def aon(self, variant=0):
        y = np.empty(self.E, dtype=np.float64)
        self._check(lib().tapgpu_aon(self._h, variant, _p(y)))
        return y

    def load_pred(self, pred, variant=0):
        """All-or-nothing loading along caller-supplied trees
        (tapgpu_load_pred): pred is [(N+1), Z] (row = node id, column =
        origin zone - 1) of the last link on each path, -1 = none."""
        p = np.ascontiguousarray(pred, dtype=np.int64)
        assert p.shape == (self.N + 1, self.Z)
        y = np.empty(self.E, dtype=np.float64)
        self._check(lib().tapgpu_load_pred(self._h, variant, _p(p), self.Z, 0, _p(y)))
        return y

    # -- Frank-Wolfe --
    def fw_init(self, warm_volumes=None):
        w = None
        if warm_volumes is not None:
            w = np.ascontiguousarray(warm_volumes, dtype=np.float64).reshape(-1)
            assert w.size == self.V * self.E
        self._check(lib().tapgpu_fw_init(self._h, _p(w)))

    def fw_run(self, max_iters=0, target_relgap=0.0):
        res = (_Result * self.V)()
        self._check(lib().tapgpu_fw_run(self._h, int(max_iters), float(target_relgap),
                                        C.cast(res, vp)))
        return [dict(iterations=r.iterations, converged=bool(r.converged), relgap=r.relgap,
                     objective=r.objective, vht=r.vht) for r in res]

    def flows(self, variant=0):
        x = np.empty(self.E, dtype=np.float64)
        c = np.empty(self.E, dtype=np.float64)
        self._check(lib().tapgpu_get_flows(self._h, variant, _p(x), _p(c)))
        return x, c

    def totals(self, variant=0):
        """(Beckmann objective, VHT) at the variant's current volumes."""
        o, v = C.c_double(), C.c_double()
        self._check(lib().tapgpu_get_totals(self._h, variant, C.byref(o), C.byref(v)))
        return o.value, v.value

    def set_links(self, capacity=None, fftime=None, variant=0):
        """Replace the variant's capacities and/or free-flow times (fw_init
        must follow)."""
        a = None if capacity is None else np.ascontiguousarray(capacity, dtype=np.float64)
        b = None if fftime is None else np.ascontiguousarray(fftime, dtype=np.float64)
        for t in (a, b):
            if t is not None and t.shape != (self.E,):
                raise ValueError("link arrays must have num_links entries")
        self._check(lib().tapgpu_set_links(self._h, variant, _p(a), _p(b)))

    def trace(self, variant=0):
        n = C.c_int(0)
        self._check(lib().tapgpu_get_trace(self._h, variant, None, 0, C.byref(n)))
        buf = (_Iter * max(n.value, 1))()
        self._check(lib().tapgpu_get_trace(self._h, variant, C.cast(buf, vp), n.value,
                                           C.byref(n)))
        return {f: np.array([getattr(buf[i], f) for i in range(n.value)]) for f in ITER_FIELDS}

    def device_state(self):
        x, c = vp(), vp()
        self._check(lib().tapgpu_device_state(self._h, C.byref(x), C.byref(c)))
        return x.value, c.value

    def gradient(self, variant=0, eps=1e-9, cg_tol=1e-12, max_cg=20000):
        """Equilibrium gradients at the current volumes (tapgpu_gradient):
        dict of per-link arrays dobj_dt0, dobj_dcap, dvht_dcap, dvht_dt0 and
        an "info" dict (cycles, nonzeros, cg_iters, cg_relres, ...)."""
        out = {k: np.empty(self.E, dtype=np.float64)
               for k in ("dobj_dt0", "dobj_dcap", "dvht_dcap", "dvht_dt0")}
        info = _GradInfo()
        self._check(lib().tapgpu_gradient(self._h, variant, float(eps), float(cg_tol), int(max_cg),
                                          _p(out["dobj_dt0"]), _p(out["dobj_dcap"]),
                                          _p(out["dvht_dcap"]), _p(out["dvht_dt0"]),
                                          C.byref(info)))
        out["info"] = {f: getattr(info, f) for f, _ in _GradInfo._fields_ if f != "reserved"}
        return out

    def greedy(self, variant=0, max_iters=0, trace_cap=10000):
        """The reference's greedy step-cost assignment (tapgpu_greedy): dict
        with volumes x, costs c, iterations, stalled, unsatisfied,
        total_cost and per-iteration unsat_trace / ms_trace."""
        x = np.empty(self.E); c = np.empty(self.E)
        un = np.zeros(trace_cap); ms = np.zeros(trace_cap)
        res = _GreedyResult()
        self._check(lib().tapgpu_greedy(self._h, variant, int(max_iters), _p(x), _p(c), _p(un),
                                        _p(ms), trace_cap, C.byref(res)))
        n = min(res.iterations, trace_cap)
        return dict(x=x, c=c, iterations=res.iterations, stalled=bool(res.stalled),
                    unsatisfied=res.unsatisfied, total_cost=res.total_cost,
                    unsat_trace=un[:n], ms_trace=ms[:n])

    STRATEGIES = {"auto": 0, "sweeps": 1, "lds": 2, "hbm": 3}

    def set_strategy(self, name):
        self._check(lib().tapgpu_set_strategy(self._h, self.STRATEGIES[name]))

    @property
    def strategy(self):
        return {1: "sweeps", 2: "lds", 3: "hbm"}[lib().tapgpu_get_strategy(self._h)]

    def set_origin_chase(self, chase):
        """Chained expansions per thread and round on the one-origin-per-
        workgroup path, in rounds that scan the frontier bitmap (-1: tune on
        the next loading)."""
        self._check(lib().tapgpu_set_origin_chase(self._h, int(chase)))

    @property
    def origin_chase(self):
        return lib().tapgpu_get_origin_chase(self._h)

    def set_origin_queue(self, entries):
        """Largest frontier the one-origin-per-workgroup paths take from
        their LDS queues (-1: as many as their LDS holds, 0: bitmap scans
        only)."""
        self._check(lib().tapgpu_set_origin_queue(self._h, int(entries)))

    @property
    def origin_queue(self):
        return lib().tapgpu_get_origin_queue(self._h)

    def set_origin_split(self, origins):
        """Origins per variant co-run by the labels-in-HBM kernel beside the
        LDS kernel (0: none, -1: tune on the next loading)."""
        self._check(lib().tapgpu_set_origin_split(self._h, int(origins)))

    @property
    def origin_split(self):
        return lib().tapgpu_get_origin_split(self._h)

    def set_origin_block(self, threads):
        """Workgroup size of the LDS path (1024, 512, 256; -1: tune with the
        queue size on the next loading)."""
        self._check(lib().tapgpu_set_origin_block(self._h, int(threads)))

    @property
    def origin_block(self):
        return lib().tapgpu_get_origin_block(self._h)

    def set_origin_delta(self, mult):
        """Bucket width of the labels-in-HBM path in mean free-flow link
        times (0: none, -1: tune on the next loading)."""
        self._check(lib().tapgpu_set_origin_delta(self._h, float(mult)))

    @property
    def origin_delta(self):
        return lib().tapgpu_get_origin_delta(self._h)

    def set_lanes(self, min_active):
        """Batched loading with variants in wavefront lanes when at least
        min_active variants are active (0: never, -1: the default)."""
        self._check(lib().tapgpu_set_lanes(self._h, int(min_active)))

    @property
    def lanes(self):
        return lib().tapgpu_get_lanes(self._h)

    def set_lanes_delta(self, mult):
        """Bucket width of the lanes kernel in mean free-flow link times (0:
        none, -1: tune on the next loading)."""
        self._check(lib().tapgpu_set_lanes_delta(self._h, float(mult)))

    @property
    def lanes_delta(self):
        return lib().tapgpu_get_lanes_delta(self._h)

    def set_link_passes(self, mode, window=0):
        """Link passes: 0 by network size, 1 one workgroup per variant, 2
        multi-block; window: step sizes per line-search window (0: 16)."""
        self._check(lib().tapgpu_set_link_passes(self._h, int(mode), int(window)))

    def set_graphs(self, mode=1):
        """Replay Frank-Wolfe iterations from captured HIP graphs (1, the
        default: iterations per replay from the iteration time; 2..8: that
        many) or issue their launches one by one (0 / False)."""
        self._check(lib().tapgpu_set_graphs(self._h, int(mode)))

    @property
    def graphs(self):
        v = C.c_int()
        self._check(lib().tapgpu_get_graphs(self._h, C.byref(v)))
        return v.value

    def set_profiling(self, on=True):
        self._check(lib().tapgpu_set_profiling(self._h, int(bool(on))))

    def stats(self):
        s = _Stats()
        self._check(lib().tapgpu_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _Stats._fields_}
