"""CPU FP64 reference for the equilibrium gradients (test infrastructure only).

The reference has no gradient code.  What fixes the numbers is mathematics:

* Beckmann objective at equilibrium, Z(theta) = min_x sum_a int_0^{x_a}
  t_a(s; theta) ds.  By the envelope theorem dZ/dtheta is the partial
  derivative of the integral at the equilibrium volumes, for BPR
  t = t0 (1 + B (x/c)^p):
      dZ/dt0 = x (1 + B (x/c)^p / (p+1)),
      dZ/dc  = -t0 B p/(p+1) (x/c)^(p+1).
* VHT = sum_a x_a t_a(x_a) at equilibrium.  The equilibrium moves inside the
  span S of path differences p - q of each OD pair's active paths (paths
  with flow) with S'(J dx + r) = 0, J = diag(t'), r = dt/dtheta dtheta.  So
  dVHT/dtheta_a = (x_a - z_a) dt_a/dtheta_a, z = S w with
  (S'JS) w = S'(x o t').  (The t term of dVHT/dx drops out: t'S = 0 at
  equilibrium, equal costs on the active paths of an OD pair.)

This module builds S from PATHS: the active path flows come from the
path-based solver in equilibrium.c, S is spanned by the differences of the
paths of each OD pair, and the reduced system is solved densely (small
networks only).  The GPU engine instead builds S from fundamental cycles of
the shortest-path trees of every origin and solves by conjugate gradients;
the two constructions share nothing but the definition.

Both are pinned by central differences of re-solved equilibria (`fd_gradient`).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _Keep, capacity_fixed, f64p, i64p, lib


def _gp_lib():
    L = lib()
    if not getattr(L, "_gp_ready", False):
        L.ora_gp_create.restype = C.c_void_p
        L.ora_gp_create.argtypes = [C.c_void_p, C.c_void_p]
        L.ora_gp_destroy.argtypes = [C.c_void_p]
        L.ora_gp_set_params.argtypes = [C.c_void_p, f64p, f64p]
        L.ora_gp_relgap.restype = C.c_double
        L.ora_gp_relgap.argtypes = [C.c_void_p]
        L.ora_gp_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double)]
        L.ora_gp_flows.argtypes = [C.c_void_p, f64p, C.c_void_p]
        L.ora_gp_count.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.ora_gp_paths.argtypes = [C.c_void_p, i64p, f64p, i64p, i64p, i64p, i64p, f64p]
        L.ora_gp_num_od.restype = C.c_int64
        L.ora_gp_num_od.argtypes = [C.c_void_p]
        L._gp_ready = True
    return L


class PathEquilibrium:
    """Path-based user equilibrium of (net, dem) via gradient projection."""

    def __init__(self, net, dem):
        self.net = net
        self._k = _Keep(net, dem)
        self.h = _gp_lib().ora_gp_create(C.byref(self._k.net), C.byref(self._k.dem))
        if not self.h:
            raise RuntimeError("no path for a positive demand")
        self.t0 = np.array(net.fftime, dtype=np.float64)
        self.cap = capacity_fixed(net)

    def close(self):
        if self.h:
            _gp_lib().ora_gp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, t0=None, cap=None):
        if t0 is not None:
            self.t0 = np.ascontiguousarray(t0, dtype=np.float64)
        if cap is not None:
            self.cap = np.ascontiguousarray(cap, dtype=np.float64)
        _gp_lib().ora_gp_set_params(self.h, self.t0, self.cap)

    def solve(self, tol=1e-14, max_sweeps=20000):
        rg = C.c_double(0)
        n = _gp_lib().ora_gp_run(self.h, max_sweeps, tol, C.byref(rg))
        if n < 0:
            raise RuntimeError("no path for a positive demand")
        return n, rg.value

    def relgap(self):
        return _gp_lib().ora_gp_relgap(self.h)

    def flows(self):
        E = self.net.num_links
        x = np.zeros(max(E, 1)); t = np.zeros(max(E, 1))
        _gp_lib().ora_gp_flows(self.h, x, t.ctypes.data)
        return x[:E], t[:E]

    def paths(self):
        L = _gp_lib()
        npth, nl = C.c_int64(0), C.c_int64(0)
        L.ora_gp_count(self.h, C.byref(npth), C.byref(nl))
        nod = L.ora_gp_num_od(self.h)
        od = np.zeros(max(npth.value, 1), np.int64)
        flow = np.zeros(max(npth.value, 1))
        off = np.zeros(npth.value + 1, np.int64)
        links = np.zeros(max(nl.value, 1), np.int64)
        oo = np.zeros(max(nod, 1), np.int64); odd = np.zeros(max(nod, 1), np.int64)
        dd = np.zeros(max(nod, 1))
        L.ora_gp_paths(self.h, od, flow, off, links, oo, odd, dd)
        n = npth.value
        return dict(od=od[:n], flow=flow[:n], off=off, links=links[:nl.value],
                    od_origin=oo[:nod], od_dest=odd[:nod], od_demand=dd[:nod])

    def objective(self):
        x, _ = self.flows()
        return float(np.sum(beckmann_terms(self.t0, self.cap, self.net.B, self.net.power, x)))

    def vht(self):
        x, t = self.flows()
        return float(np.dot(x, t))


# Curvatures t' below this fraction of the largest count as zero: a detour of
# links with constant cost (parallel centroid connectors, t' ~ 1e-15) has an
# undetermined flow split, and the minimum-norm solution leaves it unchanged.
CURV_FLOOR = 1e-8


def beckmann_terms(t0, cap, B, p, x):
    """int_0^x t0 (1 + B (s/c)^p) ds per link."""
    q = x / cap
    return t0 * x * (1.0 + B * q ** p / (p + 1.0))


def link_partials(t0, cap, B, p, x):
    """Per-link closed forms at volumes x: envelope gradients of the Beckmann
    objective, dt/dc, dt/dt0 and t' = dt/dx."""
    q = x / cap
    Q = q ** p
    dz_dt0 = x * (1.0 + B * Q / (p + 1.0))
    dz_dc = -t0 * B * p / (p + 1.0) * q * Q
    dt_dc = -t0 * B * p * Q / cap
    dt_dt0 = 1.0 + B * Q
    with np.errstate(divide="ignore", invalid="ignore"):
        dt_dx = np.where(x > 0, t0 * B * p * q ** (p - 1.0) / cap,
                         np.where(p == 1.0, t0 * B / cap, 0.0))
    return dict(dobj_dt0=dz_dt0, dobj_dcap=dz_dc, dt_dcap=dt_dc, dt_dt0=dt_dt0, dt_dx=dt_dx)


def reference_gradient(net, x, paths, t0=None, cap=None, flow_tol=0.0, rank_tol=1e-9):
    """Gradients at equilibrium volumes x with the active path set `paths`
    (from PathEquilibrium.paths(); a path is active if its flow > flow_tol).

    S is spanned by the link-incidence differences (path - heaviest path) of
    every OD pair with two or more active paths; an orthonormal basis U of S
    comes from the SVD of that matrix (singular values below rank_tol times
    the largest are dropped: different OD pairs repeat the same detours), and
    z = U (U'JU)^-1 U'(x o J) is a dense direct solve.  Returns per-link
    arrays dobj_dt0, dobj_dcap, dvht_dcap, dvht_dt0, z, and "rank"."""
    t0 = np.asarray(net.fftime if t0 is None else t0, dtype=np.float64)
    cap = capacity_fixed(net) if cap is None else np.asarray(cap, dtype=np.float64)
    P = link_partials(t0, cap, np.asarray(net.B, float), np.asarray(net.power, float), x)
    E = net.num_links
    J = P["dt_dx"].copy()
    J[J < CURV_FLOOR * J.max()] = 0.0     # constant-cost detours: split left as is
    g = x * J
    act = np.nonzero(paths["flow"] > flow_tol)[0]
    od = paths["od"][act]
    off, links, flow = paths["off"], paths["links"], paths["flow"]
    cols = []
    for k in np.unique(od):
        ps = act[od == k]
        if ps.size < 2:
            continue
        p0 = ps[np.argmax(flow[ps])]
        base = np.zeros(E)
        np.add.at(base, links[off[p0]:off[p0 + 1]], 1.0)
        for p in ps:
            if p == p0:
                continue
            v = -base.copy()
            np.add.at(v, links[off[p]:off[p + 1]], 1.0)
            cols.append(v)
    z = np.zeros(E)
    rank = 0
    if cols:
        K = np.array(cols).T
        U, sv, _ = np.linalg.svd(K, full_matrices=False)
        rank = int((sv > rank_tol * sv[0]).sum())
        U = U[:, :rank]
        H = U.T @ (J[:, None] * U)
        u = np.linalg.lstsq(H, U.T @ g, rcond=1e-12)[0]
        z = U @ u
    return dict(dobj_dt0=P["dobj_dt0"], dobj_dcap=P["dobj_dcap"],
                dvht_dcap=(x - z) * P["dt_dcap"], dvht_dt0=(x - z) * P["dt_dt0"], z=z,
                rank=rank)


def fd_gradient(net, dem, links, rel_step=1e-4, tol=1e-14, max_sweeps=20000, which="cap"):
    """Central differences of re-solved equilibria for the given links:
    returns (dobj, dvht) arrays, each (F(theta+h) - F(theta-h)) / 2h with
    h = rel_step * theta_a, theta = capacity ("cap") or free-flow time ("t0").
    Each re-solve warm-starts from the previous equilibrium's paths."""
    eq = PathEquilibrium(net, dem)
    eq.solve(tol, max_sweeps)
    t0 = eq.t0.copy(); cap = eq.cap.copy()
    dobj = np.zeros(len(links)); dvht = np.zeros(len(links))
    for i, a in enumerate(links):
        vals = []
        for sgn in (1.0, -1.0):
            tt, cc = t0.copy(), cap.copy()
            th = cc if which == "cap" else tt
            h = rel_step * th[a]
            th[a] += sgn * h
            eq.set_params(t0=tt, cap=cc)
            eq.solve(tol, max_sweeps)
            vals.append((eq.objective(), eq.vht(), h))
        h = vals[0][2]
        dobj[i] = (vals[0][0] - vals[1][0]) / (2 * h)
        dvht[i] = (vals[0][1] - vals[1][1]) / (2 * h)
    eq.close()
    return dobj, dvht
