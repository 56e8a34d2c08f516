"""Upgrade-subset variants as one batched engine (the MPI fan-out of
trunk/tap/tap_mpi_main.c:603-844 as a single GPU launch per FW step).

Each variant is the base network with a subset of modifications applied:
ADD appends a link record, CHANGE overwrites the whole record of an
existing (init, term) link (tap_mpi_main.c:201-240), then links are sorted
by (init, term) (:753).  The batch runs on the *union* topology (base links
plus every ADD link, sorted); a variant's own link array is exactly the union
restricted to its present links, in the same order, so sums over its links
are taken in the same order as in a solve of that variant alone.
"""
from __future__ import annotations

import numpy as np

from . import synth

FIELDS = ("capacity", "fftime", "B", "power")


def union_network(base: synth.Network, mods) -> synth.Network:
    """Base links plus every ADD link of the modification list."""
    adds = sorted({m.change_id for m in mods if m.modtype == "ADD"})
    u = synth.apply_mods(base, [m for m in mods if m.modtype == "ADD"], adds)
    keys = list(zip(u.init.tolist(), u.term.tolist()))
    if len(set(keys)) != len(keys):
        raise ValueError("duplicate (init, term) links in the union network")
    return u


def batch_arrays(base: synth.Network, mods, subsets):
    """(union, present[V,E] uint8, {field: [V,E] float64}) for the subsets."""
    u = union_network(base, mods)
    E, V = u.num_links, len(subsets)
    index = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(u.init, u.term))}
    base_keys = set(zip(base.init.tolist(), base.term.tolist()))
    base_mask = np.array([(int(a), int(b)) in base_keys for a, b in zip(u.init, u.term)],
                         dtype=np.uint8)
    present = np.tile(base_mask, (V, 1))
    rec = {f: np.tile(getattr(u, f), (V, 1)) for f in FIELDS}
    by_id = {}
    for m in mods:
        by_id.setdefault(m.change_id, []).append(m)
    for v, chosen in enumerate(subsets):
        for cid in chosen:
            for m in by_id[cid]:
                k = index.get((m.link["init"], m.link["term"]))
                if k is None:
                    raise KeyError(f"link not found for change {cid}")
                if m.modtype == "ADD":
                    present[v, k] = 1
                for f in FIELDS:
                    rec[f][v, k] = m.link[f]
    return u, present, rec


def variant_network(base: synth.Network, mods, chosen) -> synth.Network:
    return synth.apply_mods(base, mods, chosen)


def make_engine(base, dem, mods, subsets):
    from .engine import TapEngine
    u, present, rec = batch_arrays(base, mods, subsets)
    eng = TapEngine(u, dem, num_variants=len(subsets), present=present,
                    capacity=rec["capacity"], fftime=rec["fftime"], B=rec["B"],
                    power=rec["power"])
    return eng, u, present


def workload(n_variants=512, scale=1.0, budget=40.0):
    """The batched workload: metro network, 32 upgrades, budgeted subsets."""
    base, dem = synth.metro(scale=scale)
    mods = synth.upgrades(base)
    subsets = synth.budget_subsets(mods, n_variants=n_variants, budget=budget)
    return base, dem, mods, subsets
