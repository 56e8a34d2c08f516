"""Batched-variant construction against the reference's modification rules
(trunk/tap/tap_mpi_main.c:201-240: ADD appends, CHANGE overwrites the whole
record, then links are re-sorted)."""
from __future__ import annotations

import numpy as np

from traffic_assignment_amd import synth, tntp, variants


def _small():
    base, dem = synth.metro(scale=0.3)
    mods = synth.upgrades(base, n_change=6, n_add=4)
    subsets = synth.budget_subsets(mods, n_variants=12, budget=25.0)
    return base, dem, mods, subsets


def test_union_restricted_to_present_is_the_variant_network():
    base, dem, mods, subsets = _small()
    u, present, rec = variants.batch_arrays(base, mods, subsets)
    for v, chosen in enumerate(subsets):
        net = variants.variant_network(base, mods, chosen)
        m = present[v].astype(bool)
        assert np.array_equal(u.init[m], net.init) and np.array_equal(u.term[m], net.term)
        for f in variants.FIELDS:
            assert np.array_equal(rec[f][v][m], getattr(net, f)), f


def test_budget_and_change_semantics():
    base, dem, mods, subsets = _small()
    cost = {m.change_id: m.project_cost for m in mods}
    for chosen in subsets:
        assert sum(cost[c] for c in chosen) <= 25.0 + 1e-12
        assert len(chosen) >= 1
    # CHANGE rows scale capacity by 1.5 on both directions of a road
    ch = [m for m in mods if m.modtype == "CHANGE"]
    k = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(base.init, base.term))}
    for m in ch:
        i = k[(m.link["init"], m.link["term"])]
        assert m.link["capacity"] == base.capacity[i] * 1.5
    # uneven link counts across variants
    u, present, _ = variants.batch_arrays(base, mods, subsets)
    assert len(set(present.sum(axis=1).tolist())) > 1


def test_mods_file_round_trip(tmp_path):
    base, dem, mods, subsets = _small()
    p = tmp_path / "mods.txt"
    tntp.write_mods(str(p), mods)
    back = tntp.read_mods(str(p))
    assert [(m.change_id, m.modtype, m.project_cost) for m in back] == \
           [(m.change_id, m.modtype, m.project_cost) for m in mods]
    assert all(a.link == b.link for a, b in zip(back, mods))
