"""The network-modification driver (tap_variants_hip, the flags and output of
trunk/tap/tap_mpi_main.c) against the oracle: every enumerated variant of a
modification file in the reference's format (SiouxFalls_mods.txt layout:
multi-row change ids, CHANGE and two-row bidirectional ADD) is solved in one
batch, and each "id = <id> VHT = <cost>" line and flow file must match the
oracle's solve of that variant network built by the reference's rules."""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, network

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "traffic_assignment_amd", "tap_variants_hip")
TARGET = 1e-5


def _mods(net):
    from traffic_assignment_amd import synth

    def rec(k, **kw):
        r = {f: (int(getattr(net, f)[k]) if f in ("init", "term", "ltype") else
                 float(getattr(net, f)[k]))
             for f in ("init", "term", "capacity", "length", "fftime", "B", "power", "speed",
                       "toll", "ltype")}
        r.update(kw)
        return r

    def key(a, b):
        return int(np.nonzero((net.init == a) & (net.term == b))[0][0])

    k1, k2, k3 = key(1, 2), key(2, 1), key(7, 8)
    new = dict(capacity=9000.0, length=3.0, fftime=2.5, B=0.15, power=4.0, speed=0.0, toll=0.0,
               ltype=1)
    return [
        synth.NetMod("sf_change_1", "CHANGE", rec(k1, capacity=0.5 * net.capacity[k1]), 1.0),
        synth.NetMod("sf_change_1", "CHANGE", rec(k2, capacity=0.5 * net.capacity[k2]), 1.0),
        synth.NetMod("sf_change_2", "CHANGE", rec(k3, fftime=1.25 * net.fftime[k3]), 2.0),
        synth.NetMod("sf_add_1", "ADD", dict(new, init=1, term=13), 3.0),
        synth.NetMod("sf_add_1", "ADD", dict(new, init=13, term=1), 3.0),
        synth.NetMod("sf_add_2", "ADD", dict(new, init=20, term=6, fftime=4.0), 2.0),
        synth.NetMod("sf_change_3", "CHANGE", rec(key(12, 13), capacity=2 * net.capacity[key(12, 13)],
                                                  B=0.3), 1.5),
    ]


def _run(tmp_path, flags, tag):
    from traffic_assignment_amd import tntp
    net, dem = network("grid5")
    mods = _mods(net)
    nf, tf, mf = tmp_path / "net.txt", tmp_path / "trips.txt", tmp_path / "mods.txt"
    tntp.write_net(str(nf), net)
    tntp.write_trips(str(tf), dem)
    tntp.write_mods(str(mf), mods)
    base = str(tmp_path / f"{tag}_flows_")
    r = subprocess.run([CLI, "-r", str(TARGET), *flags, str(nf), str(tf), str(mf), base],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ids = re.findall(r"^id = (\S+) VHT = ([-0-9.]+)$", r.stderr, flags=re.M)
    return net, dem, mods, base, ids, r.stderr


def _check(oracle_lib, net, dem, mods, base, ids):
    from traffic_assignment_amd import synth, tntp
    for vid, vht in ids:
        chosen = re.split(r"_and_|\+", vid)
        vnet = synth.apply_mods(net, mods, chosen)
        oracle_lib.lib().ora_set_pow_mode(1)        # the engine's correctly rounded x^4
        try:
            o = oracle_lib.fw_solve(vnet, dem, target_relgap=TARGET, nthreads=4)
        finally:
            oracle_lib.lib().ora_set_pow_mode(0)
        init, term, vol, cost, hdr = tntp.read_flows(base + vid + ".txt")
        assert np.array_equal(init, vnet.init) and np.array_equal(term, vnet.term), vid
        assert int(hdr["iterations"]) == o["iterations"], vid
        np.testing.assert_allclose(vol, o["x"], rtol=1e-9, atol=1e-6)
        np.testing.assert_allclose(cost, o["c"], rtol=1e-9, atol=1e-9)
        assert abs(float(vht) - o["vht"]) <= 1e-9 * abs(o["vht"]) + 1e-6, (vid, vht, o["vht"])


def test_single_changes(engine_mod, oracle_lib, tmp_path):
    net, dem, mods, base, ids, err = _run(tmp_path, [], "single")
    assert [i for i, _ in ids] == ["sf_add_1", "sf_add_2", "sf_change_1", "sf_change_2",
                                   "sf_change_3"]
    assert "[add sf_add_1] initnode 1 termnode 13" in err
    assert "[change sf_change_1] initnode 1 termnode 2 old capacity" in err
    _check(oracle_lib, net, dem, mods, base, ids)


def test_pairs_subsets_and_list(engine_mod, oracle_lib, tmp_path):
    net, dem, mods, base, ids, _ = _run(tmp_path, ["-p"], "pairs")
    assert len(ids) == 10 and ids[0][0] == "sf_add_1_and_sf_add_2"
    _check(oracle_lib, net, dem, mods, base, ids)
    net, dem, mods, base, ids, err = _run(tmp_path, ["-s"], "subsets")
    assert "running power set of modifications: 16 total" in err
    assert len(ids) == 16 and ids[0][0] == "sf_add_1+sf_add_2+sf_change_1"
    _check(oracle_lib, net, dem, mods, base, ids)
    net, dem, mods, base, ids, err = _run(tmp_path, ["-l", "sf_change_3,sf_add_1"], "list")
    assert [i for i, _ in ids] == ["sf_change_3+sf_add_1"]
    _check(oracle_lib, net, dem, mods, base, ids)
    # -c skips variants whose flow file exists
    r = subprocess.run([CLI, "-c", "-r", str(TARGET), "-l", "sf_change_3,sf_add_1",
                        str(tmp_path / "net.txt"), str(tmp_path / "trips.txt"),
                        str(tmp_path / "mods.txt"), base],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "exists, skipping" in r.stderr
    assert "VHT" not in r.stderr
