"""The C host library's TNTP readers against the Python ones and the
reference's own files (no GPU call is made)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, network

LIB = os.path.join(ROOT, "traffic_assignment_amd", "libtap_hip.so")
BUILD = os.path.join(ROOT, "tests", "_build")


@pytest.fixture(scope="module")
def parse_check():
    if not os.path.exists(LIB):
        pytest.skip("libtap_hip.so not built")
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, "parse_check")
    subprocess.run(["gcc", "-O1", "-I", os.path.join(ROOT, "include"), "-o", exe,
                    os.path.join(ROOT, "tests", "native", "parse_check.c"),
                    "-L", os.path.dirname(LIB), "-ltap_hip", "-ltapgpu",
                    "-Wl,-rpath," + os.path.dirname(LIB)], check=True)
    return lambda kind, path: subprocess.run([exe, kind, path], capture_output=True, text=True,
                                             check=True).stdout.split("\n")


def test_net_and_trips(parse_check, tmp_path):
    from traffic_assignment_amd import tntp
    net, dem = network("metro", 0.3)
    nf, tf = tmp_path / "net.txt", tmp_path / "trips.txt"
    tntp.write_net(str(nf), net)
    tntp.write_trips(str(tf), dem)
    out = parse_check("net", str(nf))
    assert out[0].split() == [str(net.num_zones), str(net.num_nodes), str(net.first_thru_node),
                              str(net.num_links)]
    rows = [r.split() for r in out[1:] if r]
    assert np.array_equal([int(r[0]) for r in rows], net.init)
    assert np.array_equal([float.fromhex(r[2]) for r in rows], net.capacity)
    assert np.array_equal([float.fromhex(r[4]) for r in rows], net.fftime)
    out = parse_check("trips", str(tf))
    rows = [r.split() for r in out[1:] if r]
    assert len(rows) == dem.nnz
    assert np.array_equal([int(r[1]) for r in rows], dem.dest)
    assert np.array_equal([float.fromhex(r[2]) for r in rows], dem.demand)


def test_flows_and_mod_files(parse_check, tmp_path):
    from traffic_assignment_amd import synth, tntp
    net, _ = network("metro", 0.3)
    rng = np.random.default_rng(1)
    vol = rng.uniform(0, 5000, net.num_links)
    cost = rng.uniform(0.5, 9, net.num_links)
    ff = tmp_path / "flows.txt"
    tntp.write_flows(str(ff), net, vol, cost, header="~ iterations = 3 objvalue = 1.5 relgap = 0.1\n")
    rows = [r.split() for r in parse_check("flows", str(ff))[1:] if r]
    assert np.allclose([float.fromhex(r[2]) for r in rows], vol, rtol=1e-15, atol=1e-12)
    mods = synth.upgrades(net, n_change=3, n_add=2)
    mf = tmp_path / "mods.txt"
    tntp.write_mods(str(mf), mods)
    rows = [r.split() for r in parse_check("mods", str(mf))[1:] if r]
    assert [r[0] for r in rows] == [m.change_id for m in mods]
    assert [int(r[1]) for r in rows] == [1 if m.modtype == "ADD" else 2 for m in mods]


def test_reference_mod_files(parse_check):
    """The reference's own modification and trip-modification files."""
    for name, kind, col in (("SiouxFalls_mods.txt", "mods", None),
                            ("ChicagoRegional_mods.txt", "mods", None),
                            ("ChicagoRegional_tripmods.txt", "tripmods", None)):
        path = os.path.join("/root/reference/trunk/tap", name)
        if not os.path.exists(path):
            pytest.skip("reference files are not present on this machine")
        out = parse_check(kind, path)
        n = int(out[0])
        body = [l for l in open(path).read().splitlines() if l.strip() and not l.startswith("~")]
        assert n == len(body) and n > 0


REF_OBJ = os.path.join(ROOT, "oracle", "_ref")


def test_trip_modifications_match_the_reference(tmp_path):
    """apply_trip_modification (trunk/tap/tap_functions.c:521-598): P2P,
    UNIFORM from an origin, UNIFORM to a destination, ALL, and the error
    rows (both ends nonzero in UNIFORM, a zone out of range), applied in
    file order to the same demand: the library and the reference's own
    objects give bit-identical demands, return codes and stderr lines."""
    objs = [os.path.join(REF_OBJ, o) for o in ("parsetapfiles.o", "tap_functions.o")]
    if not os.path.exists(LIB) or not all(os.path.exists(o) for o in objs):
        pytest.skip("library or reference objects not built")
    from traffic_assignment_amd import tntp
    net, dem = network("metro", 0.3)
    Z = dem.num_zones
    tf = tmp_path / "trips.txt"
    tntp.write_trips(str(tf), dem)
    o1 = int(dem.dest[0] if dem.start[1] > 0 else 2)
    mods = [("p1", "P2P", 1, int(dem.dest[dem.start[0]]), 1.5),
            ("p2", "P2P", 3, 1, 0.25),                     # pair absent or present
            ("u1", "UNIFORM", 5, 0, 1.1),
            ("u2", "UNIFORM", 0, o1, 0.9),
            ("u3", "UNIFORM", 0, Z, 2.0),
            ("a1", "ALL", 0, 0, 1.05),
            ("bad", "UNIFORM", 4, 6, 1.2),                 # both nonzero: -1
            ("range", "P2P", Z + 1, 2, 3.0)]               # beyond num_zones: -1
    mf = tmp_path / "tripmods.txt"
    mf.write_text("~ChangeId\tType\tOrig\tDest\tMultiplier\n" +
                  "".join(f"{a}\t{b}\t{c}\t{d}\t{e}\n" for a, b, c, d, e in mods))
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(ROOT, "tests", "native", "tripmod_check.c")
    inc = ["-I", os.path.join(ROOT, "include")]
    mine, ref = os.path.join(BUILD, "tripmod_lib"), os.path.join(BUILD, "tripmod_ref")
    subprocess.run(["gcc", "-O1", *inc, "-o", mine, src, "-L", os.path.dirname(LIB), "-ltap_hip",
                    "-ltapgpu", "-Wl,-rpath," + os.path.dirname(LIB)], check=True)
    subprocess.run(["gcc", "-O1", *inc, "-o", ref, src, *objs, "-lm"], check=True)
    a = subprocess.run([mine, str(tf), str(mf)], capture_output=True, text=True, check=True)
    b = subprocess.run([ref, str(tf), str(mf)], capture_output=True, text=True, check=True)
    assert a.stdout == b.stdout
    assert a.stderr == b.stderr
    rcs = [int(x) for x in a.stdout.split("\n")[:len(mods)]]
    assert rcs == [0, 0, 0, 0, 0, 0, -1, -1]
    assert "[tripchange u2] dest" in a.stderr and "[tripchange a1] ALL" in a.stderr
    # the rules themselves, on the parsed demand
    base = {(int(o), int(d)): v for o, d, v in zip(np.repeat(np.arange(1, Z + 1), np.diff(dem.start)),
                                                   dem.dest, dem.demand)}
    got = {}
    for line in a.stdout.split("\n")[len(mods):]:
        if line:
            o, d, v = line.split()
            got[(int(o), int(d))] = float.fromhex(v)
    exp = dict(base)
    for _, kind, o, d, mlt in mods[:6]:
        for key in exp:
            if (kind == "P2P" and key == (o, d)) or (kind == "UNIFORM" and o and key[0] == o) or \
                    (kind == "UNIFORM" and d and key[1] == d) or kind == "ALL":
                exp[key] = exp[key] * mlt
    assert got == exp
