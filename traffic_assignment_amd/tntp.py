"""Bar-Gera TNTP text files: net, trips, flows and network-modification files.

The layouts are the ones the reference reads and writes:
net   -- trunk/tap/parsetapfiles.c:62-146 (metadata tags, one link per line,
         init term capacity length fftime B power speed toll type)
trips -- trunk/tap/parsetapfiles.c:181-311 ("Origin o" then "d : v;" items)
flows -- written by print_flow, trunk/tap/tap_functions.c:431-445, read by
         trunk/tap/parsetapfiles.c:344-423
mods  -- trunk/tap/parsetapfiles.c:457-516 (tab separated, one row per link)

Floats are written with Python's shortest round-trip repr so that the C
library's strtod reproduces the generator's doubles bit for bit.
"""
from __future__ import annotations

import io
import re

import numpy as np

from .synth import Demand, NetMod, Network

_TAG = re.compile(r"<([^>]*)>\s*(.*)")


def _fmt(x: float) -> str:
    return repr(float(x))


def write_net(path_or_fp, net: Network) -> None:
    out = io.StringIO()
    out.write(f"<NUMBER OF ZONES> {net.num_zones}\n")
    out.write(f"<NUMBER OF NODES> {net.num_nodes}\n")
    out.write(f"<FIRST THRU NODE> {net.first_thru_node}\n")
    out.write(f"<NUMBER OF LINKS> {net.num_links}\n")
    out.write("<END OF METADATA>\n\n\n")
    out.write("~\tInit node\tTerm node\tCapacity\tLength\tFree Flow Time\tB\tPower\t"
              "Speed limit\tToll\tType\t;\n")
    for k in range(net.num_links):
        out.write(f"\t{net.init[k]}\t{net.term[k]}\t{_fmt(net.capacity[k])}\t"
                  f"{_fmt(net.length[k])}\t{_fmt(net.fftime[k])}\t{_fmt(net.B[k])}\t"
                  f"{_fmt(net.power[k])}\t{_fmt(net.speed[k])}\t{_fmt(net.toll[k])}\t"
                  f"{int(net.ltype[k])}\t;\n")
    _emit(path_or_fp, out.getvalue())


def write_trips(path_or_fp, dem: Demand) -> None:
    out = io.StringIO()
    out.write(f"<NUMBER OF ZONES> {dem.num_zones}\n")
    out.write(f"<TOTAL OD FLOW> {_fmt(dem.total())}\n")
    out.write("<END OF METADATA>\n\n\n")
    for o in range(1, dem.num_zones + 1):
        a, b = int(dem.start[o - 1]), int(dem.start[o])
        out.write(f"Origin {o}\n")
        items = [f"{int(dem.dest[i]):5d} : {_fmt(dem.demand[i])};" for i in range(a, b)]
        for j in range(0, len(items), 5):
            out.write("    " + "  ".join(items[j:j + 5]) + "\n")
        out.write("\n")
    _emit(path_or_fp, out.getvalue())


def write_flows(path_or_fp, net: Network, volume, cost, header: str = "") -> None:
    """Flow table in print_flow's layout (%.15f volumes and costs)."""
    out = io.StringIO()
    out.write(header)
    out.write(f"<NUMBER OF NODES>\t{net.num_nodes}\n")
    out.write(f"<NUMBER OF LINKS>\t{net.num_links}\n")
    out.write("<END OF METADATA>\n\n")
    out.write("~\tTail\tHead\t:\tVolume\tCost\t;\n")
    for k in range(net.num_links):
        out.write(f"\t{net.init[k]}\t{net.term[k]}\t:\t{volume[k]:.15f}\t{cost[k]:.15f};\n")
    _emit(path_or_fp, out.getvalue())


def write_mods(path_or_fp, mods) -> None:
    out = io.StringIO()
    out.write("~\tChangeId\tChangeType\tInitnode\tTermnode\tCapacity\tLength\tFreeFlowTime"
              "\tB\tPower\tSpeedlimit\tToll\tType\tProjectCost\n")
    for m in mods:
        L = m.link
        out.write(f"{m.change_id}\t{m.modtype}\t{L['init']}\t{L['term']}\t{_fmt(L['capacity'])}\t"
                  f"{_fmt(L['length'])}\t{_fmt(L['fftime'])}\t{_fmt(L['B'])}\t{_fmt(L['power'])}\t"
                  f"{_fmt(L['speed'])}\t{_fmt(L['toll'])}\t{int(L['ltype'])}\t"
                  f"{_fmt(m.project_cost)}\n")
    _emit(path_or_fp, out.getvalue())


def _emit(path_or_fp, text):
    if hasattr(path_or_fp, "write"):
        path_or_fp.write(text)
    else:
        with open(path_or_fp, "w") as f:
            f.write(text)


def _lines(path_or_fp):
    if hasattr(path_or_fp, "read"):
        return path_or_fp.read().splitlines()
    with open(path_or_fp) as f:
        return f.read().splitlines()


def read_net(path_or_fp) -> Network:
    meta = {}
    rows = []
    for line in _lines(path_or_fp):
        if not line.strip() or line.startswith("~"):
            continue
        if line.startswith("<"):
            m = _TAG.match(line)
            if m:
                meta[m.group(1).strip().upper()] = m.group(2).strip()
            continue
        tok = line.replace(";", " ").split()
        rows.append(tok[:10])
    E = int(meta["NUMBER OF LINKS"])
    if len(rows) != E:
        raise ValueError(f"num_links {E} but {len(rows)} links")
    a = np.array(rows, dtype=object)
    f = lambda j: a[:, j].astype(np.float64)
    net = Network(num_zones=int(meta["NUMBER OF ZONES"]), num_nodes=int(meta["NUMBER OF NODES"]),
                  first_thru_node=int(meta["FIRST THRU NODE"]),
                  init=a[:, 0].astype(np.int64), term=a[:, 1].astype(np.int64),
                  capacity=f(2), length=f(3), fftime=f(4), B=f(5), power=f(6), speed=f(7),
                  toll=f(8), ltype=a[:, 9].astype(np.int64).astype(np.int16))
    return net


def read_trips(path_or_fp) -> Demand:
    Z = None
    per = {}
    cur = None
    for line in _lines(path_or_fp):
        if not line.strip() or line.startswith("~"):
            continue
        if line.startswith("<"):
            m = _TAG.match(line)
            if m and m.group(1).strip().upper() == "NUMBER OF ZONES":
                Z = int(m.group(2))
            continue
        if line.startswith("Origin"):
            cur = int(line.split()[1])
            per[cur] = []
            continue
        for item in line.split(";"):
            if ":" not in item:
                continue
            d, v = item.split(":")
            per[cur].append((int(d), float(v)))
    start = np.zeros(Z + 1, dtype=np.int64)
    dest, val = [], []
    for o in range(1, Z + 1):
        for d, v in per.get(o, []):
            dest.append(d); val.append(v)
        start[o] = len(dest)
    return Demand(Z, start, np.asarray(dest, dtype=np.int64), np.asarray(val, dtype=np.float64))


def read_flows(path_or_fp):
    """(init, term, volume, cost) arrays plus the '~ key = value' header."""
    init, term, vol, cost = [], [], [], []
    header = {}
    for line in _lines(path_or_fp):
        if line.startswith("~"):
            for k, v in re.findall(r"(\w+) = ([-+0-9.eE]+)", line):
                header[k] = float(v)
            continue
        if not line.strip() or line.startswith("<"):
            continue
        t = line.replace(";", " ").replace(":", " ").split()
        init.append(int(t[0])); term.append(int(t[1]))
        vol.append(float(t[2])); cost.append(float(t[3]))
    return (np.asarray(init), np.asarray(term), np.asarray(vol), np.asarray(cost), header)


def read_mods(path_or_fp):
    mods = []
    for line in _lines(path_or_fp):
        if not line.strip() or line.startswith("~"):
            continue
        t = line.rstrip("\n").split("\t")
        rec = dict(init=int(t[2]), term=int(t[3]), capacity=float(t[4]), length=float(t[5]),
                   fftime=float(t[6]), B=float(t[7]), power=float(t[8]), speed=float(t[9]),
                   toll=float(t[10]), ltype=int(t[11]))
        mods.append(NetMod(t[0], t[1], rec, float(t[12])))
    return mods
