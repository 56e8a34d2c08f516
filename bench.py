"""Benchmark: TAP solves/sec to relative gap 1e-4 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batched]

Default workload (batched): 512 upgrade-subset variants of the metro network
(jittered 100x100 grid + 1,000 zones, 43,600 links; 32 candidate upgrades,
project-cost budget 40), all in one engine context.  One step = one
Frank-Wolfe iteration of the whole batch (every variant's all-or-nothing
loading in one launch sequence, then the link passes), inputs resident in
HBM.  W untimed steps, then K timed steps; then, timed on its own, one cold
solve of the whole batch to relative gap 1e-4 -- value = variant solves/sec
of that solve.  The same line carries the plain path (a sample of the
variants solved one at a time on the same GPU), the roofline of the
dominant kernel (HIP events around every launch of it in the K timed
steps), the reference's CPU path timed on a bounded sample on this host,
and the shortest-path GB/s of the labels-in-HBM kernel on the regional
network.

Other workloads: metro (one plain solve per step), regional, grid5,
gradient (equilibrium gradients at relgap 1e-6).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TAP solves/sec to rel-gap 1e-4 (batched vs plain, 1 GPU); SSSP HBM GB/s"
PEAK_HBM_GBS = 8000.0
TARGET = 1e-4
# stop rules of the plain workloads: (max iterations, target relgap); grid5 is
# the small CPU-runnable case, run for 200 iterations or to relgap 1e-5
STOP = {"grid5": (200, 1e-5)}

DESCR = {
    "grid5": "grid5: 24 nodes / 76 links / 24 zones",
    "metro": "metro: 11,000 nodes / 43,600 links / 1,000 zones, gravity demand 400k",
    "regional": "regional: 94,000 nodes / 386,800 links / 4,000 zones",
}


def sweep_bytes(st) -> float:
    """Algorithmic bytes of the shortest-path sweeps (worklist form): per
    processed (node, 64-origin group) item 8 B of frontier bookkeeping and
    its own label row (64 x 8 B); per scanned in-link 12 B of link data and
    the tail's label row (64 x 8 B); per lowered label 12 B (dist + pred);
    per marked out-link 4 B."""
    F, Es, U, Eo = st["sssp_items"], st["sssp_inlinks"], st["sssp_updates"], st["sssp_outlinks"]
    return 8.0 * F + 12.0 * Es + 8.0 * 64 * (F + Es) + 12.0 * U + 4.0 * Eo


def origin_bytes(st, num_links) -> float:
    """Algorithmic bytes of the one-origin-per-workgroup kernel, summed over
    its (origin, variant) workgroups: the same worklist formula with one
    origin per row (8 B per expansion + its label, 12 B of link data and a
    label per relaxed link, 12 B per lowered label -- for the labels-in-HBM
    kernel, per atomic min issued), plus predecessor
    selection (per link: 16 B of reverse-CSR data and the tail's label) and
    loading (per walk step 2 B of tree data and 8 B of flow
    read-modify-write; per tree link with flow 8 B of flow, 2 B of tree data
    and a 16 B link-volume read-modify-write)."""
    F, Es, U = st["origin_expansions"], st["origin_scanned"], st["origin_updates"]
    relax = 8.0 * F + 12.0 * Es + 8.0 * (F + Es) + 12.0 * U
    pred = 24.0 * num_links * st["origin_blocks"]
    load = 10.0 * st["origin_walk_steps"] + 26.0 * st["origin_loaded"]
    return relax + pred + load


def host_cpu():
    """(cores this process may use, CPU model from lscpu / /proc/cpuinfo)."""
    cores = len(os.sched_getaffinity(0))
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if not model:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except Exception:
            pass
    return cores, model


def cpu_baseline(net, dem, gpu_iters: int, cores: int, budget_s: float = 20.0):
    """The reference's tap_frankwolfe_pthread (compiled from its sources into
    oracle/_ref) on this host with `cores` threads: time lo = 2 iterations
    and hi = lo + k, k chosen for about budget_s of CPU work (2 <= k <= 20);
    per-iteration time x the GPU's iteration count + setup = one solve
    (extrapolated).  Without the reference build, the oracle's restatement
    (bit-identical to it on the golden vectors) is timed instead."""
    from traffic_assignment_amd import tntp
    ref = os.path.join(ROOT, "oracle", "_ref", "tap_frankwolfe_pthread")
    lo = 2
    with tempfile.TemporaryDirectory() as td:
        nf, tf = os.path.join(td, "net.txt"), os.path.join(td, "trips.txt")
        tntp.write_net(nf, net)
        tntp.write_trips(tf, dem)
        if os.path.exists(ref):
            kind = "reference"

            def run(it):
                t = time.perf_counter()
                subprocess.run([ref, "-n", str(cores), "-i", str(it), nf, tf],
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True)
                return time.perf_counter() - t
        else:
            kind = "port"
            import oracle

            def run(it):
                t = time.perf_counter()
                oracle.fw_solve(net, dem, max_iters=it, nthreads=cores)
                return time.perf_counter() - t
        t_lo = run(lo)
        hi = lo + int(min(20, max(2, budget_s / max(t_lo / lo, 1e-6))))
        t_hi = run(hi)
    per_iter = (t_hi - t_lo) / (hi - lo)
    setup = max(t_lo - lo * per_iter, 0.0)
    solve = setup + per_iter * gpu_iters
    return {"value": 1.0 / solve, "unit": "solves/s", "cores": cores, "kind": kind,
            "sample": (f"{DESCR.get(getattr(net, 'params', {}).get('workload', ''), 'workload')}: "
                       f"{lo} and {hi} FW iterations timed on {cores} threads, "
                       f"{per_iter * 1e3:.1f} ms/iteration, extrapolated to the GPU's "
                       f"{gpu_iters} iterations + {setup:.2f} s setup"),
            "seconds_per_iteration": per_iter, "measured_seconds": t_lo + t_hi}


def cpu_baselines(net, dem, gpu_iters: int):
    """All host cores (the reported baseline) and one thread beside it."""
    cores, model = host_cpu()
    allc = cpu_baseline(net, dem, gpu_iters, cores, 20.0)
    serial = cpu_baseline(net, dem, gpu_iters, 1, 10.0)
    allc["cpu_model"] = model
    allc["serial"] = {k: serial[k] for k in ("value", "seconds_per_iteration", "sample")}
    return allc


def load_traffic(workload, kernel):
    """Calibrated HBM bytes per launch of `kernel` on `workload` from the
    rocprof counter runs (profiles/traffic.json), or None when that kernel
    was not measured there."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
    except Exception:
        return None
    for key in (workload, f"{workload}:{kernel}"):
        e = d.get(key, {})
        if e.get("kernel") == kernel:
            return e.get("hbm_bytes_per_launch")
    return None


def lanes_bytes(st, num_nodes, in_degree) -> float:
    """Algorithmic bytes of the lanes kernel (one workgroup per origin and
    group of 64 variants), from its counters.  Relaxation, per wavefront
    expansion (a node in all 64 lanes): its label row 512 B; per out-link
    scanned: 4 B head, the cost row 512 B and the head's label row 512 B; per
    lane atomic min (a candidate below the label read) 8 B.  Label rows set at the start of each task: 512 B per
    node.  Loading, per lane and walk step: 8 B memo, 4 B tail, 16 B volume
    read-modify-write; per predecessor selection (lane): 8 B own label and
    8 B memo, and per in-link 8 B reverse-CSR data, 8 B cost, 8 B tail label
    (in-links at the network's mean in-degree)."""
    F, Es, U = st["origin_expansions"], st["origin_scanned"], st["origin_updates"]
    relax = 512.0 * F + (4.0 + 1024.0) * Es + 8.0 * U
    init = 512.0 * num_nodes * st["lanes_tasks"]
    load = 28.0 * st["origin_loaded"] + st["lanes_pred_scans"] * (16.0 + 24.0 * in_degree)
    return relax + init + load


def loading_roofline(acc, path, num_nodes, num_links, workload=None):
    """Roofline of the loading kernel from accumulated stats (HIP events
    around each of its launches); traffic from the counter runs of that
    kernel on `workload`."""
    launches = acc.get("origin_launches", 0)
    avg_ms = acc.get("origin_kernel_ms", 0.0) / max(launches, 1)
    if acc.get("lanes_launches", 0) and acc.get("lanes_launches") == launches:
        kernel = "k_origin_lanes"
        total = lanes_bytes(acc, num_nodes, num_links / max(num_nodes, 1))
    else:
        kernel = "k_origin_hbm" if path == "hbm" else "k_origin_lds"
        total = origin_bytes(acc, num_links)
    bpl = total / max(launches, 1)
    achieved = bpl / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
            "traffic": load_traffic(workload, kernel) if workload else None,
            "avg_launch_us": round(avg_ms * 1e3, 3),
            "launches": int(launches), "algorithmic_bytes_per_launch": round(bpl, 1)}


def _acc(acc, st):
    for k, v in st.items():
        acc[k] = acc.get(k, 0) + v


def solve_batch(eng, target, tag, chunk=25, warm=None):
    """fw_init (from warm volumes when given), then chunks of FW iterations
    until every variant has converged (a progress line per chunk); the
    per-variant results."""
    eng.fw_init(warm)
    done = 0
    while True:
        done += chunk
        res = eng.fw_run(max_iters=done, target_relgap=target)
        running = sum(1 for r in res if not r["converged"] and r["iterations"] >= done)
        print(f"{tag}: {done} iterations, {running} variants running", file=sys.stderr, flush=True)
        if running == 0:
            return res


def regional_sssp(args):
    """Shortest-path GB/s of the labels-in-HBM kernel on the regional network
    (the SSSP roofline case): a few FW iterations with HIP events around
    every launch of the loading kernel."""
    import torch
    from traffic_assignment_amd import synth
    from traffic_assignment_amd.engine import TapEngine
    net, dem = synth.make("regional")
    with TapEngine(net, dem) as eng:
        eng.fw_init()                       # bucket-width tuning happens here
        eng.set_profiling(True)
        torch.cuda.synchronize()
        eng.fw_run(max_iters=1 + args.regional_iters)
        st = eng.stats()
        path = eng.strategy
    roof = loading_roofline(st, path, net.num_nodes, net.num_links, "regional")
    roof["workload"] = DESCR["regional"]
    roof["iterations_timed"] = args.regional_iters
    roof["delta_mean_t0"] = st.get("origin_delta", 0.0)
    return roof


def eng_can_block(eng, block):
    """A fixed LDS workgroup size applies to the LDS path only."""
    return block > 0 and eng.strategy == "lds"


def warm_started(args, eng, base, dem, mods, subsets, u, present):
    """The batch again, every variant warm-started from the base network's
    equilibrium (a feasible flow of every variant: variants only change link
    records or add links, which start empty), against the same plain solves
    warm-started the same way.  The use the reference's warm start (-w
    flows file) serves: fewer FW iterations per upgrade variant."""
    import torch
    from traffic_assignment_amd import variants as V
    from traffic_assignment_amd.engine import TapEngine

    t = time.perf_counter()
    with TapEngine(base, dem) as be:
        be.fw_init()
        be.fw_run(target_relgap=TARGET)
        xb, _ = be.flows()
    base_s = time.perf_counter() - t
    index = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(u.init, u.term))}
    xu = np.zeros(u.num_links)
    xu[[index[(int(a), int(b))] for a, b in zip(base.init, base.term)]] = xb
    Vn = len(subsets)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    res = solve_batch(eng, TARGET, "warm batch solve", warm=np.tile(xu, (Vn, 1)))
    torch.cuda.synchronize()
    batch_s = time.perf_counter() - t1
    assert all(r["converged"] for r in res), "a warm-started variant did not converge"
    iters = [r["iterations"] for r in res]
    sample = list(range(0, Vn, max(1, Vn // args.plain_sample)))[:args.plain_sample]
    plain_t = 0.0
    for v in sample:
        net = V.variant_network(base, mods, subsets[v])
        xv = xu[present[v].astype(bool)]
        with TapEngine(net, dem) as pe:
            pe.fw_init()            # untimed: schedule tuning and first-launch costs
            pe.fw_run(max_iters=1)
            torch.cuda.synchronize()
            t = time.perf_counter()
            pe.fw_init(xv)
            r = pe.fw_run(target_relgap=TARGET)[0]
            torch.cuda.synchronize()
            plain_t += time.perf_counter() - t
            assert r["iterations"] == iters[v], "warm plain and batched iteration counts differ"
    plain = len(sample) / plain_t
    batched = Vn / batch_s
    print(f"warm start: batched {batched:.3f}, plain {plain:.3f} solves/s", file=sys.stderr, flush=True)
    return {"batched_solves_per_s": round(batched, 5), "plain_solves_per_s": round(plain, 5),
            "batched_vs_plain": round(batched / plain, 3), "batch_solve_s": round(batch_s, 3),
            "base_solve_s": round(base_s, 3), "plain_sample": len(sample),
            "iterations_min_median_max": [min(iters), int(np.median(iters)), max(iters)]}


def bench_batched(args):
    """Batched form: --variants upgrade subsets of the metro network in one
    engine context.  K timed steps = K Frank-Wolfe iterations of the whole
    batch; then one timed cold solve of the batch to relgap 1e-4 (value);
    then a sample of the variants solved one at a time (plain)."""
    import torch
    from traffic_assignment_amd import variants as V
    from traffic_assignment_amd.engine import TapEngine

    base, dem, mods, subsets = V.workload(n_variants=args.variants)
    eng, u, present = V.make_engine(base, dem, mods, subsets)
    if args.strategy != "auto":
        eng.set_strategy(args.strategy)
    if args.lanes is not None:
        eng.set_lanes(args.lanes)
    eng.set_lanes_delta(args.lanes_delta)
    # warmup: the initial loading and W iterations (schedule tuning included)
    eng.fw_init()
    eng.fw_run(max_iters=args.warmup, target_relgap=TARGET)
    print(f"warmup: {args.warmup} iterations", file=sys.stderr, flush=True)
    eng.set_profiling(True)
    acc = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for step in range(args.steps):
        eng.fw_run(max_iters=args.warmup + step + 1, target_relgap=TARGET)
        _acc(acc, eng.stats())
        if step % 5 == 4:
            print(f"step {step + 1}/{args.steps}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    eng.set_profiling(False)
    path = eng.strategy
    lanes_min = eng.lanes
    roof = loading_roofline(acc, path, u.num_nodes, u.num_links, "batched")
    if args.no_solve:           # profiling runs: the K steps only
        eng.close()
        print(json.dumps({"metric": METRIC, "value": None, "unit": "solves/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": round(1e3 * elapsed / args.steps, 3),
                          "profiling_run": True, "roofline": roof}))
        return
    # one cold solve of the whole batch to the target
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    res = solve_batch(eng, TARGET, "batch solve")
    torch.cuda.synchronize()
    batch_s = time.perf_counter() - t1
    iters = [r["iterations"] for r in res]
    assert all(r["converged"] for r in res), "a variant did not converge"
    batched = len(subsets) / batch_s
    warm = None if args.no_warm else warm_started(args, eng, base, dem, mods, subsets, u, present)
    eng.close()
    # plain: single-variant engines, one after another (same GPU, default path)
    sample = list(range(0, len(subsets), max(1, len(subsets) // args.plain_sample)))[:args.plain_sample]
    plain_t, plain_it = 0.0, []
    for v in sample:
        net = V.variant_network(base, mods, subsets[v])
        with TapEngine(net, dem) as pe:
            pe.fw_init()            # untimed: schedule tuning and first-launch costs
            pe.fw_run(max_iters=1)
            torch.cuda.synchronize()
            t = time.perf_counter()
            pe.fw_init()
            r = pe.fw_run(target_relgap=TARGET)[0]
            torch.cuda.synchronize()
            plain_t += time.perf_counter() - t
            plain_it.append(r["iterations"])
            assert r["iterations"] == iters[v], "plain and batched iteration counts differ"
    plain = len(sample) / plain_t
    print(f"plain: {len(sample)} variants, {plain:.3f} solves/s", file=sys.stderr, flush=True)
    sssp = None if args.no_regional else regional_sssp(args)
    cpu = None
    if not args.no_cpu:
        net0 = V.variant_network(base, mods, subsets[0])
        c1 = cpu_baselines(net0, dem, int(iters[0]))
        cpu = dict(c1, sample="variant 0 of the batch: " + c1["sample"] + "; per-variant rate "
                   "(the variants are independent solves)")
    out = {
        "metric": METRIC, "value": round(batched, 5), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, traffic_assignment_amd/synth.py)",
        "config": {"workload": f"batched: {len(subsets)} upgrade subsets of metro (11,000 nodes / "
                               "43,600 links / 1,000 zones), 32 candidate upgrades, budget 40",
                   "step": f"one Frank-Wolfe iteration of all {len(subsets)} variants",
                   "target_relgap": TARGET,
                   "engine_path": f"lanes kernel while >= {lanes_min} variants run, then {path}",
                   "lanes_min_active": lanes_min,
                   "iterations_min_median_max": [min(iters), int(np.median(iters)), max(iters)],
                   "parallelism": "1 GPU"},
        "batch_solve_s": round(batch_s, 3),
        "plain_solves_per_s": round(plain, 5), "plain_sample": len(sample),
        "batched_vs_plain": round(batched / plain, 3),
        "phase_ms_per_step": {k: round(acc.get(k, 0.0) / args.steps, 3) for k in
                              ("aon_ms", "link_ms", "total_ms")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "sssp_regional": sssp,
        "warm_start": warm,
    }
    if sssp:
        out["sssp_gbps"] = sssp["achieved"]
    if cpu:
        out["gpu_vs_cpu"] = round(batched / cpu["value"], 2)
    print(json.dumps(out))


def bench_gradient(args):
    """Gradient form: one cold FW solve of --grad-network to relgap 1e-6,
    then --steps timed gradient calls (envelope d Beckmann/d(c, t0) and
    implicit d VHT/d(c, t0) for every link), against warm re-solves after a
    1e-4 capacity change -- what one central finite difference per link
    costs."""
    import torch
    from traffic_assignment_amd import synth
    from traffic_assignment_amd.engine import TapEngine

    target = 1e-6
    net, dem = synth.make(args.grad_network)

    def solve(eng):
        done, r = 0, None
        while True:         # chunks with a progress line each
            done += 200
            r = eng.fw_run(max_iters=min(done, 100000), target_relgap=target)[0]
            print(f"{r['iterations']} iterations, relgap {r['relgap']:.3e}", file=sys.stderr,
                  flush=True)
            if r["converged"] or done >= 100000:
                return r

    eng = TapEngine(net, dem)
    torch.cuda.synchronize()
    t = time.perf_counter()
    eng.fw_init()
    r = solve(eng)
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - t
    x, _ = eng.flows()
    eps = max(1e-9, 20 * r["relgap"])
    for _ in range(args.warmup):
        eng.gradient(eps=eps, cg_tol=1e-10)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        g = eng.gradient(eps=eps, cg_tol=1e-10)
    torch.cuda.synchronize()
    grad_s = (time.perf_counter() - t) / args.steps
    eng.close()
    cap = np.array(net.capacity, dtype=np.float64)
    k = int(np.argmax(x / cap))
    warm = []
    for sgn in (1.0, -1.0):
        cp = cap.copy()
        cp[k] *= 1.0 + sgn * 1e-4
        with TapEngine(net, dem, num_variants=1, capacity=cp[None, :]) as e2:
            e2.fw_init(x)
            torch.cuda.synchronize()
            t = time.perf_counter()
            solve(e2)
            torch.cuda.synchronize()
            warm.append(time.perf_counter() - t)
    resolve = float(np.mean(warm))
    info = g["info"]
    out = {
        "metric": "equilibrium gradients/sec (d Beckmann, d VHT w.r.t. capacity and t0, all links)",
        "value": round(1.0 / grad_s, 4), "unit": "gradients/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * grad_s, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, traffic_assignment_amd/synth.py)",
        "config": {"workload": "gradient form on " + DESCR[args.grad_network] + " at relgap 1e-6",
                   "fw_iterations": r["iterations"], "relgap": r["relgap"], "eps": eps},
        "solve_to_1e-6_s": round(solve_s, 3),
        "warm_resolve_s": round(resolve, 4),
        "fd_64x2_warm_resolves_s": round(128 * resolve, 2),
        "fd_64x2_cold_resolves_s": round(128 * solve_s, 1),
        "gradient_vs_64x2_warm_resolves": round(128 * resolve / grad_s, 1),
        "gradient_info": info,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="batched",
                    choices=["grid5", "metro", "regional", "batched", "gradient"])
    ap.add_argument("--variants", type=int, default=512, help="batched: number of variants")
    ap.add_argument("--grad-network", default="metro", choices=["grid5", "metro", "regional"],
                    help="gradient: network solved to relgap 1e-6")
    ap.add_argument("--plain-sample", type=int, default=16,
                    help="batched: variants solved one by one for the plain comparison")
    ap.add_argument("--lanes", type=int, default=None,
                    help="batched: lanes kernel from this many active variants (0: off)")
    ap.add_argument("--lanes-delta", type=float, default=-1.0,
                    help="batched: lanes kernel bucket width in mean free-flow times (-1: tuned)")
    ap.add_argument("--no-solve", action="store_true",
                    help="batched: time the K steps only (profiling runs)")
    ap.add_argument("--no-warm", action="store_true",
                    help="batched: skip the warm-started batch and plain solves")
    ap.add_argument("--no-regional", action="store_true",
                    help="batched: skip the regional shortest-path measurement")
    ap.add_argument("--regional-iters", type=int, default=2,
                    help="batched: timed FW iterations of the regional network")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--strategy", default="auto", choices=["auto", "sweeps", "lds", "hbm"])
    ap.add_argument("--chase", type=int, default=-1,
                    help="LDS path: chained expansions per bitmap-scan round (-1: tuned)")
    ap.add_argument("--block", type=int, default=-1,
                    help="LDS path: threads per workgroup, 1024/512/256 (-1: tuned with the queue size)")
    ap.add_argument("--delta", type=float, default=-1.0,
                    help="HBM path: bucket width in mean free-flow times (-1: tuned)")
    ap.add_argument("--max-iters", type=int, default=0,
                    help="cap FW iterations per step (profiling runs; 0 = solve to 1e-4)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU "
              "sharding of this path (by origin or by variant) is not implemented", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.set_device(0)
    from traffic_assignment_amd import synth
    from traffic_assignment_amd.engine import TapEngine

    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "gradient":
        return bench_gradient(args)
    net, dem = synth.make(args.workload)
    eng = TapEngine(net, dem)          # graph, link data and demand -> HBM
    if args.strategy != "auto":
        eng.set_strategy(args.strategy)
    eng.set_origin_chase(args.chase)
    eng.set_origin_delta(args.delta)
    if eng_can_block(eng, args.block):
        eng.set_origin_block(args.block)
    max_iters, target = STOP.get(args.workload, (args.max_iters, TARGET))
    if args.max_iters:
        max_iters = args.max_iters
    for _ in range(args.warmup):
        eng.fw_init()
        eng.fw_run(max_iters=max_iters, target_relgap=target)
    # timed steps: complete solves as a caller runs them (iterations replayed
    # from captured graphs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        eng.fw_init()
        r = eng.fw_run(max_iters=max_iters, target_relgap=target)[0]
        iters.append(r["iterations"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # one more solve with profiling (launch by launch, HIP events around
    # every loading launch): the kernel's time and counters
    eng.set_profiling(True)
    eng.fw_init()
    t1 = time.perf_counter()
    eng.fw_run(max_iters=max_iters, target_relgap=target)
    torch.cuda.synchronize()
    profiled_s = time.perf_counter() - t1
    acc = dict(eng.stats())
    eng.set_profiling(False)
    eng_path = eng.strategy
    queue = eng.origin_queue
    eng.close()

    value = args.steps / elapsed
    if acc.get("origin_launches", 0) > 0:
        kernel = "k_origin_hbm" if eng_path == "hbm" else "k_origin_lds"
        launches = acc["origin_launches"]
        avg_ms = acc["origin_kernel_ms"] / launches
        bytes_per_launch = origin_bytes(acc, net.num_links) / launches
    else:
        kernel = "k_sssp_sweep"
        launches = acc["sweep_launches"]
        avg_ms = acc["sweep_kernel_ms"] / max(launches, 1)
        bytes_per_launch = sweep_bytes(acc) / max(launches, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    roof = {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 2),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
            "traffic": load_traffic(args.workload, kernel),
            "avg_launch_us": round(avg_ms * 1e3, 3), "launches": int(launches),
            "algorithmic_bytes_per_launch": round(bytes_per_launch, 1)}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines(net, dem, int(np.median(iters)))
    out = {
        "metric": METRIC, "value": round(value, 5), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, traffic_assignment_amd/synth.py)",
        "config": {"workload": DESCR[args.workload], "target_relgap": target, "max_iters": max_iters,
                   "engine_path": eng_path,
                   "iterations": iters, "nodes": net.num_nodes, "links": net.num_links,
                   "zones": net.num_zones, "parallelism": "1 GPU"},
        "sssp_gbps": round(achieved, 2),
        "schedule": {"origin_chase": acc.get("origin_chase", 0), "origin_delta": acc.get("origin_delta", 0.0),
                     "origin_block": acc.get("origin_block", 0), "origin_queue": queue},
        "origin_phase_us_per_workgroup": {
            k: round(1e3 * acc.get("origin_" + k + "_ms", 0.0) / max(acc.get("origin_blocks", 0), 1), 2)
            for k in ("relax", "pred", "load")},
        "profiled_solve": {"s": round(profiled_s, 4),
                           "phase_ms": {k: round(acc[k], 3) for k in ("aon_ms", "link_ms", "total_ms")}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_vs_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
