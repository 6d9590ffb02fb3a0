"""MD-step throughput on one MI355X: ns/day of the HPS slab (100 chains x 140
beads = 14,000 beads, periodic 15 x 15 x 150 nm, dt 10 fs, 300 K).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hps|go|replicas]

Protocol: build the seeded slab, relax it on the device (ocm_minimize), draw
Maxwell-Boltzmann velocities, run W untimed steps, then time exactly K steps
(one ocm_step call; it returns after the device finished).  A second phase
of the same length runs with HIP-event timing around every pair-kernel launch
(on the context's stream) to give the kernel's average duration, from which
the roofline figure is computed.  The CPU baseline is the cell-list FP64 CPU
step (oracle/liboracle.so, OpenMP) on a bounded number of steps of the same
relaxed configuration.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "ns/day (single system + 256-replica aggregate); p50/p99 job latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def build_workload(name):
    from opencafemol_amd import configs
    if name == "hps":
        s, it, x = configs.hps_slab()
        cfg = {"workload": "hps_slab_100x140", "beads": 14000, "chains": 100, "box_nm": [15.0, 15.0, 150.0],
               "slab_nm": 20.0, "cutoff_ah_nm": 2.0, "cutoff_dh_nm": 3.5, "dt_ps": 0.01, "temperature_K": 300.0}
        return [s], [it], [x], cfg
    if name == "go":
        s, it, x = configs.go_chain(62, seed=0)
        cfg = {"workload": "go_chain_62", "beads": 62, "dt_ps": 0.01, "temperature_K": 300.0}
        return [s], [it], [x], cfg
    if name == "replicas":
        reps = configs.go_replicas(256, seed=0)
        cfg = {"workload": "go_replicas_256", "beads": int(sum(r[0].n for r in reps)), "replicas": 256,
               "dt_ps": 0.01, "temperature_K": [280.0, 400.0]}
        return [r[0] for r in reps], [r[1] for r in reps], [r[2] for r in reps], cfg
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(system, integ, x, v, dt_ps, budget_s=10.0):
    from oracle import pyoracle
    threads = min(16, os.cpu_count() or 1)
    t = time.perf_counter()
    pyoracle.baoab_cells(system, integ, x, v, 0, 2, threads)
    per = (time.perf_counter() - t) / 2
    n = max(2, int(budget_s / max(per, 1e-6)))
    t = time.perf_counter()
    pyoracle.baoab_cells(system, integ, x, v, 0, n, threads)
    el = time.perf_counter() - t
    sps = n / el
    return {"value": sps * dt_ps * 86400.0 / 1000.0, "unit": "ns/day", "cores": threads, "kind": "port",
            "sample": f"{n} BAOAB steps of the same relaxed system, FP64 cell-list CPU step "
                      f"(oracle/oracle.cpp oracle_baoab_cells), OpenMP {threads} threads"}


def replicas_bench(args):
    """256 Go chains (50-500 beads, 280-400 K ladder) packed in one batch vs the
    same replicas stepped one at a time through single-system contexts."""
    from opencafemol_amd import abi
    from opencafemol_amd.context import Batch, Context
    systems, integs, xs, cfg = build_workload("replicas")
    kB = 1.380649e-23 * 6.02214076e23 / 1000.0
    rng = np.random.default_rng(5)
    vs = [rng.normal(size=(s.n, 3)) * np.sqrt(kB * it.temperature / s.masses)[:, None] for s, it in zip(systems, integs)]
    b = Batch(systems, integs, use_graph=0 if args.no_graph else 1)
    for k in range(len(systems)):
        b.set_positions(k, xs[k])
        b.set_velocities(k, vs[k])
    code = b.step(args.warmup)
    if code != abi.OCM_OK:
        raise SystemExit(f"warmup failed with status {code}")
    t0 = time.perf_counter()
    code = b.step(args.steps)
    el = time.perf_counter() - t0
    if code != abi.OCM_OK:
        raise SystemExit(f"timed run failed with status {code}")
    dt_ps = integs[0].dt
    agg = len(systems) * args.steps / el * dt_ps * 86400.0 / 1000.0
    log(f"batched: {args.steps} steps of {len(systems)} replicas in {el:.3f} s -> {agg:.0f} ns/day aggregate")
    b.close()
    serial_steps = max(200, args.steps // 10)
    tot = 0.0
    for k in range(len(systems)):
        c = Context(systems[k], integs[k], use_graph=0 if args.no_graph else 1)
        c.set_positions(xs[k])
        c.set_velocities(vs[k])
        c.step(200)
        t = time.perf_counter()
        c.step(serial_steps)
        tot += time.perf_counter() - t
        c.close()
    serial = len(systems) * serial_steps / tot * dt_ps * 86400.0 / 1000.0
    log(f"serial: {len(systems)} replicas x {serial_steps} steps one at a time in {tot:.2f} s -> {serial:.0f} ns/day")
    out = {"metric": "aggregate ns/day, 256 Go replicas (50-500 beads, 280-400 K)", "value": agg, "unit": "ns/day",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1000,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
           "data": "synthetic", "config": cfg, "serial_plain_ns_day": serial, "batched_over_serial": agg / serial}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--workload", default="hps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--minimize-iters", type=int, default=3000)
    ap.add_argument("--no-graph", action="store_true", help="eager launches (for profilers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (independent systems shard "
              f"across GPUs without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "replicas":
        replicas_bench(args)
        return

    from opencafemol_amd import abi
    from opencafemol_amd.context import Batch

    systems, integs, xs, cfg = build_workload(args.workload)
    log(f"workload {cfg['workload']}: {sum(s.n for s in systems)} beads in {len(systems)} system(s)")
    b = Batch(systems, integs, use_graph=0 if args.no_graph else 1)
    for k, x in enumerate(xs):
        b.set_positions(k, x)
    t = time.perf_counter()
    b.minimize(tolerance=10.0, max_iterations=args.minimize_iters)
    log(f"relaxed in {time.perf_counter() - t:.1f} s")
    if len(systems) == 1:
        abi.check(b.lib.ocm_set_velocities_to_temperature(b.h, integs[0].temperature, 2024))
    else:
        rng = np.random.default_rng(5)
        kB = 1.380649e-23 * 6.02214076e23 / 1000.0
        for k, (s, it) in enumerate(zip(systems, integs)):
            v = rng.normal(size=(s.n, 3)) * np.sqrt(kB * it.temperature / s.masses)[:, None]
            b.set_velocities(k, v)
    code = b.step(args.warmup)
    if code != abi.OCM_OK:
        raise SystemExit(f"warmup failed with status {code}")
    log(f"warmup {args.warmup} steps done")

    t0 = time.perf_counter()
    code = b.step(args.steps)
    t1 = time.perf_counter()
    if code != abi.OCM_OK:
        raise SystemExit(f"timed run failed with status {code}")
    el = t1 - t0
    ms_per_step = el / args.steps * 1000.0
    dt_ps = integs[0].dt
    sps = args.steps / el
    ns_day = sps * dt_ps * 86400.0 / 1000.0 * len(systems)
    log(f"timed {args.steps} steps in {el:.3f} s: {ms_per_step * 1000:.2f} us/step, {ns_day:.1f} ns/day")

    # kernel timing phase (HIP events around each pair launch on the context stream)
    kt_steps = min(args.steps, 4000)
    b.set_timing(True)
    code = b.step(kt_steps)
    b.set_timing(False)
    pair_ms, pair_n = b.timing(0)
    bond_ms, bond_n = b.timing(1)
    int_ms, int_n = b.timing(2)
    nl_ms, nl_n = b.timing(3)
    pk_ms, pk_n = b.timing(4)
    entries, members, rebuilds = b.list_stats()
    avg_us = pair_ms / max(pair_n, 1) * 1000.0
    bytes_per_launch = 4 * entries + 64 * members
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    log(f"pair kernel {avg_us:.2f} us/launch over {pair_n} launches; list {entries} entries, {members} rows; "
        f"{achieved:.0f} GB/s algorithmic")
    kernels = {"pair_us": avg_us, "bonded_us": bond_ms / max(bond_n, 1) * 1000, "integrator_us":
               int_ms / max(int_n, 1) * 1000, "list_build_us": nl_ms / max(nl_n, 1) * 1000,
               "pack_us": pk_ms / max(pk_n, 1) * 1000, "list_entries": entries, "list_rows": members,
               "entries_per_s": entries / (avg_us * 1e-6)}

    cpu = None
    if not args.no_cpu and len(systems) == 1:
        st = b.state(0, abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
        cpu = cpu_baseline(systems[0], integs[0], st["positions"], st["velocities"], dt_ps)
        log(f"cpu baseline {cpu['value']:.2f} ns/day on {cpu['cores']} threads")
    b.close()

    out = {
        "metric": METRIC,
        "value": ns_day,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": cfg,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
