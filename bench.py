"""MD-step throughput on one MI355X: ns/day of the HPS slab (100 chains x 140
beads = 14,000 beads, periodic 15 x 15 x 150 nm, dt 10 fs, 300 K).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload hps|condensate|go|replicas|serve|copy]

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
# HBM bytes per list entry of the pair kernel, from the FETCH_SIZE / WRITE_SIZE
# counter runs summarized in profiles/ (calibrated against a copy kernel)
TRAFFIC_FILES = {"hps_slab_100x140": os.path.join(ROOT, "profiles", "k_pair_hps_slab_14000_traffic.json")}
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
    if name == "condensate":
        s, it, x = configs.hps_condensate()
        cfg = {"workload": "hps_condensate_1000x1000", "beads": 1000000, "chains": 1000,
               "box_nm": [60.0, 60.0, 600.0], "start": "serpentine lattice, 0.7 nm", "cutoff_ah_nm": 2.0,
               "cutoff_dh_nm": 3.5, "skin_nm": 0.3, "rebuild_interval": 20, "dt_ps": 0.01, "temperature_K": 300.0}
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


def list_options(args, base=None):
    """Neighbor-list options given on the command line, over `base`."""
    lopt = dict(base or {})
    if args.skin is not None:
        lopt["skin"] = args.skin
    if args.rebuild_interval is not None:
        lopt["rebuild_interval"] = args.rebuild_interval
    if args.prune_interval is not None:
        lopt["prune_interval"] = args.prune_interval
    return lopt


def replicas_bench(args):
    """256 Go chains (50-500 beads, 280-400 K ladder) packed in one batch vs the
    same replicas stepped one at a time through single-system contexts."""
    from opencafemol_amd import abi
    from opencafemol_amd.context import Batch, Context
    systems, integs, xs, cfg = build_workload("replicas")
    kB = 1.380649e-23 * 6.02214076e23 / 1000.0
    rng = np.random.default_rng(5)
    vs = [rng.normal(size=(s.n, 3)) * np.sqrt(kB * it.temperature / s.masses)[:, None] for s, it in zip(systems, integs)]
    lopt = list_options(args)
    b = Batch(systems, integs, use_graph=0 if args.no_graph else 1, **lopt)
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
    out = {"metric": "aggregate ns/day, 256 Go replicas (50-500 beads, 280-400 K)", "value": agg, "unit": "ns/day",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1000,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
           "data": "synthetic", "config": cfg}
    if args.no_serial:
        print(json.dumps(out))
        return
    serial_steps = max(200, args.steps // 10)
    tot = 0.0
    for k in range(len(systems)):
        c = Context(systems[k], integs[k], use_graph=0 if args.no_graph else 1, **lopt)
        c.set_positions(xs[k])
        c.set_velocities(vs[k])
        c.step(200)
        t = time.perf_counter()
        c.step(serial_steps)
        tot += time.perf_counter() - t
        c.close()
    serial = len(systems) * serial_steps / tot * dt_ps * 86400.0 / 1000.0
    log(f"serial: {len(systems)} replicas x {serial_steps} steps one at a time in {tot:.2f} s -> {serial:.0f} ns/day")
    out.update({"serial_plain_ns_day": serial, "batched_over_serial": agg / serial})
    print(json.dumps(out))


def _pct(v, q):
    """Nearest-rank percentile (as the server reports)."""
    v = sorted(v)
    if not v:
        return 0.0
    k = int(np.ceil(q * len(v)))
    return v[min(len(v) - 1, max(k, 1) - 1)]


def serve_bench(args):
    """Served job stream: 60% Go chains (50-500 beads), 40% HPS slabs
    (2,000-20,000 beads), one frame per 10,000 steps.  Three runs of the same
    jobs: (1) serial plain execution, one context per job, back to back;
    (2) the server with every job submitted at once (capacity); (3) the server
    fed Poisson arrivals at `--serve-load` times the serial service rate, with
    job latency (submit -> completion) against a FIFO of the serial run's
    measured job times on the same arrivals."""
    from opencafemol_amd import abi, configs
    from opencafemol_amd.context import Batch, Context, Server
    lo, hi = (int(v) for v in args.serve_steps.split(","))
    t = time.perf_counter()
    jobs = configs.served_job_mix(args.serve_jobs, seed=args.seed, steps=(lo, hi))
    hps = [j for j in jobs if j["kind"] == "hps"]
    log(f"{len(jobs)} jobs ({len(hps)} HPS, {len(jobs) - len(hps)} Go), "
        f"{sum(j['system'].n for j in jobs)} beads, built in {time.perf_counter() - t:.1f} s")
    if hps:   # pre-relax the packed slabs in one batch
        t = time.perf_counter()
        with Batch([j["system"] for j in hps], [j["integrator"] for j in hps]) as b:
            for k, j in enumerate(hps):
                b.set_positions(k, j["positions"])
            b.minimize(tolerance=10.0, max_iterations=args.minimize_iters)
            for k, j in enumerate(hps):
                j["positions"] = b.state(k, abi.STATE_POSITIONS)["positions"]
        log(f"relaxed {len(hps)} slabs in {time.perf_counter() - t:.1f} s")
    sim_ns = [j["nsteps"] * j["integrator"].dt * 1e-3 for j in jobs]
    opt = {"use_graph": 0 if args.no_graph else 1}

    def submit(srv, j):
        return srv.submit(j["system"], j["integrator"], j["positions"], None, nsteps=j["nsteps"],
                          frame_interval=j["frame_interval"])

    # warm-up: library load, first kernels
    with Server(**opt) as srv:
        srv.drain([submit(srv, dict(jobs[0], nsteps=jobs[0]["frame_interval"]))])

    # (1) serial plain execution
    serial_t, serial_x, n_frames = [], [], 0
    for j in jobs:
        t = time.perf_counter()
        with Context(j["system"], j["integrator"], **opt) as c:
            c.set_positions(j["positions"])
            c.set_velocities_to_temperature(j["integrator"].temperature, j["integrator"].seed)
            done = 0
            while done < j["nsteps"]:
                if c.step(j["frame_interval"]) != abi.OCM_OK:
                    raise SystemExit("serial run failed")
                done += j["frame_interval"]
                c.get_state(abi.STATE_POSITIONS)
                n_frames += 1
            serial_x.append(c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES))
        serial_t.append(time.perf_counter() - t)
    serial_total = sum(serial_t)
    serial_ns_day = sum(sim_ns) / serial_total * 86400.0
    log(f"serial: {serial_total:.2f} s -> {serial_ns_day:.0f} ns/day aggregate")

    # (2) capacity: everything submitted at once
    with Server(**opt) as srv:
        t = time.perf_counter()
        ids = [submit(srv, j) for j in jobs]
        res = srv.drain(ids)
        sat_s = time.perf_counter() - t
        sat_stats = srv.stats()
    identical = all(np.array_equal(res[i]["positions"], sx["positions"]) and
                    np.array_equal(res[i]["velocities"], sx["velocities"]) for i, sx in zip(ids, serial_x))
    sat_ns_day = sum(sim_ns) / sat_s * 86400.0
    log(f"served, all at once: {sat_s:.2f} s -> {sat_ns_day:.0f} ns/day aggregate "
        f"({sat_stats['repacks']} batches, identical to serial: {identical})")

    # (3) Poisson stream
    rng = np.random.default_rng(args.seed + 1)
    rate = args.serve_load * len(jobs) / serial_total       # jobs per second
    arrivals = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / rate, len(jobs) - 1))])
    order = rng.permutation(len(jobs))
    with Server(**opt) as srv:
        t0 = time.perf_counter()
        ids, pending, results, k = [], set(), {}, 0
        while k < len(jobs) or pending:
            now = time.perf_counter() - t0
            while k < len(jobs) and arrivals[k] <= now:
                jid = submit(srv, jobs[order[k]])
                ids.append(jid)
                pending.add(jid)
                k += 1
            wait_ms = 20 if k >= len(jobs) else max(0, min(20, int((arrivals[k] - now) * 1000)))
            if srv.next_frame(timeout_ms=wait_ms) is None:
                for jid in list(pending):
                    r = srv.wait(jid, timeout_ms=0)
                    if r is not None:
                        results[jid] = r
                        pending.discard(jid)
        stream_s = time.perf_counter() - t0
        st = srv.stats()
    lat = [(results[i]["end_ms"] - results[i]["submit_ms"]) / 1000.0 for i in ids]
    fifo_end, fifo_lat = 0.0, []
    for k2, jidx in enumerate(order):
        start = max(arrivals[k2], fifo_end)
        fifo_end = start + serial_t[jidx]
        fifo_lat.append(fifo_end - arrivals[k2])
    stream = {"arrival_rate_per_s": rate, "load_vs_serial": args.serve_load,
              "latency_p50_s": _pct(lat, 0.5), "latency_p99_s": _pct(lat, 0.99),
              "serial_fifo_latency_p50_s": _pct(fifo_lat, 0.5), "serial_fifo_latency_p99_s": _pct(fifo_lat, 0.99),
              "wait_p50_s": st["wait_p50_ms"] / 1000.0, "wait_p99_s": st["wait_p99_ms"] / 1000.0,
              "makespan_s": stream_s, "serial_fifo_makespan_s": fifo_end,
              "aggregate_ns_day": sum(sim_ns) / stream_s * 86400.0, "batches": st["repacks"]}
    log(f"stream: p50 {stream['latency_p50_s']:.2f} s, p99 {stream['latency_p99_s']:.2f} s "
        f"(serial FIFO p50 {stream['serial_fifo_latency_p50_s']:.2f} s, p99 {stream['serial_fifo_latency_p99_s']:.2f} s)")
    cfg = {"workload": "served_job_stream", "jobs": len(jobs), "hps_jobs": len(hps),
           "go_beads": [50, 500], "hps_beads": [2000, 20000], "steps_per_job": [lo, hi], "frame_interval": 10000,
           "frames": n_frames, "total_beads": int(sum(j["system"].n for j in jobs)), "dt_ps": 0.01,
           "temperature_K": 300.0}
    out = {"metric": "served job stream: aggregate ns/day (all jobs submitted at once) and p50/p99 job latency "
                     "(Poisson arrivals)", "value": sat_ns_day, "unit": "ns/day", "n_gpus": 1,
           "steps": int(sum(j["nsteps"] for j in jobs)), "warmup": 0, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
           "data": "synthetic", "config": cfg, "serial_plain_ns_day": serial_ns_day,
           "served_over_serial": sat_ns_day / serial_ns_day, "identical_to_serial": bool(identical),
           "stream": stream}
    print(json.dumps(out))


def copy_bench(args):
    """Plain device copy of a known size (256 MiB read + 256 MiB written per
    launch, torch's elementwise kernel): the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters of the pair kernel to bytes."""
    import torch
    n = 64 * 2**20
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    t = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    t.record()
    for _ in range(args.steps):
        torch.neg(src, out=dst)
    e.record()
    torch.cuda.synchronize()
    ms = t.elapsed_time(e) / args.steps
    gbs = 2 * 4 * n / (ms * 1e-3) / 1e9
    print(json.dumps({"workload": "copy_256MiB", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "ms_per_launch": ms, "GB/s": gbs}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--workload", default="hps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serial", action="store_true",
                    help="replicas workload: skip the one-at-a-time comparison (the batched timing is unchanged)")
    ap.add_argument("--minimize-iters", type=int, default=3000)
    ap.add_argument("--no-graph", action="store_true", help="eager launches (for profilers)")
    ap.add_argument("--skin", type=float, default=None, help="neighbor-list skin (nm); default: library's")
    ap.add_argument("--rebuild-interval", type=int, default=None, help="steps between list rebuilds")
    ap.add_argument("--prune-interval", type=int, default=None,
                    help="steps between inner-list prunes (<= 1: off); default: library's")
    ap.add_argument("--serve-jobs", type=int, default=24)
    ap.add_argument("--serve-steps", default="10000,50000", help="per-job step range (multiples of 10000)")
    ap.add_argument("--serve-load", type=float, default=0.8, help="arrival rate / serial service rate")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported; this build measures one "
              f"GPU's share (independent systems and jobs would shard across GPUs without communication)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "replicas":
        replicas_bench(args)
        return
    if args.workload == "serve":
        serve_bench(args)
        return
    if args.workload == "copy":
        copy_bench(args)
        return

    from opencafemol_amd import abi
    from opencafemol_amd.context import Batch

    systems, integs, xs, cfg = build_workload(args.workload)
    log(f"workload {cfg['workload']}: {sum(s.n for s in systems)} beads in {len(systems)} system(s)")
    # the condensate configuration states its list parameters
    lopt = list_options(args, {"skin": 0.3, "rebuild_interval": 20} if args.workload == "condensate" else None)
    b = Batch(systems, integs, use_graph=0 if args.no_graph else 1, **lopt)
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
    b.set_timing(2)   # pair-kernel duration stamps on the device clock; graphs kept (captured in warmup)
    code = b.step(args.warmup)
    if code != abi.OCM_OK:
        raise SystemExit(f"warmup failed with status {code}")
    log(f"warmup {args.warmup} steps done")

    rebuilds0 = b.list_stats()[2]
    ms0, n0 = b.timing(5)
    t0 = time.perf_counter()
    code = b.step(args.steps)
    t1 = time.perf_counter()
    ms1, n1 = b.timing(5)
    b.set_timing(0)
    rebuilds_timed = b.list_stats()[2] - rebuilds0
    stamp_ms, stamp_n = ms1 - ms0, n1 - n0
    if code != abi.OCM_OK:
        raise SystemExit(f"timed run failed with status {code}")
    el = t1 - t0
    ms_per_step = el / args.steps * 1000.0
    dt_ps = integs[0].dt
    sps = args.steps / el
    ns_day = sps * dt_ps * 86400.0 / 1000.0 * len(systems)
    log(f"timed {args.steps} steps in {el:.3f} s: {ms_per_step * 1000:.2f} us/step, {ns_day:.1f} ns/day")

    # pair kernel duration over the timed region: first block start to last
    # block end of every launch, on the device wall clock
    entries, members, rebuilds = b.list_stats()
    inner_entries, prune = b.inner_list_stats()
    avg_us = stamp_ms / max(stamp_n, 1) * 1000.0
    # entries one force-pass launch reads: the inner rows when the inner list
    # is on (a separate prune kernel rewrites them from the outer rows every
    # few steps), the outer rows otherwise
    steps_per_rebuild = args.steps / max(rebuilds_timed, 1)
    read = float(inner_entries if prune > 1 else entries)
    bytes_per_launch = 4 * read + 64 * members
    achieved = bytes_per_launch / (avg_us * 1e-6) / 1e9
    log(f"pair kernel {avg_us:.2f} us/launch over {stamp_n} launches of the timed region; list {entries} "
        f"entries (inner {inner_entries}, prune interval {prune}), {members} rows; {achieved:.0f} GB/s algorithmic")

    # breakdown phase: HIP events around every launch (eager; adds launch latency to each figure)
    kt_steps = min(args.steps, 4000)
    b.set_timing(1)
    code = b.step(kt_steps)
    b.set_timing(0)
    pair_ms, pair_n = b.timing(0)
    bond_ms, bond_n = b.timing(1)
    int_ms, int_n = b.timing(2)
    nl_ms, nl_n = b.timing(3)
    pk_ms, pk_n = b.timing(4)
    pr_ms, pr_n = b.timing(6)
    kernels = {"pair_us": avg_us, "pair_us_event_bracketed": pair_ms / max(pair_n, 1) * 1000,
               "bonded_us": bond_ms / max(bond_n, 1) * 1000, "integrator_us":
               int_ms / max(int_n, 1) * 1000, "list_build_us": nl_ms / max(nl_n, 1) * 1000,
               "pack_us": pk_ms / max(pk_n, 1) * 1000, "prune_us_event_bracketed": pr_ms / max(pr_n, 1) * 1000,
               "prunes_per_step": pr_n / max(kt_steps, 1), "list_entries": entries, "list_rows": members,
               "inner_list_entries": inner_entries, "prune_interval": prune,
               "entries_read_per_launch": read, "entries_per_s": read / (avg_us * 1e-6),
               "list_rebuilds_timed": rebuilds_timed, "steps_per_rebuild": steps_per_rebuild}

    traffic = None
    tf = TRAFFIC_FILES.get(cfg["workload"])
    if tf and os.path.exists(tf):
        traffic = json.load(open(tf))["hbm_bytes_per_entry"] * read

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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
