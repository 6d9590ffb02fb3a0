"""MD-step throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|hps|condensate|go|replicas|serve|copy]

The default (`all`) prints one JSON line whose `value` is the ns/day of the
HPS slab (100 chains x 140 beads = 14,000 beads, periodic 15 x 15 x 150 nm,
dt 10 fs, 300 K), with more keys for the metric's other parts:
  - `go_chain_62`: ns/day of the 62-bead Go chain alone (20,000 timed steps;
    GPU only);
  - `condensate_1m`: ns/day of the 1,000,000-bead condensate (1,000 timed
    steps, 0.3 nm skin, rebuilds every 20 steps at most; GPU only), with the
    force pass's roofline figures;
  - `replicas_256`: aggregate ns/day of 256 Go replicas (50-500 beads,
    280-400 K ladder) packed in one batch, against the same replicas stepped
    one at a time through single-system contexts;
  - `served_stream`: a served job stream (60% Go chains of 50-500 beads, 40%
    HPS slabs of 2,000-20,000 beads, 1e5-1e6 steps per job, a frame every
    1e4 steps): aggregate ns/day with every job submitted at once, and p50/p99
    job latency under Poisson arrivals, against serial plain execution.

Slab protocol: build the seeded slab, relax it on the device (ocm_minimize),
draw Maxwell-Boltzmann velocities, equilibrate (2,000 untimed steps, stated
in the line's config), run W untimed steps, capture the step
graphs the timed call will launch (ocm_prepare_steps, no stepping), then time
exactly K steps (one ocm_step call; it returns after the device finished).
The pair kernel times its own launches on the device clock inside the graphs;
a second phase with HIP events around every launch gives the other kernels.
The CPU baseline is the FP64 Verlet-list CPU step (oracle/fast_step.cpp in
oracle/liboracle.so: SIMD pair rows, OpenMP) on a bounded number of steps of
the same relaxed configuration, on all usable host threads and on one.
Batches whose every system has at most 1,024 beads (the Go chain, the
replicas) step resident: one workgroup per system runs whole chunks.
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
# Counter summaries of the force pass (profiles/): HBM bytes per inner-list
# entry (FETCH_SIZE / WRITE_SIZE, calibrated against a copy kernel) and VALU
# wave-instructions per entry (SQ_INSTS_VALU), on the slab and on the
# condensate.
TRAFFIC_FILES = {"hps_slab_100x140": os.path.join(ROOT, "profiles", "k_pair_hps_slab_14000_traffic.json"),
                 "hps_condensate_1000x1000": os.path.join(ROOT, "profiles",
                                                          "k_pair_hps_condensate_1000000_traffic.json")}
VALU_FILES = {"hps_slab_100x140": os.path.join(ROOT, "profiles", "k_pair_hps_slab_14000_valu.json"),
              "hps_condensate_1000x1000": os.path.join(ROOT, "profiles", "k_pair_hps_condensate_1000000_valu.json")}
HBM_PEAK_GBS = 8000.0
# VALU issue ceiling: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per
# SIMD every 4 cycles (16 lanes wide); x 64 lanes x 2 flops per FMA this is
# the published 78.6 TFLOP/s of non-packed vector math.
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4   # G wave-instructions/s (614.4)
VALU_CLOCK_HZ = 2.4e9
VALU_SIMDS = 256 * 4
KB = 1.380649e-23 * 6.02214076e23 / 1000.0   # kJ/mol/K


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def ns_per_day(steps_per_s, dt_ps):
    return steps_per_s * dt_ps * 86400.0 / 1000.0


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


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


CPU_SKIN_NM = 0.3


def cpu_rate(system, integ, x, v, threads, budget_s):
    """Steps/s of the FP64 Verlet-list CPU step on `threads` threads over
    about budget_s seconds (list set-up and builds inside the timed call),
    the number of steps timed and the list builds among them."""
    from oracle import pyoracle
    t = time.perf_counter()
    pyoracle.baoab_fast(system, integ, x, v, 0, 2, threads, CPU_SKIN_NM)
    per = (time.perf_counter() - t) / 2
    n = max(2, int(budget_s / max(per, 1e-6)))
    t = time.perf_counter()
    _, _, builds = pyoracle.baoab_fast(system, integ, x, v, 0, n, threads, CPU_SKIN_NM)
    return n / (time.perf_counter() - t), n, builds


def cpu_baseline(system, integ, x, v, dt_ps, budget_s=10.0):
    # the GPU box gives one GPU a share of 16 host threads (OMP_NUM_THREADS
    # there, whatever os.cpu_count() shows); the step is timed on all of them
    # and on 1, and the faster is the baseline (a small system runs faster
    # on one thread than across sixteen)
    model, nproc, usable = host_cpu()
    threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16), usable))
    sps, n, nb = cpu_rate(system, integ, x, v, threads, budget_s)
    sps1, n1, nb1 = cpu_rate(system, integ, x, v, 1, budget_s)
    best_t, best_sps, best_n, best_nb = (threads, sps, n, nb) if sps >= sps1 else (1, sps1, n1, nb1)
    return {"value": ns_per_day(best_sps, dt_ps), "unit": "ns/day", "cores": best_t, "kind": "port",
            "sample": f"{best_n} BAOAB steps of the same relaxed system ({best_nb} list builds among them), FP64 "
                      f"Verlet-list CPU step (oracle/fast_step.cpp oracle_baoab_fast: skin {CPU_SKIN_NM} nm, "
                      f"AVX2 SIMD rows, OpenMP {best_t} threads); the faster of {threads} threads and 1",
            f"value_{threads}_threads": ns_per_day(sps, dt_ps), "value_1_core": ns_per_day(sps1, dt_ps),
            "sample_1_core": f"{n1} steps on 1 thread ({nb1} list builds)",
            "host_cpu": model, "nproc": nproc, "usable_cpus": usable}


def list_options(args, base=None):
    """Neighbor-list options given on the command line, over `base`."""
    lopt = dict(base or {})
    if args.skin is not None:
        lopt["skin"] = args.skin
    if args.rebuild_interval is not None:
        lopt["rebuild_interval"] = args.rebuild_interval
    if args.prune_interval is not None:
        lopt["prune_interval"] = args.prune_interval
    if args.resident is not None:
        lopt["resident"] = args.resident
    return lopt


def mb_velocities(systems, integs, seed=5):
    rng = np.random.default_rng(seed)
    return [rng.normal(size=(s.n, 3)) * np.sqrt(KB * it.temperature / s.masses)[:, None]
            for s, it in zip(systems, integs)]


def replicas_run(args, steps, warmup, serial_steps):
    """256 Go chains (50-500 beads, 280-400 K ladder) packed in one batch vs the
    same replicas stepped one at a time through single-system contexts
    (serial_steps timed steps each, after 200 untimed ones; 0: skip)."""
    from opencafemol_amd import abi
    from opencafemol_amd.context import Batch, Context
    systems, integs, xs, cfg = build_workload("replicas")
    vs = mb_velocities(systems, integs)
    lopt = list_options(args)
    graph = 0 if args.no_graph else 1
    with Batch(systems, integs, use_graph=graph, **lopt) as b:
        for k in range(len(systems)):
            b.set_positions(k, xs[k])
            b.set_velocities(k, vs[k])
        if b.step(warmup) != abi.OCM_OK:
            raise SystemExit("replica batch warmup failed")
        b.prepare(steps)
        t0 = time.perf_counter()
        code = b.step(steps)
        el = time.perf_counter() - t0
        if code != abi.OCM_OK:
            raise SystemExit(f"replica batch failed with status {code}")
        shape = b.launch_shape()
        resident = b.resident_threads()
    dt_ps = integs[0].dt
    agg = ns_per_day(len(systems) * steps / el, dt_ps)
    log(f"replicas batched: {steps} steps of {len(systems)} replicas in {el:.3f} s -> {agg:.0f} ns/day aggregate")
    out = {"aggregate_ns_day": agg, "steps": steps, "warmup": warmup, "ms_per_step": el / steps * 1000.0,
           "timed_s": el, "launch_shape": shape, "resident_threads": resident, "config": cfg}
    if serial_steps <= 0:
        return out
    tot = 0.0
    for k in range(len(systems)):
        with Context(systems[k], integs[k], use_graph=graph, **lopt) as c:
            c.set_positions(xs[k])
            c.set_velocities(vs[k])
            c.step(200)
            c.prepare(serial_steps)
            t = time.perf_counter()
            c.step(serial_steps)
            tot += time.perf_counter() - t
        if (k + 1) % 64 == 0:
            log(f"replicas serial: {k + 1} of {len(systems)} done")
    serial = ns_per_day(len(systems) * serial_steps / tot, dt_ps)
    log(f"replicas serial: {len(systems)} x {serial_steps} steps one at a time in {tot:.2f} s -> {serial:.0f} ns/day")
    out.update({"serial_plain_ns_day": serial, "serial_steps_each": serial_steps,
                "batched_over_serial": agg / serial})
    return out


def _pct(v, q):
    """Nearest-rank percentile (as the server reports)."""
    v = sorted(v)
    if not v:
        return 0.0
    k = int(np.ceil(q * len(v)))
    return v[min(len(v) - 1, max(k, 1) - 1)]


def serve_run(args, n_jobs, steps_range, serial_frames, load):
    """Served job stream: 60% Go chains (50-500 beads), 40% HPS slabs
    (2,000-20,000 beads), a frame per 10,000 steps.

    Serial plain execution (one context per job, back to back, a frame read
    after every 10,000 steps) is timed on each job's first `serial_frames`
    frame stretches after the first one, and each job's serial time is
    its setup time plus its number of stretches times the measured stretch
    (0: run every job whole).  The server then runs every job in full twice:
    all submitted at once (capacity), and under Poisson arrivals at `load`
    times the serial service rate, whose job latency (submit -> completion)
    is compared with a FIFO of the serial job times on the same arrivals."""
    from opencafemol_amd import abi, configs
    from opencafemol_amd.context import Batch, Context, Server
    lo, hi = steps_range
    t = time.perf_counter()
    jobs = configs.served_job_mix(n_jobs, seed=args.seed, steps=(lo, hi))
    hps = [j for j in jobs if j["kind"] == "hps"]
    log(f"served stream: {len(jobs)} jobs ({len(hps)} HPS, {len(jobs) - len(hps)} Go), "
        f"{sum(j['system'].n for j in jobs)} beads, {sum(j['nsteps'] for j in jobs)} steps, "
        f"built in {time.perf_counter() - t:.1f} s")
    if hps:   # pre-relax the packed slabs in one batch
        t = time.perf_counter()
        with Batch([j["system"] for j in hps], [j["integrator"] for j in hps]) as b:
            for k, j in enumerate(hps):
                b.set_positions(k, j["positions"])
            b.minimize(tolerance=10.0, max_iterations=args.minimize_iters)
            for k, j in enumerate(hps):
                j["positions"] = b.state(k, abi.STATE_POSITIONS)["positions"]
        log(f"served stream: relaxed {len(hps)} slabs in {time.perf_counter() - t:.1f} s")
    sim_ns = [j["nsteps"] * j["integrator"].dt * 1e-3 for j in jobs]
    opt = {"use_graph": 0 if args.no_graph else 1}

    def submit(srv, j):
        return srv.submit(j["system"], j["integrator"], j["positions"], None, nsteps=j["nsteps"],
                          frame_interval=j["frame_interval"])

    # warm-up: library load, first kernels
    with Server(**opt) as srv:
        srv.drain([submit(srv, dict(jobs[0], nsteps=jobs[0]["frame_interval"]))])

    # serial plain execution
    serial_t = []
    for j in jobs:
        fi = j["frame_interval"]
        stretches = j["nsteps"] // fi
        t = time.perf_counter()
        with Context(j["system"], j["integrator"], **opt) as c:
            c.set_positions(j["positions"])
            c.set_velocities_to_temperature(j["integrator"].temperature, j["integrator"].seed)
            if c.step(fi) != abi.OCM_OK:
                raise SystemExit("serial run failed")
            c.get_state(abi.STATE_POSITIONS)
            setup_and_first = time.perf_counter() - t
            n = stretches - 1 if serial_frames <= 0 else min(serial_frames, stretches - 1)
            t1 = time.perf_counter()
            for _ in range(n):
                if c.step(fi) != abi.OCM_OK:
                    raise SystemExit("serial run failed")
                c.get_state(abi.STATE_POSITIONS)
            per = (time.perf_counter() - t1) / max(n, 1)
        serial_t.append(setup_and_first + (stretches - 1) * per)
    serial_total = sum(serial_t)
    serial_ns_day = sum(sim_ns) / serial_total * 86400.0
    log(f"served stream: serial plain execution {serial_total:.2f} s -> {serial_ns_day:.0f} ns/day aggregate")

    # capacity: everything submitted at once
    with Server(**opt) as srv:
        t = time.perf_counter()
        ids = [submit(srv, j) for j in jobs]
        frames = [0]
        res = srv.drain(ids, on_frame=lambda *a: frames.__setitem__(0, frames[0] + 1))
        sat_s = time.perf_counter() - t
        sat_stats = srv.stats()
    ok = all(r["status"] == abi.OCM_OK for r in res.values())
    sat_ns_day = sum(sim_ns) / sat_s * 86400.0
    log(f"served stream: all at once {sat_s:.2f} s -> {sat_ns_day:.0f} ns/day aggregate "
        f"({sat_stats['repacks']} batches, {frames[0]} frames)")

    # Poisson stream
    rng = np.random.default_rng(args.seed + 1)
    rate = load * len(jobs) / serial_total       # jobs per second
    arrivals = np.concatenate([[0.0], np.cumsum(rng.exponential(1.0 / rate, len(jobs) - 1))])
    order = rng.permutation(len(jobs))
    with Server(**opt) as srv:
        t0 = time.perf_counter()
        ids, pending, results, k = [], set(), {}, 0
        last_log = t0
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
            if time.perf_counter() - last_log > 30:
                last_log = time.perf_counter()
                log(f"served stream: Poisson phase {last_log - t0:.0f} s, {len(results)} of {len(jobs)} done")
        stream_s = time.perf_counter() - t0
        st = srv.stats()
    ok = ok and all(r["status"] == abi.OCM_OK for r in results.values())
    lat = [(results[i]["end_ms"] - results[i]["submit_ms"]) / 1000.0 for i in ids]
    fifo_end, fifo_lat = 0.0, []
    for k2, jidx in enumerate(order):
        start = max(arrivals[k2], fifo_end)
        fifo_end = start + serial_t[jidx]
        fifo_lat.append(fifo_end - arrivals[k2])
    out = {"aggregate_ns_day": sat_ns_day, "serial_plain_ns_day": serial_ns_day,
           "served_over_serial": sat_ns_day / serial_ns_day,
           "latency_p50_s": _pct(lat, 0.5), "latency_p99_s": _pct(lat, 0.99),
           "serial_fifo_latency_p50_s": _pct(fifo_lat, 0.5), "serial_fifo_latency_p99_s": _pct(fifo_lat, 0.99),
           "wait_p50_s": st["wait_p50_ms"] / 1000.0, "wait_p99_s": st["wait_p99_ms"] / 1000.0,
           "stream_aggregate_ns_day": sum(sim_ns) / stream_s * 86400.0, "stream_makespan_s": stream_s,
           "serial_fifo_makespan_s": fifo_end, "arrival_rate_per_s": rate, "load_vs_serial": load,
           "capacity_s": sat_s, "batches": sat_stats["repacks"], "repack_ms": sat_stats["repack_ms"],
           "busy_ms": sat_stats["busy_ms"], "all_jobs_ok": bool(ok),
           "serial_timing": ("every job whole" if serial_frames <= 0 else
                             f"setup + first stretch, then {serial_frames} timed 10,000-step stretches "
                             f"per job, extrapolated to the job's length"),
           "config": {"workload": "served_job_stream", "jobs": len(jobs), "hps_jobs": len(hps),
                      "go_beads": [50, 500], "hps_beads": [2000, 20000], "steps_per_job": [lo, hi],
                      "frame_interval": 10000, "total_beads": int(sum(j["system"].n for j in jobs)),
                      "total_steps": int(sum(j["nsteps"] for j in jobs)), "dt_ps": 0.01,
                      "temperature_K": 300.0}}
    log(f"served stream: p50 {out['latency_p50_s']:.2f} s, p99 {out['latency_p99_s']:.2f} s "
        f"(serial FIFO p50 {out['serial_fifo_latency_p50_s']:.2f} s, p99 {out['serial_fifo_latency_p99_s']:.2f} s)")
    return out


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


def single_run(args, workload):
    """One system (or the replica batch): relax, warm up, time K steps; the
    pair kernel's own device-clock durations give the roofline figures."""
    from opencafemol_amd import abi
    from opencafemol_amd.context import Batch

    systems, integs, xs, cfg = build_workload(workload)
    log(f"workload {cfg['workload']}: {sum(s.n for s in systems)} beads in {len(systems)} system(s)")
    # the condensate configuration states its list parameters
    lopt = list_options(args, {"skin": 0.3, "rebuild_interval": 20} if workload == "condensate" else None)
    b = Batch(systems, integs, use_graph=0 if args.no_graph else 1, **lopt)
    for k, x in enumerate(xs):
        b.set_positions(k, x)
    t = time.perf_counter()
    b.minimize(tolerance=10.0, max_iterations=args.minimize_iters)
    log(f"relaxed in {time.perf_counter() - t:.1f} s")
    if len(systems) == 1:
        abi.check(b.lib.ocm_set_velocities_to_temperature(b.h, integs[0].temperature, 2024))
    else:
        for k, v in enumerate(mb_velocities(systems, integs)):
            b.set_velocities(k, v)
    # equilibration from the packed start (the slab's list shrinks by ~13% over
    # the first thousand steps), untimed, before the warm-up
    eq = args.equilibrate if args.equilibrate is not None else (200 if workload == "condensate" else 2000)
    if eq > 0 and b.step(eq) != abi.OCM_OK:
        raise SystemExit("equilibration failed")
    cfg = dict(cfg, equilibration_steps=eq)
    b.set_timing(2)   # pair-kernel duration stamps on the device clock, inside the step graphs
    code = b.step(args.warmup)
    if code != abi.OCM_OK:
        raise SystemExit(f"warmup failed with status {code}")
    # capture the graphs of the timed call now: its timed region then holds
    # stepping only
    t = time.perf_counter()
    b.prepare(args.steps)
    log(f"warmup {args.warmup} steps done; timed call's graphs captured in {time.perf_counter() - t:.3f} s")

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
    ns_day = ns_per_day(args.steps / el, dt_ps) * len(systems)
    log(f"timed {args.steps} steps in {el:.4f} s: {ms_per_step * 1000:.2f} us/step, {ns_day:.1f} ns/day")

    resident = b.resident_threads()
    if resident:
        # one workgroup per system steps whole chunks (k_resident): no pair
        # kernel to stamp or break down
        cpu = None
        if not args.no_cpu and len(systems) == 1:
            st = b.state(0, abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
            cpu = cpu_baseline(systems[0], integs[0], st["positions"], st["velocities"], dt_ps)
            log(f"cpu baseline {cpu['value']:.2f} ns/day on {cpu['cores']} threads")
        b.close()
        return {"metric": METRIC, "value": ns_day, "unit": "ns/day", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "fp32+fp64", "data": "synthetic", "config": cfg,
                "roofline": None, "cpu_baseline": cpu, "timed_s": el,
                "kernels": {"resident_threads": resident, "us_per_step": ms_per_step * 1000.0,
                            "note": "resident stepping: k_resident runs each chunk in one workgroup per system"}}

    # pair kernel duration over the timed region: first block start to last
    # block end of every launch, on the device wall clock
    entries, members, _ = b.list_stats()
    shape = b.launch_shape()
    inner_padded, prune = b.inner_list_stats()
    counts = b.list_counts()
    avg_us = stamp_ms / max(stamp_n, 1) * 1000.0
    inner_on = prune > 1
    # the entries one force-pass launch reads: the inner rows when the inner
    # list is on (the prune kernel rewrites them from the outer rows every few
    # steps), the outer rows otherwise
    read_padded = float(inner_padded if inner_on else entries)
    listed_pairs = (counts["inner_valid"] if inner_on else counts["outer_valid"]) / 2.0
    n_beads = sum(s.n for s in systems)
    # algorithmic bytes of the nonbonded pass: 4 B per unique listed pair + 52 B per bead
    alg_bytes = 4.0 * listed_pairs + 52.0 * n_beads
    achieved = alg_bytes / (avg_us * 1e-6) / 1e9
    # the kernel's own access pattern: 4 B per directed padded entry + 64 B per row
    directed_bytes = 4.0 * read_padded + 64.0 * members
    log(f"pair kernel {avg_us:.2f} us/launch over {stamp_n} launches of the timed region; "
        f"{listed_pairs:.0f} unique listed pairs, {n_beads} beads: {achieved:.0f} GB/s algorithmic")

    # breakdown phase: HIP events around every launch (eager; adds launch latency to each figure)
    kt_steps = min(args.steps, 4000)
    b.set_timing(1)
    code = b.step(kt_steps)
    b.set_timing(0)
    pair_ms, pair_n = b.timing(0)
    int_ms, int_n = b.timing(2)
    nl_ms, nl_n = b.timing(3)
    pr_ms, pr_n = b.timing(6)
    kernels = {"pair_us": avg_us, "pair_us_event_bracketed": pair_ms / max(pair_n, 1) * 1000,
               "integrator_us_event_bracketed": int_ms / max(int_n, 1) * 1000,
               "list_build_us_event_bracketed": nl_ms / max(nl_n, 1) * 1000,
               "prune_us_event_bracketed": pr_ms / max(pr_n, 1) * 1000,
               "prunes_per_step": pr_n / max(kt_steps, 1), "list_entries_padded": entries, "list_rows": members,
               "list_entries": counts["outer_valid"], "inner_list_entries": counts["inner_valid"],
               "inner_list_entries_padded": inner_padded, "prune_interval": prune,
               "unique_listed_pairs_read": listed_pairs, "entries_read_per_launch": read_padded,
               "pairs_per_s": listed_pairs / (avg_us * 1e-6),
               "directed_entry_bytes_per_launch": directed_bytes,
               "directed_entry_GBs": directed_bytes / (avg_us * 1e-6) / 1e9,
               "list_rebuilds_timed": rebuilds_timed, "steps_per_rebuild": args.steps / max(rebuilds_timed, 1),
               "launch_shape": shape}

    traffic = None
    tf = TRAFFIC_FILES.get(cfg["workload"])
    if tf and os.path.exists(tf):
        traffic = json.load(open(tf))["hbm_bytes_per_entry"] * read_padded
    valu = None
    vf = VALU_FILES.get(cfg["workload"])
    if vf and os.path.exists(vf):
        vj = json.load(open(vf))
        insts = vj["valu_insts_per_entry"] * read_padded
        # SIMD issue cycles: 4 per wave64 instruction, 8 for an FP64 add/mul/fma
        cycles = vj["valu_issue_cycles_per_entry"] * read_padded
        simd_cycles = avg_us * 1e-6 * VALU_CLOCK_HZ * VALU_SIMDS
        valu = {"bound": "valu", "achieved": insts / (avg_us * 1e-6) / 1e9, "peak": VALU_PEAK_GINST,
                "unit": "G wave-instructions/s", "frac": insts / (avg_us * 1e-6) / 1e9 / VALU_PEAK_GINST,
                "issue_cycle_frac": cycles / simd_cycles,
                "insts_per_launch": insts, "insts_per_entry": vj["valu_insts_per_entry"],
                "f64_share": vj["f64_share_of_valu_insts"], "source": os.path.relpath(vf, ROOT)}

    cpu = None
    if not args.no_cpu and len(systems) == 1:
        st = b.state(0, abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
        cpu = cpu_baseline(systems[0], integs[0], st["positions"], st["velocities"], dt_ps)
        log(f"cpu baseline {cpu['value']:.2f} ns/day on {cpu['cores']} threads, "
            f"{cpu['value_1_core']:.2f} on 1 ({cpu['host_cpu']}, nproc {cpu['nproc']})")
    b.close()

    return {
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "accounting": "4 B per unique listed pair + 52 B per bead, per force-pass launch"},
        "valu_roofline": valu,
        "cpu_baseline": cpu,
        "timed_s": el,
        "kernels": kernels,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--workload", default="all")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serial", action="store_true",
                    help="replicas workload: skip the one-at-a-time comparison (the batched timing is unchanged)")
    ap.add_argument("--minimize-iters", type=int, default=3000)
    ap.add_argument("--equilibrate", type=int, default=None,
                    help="untimed steps after relaxation, before the warm-up (default 2000; 200 for the condensate)")
    ap.add_argument("--no-graph", action="store_true", help="eager launches (for profilers)")
    ap.add_argument("--skin", type=float, default=None, help="neighbor-list skin (nm); default: library's")
    ap.add_argument("--rebuild-interval", type=int, default=None, help="steps between list rebuilds")
    ap.add_argument("--prune-interval", type=int, default=None,
                    help="steps between inner-list prunes (<= 1: off); default: library's")
    ap.add_argument("--resident", type=int, default=None,
                    help="resident stepping of small systems: -1 where the batch qualifies (library default), 0 never")
    ap.add_argument("--replica-steps", type=int, default=4000, help="timed steps of the replica batch")
    ap.add_argument("--replica-serial-steps", type=int, default=1000, help="timed steps per replica alone")
    ap.add_argument("--serve-jobs", type=int, default=8)
    ap.add_argument("--serve-steps", default="100000,1000000", help="per-job step range (multiples of 10000)")
    ap.add_argument("--serve-serial-frames", type=int, default=2,
                    help="10,000-step stretches timed per job for the serial baseline (0: every job whole)")
    ap.add_argument("--serve-load", type=float, default=0.8, help="arrival rate / serial service rate")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported; this build measures one "
              f"GPU's share (independent systems and jobs would shard across GPUs without communication)",
              file=sys.stderr)
        sys.exit(2)
    steps_range = tuple(int(v) for v in args.serve_steps.split(","))

    if args.workload == "copy":
        copy_bench(args)
        return
    if args.workload == "replicas":
        r = replicas_run(args, args.steps, args.warmup, 0 if args.no_serial else args.replica_serial_steps)
        print(json.dumps({"metric": "aggregate ns/day, 256 Go replicas (50-500 beads, 280-400 K)",
                          "value": r["aggregate_ns_day"], "unit": "ns/day", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64", "data": "synthetic",
                          "config": r.pop("config"), "replicas_256": r}))
        return
    if args.workload == "serve":
        r = serve_run(args, args.serve_jobs, steps_range, args.serve_serial_frames, args.serve_load)
        print(json.dumps({"metric": "served job stream: aggregate ns/day (all jobs submitted at once) and "
                                    "p50/p99 job latency (Poisson arrivals)", "value": r["aggregate_ns_day"],
                          "unit": "ns/day", "n_gpus": 1, "steps": r["config"]["total_steps"], "warmup": 0,
                          "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "fp32+fp64", "data": "synthetic", "config": r.pop("config"),
                          "served_stream": r}))
        return
    if args.workload in ("hps", "condensate", "go"):
        print(json.dumps(single_run(args, args.workload)))
        return
    if args.workload != "all":
        raise SystemExit(f"unknown workload {args.workload}")

    out = single_run(args, "hps")
    # the 62-bead Go chain alone (GPU rate only; its CPU baseline is the go workload's own line)
    go_args = argparse.Namespace(**dict(vars(args), steps=20000, warmup=2000, no_cpu=True))
    go = single_run(go_args, "go")
    out["go_chain_62"] = {"ns_day": go["value"], "us_per_step": go["ms_per_step"] * 1000.0,
                          "steps": go["steps"], "resident_threads": go["kernels"].get("resident_threads", 0),
                          "config": go["config"]}
    # the 1,000,000-bead condensate (its stated list parameters; GPU only)
    c3_args = argparse.Namespace(**dict(vars(args), steps=1000, warmup=200, no_cpu=True))
    c3 = single_run(c3_args, "condensate")
    out["condensate_1m"] = {"ns_day": c3["value"], "us_per_step": c3["ms_per_step"] * 1000.0,
                            "steps": c3["steps"], "roofline": c3["roofline"],
                            "valu_roofline": c3["valu_roofline"],
                            "pair_us": c3["kernels"].get("pair_us"), "pairs_per_s": c3["kernels"].get("pairs_per_s"),
                            "config": c3["config"]}
    out["replicas_256"] = replicas_run(args, args.replica_steps, 500,
                                       0 if args.no_serial else args.replica_serial_steps)
    out["served_stream"] = serve_run(args, args.serve_jobs, steps_range, args.serve_serial_frames,
                                     args.serve_load)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
