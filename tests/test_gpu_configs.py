"""The benchmark configurations at their stated sizes, on the device, against
the FP64 oracle.

- Condensate: 1,000 chains x 1,000 beads = 1,000,000 beads in a periodic
  60 x 60 x 600 nm box, list skin 0.3 nm rebuilt every 20 steps: forces and
  energies of every bead against the cell-list oracle (itself pinned against
  the brute-force oracle in test_oracle.py), before and after 200 steps, and
  both neighbor-pair sets against a periodic k-d tree.
- Replica ensemble: 256 Go chains of 50-500 beads on the 280-400 K ladder in
  one batch: step-0 forces and energies of every replica, each replica's
  kinetic temperature over 10,000 steps against its ladder temperature, and
  8 replicas stepped alone against their bits in the batch.
- Open boundaries at scale: a 21,000-bead HPS cluster with no box, whose
  pair groups use a cell grid over their bounding box.
- Served stream: 10 jobs of the 60% Go / 40% HPS mix (HPS slabs of 2,520 to
  17,640 beads), 2,000-6,000 steps each with a frame every 1,000: every frame
  and final state from the server equal the job run alone.
Each case prints its name before touching the device.
"""
import sys

import numpy as np
import pytest

from opencafemol_amd import abi, configs
from opencafemol_amd.context import Batch, Context
from tests.tolerances import assert_energies_close, assert_forces_close

pytestmark = pytest.mark.gpu

KB = 1.380649e-23 * 6.02214076e23 / 1000.0


@pytest.fixture(autouse=True)
def _announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()
    yield


def _pair_keys(pairs, n):
    pairs = np.sort(np.asarray(pairs, dtype=np.int64).reshape(-1, 2), axis=1)
    return np.unique(pairs[:, 0] * np.int64(n) + pairs[:, 1])


def _check_pairs_against_tree(c, s, x):
    """Both nonbonded lists' pairs within the cutoff against a periodic k-d
    tree of the participating beads, less the exclusions.  Pairs within 1e-12
    of the cutoff may fall either way (the tree's distance arithmetic is not
    the kernel's)."""
    from scipy.spatial import cKDTree
    box = np.array(s.box) if s.box is not None else None
    xw = np.mod(x, box) if box is not None else x
    checked = 0
    for fi, f in enumerate(s.forces):
        if f.kind not in (abi.LJ_REPULSIVE, abi.DEBYE_HUCKEL):
            continue
        got = c.neighbor_pairs(fi)
        g = _pair_keys(got, s.n)
        assert len(g) == len(got)
        idx = np.arange(s.n) if f.category is None else np.where(f.category >= 0)[0]
        tree = cKDTree(xw[idx], boxsize=box)
        ex = _pair_keys(f.exclusions, s.n) if f.exclusions is not None and len(f.exclusions) else np.zeros(0, np.int64)
        inner = np.setdiff1d(_pair_keys(idx[tree.query_pairs(f.cutoff * (1 - 1e-12), output_type="ndarray")], s.n), ex)
        outer = np.setdiff1d(_pair_keys(idx[tree.query_pairs(f.cutoff * (1 + 1e-12), output_type="ndarray")], s.n), ex)
        assert np.isin(inner, g).all(), (f.name, int((~np.isin(inner, g)).sum()))
        assert np.isin(g, outer).all(), (f.name, int((~np.isin(g, outer)).sum()))
        checked += len(g)
        sys.__stderr__.write(f"[gpu-case] {f.name}: {len(g)} pairs equal the k-d tree's\n")
    return checked


def test_condensate_one_million_beads(oracle):
    s, it, x = configs.hps_condensate()
    assert s.n == 1_000_000
    with Context(s, it, skin=0.3, rebuild_interval=20) as c:
        c.set_positions(x)
        st = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY)
        Fo, Eo = oracle.evaluate_cells(s, x)
        assert_forces_close(st["forces"], Fo)
        assert_energies_close(st["energy_by_group"], Eo)
        assert _check_pairs_against_tree(c, s, x) > 40_000_000
        # relax the lattice start a little, then 200 steps (10 list lifetimes)
        c.minimize(tolerance=100.0, max_iterations=30)
        c.set_velocities_to_temperature(it.temperature, 11)
        assert c.step(200) == abi.OCM_OK
        st = c.get_state(abi.STATE_POSITIONS | abi.STATE_FORCES | abi.STATE_ENERGY)
        assert st["step"] == 200
        Fo, Eo = oracle.evaluate_cells(s, st["positions"])
        assert_forces_close(st["forces"], Fo)
        assert_energies_close(st["energy_by_group"], Eo)


def test_open_boundary_cluster_cell_grid(oracle):
    """21,000 HPS beads with open boundaries: the pair groups get a cell grid
    over the members' bounding box, laid out again at every build.  Pair
    sets against a k-d tree and forces against the oracle, at the start and
    after 400 steps (the grid follows the moving beads)."""
    s, it, x = configs.hps_open_cluster()
    assert s.n >= 20_000 and s.box is None
    with Context(s, it) as c:
        c.set_positions(x)
        st = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY)
        Fo, Eo = oracle.evaluate_cells(s, x)
        assert_forces_close(st["forces"], Fo)
        assert_energies_close(st["energy_by_group"], Eo)
        assert _check_pairs_against_tree(c, s, x) > 1_000_000
        c.minimize(tolerance=100.0, max_iterations=50)
        c.set_velocities_to_temperature(it.temperature, 3)
        assert c.step(400) == abi.OCM_OK
        st = c.get_state(abi.STATE_POSITIONS | abi.STATE_FORCES | abi.STATE_ENERGY)
        Fo, Eo = oracle.evaluate_cells(s, st["positions"])
        assert_forces_close(st["forces"], Fo)
        assert_energies_close(st["energy_by_group"], Eo)
        assert _check_pairs_against_tree(c, s, st["positions"]) > 1_000_000


def test_replicas_256_forces_and_ladder_temperatures(oracle):
    reps = configs.go_replicas(256, seed=0)
    systems = [r[0] for r in reps]
    integs = [r[1] for r in reps]
    rng = np.random.default_rng(21)
    # start 0.02 nm off the native structures: at a native geometry every
    # bead's net force is ~1e-13 and a relative force error is undefined
    reps = [(s, it, x + rng.normal(scale=0.02, size=x.shape)) for s, it, x in reps]
    vs = [rng.normal(size=(s.n, 3)) * np.sqrt(KB * it.temperature / s.masses)[:, None] for s, it, _ in reps]
    temps = np.array([it.temperature for it in integs])
    assert temps[0] == pytest.approx(280.0) and temps[-1] == pytest.approx(400.0)
    sizes = np.array([s.n for s in systems])
    assert sizes.min() >= 50 and sizes.max() <= 500
    sample = [0, 37, 80, 128, 170, 201, 230, 255]
    n_eq, n_blk, blk = 2000, 40, 200
    ke = np.zeros((len(reps), n_blk))
    with Batch(systems, integs) as b:
        for k, (s, it, x) in enumerate(reps):
            b.set_positions(k, x)
            b.set_velocities(k, vs[k])
        for k, (s, it, x) in enumerate(reps):
            st = b.state(k, abi.STATE_FORCES | abi.STATE_ENERGY)
            Fo, Eo = oracle.evaluate(s, x)
            assert_forces_close(st["forces"], Fo)
            assert_energies_close(st["energy_by_group"], Eo)
        assert b.step(n_eq) == abi.OCM_OK
        for q in range(n_blk):
            assert b.step(blk) == abi.OCM_OK
            for k, s in enumerate(systems):
                v = b.state(k, abi.STATE_VELOCITIES)["velocities"]
                ke[k, q] = 0.5 * np.sum(s.masses[:, None] * v * v)
        final = {k: b.state(k, abi.STATE_POSITIONS | abi.STATE_VELOCITIES) for k in sample}
    # kinetic temperature per replica.  BAOAB at 10 fs sits about 2% below the
    # bath for these chains (the stiff bonds' O(dt^2) bias; the FP64 oracle's
    # BAOAB gives 0.984 and 0.979 for 60 and 200 beads), so the ladder is
    # checked as T_kin / T_bath = 0.98 within 0.02 plus 5 standard errors
    # (samples 2 ps apart, about 30 effective of the 40)
    ratio = 2.0 * ke.mean(axis=1) / (3.0 * sizes * KB * temps)
    sigma = np.sqrt(2.0 / (3.0 * sizes)) / np.sqrt(30.0)
    bad = np.where(np.abs(ratio - 0.98) > 0.02 + 5.0 * sigma)[0]
    assert len(bad) == 0, [(int(k), float(ratio[k]), float(temps[k])) for k in bad]
    assert 0.96 < ratio.mean() < 1.0
    # the ladder itself: measured temperature rises with the bath's
    slope = np.polyfit(temps, ratio * temps, 1)[0]
    assert 0.95 < slope < 1.01
    # alone == in the batch
    for k in sample:
        s, it, x = reps[k]
        with Context(s, it) as c:
            c.set_positions(x)
            c.set_velocities(vs[k])
            assert c.step(n_eq + n_blk * blk) == abi.OCM_OK
            st = c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
        assert np.array_equal(st["positions"], final[k]["positions"]), k
        assert np.array_equal(st["velocities"], final[k]["velocities"]), k


def test_served_mix_jobs_alone_equal_in_mix():
    from tests.test_gpu_server import _assert_same, _serve
    mix = configs.served_job_mix(10, seed=3, steps=(2000, 6000), frame_interval=1000)
    hps = [j for j in mix if j["kind"] == "hps"]
    assert len(hps) >= 2 and all(2000 <= j["system"].n <= 20000 for j in hps)
    with Batch([j["system"] for j in hps], [j["integrator"] for j in hps]) as b:
        for k, j in enumerate(hps):
            b.set_positions(k, j["positions"])
        b.minimize(tolerance=50.0, max_iterations=300)
        for k, j in enumerate(hps):
            j["positions"] = b.state(k, abi.STATE_POSITIONS)["positions"]
    jobs = [(j["system"], j["integrator"], j["positions"], None, j["nsteps"], j["frame_interval"]) for j in mix]
    ids, frames, results, stats = _serve(jobs, stagger=4, quantum_steps=1500)
    _assert_same(jobs, ids, frames, results)
    assert stats["completed"] == len(jobs) and stats["failed"] == 0
    assert stats["frames"] == sum(n // fi for *_, n, fi in jobs)


def test_velocities_to_temperature_match_oracle_stream(oracle):
    """ocm_set_velocities_to_temperature: v_i = sqrt(kT/m_i) g, g the FP64
    Box-Muller normals of Philox stream 1 at step 0 for bead i."""
    s, it, x = configs.go_chain(80, seed=4)
    s.masses[::7] = 0.0   # fixed beads get 0 (1/m = 0)
    with Context(s, it) as c:
        c.set_positions(x)
        c.set_velocities_to_temperature(350.0, 977)
        v = c.get_state(abi.STATE_VELOCITIES)["velocities"]
    kT = KB * 350.0
    for i in range(s.n):
        g = oracle.gauss3(977, 0, i, 1)
        want = np.sqrt(kT / s.masses[i]) * g if s.masses[i] > 0 else np.zeros(3)
        assert np.allclose(v[i], want, rtol=1e-12, atol=0.0), (i, v[i], want)
