"""Resident stepping: batches whose every system has at most 1024 beads step
one workgroup per system, whole chunks on chip, with an in-kernel list.  The
per-pair, per-term and per-bead arithmetic is the standard path's, and a
pair's contribution to j is the negation of its contribution to i, so every
state must equal the standard path's bit for bit: Go chains and replica
batches, periodic HPS slabs up to 1,000 beads, the membrane kinds, every bonded kind of AICG2+, fixed beads, several chunks and
any grouping of steps into calls.  State queries after resident steps
(forces of group subsets, energies) are checked against the oracle."""
import sys

import numpy as np
import pytest

from opencafemol_amd import abi, configs
from opencafemol_amd.context import Batch, Context
from tests.tolerances import assert_energies_close, assert_forces_close

pytestmark = pytest.mark.gpu

WHAT = abi.STATE_POSITIONS | abi.STATE_VELOCITIES | abi.STATE_FORCES


@pytest.fixture(autouse=True)
def _announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()
    yield


def _thermal(x, seed, scale=0.02):
    return x + np.random.default_rng(500 + seed).normal(scale=scale, size=x.shape)


def _relaxed(s, it, x, iters=300):
    with Context(s, it) as c:
        c.set_positions(x)
        c.minimize(tolerance=50.0, max_iterations=iters)
        return c.get_state(abi.STATE_POSITIONS)["positions"]


def _run_batch(members, nsteps, resident, splits=None, vseed=7):
    """Steps a batch and returns each member's state and the workgroup size."""
    with Batch([m[0] for m in members], [m[1] for m in members], resident=resident) as b:
        for k, (s, it, x) in enumerate(members):
            b.set_positions(k, x)
            v = np.random.default_rng(vseed + k).normal(scale=0.3, size=x.shape)
            b.set_velocities(k, v)
        for n in (splits or [nsteps]):
            assert b.step(n) == abi.OCM_OK
        return [b.state(k, WHAT) for k in range(len(members))], b.resident_threads()


def _same(a, b, label=""):
    for key in ("positions", "velocities", "forces"):
        assert np.array_equal(a[key], b[key]), (label, key, np.abs(a[key] - b[key]).max())


def _both(members, nsteps, splits=None):
    std, t0 = _run_batch(members, nsteps, 0, splits)
    res, t1 = _run_batch(members, nsteps, -1, splits)
    assert t0 == 0 and t1 == 1024
    for k, (a, b) in enumerate(zip(std, res)):
        _same(a, b, k)
        assert a["step"] == b["step"] == nsteps
    return res


def test_go_chain_several_chunks():
    # 2,500 steps: a full 2,000-step launch and a partial one
    s, it, x = configs.go_chain(62, seed=3)
    _both([(s, it, _thermal(x, 3))], 2500)


def test_replica_batch():
    reps = configs.go_replicas(24, n_min=50, n_max=500, seed=5)
    reps = [(s, it, _thermal(x, k)) for k, (s, it, x) in enumerate(reps)]
    _both(reps, 600)


def test_hps_slab_periodic_1000_beads():
    s, it, x = configs.hps_slab(n_chains=20, chain_len=50, box=(9.0, 9.0, 30.0), slab_z=7.0, seed=2)
    assert 512 < s.n <= 1024
    _both([(s, it, _relaxed(s, it, x))], 400)


def test_hps_small_and_membrane_and_aicg2_in_one_batch():
    hps = configs.hps_slab(n_chains=12, chain_len=30, box=(8.0, 8.0, 30.0), slab_z=6.0, seed=6)
    hps = (hps[0], hps[1], _relaxed(*hps))
    mem = configs.lipid_patch(64, box=(6.0, 6.0, 6.0), seed=1)
    aic = configs.aicg2_chain(56, seed=4)
    cap = configs.capped_go_chain(50, seed=2, capping_ratio=1.0)
    members = [hps, mem, (aic[0], aic[1], _thermal(aic[2], 1)), (cap[0], cap[1], _thermal(cap[2], 2))]
    _both(members, 500)


def test_fixed_beads_and_step_grouping():
    s, it, x = configs.go_chain(80, seed=9)
    s.masses[[0, 33, 79]] = 0.0
    x = _thermal(x, 9)
    whole, _ = _run_batch([(s, it, x)], 300, -1)
    split, t = _run_batch([(s, it, x)], 300, -1, splits=[1, 7, 92, 200])
    assert t == 1024
    _same(whole[0], split[0])
    std, _ = _run_batch([(s, it, x)], 300, 0, splits=[300])
    _same(whole[0], std[0])
    assert np.array_equal(whole[0]["positions"][[0, 33, 79]], x[[0, 33, 79]])


def test_over_1024_beads_steps_standard():
    s, it, x = configs.hps_slab_of_size(1400, seed=4)
    with Context(s, it) as c:
        assert c.resident_threads() == 0
    with Batch([configs.go_chain(60)[0], s], [configs.go_chain(60)[1], it]) as b:
        assert b.resident_threads() == 0


def test_state_queries_after_resident_steps(oracle):
    # forces of a group subset and energies read the standard path's lists,
    # which resident steps leave behind: they are rebuilt first
    s, it, x = configs.hps_slab(n_chains=12, chain_len=30, box=(8.0, 8.0, 30.0), slab_z=6.0, seed=8)
    x = _relaxed(s, it, x)
    with Context(s, it) as c:
        assert c.resident_threads() == 1024
        c.set_positions(x)
        c.set_velocities_to_temperature(it.temperature, 3)
        assert c.step(800) == abi.OCM_OK
        full = c.get_state(abi.STATE_POSITIONS | abi.STATE_FORCES | abi.STATE_ENERGY)
        Fo, Eo = oracle.evaluate(s, full["positions"])
        assert_forces_close(full["forces"], Fo)
        assert_energies_close(full["energy_by_group"], Eo)
        groups = sorted({f.group for f in s.forces})
        sub = 1 << groups[-1]
        part = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY, sub)
        Fs, Es = oracle.evaluate(s, full["positions"], sub)
        assert_forces_close(part["forces"], Fs)
        assert_energies_close(part["energy_by_group"], Es)
        # stepping on after the queries continues the same trajectory
        assert c.step(200) == abi.OCM_OK
        after = c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
    with Context(s, it, resident=0) as c:
        c.set_positions(x)
        c.set_velocities_to_temperature(it.temperature, 3)
        assert c.step(1000) == abi.OCM_OK
        ref = c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
    assert after["step"] == ref["step"] == 1000
    assert np.array_equal(after["positions"], ref["positions"])
    assert np.array_equal(after["velocities"], ref["velocities"])


def test_replicas_256_resident_equals_standard():
    # the benchmark batch at full size (256 Go replicas of 50-500 beads,
    # 280-400 K ladder): every replica's state after 600 steps, resident and
    # standard, bit for bit
    reps = configs.go_replicas(256, seed=0)
    rng = np.random.default_rng(11)
    reps = [(s, it, x + rng.normal(scale=0.02, size=x.shape)) for s, it, x in reps]
    std, t0 = _run_batch(reps, 600, 0)
    res, t1 = _run_batch(reps, 600, -1)
    assert t0 == 0 and t1 == 1024
    for k, (a, b) in enumerate(zip(std, res)):
        _same(a, b, k)


def test_far_apart_chains_use_the_fp64_build():
    # an open system whose beads lie more than 1,000 nm from bead 0 (one chain
    # moved 1,500 nm away): the in-kernel build tests its pairs in FP64
    # instead of the FP32 relative coordinates; the bits still match the
    # standard path's
    s, it, x = configs.hps_open_cluster(n_chains=6, chain_len=40, extent=(6.0, 6.0), seed=3)
    assert s.n <= 1024
    x = x.copy()
    x[:40, 0] += 1500.0   # the first chain, far from the rest
    x = _relaxed(s, it, x, iters=200)
    _both([(s, it, x)], 300)
