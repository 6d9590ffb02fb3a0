"""Edge cases of the device step against the oracle: systems with no pair
forces, no bonded terms or no forces at all, a nonbonded force no bead takes
part in, every pair excluded, fixed (massless) beads, and a batch mixing
all of these with periodic and open systems (each system's bits equal to
its run alone)."""
import sys

import numpy as np
import pytest

from opencafemol_amd import abi, configs
from opencafemol_amd.context import Batch, Context
from opencafemol_amd.model import ForceSpec, IntegratorSpec, SystemSpec
from tests.tolerances import assert_energies_close, assert_forces_close

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()
    yield


def _bonded_only(n=40, seed=1):
    s, it, x = configs.go_chain(n, seed=seed)
    s2 = SystemSpec(masses=s.masses)
    for f in s.forces:
        if f.kind in abi.BONDED_ARITY:
            s2.add(f)
    return s2, it, x


def _gas(n=600, box=6.0, seed=2):
    """Beads with excluded volume only, in a periodic box, no bonds."""
    rng = np.random.default_rng(seed)
    g = int(np.ceil(n ** (1 / 3)))
    grid = np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    x = (grid + 0.5) * (box / g) + rng.uniform(-0.05, 0.05, size=(n, 3))
    s = SystemSpec(masses=np.full(n, 100.0), box=(box, box, box))
    s.add(ForceSpec(abi.EXCLUDED_VOLUME, "ExcludedVolume", use_periodic=True, particle_params=np.full((n, 1), 0.15),
                    globals=np.array([2.5]), cutoff=0.6))
    return s, IntegratorSpec(300.0, 0.5, 0.01, seed), x


def _single_bead():
    s = SystemSpec(masses=np.array([50.0]))
    return s, IntegratorSpec(300.0, 1.0, 0.01, 9), np.array([[0.1, 0.2, 0.3]])


def _check(s, it, x, oracle):
    with Context(s, it) as c:
        c.set_positions(x)
        st = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY)
        Fo, Eo = oracle.evaluate(s, x)
        if np.abs(Fo).max() > 0:
            assert_forces_close(st["forces"], Fo)
        else:
            assert np.array_equal(st["forces"], np.zeros_like(Fo))
        assert_energies_close(st["energy_by_group"], Eo)
        assert c.step(300) == abi.OCM_OK
        st = c.get_state(abi.STATE_POSITIONS | abi.STATE_FORCES | abi.STATE_ENERGY)
        assert np.isfinite(st["positions"]).all() and st["step"] == 300
        Fo, Eo = oracle.evaluate(s, st["positions"])
        if np.abs(Fo).max() > 0:
            assert_forces_close(st["forces"], Fo)
        assert_energies_close(st["energy_by_group"], Eo)
        return st


def test_bonded_terms_only(oracle):
    s, it, x = _bonded_only()
    _check(s, it, x + np.random.default_rng(0).normal(scale=0.02, size=x.shape), oracle)


def test_pair_force_only(oracle):
    s, it, x = _gas()
    _check(s, it, x, oracle)


def test_single_bead_without_forces(oracle):
    s, it, x = _single_bead()
    st = _check(s, it, x, oracle)
    assert not np.array_equal(st["positions"], x)   # the thermostat moves it


def test_force_no_bead_takes_part_and_all_pairs_excluded(oracle):
    s, it, x = configs.go_chain(50, seed=6)
    x = x + np.random.default_rng(1).normal(scale=0.02, size=x.shape)
    n = s.n
    # Debye-Hueckel over no participating bead
    s.add(ForceSpec(abi.DEBYE_HUCKEL, "DebyeHuckel2", particle_params=np.zeros((n, 1)),
                    globals=np.array([1.0, 1.0, 0.0]), cutoff=2.0, category=np.full(n, -1, np.int32),
                    n_categories=1))
    # excluded volume with every pair excluded
    ii, jj = np.triu_indices(n, 1)
    s.add(ForceSpec(abi.EXCLUDED_VOLUME, "ExcludedVolume2", particle_params=np.full((n, 1), 0.2),
                    globals=np.array([1.0]), cutoff=0.8, exclusions=np.stack([ii, jj], 1).astype(np.int32)))
    _check(s, it, x, oracle)


def test_fixed_beads_stay_put(oracle):
    s, it, x = configs.go_chain(60, seed=3)
    x = x + np.random.default_rng(2).normal(scale=0.02, size=x.shape)
    fixed = np.zeros(s.n, bool)
    fixed[[0, 17, 59]] = True
    s.masses[fixed] = 0.0
    with Context(s, it) as c:
        c.set_positions(x)
        c.set_velocities_to_temperature(it.temperature, 5)
        assert c.step(500) == abi.OCM_OK
        st = c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES | abi.STATE_FORCES)
    assert np.array_equal(st["positions"][fixed], x[fixed])
    assert np.array_equal(st["velocities"][fixed], np.zeros((fixed.sum(), 3)))
    assert not np.array_equal(st["positions"][~fixed], x[~fixed])
    Fo, _ = oracle.evaluate(s, st["positions"])
    assert_forces_close(st["forces"], Fo)


def test_batch_of_everything_alone_equals_in_mix():
    rng = np.random.default_rng(4)
    go = configs.go_chain(80, seed=9)
    go = (go[0], go[1], go[2] + rng.normal(scale=0.02, size=go[2].shape))
    members = [_bonded_only(), _gas(), _single_bead(), go, configs.hps_slab_of_size(1400, seed=4),
               configs.lipid_patch(64, box=(6.0, 6.0, 6.0), seed=1)]
    members = [(s, IntegratorSpec(it.temperature, it.friction, it.dt, 100 + k), x)
               for k, (s, it, x) in enumerate(members)]
    # relax the packed HPS slab first (alone), so the overlap clamp never fires
    s, it, x = members[4]
    with Context(s, it) as c:
        c.set_positions(x)
        c.minimize(tolerance=50.0, max_iterations=300)
        members[4] = (s, it, c.get_state(abi.STATE_POSITIONS)["positions"])
    nsteps = 400
    with Batch([m[0] for m in members], [m[1] for m in members]) as b:
        for k, (s, it, x) in enumerate(members):
            b.set_positions(k, x)
        assert b.step(nsteps) == abi.OCM_OK
        mixed = [b.state(k, abi.STATE_POSITIONS | abi.STATE_VELOCITIES) for k in range(len(members))]
    for k, (s, it, x) in enumerate(members):
        with Context(s, it) as c:
            c.set_positions(x)
            assert c.step(nsteps) == abi.OCM_OK
            st = c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
        assert np.array_equal(st["positions"], mixed[k]["positions"]), k
        assert np.array_equal(st["velocities"], mixed[k]["velocities"]), k
