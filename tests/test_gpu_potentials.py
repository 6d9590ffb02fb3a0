"""The membrane and capped potentials on the device against the FP64 oracle:
WCA, iSoLF attraction and the uniform (table) WCA and LJ-attractive kinds on a
periodic three-bead lipid patch; capped Go contacts and cosine dihedrals on a
Go chain.  Forces within 1e-4 per bead, energies within 1e-6 of the total,
neighbor-pair sets equal, a short trajectory that stays finite, and one
batch of both systems stepping to the bits each reaches alone.
"""
import sys

import numpy as np
import pytest

from opencafemol_amd import abi, configs
from opencafemol_amd.context import Batch, Context
from tests.tolerances import assert_energies_close, assert_forces_close

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()
    yield


def _thermal(x, seed, scale):
    return x + np.random.default_rng(500 + seed).normal(scale=scale, size=x.shape)


@pytest.mark.parametrize("seed", [0, 1])
def test_lipid_patch_forces_energies_pairs(oracle, seed):
    s, it, x = configs.lipid_patch(seed=seed)
    x = _thermal(x, seed, 0.03)
    with Context(s, it) as c:
        c.set_positions(x)
        st = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY)
        Fo, Eo = oracle.evaluate(s, x)
        assert_forces_close(st["forces"], Fo)
        assert_energies_close(st["energy_by_group"], Eo)
        for fi, f in enumerate(s.forces):
            if f.kind < abi.EXCLUDED_VOLUME:
                continue
            got = c.neighbor_pairs(fi)
            want = oracle.neighbor_pairs(s, fi, x)
            assert np.array_equal(got, want), f.name
            assert len(want) > 0, f.name


@pytest.mark.parametrize("cap", [0.8, 1.0])
def test_capped_go_cosine_chain_forces_energies(oracle, cap):
    s, it, x = configs.capped_go_chain(80, seed=3, capping_ratio=cap)
    x = _thermal(x, 3, 0.02)
    with Context(s, it) as c:
        c.set_positions(x)
        st = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY)
    Fo, Eo = oracle.evaluate(s, x)
    assert_forces_close(st["forces"], Fo)
    assert_energies_close(st["energy_by_group"], Eo)


def test_membrane_and_capped_systems_step_alone_and_batched(oracle):
    lip = configs.lipid_patch(seed=2)
    cap = configs.capped_go_chain(70, seed=4, capping_ratio=0.9)
    cases = []
    for s, it, x in (lip, cap):
        with Context(s, it) as c:
            c.set_positions(x)
            c.minimize(tolerance=20.0, max_iterations=200)
            x = c.get_state(abi.STATE_POSITIONS)["positions"]
        cases.append((s, it, x))
    alone = []
    for s, it, x in cases:
        with Context(s, it) as c:
            c.set_positions(x)
            c.set_velocities_to_temperature(it.temperature, 13)
            assert c.step(1000) == abi.OCM_OK
            st = c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES | abi.STATE_FORCES | abi.STATE_ENERGY)
            assert np.all(np.isfinite(st["positions"]))
            Fo, Eo = oracle.evaluate(s, st["positions"])
            assert_forces_close(st["forces"], Fo)
            assert_energies_close(st["energy_by_group"], Eo)
            alone.append(st)
    with Batch([c[0] for c in cases], [c[1] for c in cases]) as b:
        for k, (s, it, x) in enumerate(cases):
            b.set_positions(k, x)
            b.set_velocities(k, alone_velocities(s, it, x))
        assert b.step(1000) == abi.OCM_OK
        for k in range(len(cases)):
            st = b.state(k, abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
            assert np.array_equal(st["positions"], alone[k]["positions"]), k
            assert np.array_equal(st["velocities"], alone[k]["velocities"]), k


def alone_velocities(s, it, x):
    """The Maxwell-Boltzmann velocities ocm_set_velocities_to_temperature(T, 13)
    drew for the alone runs, read back from a fresh context."""
    with Context(s, it) as c:
        c.set_positions(x)
        c.set_velocities_to_temperature(it.temperature, 13)
        return c.get_state(abi.STATE_VELOCITIES)["velocities"]
