"""HIP path vs the FP64 oracle, called through the C ABI (libocm.so).

Every case prints its name to the real stderr before touching the GPU so a
failing run names the case even if the process dies.
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


def thermal(x, seed, scale=0.02):
    """Native geometry plus N(0, scale nm) displacements.  At the exact native
    geometry every net force is ~1e-13 and a relative force error is undefined;
    thermal configurations are the inputs the step actually sees."""
    return x + np.random.default_rng(1000 + seed).normal(scale=scale, size=x.shape)


def _ctx(system, integ, x, **opt):
    c = Context(system, integ, **opt)
    c.set_positions(x)
    return c


def _check_forces_energies(system, integ, x, oracle, groups=0xFFFFFFFF, **opt):
    with _ctx(system, integ, x, **opt) as c:
        st = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY, groups)
        Fo, Eo = oracle.evaluate(system, x, groups)
        assert_forces_close(st["forces"], Fo)
        assert_energies_close(st["energy_by_group"], Eo)
        return st


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_go_chain_forces_energies(oracle, seed):
    s, it, x = configs.go_chain(62, seed=seed)
    _check_forces_energies(s, it, thermal(x, seed), oracle)


@pytest.mark.parametrize("seed", [0, 5])
def test_aicg2_chain_all_bonded_kinds(oracle, seed):
    s, it, x = configs.aicg2_chain(56, seed=seed)
    kinds = {f.kind for f in s.forces}
    assert {1, 2, 3, 4, 5, 6, 7, 16, 17} <= kinds
    _check_forces_energies(s, it, thermal(x, seed), oracle)


def test_aicg2_perturbed_outside_spline_range(oracle):
    # angles pushed beyond the FLA knot range exercise the linear walls
    s, it, x = configs.aicg2_chain(40, seed=11)
    rng = np.random.default_rng(4)
    x = x + rng.normal(scale=0.08, size=x.shape)
    _check_forces_energies(s, it, x, oracle)


def _small_hps(seed=0):
    return configs.hps_slab(n_chains=12, chain_len=30, box=(8.0, 8.0, 30.0), slab_z=6.0, seed=seed)


@pytest.mark.parametrize("seed", [0, 1])
def test_hps_small_forces_energies(oracle, seed):
    s, it, x = _small_hps(seed)
    _check_forces_energies(s, it, x, oracle)


def test_group_subsets(oracle):
    s, it, x = _small_hps(2)
    for g in range(len(s.group_ids)):
        _check_forces_energies(s, it, x, oracle, groups=1 << g)


@pytest.mark.parametrize("force_index", [1, 2, 3])
def test_hps_neighbor_pairs_bit_exact(oracle, force_index):
    s, it, x = _small_hps(3)
    with _ctx(s, it, x) as c:
        got = c.neighbor_pairs(force_index)
    ref = oracle.neighbor_pairs(s, force_index, x)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


def test_go_neighbor_pairs_bit_exact(oracle):
    s, it, x = configs.go_chain(120, seed=7)
    names = [f.name for f in s.forces]
    for fi in (names.index("ExcludedVolume"), names.index("DebyeHuckel")):
        with _ctx(s, it, x) as c:
            got = c.neighbor_pairs(fi)
        ref = oracle.neighbor_pairs(s, fi, x)
        assert np.array_equal(got, ref)


def test_mixed_exclusion_counts(oracle):
    """Rows with up to two excluded partners are filtered while the list is
    swept; rows with more have their entries checked after compaction.  Here
    each pair group's rows carry their chain exclusions plus up to four random
    nearby partners, so both paths run in one group: the pair sets
    must still equal the oracle's bit for bit, and forces must match."""
    import dataclasses
    s0, it, x = _small_hps(4)
    rng = np.random.default_rng(12)
    s = dataclasses.replace(s0, forces=[], group_ids={})
    box = np.array(s0.box)
    extras = {}   # one set per (cutoff, members): the two LJ halves share theirs, as they share a group

    def extra_exclusions(f):
        members = np.arange(s0.n) if f.category is None else np.where(f.category >= 0)[0]
        key = (f.cutoff, tuple(members))
        if key not in extras:
            out = []
            for i in members:
                k = int(rng.integers(0, 5))
                d = x[members] - x[i]
                d = np.linalg.norm(d - box * np.rint(d / box), axis=1)
                near = members[(d < f.cutoff) & (members != i)]
                for j in rng.choice(near, size=min(k, len(near)), replace=False):
                    out.append((min(i, j), max(i, j)))
            extras[key] = np.array(out, np.int32).reshape(-1, 2)
        return extras[key]

    for f in s0.forces:
        if f.particle_params is None:
            s.add(dataclasses.replace(f))
            continue
        ex = np.unique(np.concatenate([np.asarray(f.exclusions).reshape(-1, 2), extra_exclusions(f)]), axis=0)
        s.add(dataclasses.replace(f, exclusions=ex))
    counts = np.bincount(np.asarray(s.forces[1].exclusions).ravel(), minlength=s.n)
    assert (counts <= 2).any() and (counts > 2).any()
    for fi, f in enumerate(s.forces):
        if f.particle_params is None:
            continue
        with _ctx(s, it, x) as c:
            got = c.neighbor_pairs(fi)
        ref = oracle.neighbor_pairs(s, fi, x)
        assert np.array_equal(got, ref), f.name
    _check_forces_energies(s, it, thermal(x, 4, 0.01), oracle)


def test_hps_slab_full_size_forces(oracle):
    # 100 chains x 140 beads in a 15 x 15 x 150 nm periodic box
    s, it, x = configs.hps_slab()
    _check_forces_energies(s, it, x, oracle)


@pytest.mark.parametrize("resident", [0, -1])
def test_baoab_matches_oracle(oracle, resident):
    s, it, x = configs.go_chain(62, seed=2)
    x = thermal(x, 2)
    rng = np.random.default_rng(0)
    v = rng.normal(scale=0.1, size=x.shape)
    with _ctx(s, it, x, resident=resident) as c:
        assert (c.resident_threads() > 0) == (resident != 0)
        c.set_velocities(v)
        c.step(7)
        st = c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES)
    xo, vo = oracle.baoab(s, it, x, v, 0, 7)
    assert st["step"] == 7
    # FP32 forces differ from FP64 by up to ~1e-5 relative (~1e-2 kJ/mol/nm);
    # over 7 steps that moves beads by < 7^2 dt^2 dF / m ~ 1e-8 nm
    np.testing.assert_allclose(st["positions"], xo, rtol=0, atol=1e-7)
    np.testing.assert_allclose(st["velocities"], vo, rtol=0, atol=2e-5)


def test_step_deterministic_and_graph_equals_eager(oracle):
    # the standard path's step graphs (resident stepping captures none)
    s, it, x = configs.go_chain(90, seed=4)
    outs = []
    for graph in (1, 1, 0):
        with _ctx(s, it, x, use_graph=graph, chunk_steps=50, rebuild_interval=10, resident=0) as c:
            c.step(230)
            outs.append(c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES))
    for o in outs[1:]:
        assert np.array_equal(o["positions"], outs[0]["positions"])
        assert np.array_equal(o["velocities"], outs[0]["velocities"])


@pytest.mark.parametrize("resident", [0, -1])
def test_batch_alone_equals_in_mix(oracle, resident):
    reps = configs.go_replicas(6, n_min=50, n_max=200, seed=1)
    alone = []
    for s, it, x in reps:
        with _ctx(s, it, x, resident=resident) as c:
            c.step(150)
            alone.append(c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES | abi.STATE_ENERGY))
    with Batch([r[0] for r in reps], [r[1] for r in reps], resident=resident) as b:
        for k, (s, it, x) in enumerate(reps):
            b.set_positions(k, x)
        b.step(150)
        for k in range(len(reps)):
            st = b.state(k, abi.STATE_POSITIONS | abi.STATE_VELOCITIES | abi.STATE_ENERGY)
            assert np.array_equal(st["positions"], alone[k]["positions"])
            assert np.array_equal(st["velocities"], alone[k]["velocities"])
            assert np.array_equal(st["energy_by_group"], alone[k]["energy_by_group"])


@pytest.mark.parametrize("resident", [0, -1])
def test_batch_of_many_systems_xcd_slices(resident):
    """From 16 systems a batch lays its pair-kernel table out as 8 XCD slices
    (each XCD's workgroups walk one contiguous eighth).  Every system must
    still step to the bits it reaches alone: 20 Go replicas and 2 HPS slabs
    (periodic, with the inner list) in one batch; the same batch stepped
    resident, one workgroup per system, too."""
    reps = configs.go_replicas(20, n_min=50, n_max=160, seed=3)
    for seed in (5, 6):
        s, it, x = _small_hps(seed)
        reps.append((s, it, x))
    what = abi.STATE_POSITIONS | abi.STATE_VELOCITIES | abi.STATE_FORCES | abi.STATE_ENERGY
    alone = []
    for s, it, x in reps:
        with _ctx(s, it, x, chunk_steps=60, resident=resident) as c:
            c.step(150)
            alone.append(c.get_state(what))
    with Batch([r[0] for r in reps], [r[1] for r in reps], chunk_steps=60, resident=resident) as b:
        for k, (s, it, x) in enumerate(reps):
            b.set_positions(k, x)
        b.step(150)
        for k in range(len(reps)):
            st = b.state(k, what)
            for key in ("positions", "velocities", "forces", "energy_by_group"):
                assert np.array_equal(st[key], alone[k][key]), (k, key)


def test_rebuild_interval_does_not_change_bits(oracle):
    s, it, x = _small_hps(4)
    res = []
    for interval in (1, 7):
        with _ctx(s, it, x, rebuild_interval=interval, use_graph=0, skin=0.4, resident=0) as c:
            c.step(20)
            res.append(c.get_state(abi.STATE_POSITIONS)["positions"])
    assert np.array_equal(res[0], res[1])


def test_team_sizes_give_identical_forces(oracle):
    # every launch shape the autotuner may pick: lanes per row, and list
    # entries per lane per round
    s, it, x = _small_hps(5)
    ref = None
    shapes = [(t, 4) for t in (4, 8, 16, 32, 64)] + [(t, 2) for t in (8, 16, 32)]
    for team, u in shapes:
        with _ctx(s, it, x, team_size=team, entries_per_lane=u) as c:
            F = c.get_state(abi.STATE_FORCES)["forces"]
            shape = c.launch_shape()
        assert (shape["team"], shape["entries_per_lane"]) == (team, u)
        if ref is None:
            ref = F
            Fo, _ = oracle.evaluate(s, x)
            assert_forces_close(F, Fo)
        assert np.array_equal(F, ref), (team, u)


@pytest.mark.parametrize("resident", [0, -1])
def test_numeric_overflow_is_reported(oracle, resident):
    # two beads of a Go chain placed on top of each other: the r^-12 terms
    # overflow the fixed-point range and the step reports it
    s, it, x = configs.go_chain(30, seed=1)
    x = x.copy()
    x[10] = x[20] + 1e-4
    with _ctx(s, it, x, resident=resident) as c:
        code = c.step(3)
        assert code == abi.OCM_ERR_NUMERIC
        assert c.status(0) == abi.OCM_ERR_NUMERIC


def test_large_list_build_pairs():
    """A list with more than 2^17 rows is built from 64-row work items (each
    stages a stencil once for a whole cell).  Its pairs within the cutoff
    must be those a periodic k-d tree finds, less the exclusions; pairs
    within 1e-12 of the cutoff may fall either way, since the tree's
    distance arithmetic is not the kernel's."""
    from scipy.spatial import cKDTree
    s, it, x = configs.hps_condensate(n_chains=140, chain_len=1000, box=(30.0, 30.0, 300.0), seed=5)
    box = np.array(s.box)
    xw = np.mod(x, box)
    n = np.int64(s.n)

    def keys(pairs):
        pairs = np.sort(np.asarray(pairs, dtype=np.int64).reshape(-1, 2), axis=1)
        return np.unique(pairs[:, 0] * n + pairs[:, 1])

    with _ctx(s, it, x) as c:
        assert c.list_stats()[1] > 2 ** 17
        for fi, f in enumerate(s.forces):
            if f.kind not in (abi.LJ_REPULSIVE, abi.DEBYE_HUCKEL):
                continue
            got = c.neighbor_pairs(fi)
            g = keys(got)
            assert len(g) == len(got)
            idx = np.arange(s.n) if f.category is None else np.where(f.category >= 0)[0]
            tree = cKDTree(xw[idx], boxsize=box)
            ex = keys(f.exclusions) if f.exclusions is not None and len(f.exclusions) else np.zeros(0, np.int64)
            inner = np.setdiff1d(keys(idx[tree.query_pairs(f.cutoff * (1 - 1e-12), output_type="ndarray")]), ex)
            outer = np.setdiff1d(keys(idx[tree.query_pairs(f.cutoff * (1 + 1e-12), output_type="ndarray")]), ex)
            assert np.isin(inner, g).all(), (f.name, int((~np.isin(inner, g)).sum()))
            assert np.isin(g, outer).all(), (f.name, int((~np.isin(g, outer)).sum()))
            assert len(inner) > 100_000


@pytest.mark.parametrize("resident", [0, -1])
@pytest.mark.parametrize("case", ["go", "hps"])
def test_zero_friction_conserves_energy(case, resident):
    # with no friction BAOAB is velocity Verlet: over 3,000 steps of 2 fs the
    # total energy (FP64 energy pass + kinetic) stays put, which ties the
    # force pass to the energy pass along a trajectory on both step paths
    from opencafemol_amd.model import IntegratorSpec
    if case == "go":
        s, _, x = configs.go_chain(62, seed=3)   # the native structure: thermal motion only
    else:
        s, it0, x = _small_hps(3)
        with Context(s, it0) as c:
            c.set_positions(x)
            c.minimize(tolerance=20.0, max_iterations=500)
            x = c.get_state(abi.STATE_POSITIONS)["positions"]
    it = IntegratorSpec(300.0, 0.0, 0.002, 5)
    with Context(s, it, resident=resident) as c:
        c.set_positions(x)
        c.set_velocities_to_temperature(300.0, 5)
        assert c.resident_threads() == (1024 if resident else 0)
        st = c.get_state(abi.STATE_ENERGY)
        e0, k0 = st["potential"] + st["kinetic"], st["kinetic"]
        for _ in range(6):
            assert c.step(500) == abi.OCM_OK
            st = c.get_state(abi.STATE_ENERGY)
            e = st["potential"] + st["kinetic"]
            assert abs(e - e0) < 2e-3 * (abs(e0) + k0), (case, resident, st["step"], e, e0, k0)


@pytest.mark.parametrize("resident", [0, -1])
def test_linear_chain_trajectory_matches_oracle_over_2000_steps(oracle, resident):
    # a harmonic chain is a linear system: the FP32-force and FP32-noise
    # differences (~1e-6 relative) stay bounded instead of growing as in a
    # chaotic one, so the GPU trajectory must follow the FP64 oracle's (the
    # same Philox stream) for thousands of steps
    from opencafemol_amd.model import ForceSpec, IntegratorSpec, SystemSpec
    n = 24
    s = SystemSpec(masses=np.full(n, 60.0))
    b = np.stack([np.arange(n - 1), np.arange(1, n)], 1)
    s.add(ForceSpec(abi.HARMONIC_BOND, "HarmonicBond", indices=b, params=np.tile([0.38, 2000.0], (n - 1, 1))))
    it = IntegratorSpec(300.0, 2.0, 0.005, 31)
    x = np.stack([0.38 * np.arange(n), 0.05 * np.sin(np.arange(n)), np.zeros(n)], 1)
    v = np.random.default_rng(8).normal(scale=0.2, size=x.shape)
    with _ctx(s, it, x, resident=resident) as c:
        assert (c.resident_threads() > 0) == (resident != 0)
        c.set_velocities(v)
        got = []
        for _ in range(4):
            assert c.step(500) == abi.OCM_OK
            got.append(c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES))
    xo, vo = x, v
    for k in range(4):
        xo, vo = oracle.baoab(s, it, xo, vo, 500 * k, 500)
        assert got[k]["step"] == 500 * (k + 1)
        np.testing.assert_allclose(got[k]["positions"], xo, rtol=0, atol=1e-4)
        np.testing.assert_allclose(got[k]["velocities"], vo, rtol=0, atol=1e-2)
