"""Inner (pruned) neighbor list: most steps read the rows pruned to
cutoff + margin; a prune pass rewrites them every few steps and at every
rebuild.  Forces are fixed-point sums over the pairs inside the cutoff, so any
valid list gives the same bits: every setting of the prune interval -- off,
short, long enough that beads outrun the margin and the chunk is rolled back
-- must end in the same state as the full list."""
import sys

import numpy as np
import pytest

from opencafemol_amd import abi, configs
from opencafemol_amd.context import Batch, Context

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()
    yield


def _relaxed_hps(seed=3):
    s, it, x = configs.hps_slab(n_chains=16, chain_len=40, box=(9.0, 9.0, 30.0), slab_z=6.0, seed=seed)
    with Context(s, it) as c:
        c.set_positions(x)
        c.minimize(tolerance=50.0, max_iterations=400)
        x = c.get_state(abi.STATE_POSITIONS)["positions"]
    return s, it, x


def _run(s, it, x, nsteps, **opt):
    # the inner list belongs to the standard path (resident stepping keeps its own list)
    with Context(s, it, resident=0, **opt) as c:
        c.set_positions(x)
        c.set_velocities_to_temperature(it.temperature, 9)
        assert c.step(nsteps) == abi.OCM_OK
        st = c.get_state(abi.STATE_POSITIONS | abi.STATE_VELOCITIES | abi.STATE_FORCES)
        inner = c.inner_list_stats()
        entries = c.list_stats()[0]
    return st, inner, entries


def _same(a, b):
    for k in ("positions", "velocities", "forces"):
        assert np.array_equal(a[k], b[k]), k


def test_prune_interval_does_not_change_bits():
    s, it, x = _relaxed_hps()
    ref, inner, _ = _run(s, it, x, 600, prune_interval=0, chunk_steps=64)
    assert inner == (0, 0)
    for p in (2, 4, 7):
        st, (entries, used), outer = _run(s, it, x, 600, prune_interval=p, chunk_steps=64)
        _same(st, ref)
        assert 1 <= used <= p   # a margin overrun lowers it for a while
        # the inner rows hold only pairs near the cutoff: fewer than the outer ones
        assert 0 < entries < outer


def test_margin_overrun_rolls_back_and_keeps_bits():
    # a 64-step prune interval with a wide skin: beads cross margin/2 = 0.05 nm
    # long before the outer list goes stale, so chunks are redone with shorter
    # prune intervals
    s, it, x = _relaxed_hps(5)
    opt = dict(skin=1.0, rebuild_interval=200, chunk_steps=200)
    ref, _, _ = _run(s, it, x, 400, prune_interval=0, **opt)
    st, (entries, used), _ = _run(s, it, x, 400, prune_interval=64, **opt)
    _same(st, ref)
    assert 1 <= used < 64


def test_batch_and_go_chain_with_inner_list():
    reps = configs.go_replicas(4, n_min=60, n_max=200, seed=2)
    out = []
    for p in (0, 4):
        with Batch([r[0] for r in reps], [r[1] for r in reps], prune_interval=p, chunk_steps=50, resident=0) as b:
            for k, (s, it, x) in enumerate(reps):
                b.set_positions(k, x)
            b.step(300)
            out.append([b.state(k, abi.STATE_POSITIONS | abi.STATE_VELOCITIES) for k in range(len(reps))])
    for a, c in zip(*out):
        assert np.array_equal(a["positions"], c["positions"])
        assert np.array_equal(a["velocities"], c["velocities"])


def test_small_list_keeps_no_inner_list():
    # a 62-bead chain's whole list is read within the force pass's launch
    # latency: no inner list is kept, whatever prune_interval asks
    s, it, x = configs.go_chain(62, seed=2)
    with Context(s, it, resident=0) as c:
        c.set_positions(x)
        assert c.step(100) == abi.OCM_OK
        assert c.inner_list_stats() == (0, 0)
