"""ocm_get_neighbor_pairs_noalloc against ocm_get_neighbor_pairs.

Two calls in a row through the C ABI: a 5,600-bead HPS slab (both of its
nonbonded forces) after 40 steps, then a smaller 120-bead Go chain.  The
scratch array keeps the first call's contents, so the second call also shows
that the count starts from zero and no stale pairs come back.  The pair sets
must be equal bit for bit, in the same (i, j) order.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from opencafemol_amd import abi, configs
from opencafemol_amd.context import Context

pytestmark = pytest.mark.gpu


def _noalloc(c, force_index):
    cnt = C.c_int64(-1)
    abi.check(c.lib.ocm_get_neighbor_pairs_noalloc(c.h, force_index, None, 0, C.byref(cnt)))
    n = cnt.value
    out = np.full((max(n, 1), 2), -7, dtype=np.int32)
    abi.check(c.lib.ocm_get_neighbor_pairs_noalloc(c.h, force_index, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   n, C.byref(cnt)))
    assert cnt.value == n
    return out[:n]


def _pair_forces(s):
    return [k for k, f in enumerate(s.forces) if f.kind in (abi.EXCLUDED_VOLUME, abi.DEBYE_HUCKEL,
                                                              abi.LJ_REPULSIVE)]


def test_noalloc_pairs_equal_allocating_pairs_twice():
    cases = [("hps_slab_5600", configs.hps_slab_of_size(5600, seed=2), 40),
             ("go_chain_120", configs.go_chain(120, seed=5), 10)]
    sizes = []
    for name, (s, it, x), nsteps in cases:
        sys.__stderr__.write(f"[gpu-case] no-alloc pairs: {name}, {s.n} beads\n")
        sys.__stderr__.flush()
        with Context(s, it) as c:
            c.set_positions(x)
            c.minimize(tolerance=50.0, max_iterations=300)   # the packed start overlaps
            c.set_velocities_to_temperature(it.temperature, 3)
            assert c.step(nsteps) == abi.OCM_OK
            for fi in _pair_forces(s):
                got = _noalloc(c, fi)
                want = c.neighbor_pairs(fi)
                assert len(want) > 0
                np.testing.assert_array_equal(got, want)
                sizes.append(len(got))
                # a cap below the count is refused, with the count reported
                cnt = C.c_int64(-1)
                out = np.zeros((len(want), 2), dtype=np.int32)
                code = c.lib.ocm_get_neighbor_pairs_noalloc(c.h, fi, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                                            len(want) - 1, C.byref(cnt))
                assert code == abi.OCM_ERR_INVALID and cnt.value == len(want)
    assert sizes[-1] < sizes[0]
