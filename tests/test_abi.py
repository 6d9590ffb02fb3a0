"""The C-ABI library loads, exports every declared entry point, lays out its
structs as the ctypes mirror does, and rejects malformed descriptors with the
reference's error behaviour -- all without a GPU compute call."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from opencafemol_amd import abi, configs
from opencafemol_amd.model import ForceSpec, IntegratorSpec, SystemSpec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(ocm_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    syms = declared_symbols()
    assert syms, "no declarations found"
    assert syms == set(abi.EXPORTED)
    for s in syms:
        assert hasattr(lib, s), s


def test_struct_layout_matches_c(oracle):
    L = oracle.lib()
    L.oracle_abi_layout.argtypes = [C.POINTER(C.c_int64), C.c_int32]
    out = (C.c_int64 * 32)()
    n = L.oracle_abi_layout(out, 32)
    got = list(out[:n])
    F, S, O, T = abi.ForceDesc, abi.SystemDesc, abi.Options, abi.State
    expect = [C.sizeof(F), F.indices.offset, F.globals.offset, F.particle_params.offset, F.cutoff.offset,
              F.n_exclusions.offset, F.exclusions.offset, C.sizeof(S), S.box.offset, S.forces.offset,
              C.sizeof(abi.IntegratorDesc), C.sizeof(O), O.skin.offset, C.sizeof(T),
              T.energy_by_group.offset, T.step.offset]
    SO, SS, J = abi.ServerOptions, abi.ServerStats, abi.JobInfo
    expect += [C.sizeof(SO), SO.hbm_budget_bytes.offset, SO.quantum_steps.offset, C.sizeof(SS),
               SS.latency_p50_ms.offset, SS.budget_bytes.offset, C.sizeof(J), J.steps_done.offset,
               J.end_ms.offset]
    assert got == expect


def test_options_default_and_version():
    lib = abi.load_library()
    o = abi.Options()
    lib.ocm_options_default(C.byref(o))
    assert (o.rebuild_interval, o.skin, o.use_graph, o.chunk_steps, o.prune_interval) == (40, 0.7, 1, 200, 5)
    assert b"gfx950" in lib.ocm_version()


def _create(system):
    lib = abi.load_library()
    sd, keep = system.descriptor()
    it = IntegratorSpec(300.0, 0.5, 0.01).descriptor()
    h = C.c_void_p()
    code = lib.ocm_context_create(C.byref(sd), C.byref(it), None, C.byref(h))
    if code == abi.OCM_OK:
        lib.ocm_context_destroy(h)
    return code, lib.ocm_last_error().decode()


def test_entries_per_lane_needs_a_matching_team():
    # 2 or 1 entries per lane exist for teams of 8, 16 and 32 only: the
    # options are checked before any device is touched
    lib = abi.load_library()
    s = SystemSpec(masses=np.ones(3))
    sd, keep = s.descriptor()
    it = IntegratorSpec(300.0, 0.5, 0.01).descriptor()
    o = abi.Options()
    lib.ocm_options_default(C.byref(o))
    o.team_size, o.entries_per_lane = 64, 2
    h = C.c_void_p()
    assert lib.ocm_context_create(C.byref(sd), C.byref(it), C.byref(o), C.byref(h)) == abi.OCM_ERR_INVALID
    assert "entries_per_lane" in lib.ocm_last_error().decode()


def test_index_out_of_range_is_rejected():
    # ReadTOMLForceFieldGenerator.cpp:48-55 throws on an index beyond the system
    s = SystemSpec(masses=np.ones(3))
    s.add(ForceSpec(abi.HARMONIC_BOND, "HarmonicBond", indices=np.array([[0, 3]]), params=np.array([[0.38, 1.0]])))
    code, msg = _create(s)
    assert code == abi.OCM_ERR_INVALID and "out of range" in msg


@pytest.mark.parametrize("mutate,needle", [
    (lambda f: setattr(f, "cutoff", 0.0), "cutoff"),
    (lambda f: setattr(f, "globals", None), "globals"),
    (lambda f: setattr(f, "switch_distance", 5.0), "switch"),
])
def test_bad_nonbonded_descriptors(mutate, needle):
    s, it, x = configs.go_chain(20, seed=1)
    exv = [f for f in s.forces if f.kind == abi.EXCLUDED_VOLUME][0]
    mutate(exv)
    code, msg = _create(s)
    assert code == abi.OCM_ERR_INVALID and needle in msg


def test_participant_with_nan_parameter_is_rejected():
    s, it, x = configs.go_chain(20, seed=1)
    exv = [f for f in s.forces if f.kind == abi.EXCLUDED_VOLUME][0]
    exv.particle_params = exv.particle_params.copy()
    exv.particle_params[3] = np.nan
    code, msg = _create(s)
    assert code == abi.OCM_ERR_INVALID and "non-finite" in msg


def test_asymmetric_category_matrix_is_rejected():
    s, it, x = configs.go_chain(20, seed=1)
    exv = [f for f in s.forces if f.kind == abi.EXCLUDED_VOLUME][0]
    exv.category = np.array([0] * 10 + [1] * 10, dtype=np.int32)
    exv.n_categories = 2
    exv.allow = np.array([[1, 1], [0, 1]], dtype=np.uint8)
    code, msg = _create(s)
    assert code == abi.OCM_ERR_INVALID and "symmetric" in msg


def test_periodic_force_without_box_is_rejected():
    s, it, x = configs.go_chain(20, seed=1)
    s.forces[0].use_periodic = True
    code, msg = _create(s)
    assert code == abi.OCM_ERR_INVALID and "periodic" in msg


def test_no_device_is_reported_on_a_cpu_host():
    # a valid descriptor on a host without a gfx950 device reports NO_DEVICE
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    s, it, x = configs.go_chain(20, seed=1)
    code, msg = _create(s)
    assert code == abi.OCM_ERR_NO_DEVICE


def test_server_options_default():
    lib = abi.load_library()
    o = abi.ServerOptions()
    lib.ocm_server_options_default(C.byref(o))
    assert (o.hbm_budget_bytes, o.max_active_jobs, o.max_queued_frames, o.quantum_steps) == (0, 256, 4096, 2000)
    assert (o.engine.rebuild_interval, o.engine.chunk_steps) == (40, 200)


def test_server_reports_no_device_on_a_cpu_host():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = abi.load_library()
    h = C.c_void_p()
    assert lib.ocm_server_create(None, C.byref(h)) == abi.OCM_ERR_NO_DEVICE
    assert not h.value
    # calls on a NULL server are rejected, not dereferenced
    st = abi.ServerStats()
    assert lib.ocm_server_get_stats(None, C.byref(st)) == abi.OCM_ERR_INVALID
    assert lib.ocm_server_destroy(None) == abi.OCM_OK
