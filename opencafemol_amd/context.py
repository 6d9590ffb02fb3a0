"""Python handle over the C ABI: Context (one system), Batch (many), Server (job stream)."""
import ctypes as C

import numpy as np

from . import abi
from .abi import check


def default_options(**kw):
    lib = abi.load_library()
    o = abi.Options()
    lib.ocm_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Batch:
    """Independent systems packed into one set of device arrays."""

    def __init__(self, systems, integrators, **options):
        self.lib = abi.load_library()
        self.systems = list(systems)
        n = len(self.systems)
        sds = (abi.SystemDesc * n)()
        ids = (abi.IntegratorDesc * n)()
        self._keep = []
        for k, (s, it) in enumerate(zip(self.systems, integrators)):
            sd, keep = s.descriptor()
            self._keep.append(keep)
            sds[k] = sd
            ids[k] = it.descriptor()
        opt = default_options(**options)
        h = C.c_void_p()
        check(self.lib.ocm_batch_create(sds, ids, n, C.byref(opt), C.byref(h)))
        self.h = h
        self._keep = None  # descriptors were copied by the library

    def close(self):
        if getattr(self, "h", None):
            self.lib.ocm_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_positions(self, s, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        check(self.lib.ocm_batch_set_positions(self.h, s, x.ctypes.data_as(C.POINTER(C.c_double)), x.shape[0]))

    def set_velocities(self, s, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        check(self.lib.ocm_batch_set_velocities(self.h, s, v.ctypes.data_as(C.POINTER(C.c_double)), v.shape[0]))

    def step(self, n):
        code = self.lib.ocm_step(self.h, int(n))
        if code not in (abi.OCM_OK, abi.OCM_ERR_NUMERIC):
            check(code)
        return code

    def prepare(self, n):
        """Captures the graphs step(n) will launch, without stepping."""
        check(self.lib.ocm_prepare_steps(self.h, int(n)))

    def status(self, s):
        return self.lib.ocm_batch_system_status(self.h, s)

    def launch_shape(self):
        v = [C.c_int32() for _ in range(4)]
        check(self.lib.ocm_get_launch_shape(self.h, *[C.byref(x) for x in v]))
        return {"team": v[0].value, "entries_per_lane": v[1].value, "grid": v[2].value, "prune_team": v[3].value}

    def resident_threads(self):
        """Workgroup size of resident stepping, 0 on the standard path."""
        v = C.c_int32()
        check(self.lib.ocm_get_resident(self.h, C.byref(v)))
        return v.value

    def write_checkpoint(self, s, path):
        check(self.lib.ocm_batch_write_checkpoint(self.h, s, str(path).encode()))

    def read_checkpoint(self, s, path):
        check(self.lib.ocm_batch_read_checkpoint(self.h, s, str(path).encode()))

    def state(self, s, what=15, groups=0xFFFFFFFF):
        n = self.systems[s].n
        st = abi.State()
        out = {}
        arrs = {}
        for key, bit in (("positions", 1), ("velocities", 2), ("forces", 4)):
            if what & bit:
                arrs[key] = np.zeros((n, 3))
                setattr(st, key, arrs[key].ctypes.data_as(C.POINTER(C.c_double)))
        check(self.lib.ocm_batch_get_state(self.h, s, what, groups, C.byref(st)))
        out.update(arrs)
        if what & abi.STATE_ENERGY:
            out["energy_by_group"] = np.array(st.energy_by_group[:])
            out["potential"] = st.potential_energy
            out["kinetic"] = st.kinetic_energy
        out["step"] = st.step
        return out

    def minimize(self, tolerance=10.0, max_iterations=0):
        check(self.lib.ocm_minimize(self.h, float(tolerance), int(max_iterations)))

    def set_timing(self, mode):
        """0 off, 1 (or True) HIP events around every launch, 2 device-clock
        stamps of every pair-kernel launch inside the step graphs."""
        check(self.lib.ocm_set_timing(self.h, int(mode)))

    def timing(self, kernel):
        ms = C.c_double()
        n = C.c_int64()
        check(self.lib.ocm_get_timing(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def list_counts(self):
        """{outer_valid, outer_padded, inner_valid, inner_padded} entry counts."""
        c = (C.c_int64 * 4)()
        check(self.lib.ocm_get_list_counts(self.h, c))
        return {"outer_valid": c[0], "outer_padded": c[1], "inner_valid": c[2], "inner_padded": c[3]}

    def list_stats(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.ocm_get_list_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def inner_list_stats(self):
        """(padded inner-list entries, prune interval in use; 0 = inner list off)."""
        a, p = C.c_int64(), C.c_int32()
        check(self.lib.ocm_get_inner_list_stats(self.h, C.byref(a), C.byref(p)))
        return a.value, p.value


class Context(Batch):
    """OpenMM::Context equivalent for one system (src/Simulator.cpp:17)."""

    def __init__(self, system, integrator, **options):
        super().__init__([system], [integrator], **options)
        self.system = system

    def set_positions(self, x):
        super().set_positions(0, x)

    def set_velocities(self, v):
        super().set_velocities(0, v)

    def set_velocities_to_temperature(self, T, seed):
        check(self.lib.ocm_set_velocities_to_temperature(self.h, float(T), int(seed)))

    def get_state(self, what=15, groups=0xFFFFFFFF):
        return self.state(0, what, groups)

    def neighbor_pairs(self, force_index, cap=None):
        cnt = C.c_int64()
        if cap is None:
            check(self.lib.ocm_get_neighbor_pairs(self.h, force_index, None, 0, C.byref(cnt)))
            cap = cnt.value
        out = np.zeros((max(cap, 1), 2), dtype=np.int32)
        check(self.lib.ocm_get_neighbor_pairs(self.h, force_index, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                              cap, C.byref(cnt)))
        return out[:min(cap, cnt.value)]


class Server:
    """Continuous-batching job server (``ocm_server_*``).

    ``submit`` queues a job and returns its id; ``next_frame`` pops a queued
    trajectory frames; ``wait`` collects a job's final state.  The worker
    stops when the frame queue is full, so a caller that waits on jobs must
    also drain frames (``drain`` does both).
    """

    def __init__(self, hbm_budget_bytes=0, max_active_jobs=0, max_queued_frames=0, quantum_steps=0,
                 **engine_options):
        self.lib = abi.load_library()
        o = abi.ServerOptions()
        self.lib.ocm_server_options_default(C.byref(o))
        for k, v in engine_options.items():
            setattr(o.engine, k, v)
        if hbm_budget_bytes:
            o.hbm_budget_bytes = hbm_budget_bytes
        if max_active_jobs:
            o.max_active_jobs = max_active_jobs
        if max_queued_frames:
            o.max_queued_frames = max_queued_frames
        if quantum_steps:
            o.quantum_steps = quantum_steps
        h = C.c_void_p()
        check(self.lib.ocm_server_create(C.byref(o), C.byref(h)))
        self.h = h
        self.n_of = {}
        self.max_n = 1

    def close(self):
        if getattr(self, "h", None):
            self.lib.ocm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, system, integrator, positions, velocities=None, nsteps=1, frame_interval=0):
        sd, keep = system.descriptor()
        it = integrator.descriptor()
        x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        vp = None
        if velocities is not None:
            v = np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1, 3)
            vp = v.ctypes.data_as(C.POINTER(C.c_double))
        jid = C.c_int64()
        check(self.lib.ocm_server_submit(self.h, C.byref(sd), C.byref(it), x.ctypes.data_as(C.POINTER(C.c_double)),
                                         vp, x.shape[0], int(nsteps), int(frame_interval), C.byref(jid)))
        del keep
        self.n_of[jid.value] = x.shape[0]
        self.max_n = max(self.max_n, x.shape[0])
        return jid.value

    def next_frame(self, timeout_ms=-1):
        """(job, step, positions) or None on timeout."""
        cap = self.max_n
        buf = np.zeros((cap, 3))
        job, step, n = C.c_int64(), C.c_int64(), C.c_int32()
        code = self.lib.ocm_server_next_frame(self.h, int(timeout_ms), C.byref(job), C.byref(step),
                                              buf.ctypes.data_as(C.POINTER(C.c_double)), cap, C.byref(n))
        if code == abi.OCM_ERR_TIMEOUT:
            return None
        check(code)
        return job.value, step.value, buf[:n.value].copy()

    def wait(self, job, timeout_ms=-1):
        """{"status", "positions", "velocities", "steps", "submit_ms", "start_ms", "end_ms"} or None."""
        n = self.n_of[job]
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        info = abi.JobInfo()
        code = self.lib.ocm_server_wait(self.h, int(job), int(timeout_ms), x.ctypes.data_as(C.POINTER(C.c_double)),
                                        v.ctypes.data_as(C.POINTER(C.c_double)), n, C.byref(info))
        if code == abi.OCM_ERR_TIMEOUT:
            return None
        check(code)
        del self.n_of[job]
        return {"status": info.status, "positions": x, "velocities": v, "steps": info.steps_done,
                "submit_ms": info.submit_ms, "start_ms": info.start_ms, "end_ms": info.end_ms}

    def drain(self, jobs, on_frame=None):
        """Collects frames until every job in ``jobs`` has ended; returns {job: result}."""
        pending = set(jobs)
        results = {}
        while pending:
            f = self.next_frame(timeout_ms=20)
            if f is not None:
                if on_frame is not None:
                    on_frame(*f)
                continue
            for j in list(pending):
                r = self.wait(j, timeout_ms=0)
                if r is not None:
                    results[j] = r
                    pending.discard(j)
        while True:
            f = self.next_frame(timeout_ms=0)
            if f is None:
                break
            if on_frame is not None:
                on_frame(*f)
        return results

    def stats(self):
        st = abi.ServerStats()
        check(self.lib.ocm_server_get_stats(self.h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in abi.ServerStats._fields_}
