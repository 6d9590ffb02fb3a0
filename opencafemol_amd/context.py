"""Python handle over the C ABI: Context (one system) and Batch (many)."""
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

    def status(self, s):
        return self.lib.ocm_batch_system_status(self.h, s)

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

    def set_timing(self, on):
        check(self.lib.ocm_set_timing(self.h, int(bool(on))))

    def timing(self, kernel):
        ms = C.c_double()
        n = C.c_int64()
        check(self.lib.ocm_get_timing(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def list_stats(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.ocm_get_list_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value


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
