"""ctypes face of include/ocm_host.h (libocm_host.so): the C++ TOML and
GENESIS front ends.

`load_toml(path)` and `load_genesis(inp_path)` return a model.SystemSpec, an
IntegratorSpec, the initial positions (nm) and velocities (nm/ps or None)
exactly as the C++ readers and generators produced them.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .model import ForceSpec, IntegratorSpec, SystemSpec

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libocm_host.so")
_lib = None


def load_library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m opencafemol_amd.build`")
        abi.load_library()
        L = C.CDLL(LIB_PATH)
        L.ocmh_load_toml.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.ocmh_load_toml.restype = C.c_int
        L.ocmh_load_genesis.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.ocmh_load_genesis.restype = C.c_int
        L.ocmh_free.argtypes = [C.c_void_p]
        L.ocmh_system.argtypes = [C.c_void_p]
        L.ocmh_system.restype = C.POINTER(abi.SystemDesc)
        L.ocmh_integrator.argtypes = [C.c_void_p]
        L.ocmh_integrator.restype = C.POINTER(abi.IntegratorDesc)
        L.ocmh_positions.argtypes = [C.c_void_p]
        L.ocmh_positions.restype = C.POINTER(C.c_double)
        L.ocmh_velocities.argtypes = [C.c_void_p]
        L.ocmh_velocities.restype = C.POINTER(C.c_double)
        L.ocmh_group_of.argtypes = [C.c_void_p, C.c_char_p]
        L.ocmh_group_of.restype = C.c_int32
        L.ocmh_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _arr(ptr, n, dtype):
    if not ptr or n == 0:
        return None
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype).copy()


def load_toml(path):
    return _load(load_library().ocmh_load_toml, path)


def load_genesis(inp_path):
    return _load(load_library().ocmh_load_genesis, inp_path)


def _load(loader, path):
    L = load_library()
    h = C.c_void_p()
    code = loader(path.encode(), C.byref(h))
    if code != abi.OCM_OK:
        raise abi.OcmError(code, L.ocmh_last_error().decode())
    try:
        sd = L.ocmh_system(h).contents
        n = sd.n_particles
        box = tuple(sd.box) if sd.use_periodic_box else None
        spec = SystemSpec(masses=_arr(sd.masses, n, np.float64), box=box)
        for k in range(sd.n_forces):
            f = sd.forces[k]
            fs = ForceSpec(kind=f.kind, name=f"force{k}", use_periodic=bool(f.use_periodic))
            if f.kind in abi.BONDED_ARITY:
                ar, npar = abi.BONDED_ARITY[f.kind], abi.BONDED_NPARAMS[f.kind]
                fs.indices = (_arr(f.indices, ar * f.n_terms, np.int32) if f.n_terms else np.zeros(0, np.int32)).reshape(-1, ar)
                fs.params = (_arr(f.params, npar * f.n_terms, np.float64) if f.n_terms else np.zeros(0)).reshape(-1, npar)
            else:
                npp = abi.PAIR_NPP[f.kind]
                fs.particle_params = (_arr(f.particle_params, npp * n, np.float64).reshape(n, npp) if npp
                                      else np.zeros((n, 0)))
                fs.cutoff = f.cutoff
                fs.switch_distance = f.switch_distance
                if f.category:
                    fs.category = _arr(f.category, n, np.int32)
                    fs.n_categories = f.n_categories
                    if f.category_allow:
                        fs.allow = _arr(f.category_allow, f.n_categories ** 2, np.uint8).reshape(f.n_categories, -1)
                ne = f.n_exclusions
                fs.exclusions = _arr(f.exclusions, 2 * ne, np.int32).reshape(-1, 2) if ne else None
            fs.globals = _arr(f.globals, f.n_globals, np.float64)
            fs.group = f.group
            spec.forces.append(fs)
        it = L.ocmh_integrator(h).contents
        integ = IntegratorSpec(it.temperature, it.friction, it.dt, it.seed)
        pos = _arr(L.ocmh_positions(h), 3 * n, np.float64)
        vel = _arr(L.ocmh_velocities(h), 3 * n, np.float64)
        groups = {}
        for name in ("HarmonicBond", "GaussianBond", "GoContact", "HarmonicAngle", "FlexibleLocalAngle",
                     "GaussianDihedral", "FlexibleLocalDihedral", "ExcludedVolume", "DebyeHuckel",
                     "LennardJonesRepulsive", "LennardJonesAttractive", "CappedGoContact", "CosineDihedral",
                     "WeeksChandlerAndersen", "UniformWeeksChadlerAndersen", "UniformLennardJonesAttractive",
                     "iSoLFAttractive"):
            g = L.ocmh_group_of(h, name.encode())
            if g >= 0:
                groups[name] = g
        spec.group_ids = groups
        return spec, integ, None if pos is None else pos.reshape(n, 3), None if vel is None else vel.reshape(n, 3)
    finally:
        L.ocmh_free(h)
