"""System descriptors in OpenMM units, the Python face of ocm_system_desc.

A ``SystemSpec`` is what the reference's ``SystemGenerator::generate``
(src/SystemGenerator.cpp:34-111) hands to OpenMM: masses, an optional cuboid
box and one force per generator, each tagged with a force group.  Group ids
follow ``SystemGenerator::add_ff_generator`` (src/SystemGenerator.cpp:4-32):
one id per generator name in insertion order, with every FlexibleLocalAngle
and every FlexibleLocalDihedral generator sharing one id each.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class ForceSpec:
    kind: int
    name: str
    use_periodic: bool = False
    indices: np.ndarray = None          # bonded: (n_terms, arity) int32
    params: np.ndarray = None           # bonded: (n_terms, nparams) float64
    globals: np.ndarray = None          # kind-specific constants
    particle_params: np.ndarray = None  # nonbonded: (n, npp) float64
    category: np.ndarray = None         # nonbonded: (n,) int32, -1 non-participant
    n_categories: int = 0
    allow: np.ndarray = None            # (n_cat, n_cat) uint8
    cutoff: float = 0.0
    switch_distance: float = 0.0
    exclusions: np.ndarray = None       # (n_excl, 2) int32
    group: int = -1                     # filled by SystemSpec.add

    @property
    def n_terms(self):
        return 0 if self.indices is None else int(self.indices.shape[0])


def group_key(name):
    if "FlexibleLocalAngle" in name:
        return "FlexibleLocalAngle"
    if "FlexibleLocalDihedral" in name:
        return "FlexibleLocalDihedral"
    if "CombinatorialGoContact" in name:
        return "CombinatorialGoContact"
    return name


@dataclass
class SystemSpec:
    masses: np.ndarray
    box: tuple = None                   # nm, periodic cuboid when set
    forces: list = field(default_factory=list)
    group_ids: dict = field(default_factory=dict)

    @property
    def n(self):
        return int(len(self.masses))

    def add(self, force):
        key = group_key(force.name)
        if key not in self.group_ids:
            self.group_ids[key] = len(self.group_ids)
        force.group = self.group_ids[key]
        self.forces.append(force)
        return force

    def descriptor(self):
        """Returns (SystemDesc, keepalive); keep `keepalive` while the desc is used."""
        keep = []

        def ptr(arr, ctype, dtype):
            if arr is None:
                return None
            a = np.ascontiguousarray(arr, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        fds = (abi.ForceDesc * max(len(self.forces), 1))()
        for k, f in enumerate(self.forces):
            d = fds[k]
            d.kind = f.kind
            d.group = f.group
            d.use_periodic = int(bool(f.use_periodic))
            d.n_terms = f.n_terms
            d.indices = ptr(f.indices, C.c_int32, np.int32)
            d.params = ptr(f.params, C.c_double, np.float64)
            d.n_globals = 0 if f.globals is None else int(np.size(f.globals))
            d.globals = ptr(f.globals, C.c_double, np.float64)
            d.particle_params = ptr(f.particle_params, C.c_double, np.float64)
            d.category = ptr(f.category, C.c_int32, np.int32)
            d.n_categories = int(f.n_categories)
            d.category_allow = ptr(f.allow, C.c_uint8, np.uint8)
            d.cutoff = float(f.cutoff)
            d.switch_distance = float(f.switch_distance)
            ne = 0 if f.exclusions is None else int(np.asarray(f.exclusions).reshape(-1, 2).shape[0])
            d.n_exclusions = ne
            d.exclusions = ptr(None if ne == 0 else np.asarray(f.exclusions).reshape(-1, 2), C.c_int32, np.int32)
        keep.append(fds)
        sd = abi.SystemDesc()
        sd.n_particles = self.n
        sd.n_forces = len(self.forces)
        sd.masses = ptr(self.masses, C.c_double, np.float64)
        sd.use_periodic_box = int(self.box is not None)
        if self.box is not None:
            for c in range(3):
                sd.box[c] = float(self.box[c])
        sd.forces = C.cast(fds, C.POINTER(abi.ForceDesc))
        keep.append(sd)
        return sd, keep


@dataclass
class IntegratorSpec:
    temperature: float
    friction: float
    dt: float
    seed: int = 0

    def descriptor(self):
        return abi.IntegratorDesc(self.temperature, self.friction, self.dt, self.seed)


def all_groups_mask(system):
    m = 0
    for f in system.forces:
        m |= 1 << f.group
    return m
