"""Fixtures of the reference's own sample inputs, for GPU boxes that do not
hold the reference.

    python tests/golden/reference_inputs.py      # writes the .npz files here

Each sample is read by this repository's C++ front end (the readers that
tests/test_frontend.py and tests/test_genesis.py pin against the same files)
and its numbers -- masses, box, positions, velocities, every force's
arrays, the integrator -- are stored as a compressed .npz:
  - fus_100_hps.npz: sample/toml_input/FUS_100_HPS.toml (16,300 HPS beads,
    periodic);
  - 1srl_cg.npz: sample/genesis_input/1SRL_cg.inp with its .top/.itp/.gro.
The files hold numbers only.  `load(name)` rebuilds the SystemSpec,
IntegratorSpec, positions and velocities.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/sample"
SAMPLES = {
    "fus_100_hps": ("toml", os.path.join(REF, "toml_input", "FUS_100_HPS.toml")),
    "1srl_cg": ("genesis", os.path.join(REF, "genesis_input", "1SRL_cg.inp")),
}
FIELDS = ("indices", "params", "globals", "particle_params", "category", "allow", "exclusions")


def write(name):
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from opencafemol_amd import host_abi
    kind, path = SAMPLES[name]
    loader = host_abi.load_toml if kind == "toml" else host_abi.load_genesis
    s, it, x, v = loader(path)
    out = {"masses": s.masses, "box": np.array(s.box if s.box else [0.0, 0.0, 0.0]),
           "periodic": np.array(s.box is not None), "positions": x,
           "velocities": v if v is not None else np.zeros((0, 3)),
           "integrator": np.array([it.temperature, it.friction, it.dt, it.seed], dtype=np.float64),
           "n_forces": np.array(len(s.forces)),
           "group_names": np.array(sorted(s.group_ids, key=s.group_ids.get)),
           "group_ids": np.array(sorted(s.group_ids.values()))}
    for k, f in enumerate(s.forces):
        out[f"f{k}_meta"] = np.array([f.kind, f.group, int(f.use_periodic), f.n_categories, f.cutoff,
                                      f.switch_distance], dtype=np.float64)
        for fld in FIELDS:
            a = getattr(f, fld)
            if a is not None:
                out[f"f{k}_{fld}"] = np.asarray(a)
    dst = os.path.join(HERE, name + ".npz")
    np.savez_compressed(dst, **out)
    return dst


def load(name):
    from opencafemol_amd.model import ForceSpec, IntegratorSpec, SystemSpec
    z = np.load(os.path.join(HERE, name + ".npz"))
    s = SystemSpec(masses=z["masses"], box=tuple(z["box"]) if bool(z["periodic"]) else None)
    for k in range(int(z["n_forces"])):
        kind, group, periodic, ncat, cutoff, sw = z[f"f{k}_meta"]
        f = ForceSpec(kind=int(kind), name=f"force{k}", use_periodic=bool(periodic), n_categories=int(ncat),
                      cutoff=float(cutoff), switch_distance=float(sw))
        for fld in FIELDS:
            if f"f{k}_{fld}" in z:
                setattr(f, fld, z[f"f{k}_{fld}"])
        f.group = int(group)
        s.forces.append(f)
    s.group_ids = {str(n): int(g) for n, g in zip(z["group_names"], z["group_ids"])}
    t, fr, dt, seed = z["integrator"]
    it = IntegratorSpec(float(t), float(fr), float(dt), int(seed))
    v = z["velocities"]
    return s, it, z["positions"], (v if len(v) else None)


if __name__ == "__main__":
    for n in SAMPLES:
        print(write(n))
