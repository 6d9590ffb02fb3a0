"""Writes the sh3 system of tests/golden/sh3_native.json as a GENESIS input
set (.inp, .top with an #include'd .itp and parameter file, .gro), in the
fixed column layout the GENESIS readers take, for the C++ front-end tests.

The numbers are the ones the TOML reader produces for the same protein, put
back into GENESIS units, so both readers should build the same potentials:
  [ bonds ]         coef = 2 k (the reader halves it)
  [ angles ] 21     coef = -k (kJ/mol), eq and w in nm
  [ dihedrals ] 41  coef = -k (kJ/mol), eq in degrees, w in rad
  [ pairs ]         coef = k (kJ/mol), eq in nm
  [ atomtypes ]     rmin = 2 x radius (nm), eps in kJ/mol, one value
Residue names come from the FLA (middle bead) and FLD (central pair) entries;
the two end beads, which no such entry names, take the residue whose mass and
radius they share.
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KJ = 4.184


def _residues(d):
    n = len(d["masses"])
    res = {}
    for ff in d["local"]:
        for p in ff["parameters"]:
            if "y" in p:
                res[p["indices"][1]] = p["y"][3:6]
            if "coef" in p:
                a, b = p["coef"].split("-")
                res[p["indices"][1]], res[p["indices"][2]] = a, b
    radius = {p["index"]: p["radius"] for p in d["global"][0]["parameters"]}
    for end in (0, n - 1):
        if end in res:
            continue
        same = [r for i, r in res.items() if d["masses"][i] == d["masses"][end] and radius[i] == radius[end]]
        res[end] = same[0] if same else ("NTR" if end == 0 else "CTR")
    return [res[i] for i in range(n)], [radius[i] for i in range(n)]


def write(directory, nsteps=1000, crdout_period=100, boundary=None, timestep=0.01, iseed=12345,
          pdbfile="sh3_init.pdb", platform=None):
    """Writes sh3.inp and the files it names into `directory`; returns the
    .inp path.  `boundary` is None (NOBC) or a box (x, y, z) in Angstrom."""
    d = json.load(open(os.path.join(HERE, "sh3_native.json")))
    n = len(d["masses"])
    names, radii = _residues(d)
    os.makedirs(os.path.join(directory, "param"), exist_ok=True)
    pos_nm = np.array(d["positions_A"]) * 0.1

    types = {}
    for r, rad in zip(names, radii):
        assert types.setdefault(r, rad) == rad, f"{r} has two radii"
    eps = d["global"][0]["epsilon"] * KJ
    with open(os.path.join(directory, "param", "atom_types.itp"), "w") as f:
        f.write("[ atomtypes ]\n;name  n    mass  charge ptype       rmin      eps\n")
        for r, rad in types.items():
            f.write(f"{r:>5s}{6:3d}{100.0:8.3f}{0.0:8.3f}{'A':>6s}{2 * rad * 0.1:10.6f} {eps:9.4f}\n")

    lines = ["[ moleculetype ]", ";name            nrexcl", f"{'sh3':<13s}{3:10d}", "", "[ atoms ]",
             ";       nr type     resnr  res atom   cg   charge     mass"]
    for i, (r, m) in enumerate(zip(names, d["masses"])):
        lines.append(f"{i + 1:10d}{r:>5s}{i + 1:10d}  {r:3s} {'CA':>4s}{1:5d}{0.0:9.3f}{m:9.3f}")
    by = {(ff["interaction"], ff["potential"]): ff["parameters"] for ff in d["local"]}
    lines += ["", "[ bonds ]"]
    for p in by[("BondLength", "Harmonic")]:
        i, j = p["indices"]
        k = p["k"] * KJ * 100.0 * 2.0
        lines.append(f"{i + 1:10d}{j + 1:10d}{1:5d}{p['v0'] * 0.1:18.10E}{2 * k:18.10E}")
    lines += ["", "[ angles ]"]
    for p in by[("BondLength", "Gaussian")]:
        i, k = p["indices"]
        lines.append(f"{i + 1:10d}{i + 2:10d}{k + 1:10d}{21:5d}{p['v0'] * 0.1:15.8E}{-p['k'] * KJ:15.8E}"
                     f"{p['σ'] * 0.1:15.8E}")
    for p in by[("BondAngle", "FlexibleLocalAngle")]:
        i, j, k = p["indices"]
        lines.append(f"{i + 1:10d}{j + 1:10d}{k + 1:10d}{22:5d}")
    lines += ["", "[ dihedrals ]"]
    for p in by[("DihedralAngle", "Gaussian")]:
        i, j, k, l = p["indices"]
        lines.append(f"{i + 1:10d}{j + 1:10d}{k + 1:10d}{l + 1:10d}{41:5d}{math.degrees(p['v0']):15.8E}"
                     f"{-p['k'] * KJ:15.8E}{p['σ']:15.8E}")
    for p in by[("DihedralAngle", "FlexibleLocalDihedral")]:
        i, j, k, l = p["indices"]
        lines.append(f"{i + 1:10d}{j + 1:10d}{k + 1:10d}{l + 1:10d}{52:5d}")
    lines += ["", "[ pairs ]"]
    for p in by[("BondLength", "GoContact")]:
        i, j = p["indices"]
        lines.append(f"{i + 1:10d}{j + 1:10d}{2:10d}{p['v0'] * 0.1:15.8E}{p['k'] * KJ:15.8E}")
    with open(os.path.join(directory, "sh3.itp"), "w") as f:
        f.write("\n".join(lines) + "\n")

    with open(os.path.join(directory, "sh3.top"), "w") as f:
        f.write('; common parameters\n#include "./param/atom_types.itp" \n\n#include "sh3.itp"\n\n'
                "[ system ]\nsh3\n\n[ molecules ]\nsh3  1\n")

    with open(os.path.join(directory, "sh3.gro"), "w") as f:
        f.write(f"sh3 native, t = 0.0\n{n:12d}\n")
        for i, (r, x) in enumerate(zip(names, pos_nm)):
            f.write(f"{i + 1:5d}{r:>5s}{'CA':>5s}{i + 1:5d}{x[0]:9.4f}{x[1]:9.4f}{x[2]:9.4f}\n")
        f.write("   0.0   0.0   0.0\n")

    inp = ["[INPUT]", "grotopfile = sh3.top   # topology", "grocrdfile = sh3.gro", "", "[OUTPUT]",
           "dcdfile = sh3_md.dcd"]
    if pdbfile:
        inp.append(f"pdbfile = {pdbfile}")
    inp += ["", "[DYNAMICS]", "integrator = VVER_CG", f"nsteps = {nsteps}", f"timestep = {timestep}",
            f"crdout_period = {crdout_period}"]
    if iseed is not None:
        inp.append(f"iseed = {iseed}")
    inp += ["", "[ENSEMBLE]", f"temperature = {d['temperature']}", "gamma_t = 0.01", "", "[BOUNDARY]"]
    if boundary is None:
        inp.append("type = NOBC")
    else:
        inp += ["type = PBC"] + [f"box_size_{a} = {v}" for a, v in zip("xyz", boundary)]
    if platform:
        inp += ["", "[PLATFORM]", f"name = {platform}"]
    path = os.path.join(directory, "sh3.inp")
    with open(path, "w") as f:
        f.write("\n".join(inp) + "\n")
    return path


def residues():
    """Residue name of each bead, as written into [ atoms ]."""
    return _residues(json.load(open(os.path.join(HERE, "sh3_native.json"))))[0]
