"""Builds the sh3 AICG2+ system from tests/golden/sh3_native.json with the
unit conversions of the reference's TOML readers:
  harmonic bond  k * 4.184 * 100 * 2, v0 * 0.1   ReadTOMLForceFieldGenerator.cpp:57-64
  Go contact     k * 4.184, v0 * 0.1             :195-198
  Gaussian bond  k * 4.184, v0 * 0.1, sigma*0.1  :127-133
  FLA            k * 4.184, residue = y[3:6]     :426-480
  Gaussian dih.  k * 4.184, v0, sigma (rad)      :531-536
  FLD            k * 4.184, class from coef      :636-766
  EXV            eps * 4.184, radius * 0.1, cutoff ratio 2.0 (default)
                 x (largest + second largest radius)   :1066-1121,
                 ExcludedVolumeForceFieldGenerator.cpp:31-83
"""
import json
import os

import numpy as np

from opencafemol_amd import abi, aicg2
from opencafemol_amd.aicg2_tables import FLA_Y
from opencafemol_amd.model import ForceSpec, IntegratorSpec, SystemSpec

HERE = os.path.dirname(os.path.abspath(__file__))
KJ = 4.184


def load():
    with open(os.path.join(HERE, "golden", "sh3_native.json")) as f:
        return json.load(f)


def native_positions_nm(d=None):
    d = d or load()
    return np.array(d["positions_A"], dtype=np.float64) * 0.1


def _pairs_within(edges, n, k):
    """Topology::ignore_list_within_edge for one edge type (src/Topology.cpp:75-135)."""
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    out = set()
    for s in range(n):
        seen = [s]

        def walk(node, dist):
            if dist == 0:
                return
            for nb in adj[node]:
                if nb not in seen:
                    seen.append(nb)
                    walk(nb, dist - 1)
        walk(s, k)
        for t in seen:
            if t != s:
                out.add((min(s, t), max(s, t)))
    return sorted(out)


def build(d=None):
    d = d or load()
    n = len(d["masses"])
    sysm = SystemSpec(masses=np.array(d["masses"], dtype=np.float64))
    edges = {"bond": [], "contact": []}
    for ff in d["local"]:
        pot, params = ff["potential"], ff["parameters"]
        idx = np.array([p["indices"] for p in params], dtype=np.int32)
        if ff["interaction"] == "BondLength" and pot == "Harmonic":
            sysm.add(ForceSpec(abi.HARMONIC_BOND, "HarmonicBond", indices=idx, params=np.array(
                [[p["v0"] * 0.1, p["k"] * KJ * 100.0 * 2.0] for p in params])))
        elif pot == "GoContact":
            sysm.add(ForceSpec(abi.GO_CONTACT, "GoContact", indices=idx, params=np.array(
                [[p["k"] * KJ, p["v0"] * 0.1] for p in params])))
        elif ff["interaction"] == "BondLength" and pot == "Gaussian":
            sysm.add(ForceSpec(abi.GAUSSIAN_BOND, "GaussianBond", indices=idx, params=np.array(
                [[p["k"] * KJ, p["v0"] * 0.1, p["σ"] * 0.1] for p in params])))
        elif pot == "FlexibleLocalAngle":
            res = {}
            for p in params:
                res[p["indices"][1]] = p["y"][3:6]
            residues = [res.get(i, "ALA") for i in range(n)]
            for f in aicg2.flexible_local_angle_forces(idx, residues, params[0]["k"]):
                sysm.add(f)
        elif ff["interaction"] == "DihedralAngle" and pot == "Gaussian":
            sysm.add(ForceSpec(abi.GAUSSIAN_DIHEDRAL, "GaussianDihedral", indices=idx, params=np.array(
                [[p["k"] * KJ, p["v0"], p["σ"]] for p in params])))
        elif pot == "FlexibleLocalDihedral":
            res = {}
            for p in params:
                a, b = p["coef"].split("-")
                res[p["indices"][1]], res[p["indices"][2]] = a, b
            residues = [res.get(i, "ALA") for i in range(n)]
            for f in aicg2.flexible_local_dihedral_forces(idx, residues, params[0]["k"]):
                sysm.add(f)
        if ff.get("topology") in edges:
            edges[ff["topology"]] += [tuple(p["indices"]) for p in params]
    for g in d["global"]:
        if g["potential"] != "ExcludedVolume":
            continue
        radii = np.zeros(n)
        for p in g["parameters"]:
            radii[p["index"]] = p["radius"] * 0.1
        srt = np.sort(radii)
        cutoff = g.get("cutoff", 2.0) * (srt[-1] + srt[-2])
        excl = set()
        for et, k in g["ignore"]["particles_within"].items():
            excl |= set(_pairs_within(edges[et], n, k))
        sysm.add(ForceSpec(abi.EXCLUDED_VOLUME, "ExcludedVolume", particle_params=radii.reshape(n, 1),
                           globals=np.array([g["epsilon"] * KJ]), cutoff=cutoff,
                           exclusions=np.array(sorted(excl), dtype=np.int32)))
    integ = IntegratorSpec(d["temperature"], 0.01 / 0.04888821290839617, 0.01 * 0.04888821290839617, 12345)
    return sysm, integ


__all__ = ["build", "load", "native_positions_nm", "FLA_Y"]
