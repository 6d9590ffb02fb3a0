"""AICG2+ flexible local terms: the constants the reference's generators
compute before handing a force to OpenMM, restated for building descriptors.

* ``fla_min_energy``  FlexibleLocalAngleForceFieldGenerator.cpp:11-19 with the
                      spline evaluator of :47-68 (reference knots and y''-table)
* ``fld_min_energy``  FlexibleLocalDihedralForceFieldGenerator.cpp:7-26
* ``flexible_local_angle_forces`` / ``flexible_local_dihedral_forces``: one
  force per residue type / residue-pair class, as the TOML reader loops over
  the tables (src/input/ReadTOMLInput.cpp:235-245, :291-303).
"""
import math

import numpy as np

from . import abi
from .aicg2_tables import FLA_MAX_THETA, FLA_MIN_THETA, FLA_Y, FLA_Y2, FLD
from .model import ForceSpec

KJ_PER_KCAL = 4.184
FLA_KNOTS = (1.30900, 1.48353, 1.65806, 1.83260, 2.00713, 2.18166, 2.35619, 2.53073, 2.70526, 2.87979)


def fla_min_energy(y, y2, tmin=FLA_MIN_THETA, tmax=FLA_MAX_THETA):
    dth = (tmax - tmin) / 9.0
    rdth = 1.0 / dth
    emin = y[0]
    th = tmin
    # the sweep accumulates th += 1e-4 exactly as the reference loop does
    ths = []
    while th < tmax:
        ths.append(th)
        th += 1e-4
    th = np.array(ths)
    n = np.minimum(8, np.floor((th - tmin) * rdth).astype(np.int64))
    kn = np.array(FLA_KNOTS)
    n = np.where(kn[n + 1] < th, n + 1, np.where(th < kn[n], n - 1, n))
    y = np.asarray(y, dtype=np.float64)
    y2 = np.asarray(y2, dtype=np.float64)
    a = (kn[n + 1] - th) * rdth
    b = (th - kn[n]) * rdth
    e = a * y[n] + b * y[n + 1] + ((a * a * a - a) * y2[n] + (b * b * b - b) * y2[n + 1]) * dth * dth / 6.0
    return float(min(emin, e.min()))


def fld_min_energy(t):
    phi = -math.pi
    ph = []
    while phi < math.pi:
        ph.append(phi)
        phi += 1e-4
    p = np.array(ph)
    s1, c1 = np.sin(p), np.cos(p)
    ss, cc = s1 * s1, c1 * c1
    s2, c2 = 2 * s1 * c1, cc - ss
    s3, c3 = s1 * (3 - 4 * ss), c1 * (4 * cc - 3)
    e = t[1] * c1 + t[2] * s1 + t[3] * c2 + t[4] * s2 + t[5] * c3 + t[6] * s3 + t[0]
    return float(e.min())


_FLA_EMIN = {}
_FLD_EMIN = {}


def fla_globals(aa):
    if aa not in _FLA_EMIN:
        _FLA_EMIN[aa] = fla_min_energy(FLA_Y[aa], FLA_Y2[aa])
    return np.array(list(FLA_Y[aa]) + [FLA_MIN_THETA, FLA_MAX_THETA, _FLA_EMIN[aa]])


def fld_globals(key):
    if key not in _FLD_EMIN:
        _FLD_EMIN[key] = fld_min_energy(FLD[key])
    return np.array(list(FLD[key]) + [_FLD_EMIN[key]])


def fld_classes(first, second):
    """Residue-pair classes a dihedral whose central bond joins `first` and
    `second` is added to (ReadTOMLForceFieldGenerator.cpp:636-766): R1-GLY for
    every X-GLY, R2-PRO for every X-PRO and additionally GLY-PRO for GLY-PRO,
    otherwise the first residue's R3 class."""
    if second == "GLY":
        return [("R1", "GLY")]
    if second == "PRO":
        return [("R2", "PRO"), ("GLY", "PRO")] if first == "GLY" else [("R2", "PRO")]
    return [(first, "R3")]


def flexible_local_angle_forces(triplets, residues, k_kcal, use_periodic=False):
    """One FlexibleLocalAngle force per residue type present; the residue of a
    triplet is that of its middle bead (y = "y1_<res>")."""
    triplets = np.asarray(triplets, dtype=np.int32).reshape(-1, 3)
    out = []
    for aa in FLA_Y:
        sel = np.array([residues[t[1]] == aa for t in triplets], dtype=bool)
        if not sel.any():
            continue
        out.append(ForceSpec(abi.FLEXIBLE_LOCAL_ANGLE, f"FlexibleLocalAngle ({aa})", use_periodic=use_periodic,
                             indices=triplets[sel], params=np.full((sel.sum(), 1), k_kcal * KJ_PER_KCAL),
                             globals=fla_globals(aa)))
    return out


def flexible_local_dihedral_forces(quads, residues, k_kcal, use_periodic=False):
    quads = np.asarray(quads, dtype=np.int32).reshape(-1, 4)
    classes = [fld_classes(residues[q[1]], residues[q[2]]) for q in quads]
    out = []
    for key in FLD:
        sel = np.array([key in c for c in classes], dtype=bool)
        if not sel.any():
            continue
        out.append(ForceSpec(abi.FLEXIBLE_LOCAL_DIHEDRAL, f"FlexibleLocalDihedral ({key[0]}-{key[1]})",
                             use_periodic=use_periodic, indices=quads[sel],
                             params=np.full((sel.sum(), 1), k_kcal * KJ_PER_KCAL), globals=fld_globals(key)))
    return out
