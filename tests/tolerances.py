"""Parity bars of the HIP path against the FP64 oracle.

Forces: per-bead relative error |F_gpu - F_ref| / max(|F_ref|, 0.01 F_rms)
<= 1e-4.  The floor (1% of the RMS bead force) keeps beads whose net force
cancels to nearly zero from dividing by ~0; without it a per-bead ratio is
undefined for such beads.
Energies: the total within 1e-6 of |E_total| (the stated bar), and each
group within 1e-6 of its own |E_g| plus 1e-5 kJ/mol (the energy pass sums
terms in fixed point at 2^-32 kJ/mol, so a group of many terms carries an
absolute rounding of that order).
"""
import numpy as np

FORCE_REL = 1e-4
ENERGY_REL = 1e-6


def force_errors(Fg, Fo):
    mag = np.linalg.norm(Fo, axis=1)
    floor = 0.01 * np.sqrt(np.mean(mag ** 2)) + 1e-12
    return np.linalg.norm(Fg - Fo, axis=1) / np.maximum(mag, floor)


def assert_forces_close(Fg, Fo, rel=FORCE_REL):
    err = force_errors(Fg, Fo)
    i = int(np.argmax(err))
    assert err.max() <= rel, f"bead {i}: rel err {err.max():.3e} gpu {Fg[i]} ref {Fo[i]}"


ENERGY_ABS = 1e-5


def assert_energies_close(Eg, Eo, rel=ENERGY_REL):
    Eg = np.asarray(Eg, dtype=np.float64)
    Eo = np.asarray(Eo, dtype=np.float64)
    d = np.abs(Eg - Eo)
    bar = rel * np.abs(Eo) + ENERGY_ABS
    g = int(np.argmax(d / bar))
    assert np.all(d <= bar), f"group {g}: gpu {Eg[g]} ref {Eo[g]}"
    total = abs(np.sum(Eo))
    assert abs(np.sum(Eg) - np.sum(Eo)) <= rel * total + ENERGY_ABS, (np.sum(Eg), np.sum(Eo))
