"""Parity bars of the HIP path against the FP64 oracle.

Forces: per-bead relative error |F_gpu - F_ref| / max(|F_ref|, 0.01 F_rms)
<= 1e-4.  The floor (1% of the RMS bead force) keeps beads whose net force
cancels to nearly zero from dividing by ~0; without it a per-bead ratio is
undefined for such beads.
Energies: |E_gpu - E_ref| <= 1e-6 * sum_g |E_ref,g| per group and in total.
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


def assert_energies_close(Eg, Eo, rel=ENERGY_REL):
    scale = np.sum(np.abs(Eo)) + 1e-9
    d = np.abs(np.asarray(Eg) - np.asarray(Eo))
    g = int(np.argmax(d))
    assert d.max() <= rel * scale, f"group {g}: gpu {Eg[g]} ref {Eo[g]} (scale {scale})"
    assert abs(np.sum(Eg) - np.sum(Eo)) <= rel * scale
