"""Pins the FP64 oracle (oracle/oracle.cpp) before anything is compared to it.

Each check is chosen so that a plausible mistake in the restatement (a wrong
sign, a dropped term, a swapped index, a wrong spline convention) fails it:
published known-answer vectors, closed-form single terms, finite differences,
the sh3 sample's native geometry, library routines (scipy's natural spline)
and invariants.
"""
import math

import numpy as np
import pytest

from opencafemol_amd import abi, configs
from opencafemol_amd.aicg2_tables import FLA_MAX_THETA, FLA_MIN_THETA, FLA_Y, FLA_Y2, FLD
from opencafemol_amd.model import ForceSpec, IntegratorSpec, SystemSpec
from tests import sh3_system

KJ = 4.184


# ------------------------------------------------------------------ Philox
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(oracle, ctr, key, expect):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert oracle.philox(ctr, key) == expect


def test_gauss_stream_statistics(oracle):
    g = np.array([oracle.gauss3(1234, s, b) for s in range(40) for b in range(250)]).ravel()
    assert abs(g.mean()) < 0.02
    assert abs(g.var() - 1.0) < 0.03
    # different bead, step, stream or seed give different draws
    a = oracle.gauss3(1, 2, 3, 0)
    for other in (oracle.gauss3(1, 2, 4, 0), oracle.gauss3(1, 3, 3, 0), oracle.gauss3(1, 2, 3, 1),
                  oracle.gauss3(2, 2, 3, 0)):
        assert not np.allclose(a, other)


# ------------------------------------------------------------- splines
def test_natural_spline_matches_scipy(oracle):
    from scipy.interpolate import CubicSpline
    h = (FLA_MAX_THETA - FLA_MIN_THETA) / 9.0
    knots = FLA_MIN_THETA + h * np.arange(10)
    for aa, y in FLA_Y.items():
        y2 = oracle.natural_spline(np.array(y), h)
        cs = CubicSpline(knots, y, bc_type="natural")
        np.testing.assert_allclose(y2, cs(knots, 2), rtol=0, atol=1e-9)


def test_reference_y2_table_is_the_natural_spline(oracle):
    # the reference ships a second-derivative table next to its y-table
    # (src/util/Constants.hpp:51-73); it is the natural spline of y to the
    # table's two decimals, which pins OpenMM's natural fit as the convention
    h = (FLA_MAX_THETA - FLA_MIN_THETA) / 9.0
    worst = 0.0
    for aa in FLA_Y:
        y2 = oracle.natural_spline(np.array(FLA_Y[aa]), h)
        worst = max(worst, np.max(np.abs(y2 - np.array(FLA_Y2[aa]))))
    # natural: 0.0105; a not-a-knot or clamped fit misses by > 400
    assert worst < 0.011


def _ref_fla_eval(y, y2, th):
    knots = np.array([1.30900, 1.48353, 1.65806, 1.83260, 2.00713, 2.18166, 2.35619, 2.53073, 2.70526, 2.87979])
    dth = (FLA_MAX_THETA - FLA_MIN_THETA) / 9.0
    n = min(8, int(math.floor((th - FLA_MIN_THETA) / dth)))
    if knots[n + 1] < th:
        n += 1
    elif th < knots[n]:
        n -= 1
    a = (knots[n + 1] - th) / dth
    b = (th - knots[n]) / dth
    return a * y[n] + b * y[n + 1] + ((a ** 3 - a) * y2[n] + (b ** 3 - b) * y2[n + 1]) * dth * dth / 6


@pytest.mark.parametrize("aa", ["ALA", "GLY", "PRO", "TRP", "VAL"])
def test_fla_min_energy_is_minimum_of_reference_spline(oracle, aa):
    from scipy.optimize import minimize_scalar
    y, y2 = FLA_Y[aa], FLA_Y2[aa]
    emin = oracle.fla_min_energy(y, y2)
    grid = np.linspace(FLA_MIN_THETA, FLA_MAX_THETA - 1e-12, 4001)
    vals = np.array([_ref_fla_eval(y, y2, t) for t in grid])
    t0 = grid[np.argmin(vals)]
    res = minimize_scalar(lambda t: _ref_fla_eval(y, y2, t), bounds=(t0 - 1e-3, t0 + 1e-3), method="bounded",
                          options={"xatol": 1e-10})
    # the 1e-4 sweep lands within (1e-4)^2 * |y''| / 8 of the true minimum
    assert res.fun - 1e-12 <= emin <= res.fun + 2e-6
    assert emin <= y[0]


@pytest.mark.parametrize("key", list(FLD)[:6])
def test_fld_min_energy_is_minimum_of_fourier_sum(oracle, key):
    from scipy.optimize import minimize_scalar
    t = FLD[key]

    def f(p):
        return t[0] + sum(t[2 * n - 1] * math.cos(n * p) + t[2 * n] * math.sin(n * p) for n in (1, 2, 3))
    grid = np.linspace(-math.pi, math.pi, 20001)
    p0 = grid[np.argmin([f(p) for p in grid])]
    res = minimize_scalar(f, bounds=(p0 - 1e-3, p0 + 1e-3), method="bounded", options={"xatol": 1e-11})
    emin = oracle.fld_min_energy(t)
    assert res.fun - 1e-12 <= emin <= res.fun + 1e-6


# ------------------------------------------------------- closed forms
def _two_body(kind, params, r, globals_=None):
    s = SystemSpec(masses=np.ones(2))
    s.add(ForceSpec(kind, "t", indices=np.array([[0, 1]]), params=np.array([params]), globals=globals_))
    x = np.array([[0.1, 0.2, 0.3], [0.1 + r, 0.2, 0.3]])
    return s, x


def _energy(oracle, s, x):
    return oracle.evaluate(s, x)[1].sum()


def test_harmonic_bond_closed_form(oracle):
    s, x = _two_body(abi.HARMONIC_BOND, [0.38, 1000.0], 0.40)
    assert _energy(oracle, s, x) == pytest.approx(0.5 * 1000 * 0.02 ** 2, rel=1e-12)
    F = oracle.evaluate(s, x)[0]
    assert F[0, 0] == pytest.approx(1000 * 0.02, rel=1e-12)  # stretched bond pulls bead 0 toward +x


def test_go_contact_closed_form(oracle):
    s, x = _two_body(abi.GO_CONTACT, [2.0, 0.55], 0.55)
    assert _energy(oracle, s, x) == pytest.approx(-2.0, rel=1e-12)
    assert np.abs(oracle.evaluate(s, x)[0]).max() < 1e-9
    s, x = _two_body(abi.GO_CONTACT, [2.0, 0.55], 0.66)
    q = 0.55 / 0.66
    assert _energy(oracle, s, x) == pytest.approx(2.0 * (5 * q ** 12 - 6 * q ** 10), rel=1e-12)


def test_gaussian_bond_closed_form(oracle):
    s, x = _two_body(abi.GAUSSIAN_BOND, [-3.0, 0.6, 0.015], 0.61)
    assert _energy(oracle, s, x) == pytest.approx(-3.0 * math.exp(-(0.01) ** 2 / (2 * 0.015 ** 2)), rel=1e-12)


def _dihedral_positions(phi):
    # b1 along x, b2 along y; b3 rotated about y by phi (IUPAC: cis = 0)
    p0 = np.array([1.0, 0.0, 0.0]) * 0.38
    p1 = np.array([0.0, 0.0, 0.0])
    p2 = np.array([0.0, 0.38, 0.0])
    p3 = p2 + 0.38 * np.array([math.cos(phi), 0.0, -math.sin(phi)])
    return np.array([p0, p1, p2, p3]) + 0.5


def test_dihedral_sign_convention(oracle):
    # positions built for phi = +1.0 must read back +1.0 (IUPAC sign, cis 0)
    x = _dihedral_positions(1.0)
    assert configs.dihedrals_of(x, np.array([[0, 1, 2, 3]]))[0] == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("phi,phi0", [(1.0, 1.0), (-3.0, 3.0), (0.3, -0.2)])
def test_gaussian_dihedral_closed_form(oracle, phi, phi0):
    s = SystemSpec(masses=np.ones(4))
    s.add(ForceSpec(abi.GAUSSIAN_DIHEDRAL, "d", indices=np.array([[0, 1, 2, 3]]),
                    params=np.array([[-4.0, phi0, 0.15]])))
    E = _energy(oracle, s, _dihedral_positions(phi))
    dt = phi - phi0
    dt -= math.floor((dt + math.pi) / (2 * math.pi)) * 2 * math.pi
    assert E == pytest.approx(-4.0 * math.exp(-dt * dt / (2 * 0.15 ** 2)), rel=1e-10, abs=1e-14)


def test_flexible_local_dihedral_closed_form(oracle):
    key = ("ALA", "R3")
    t = FLD[key]
    emin = oracle.fld_min_energy(t)
    s = SystemSpec(masses=np.ones(4))
    s.add(ForceSpec(abi.FLEXIBLE_LOCAL_DIHEDRAL, "FlexibleLocalDihedral (ALA-R3)", indices=np.array([[0, 1, 2, 3]]),
                    params=np.array([[2.0]]), globals=np.array(list(t) + [emin])))
    for phi in (0.0, 1.2, -2.5):
        E = _energy(oracle, s, _dihedral_positions(phi))
        ref = 2.0 * (t[0] - emin + sum(t[2 * n - 1] * math.cos(n * phi) + t[2 * n] * math.sin(n * phi)
                                       for n in (1, 2, 3)))
        assert E == pytest.approx(ref, rel=1e-10)


def _angle_positions(theta):
    return np.array([[0.38, 0, 0], [0, 0, 0], [0.38 * math.cos(theta), 0.38 * math.sin(theta), 0]]) + 0.3


def test_harmonic_angle_closed_form(oracle):
    s = SystemSpec(masses=np.ones(3))
    s.add(ForceSpec(abi.HARMONIC_ANGLE, "a", indices=np.array([[0, 1, 2]]), params=np.array([[1.9, 80.0]])))
    assert _energy(oracle, s, _angle_positions(2.1)) == pytest.approx(0.5 * 80 * 0.2 ** 2, rel=1e-10)


def test_flexible_local_angle_knots_and_walls(oracle):
    aa = "LEU"
    y = np.array(FLA_Y[aa])
    emin = oracle.fla_min_energy(y, FLA_Y2[aa])
    g = np.array(list(y) + [FLA_MIN_THETA, FLA_MAX_THETA, emin])
    s = SystemSpec(masses=np.ones(3))
    s.add(ForceSpec(abi.FLEXIBLE_LOCAL_ANGLE, "FlexibleLocalAngle (LEU)", indices=np.array([[0, 1, 2]]),
                    params=np.array([[3.0]]), globals=g))
    h = (FLA_MAX_THETA - FLA_MIN_THETA) / 9.0
    for i in (2, 5, 7):  # inside: the spline passes through its knots
        assert _energy(oracle, s, _angle_positions(FLA_MIN_THETA + i * h)) == pytest.approx(3.0 * (y[i] - emin), rel=1e-9)
    # below the range the spline is 0 and the wall is 30 (tmin - theta) + y0
    th = FLA_MIN_THETA - 0.1
    assert _energy(oracle, s, _angle_positions(th)) == pytest.approx(3.0 * (-emin + 30 * 0.1 + y[0]), rel=1e-9)
    th = FLA_MAX_THETA + 0.05
    assert _energy(oracle, s, _angle_positions(th)) == pytest.approx(3.0 * (-emin + 30 * 0.05 + y[9]), rel=1e-9)


def _pair_system(kind, pp, cutoff, globals_=None, switch=0.0):
    s = SystemSpec(masses=np.ones(2))
    s.add(ForceSpec(kind, "p", particle_params=np.array(pp, dtype=float), globals=globals_, cutoff=cutoff,
                    switch_distance=switch))
    return s


def _at(r):
    return np.array([[0.2, 0.2, 0.2], [0.2, 0.2 + r, 0.2]])


def test_excluded_volume_closed_form(oracle):
    s = _pair_system(abi.EXCLUDED_VOLUME, [[0.2], [0.22]], 0.84, np.array([2.5]))
    r = 0.5
    ref = 2.5 * 0.42 ** 12 * (r ** -12 - 0.84 ** -12)
    assert _energy(oracle, s, _at(r)) == pytest.approx(ref, rel=1e-12)
    assert _energy(oracle, s, _at(0.8400001)) == 0.0   # beyond the cutoff


def test_debye_huckel_closed_form(oracle):
    K, lam, corr, rc = configs.debye_huckel_constants(0.1, 300.0, 3.0)
    s = _pair_system(abi.DEBYE_HUCKEL, [[1.0], [-1.0]], rc, np.array([K, lam, corr]))
    r = 1.3
    assert _energy(oracle, s, _at(r)) == pytest.approx(-K * (math.exp(-r / lam) / r - corr), rel=1e-12)
    assert _energy(oracle, s, _at(rc * (1 - 1e-9))) == pytest.approx(0.0, abs=1e-6)


def test_lennard_jones_ah_closed_forms(oracle):
    pp = [[0.8, 0.5], [0.4, 0.7]]        # eps, sigma ; mean eps 0.6, sigma 0.6
    rep = _pair_system(abi.LJ_REPULSIVE, pp, 2.0)
    att = _pair_system(abi.LJ_ATTRACTIVE, pp, 2.0, switch=1.8)
    rstar = 0.6 * 2 ** (1 / 6)
    assert _energy(oracle, rep, _at(0.6)) == pytest.approx(0.6, rel=1e-12)            # 4(1-1)+1
    assert _energy(oracle, rep, _at(rstar * 1.001)) == 0.0
    assert _energy(oracle, att, _at(0.6)) == pytest.approx(-0.6, rel=1e-12)            # inside r*: -eps
    r = 1.2
    assert _energy(oracle, att, _at(r)) == pytest.approx(4 * 0.6 * ((0.6 / r) ** 12 - (0.6 / r) ** 6), rel=1e-12)
    r = 1.9                                                                              # switched: S(1/2) = 1/2
    assert _energy(oracle, att, _at(r)) == pytest.approx(0.5 * 4 * 0.6 * ((0.6 / r) ** 12 - (0.6 / r) ** 6), rel=1e-10)
    # at r* the attractive force vanishes from both sides
    assert np.abs(oracle.evaluate(att, _at(rstar * 1.0000001))[0]).max() < 1e-4


# ------------------------------------------------- finite differences
def _fd_check(oracle, s, x, beads, h=1e-6, tol=2e-6):
    F = oracle.evaluate(s, x)[0]
    fmag = np.sqrt(np.mean(np.sum(F ** 2, axis=1)))
    for i in beads:
        for c in range(3):
            xp = x.copy(); xp[i, c] += h
            xm = x.copy(); xm[i, c] -= h
            fd = -(oracle.evaluate(s, xp)[1].sum() - oracle.evaluate(s, xm)[1].sum()) / (2 * h)
            assert abs(fd - F[i, c]) <= tol * max(abs(F[i, c]), fmag), (i, c, fd, F[i, c])


def test_finite_differences_aicg2(oracle):
    s, it, x = configs.aicg2_chain(40, seed=2)
    x = x + np.random.default_rng(1).normal(scale=0.03, size=x.shape)
    _fd_check(oracle, s, x, range(0, 40, 3))


def test_finite_differences_hps_periodic(oracle):
    s, it, x = configs.hps_slab(n_chains=6, chain_len=20, box=(6.5, 6.5, 20.0), slab_z=4.0, seed=1)
    x = x + np.random.default_rng(2).normal(scale=0.05, size=x.shape)
    _fd_check(oracle, s, x, range(0, 120, 11), h=2e-7)


# ------------------------------------------------------------ invariants
def test_newton_third_law_and_translation(oracle):
    s, it, x = configs.aicg2_chain(50, seed=4)
    F, E = oracle.evaluate(s, x)
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * np.abs(F).max()
    F2, E2 = oracle.evaluate(s, x + np.array([3.1, -2.2, 7.7]))
    np.testing.assert_allclose(E2, E, rtol=1e-9, atol=1e-9)


def test_rotation_invariance(oracle):
    s, it, x = configs.aicg2_chain(50, seed=5)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    F, E = oracle.evaluate(s, x)
    F2, E2 = oracle.evaluate(s, x @ R.T)
    np.testing.assert_allclose(E2, E, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(F2, F @ R.T, rtol=1e-7, atol=1e-7)


def test_periodic_image_invariance(oracle):
    s, it, x = configs.hps_slab(n_chains=6, chain_len=20, box=(6.5, 6.5, 20.0), slab_z=4.0, seed=3)
    shift = np.zeros_like(x)
    shift[::3, 0] = 6.5
    shift[1::3, 2] = -20.0
    F, E = oracle.evaluate(s, x)
    F2, E2 = oracle.evaluate(s, x + shift)
    np.testing.assert_allclose(E2, E, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(F2, F, rtol=1e-7, atol=1e-6)


# ------------------------------------------- sh3 native geometry (fixture)
def test_sh3_native_geometry_known_answers(oracle):
    d = sh3_system.load()
    x = sh3_system.native_positions_nm(d)
    by = {(f["interaction"], f["potential"]): f["parameters"] for f in d["local"]}
    # Gaussian dihedrals sit at their minimum: E = k exactly up to v0's 4 decimals
    s = SystemSpec(masses=np.ones(len(x)))
    p = by[("DihedralAngle", "Gaussian")]
    s.add(ForceSpec(abi.GAUSSIAN_DIHEDRAL, "d", indices=np.array([q["indices"] for q in p]),
                    params=np.array([[q["k"] * KJ, q["v0"], q["σ"]] for q in p])))
    E = oracle.evaluate(s, x)[1].sum()
    assert E == pytest.approx(sum(q["k"] * KJ for q in p), rel=2e-6)
    # a mirrored structure has the opposite dihedrals: a sign error fails here
    Em = oracle.evaluate(s, x * np.array([1, 1, -1]))[1].sum()
    assert abs(Em) < 0.5 * abs(E)
    # Go contacts sit at their minimum -k, harmonic bonds at zero
    s = SystemSpec(masses=np.ones(len(x)))
    p = by[("BondLength", "GoContact")]
    s.add(ForceSpec(abi.GO_CONTACT, "c", indices=np.array([q["indices"] for q in p]),
                    params=np.array([[q["k"] * KJ, q["v0"] * 0.1] for q in p])))
    assert oracle.evaluate(s, x)[1].sum() == pytest.approx(-sum(q["k"] * KJ for q in p), rel=1e-6)
    sysm, _ = sh3_system.build(d)
    F, Eg = oracle.evaluate(sysm, x)
    assert abs(Eg[sysm.group_ids["HarmonicBond"]]) < 0.05


# ------------------------------------------------------------ integrator
def test_baoab_zero_friction_conserves_energy(oracle):
    s, it, x = configs.go_chain(30, seed=3)
    it = IntegratorSpec(300.0, 0.0, 0.002, 7)
    v = np.random.default_rng(3).normal(scale=0.1, size=x.shape)
    e0 = oracle.evaluate(s, x)[1].sum() + oracle.kinetic(s, v)
    x1, v1 = oracle.baoab(s, it, x, v, 0, 500)
    e1 = oracle.evaluate(s, x1)[1].sum() + oracle.kinetic(s, v1)
    assert abs(e1 - e0) < 2e-3 * (abs(e0) + oracle.kinetic(s, v))


def test_baoab_equipartition(oracle):
    # harmonic chain at 300 K: <KE> = 3/2 N kT within sampling error
    n = 20
    s = SystemSpec(masses=np.full(n, 50.0))
    b = np.stack([np.arange(n - 1), np.arange(1, n)], 1)
    s.add(ForceSpec(abi.HARMONIC_BOND, "HarmonicBond", indices=b, params=np.tile([0.38, 2000.0], (n - 1, 1))))
    it = IntegratorSpec(300.0, 5.0, 0.005, 99)
    x = np.stack([0.38 * np.arange(n), np.zeros(n), np.zeros(n)], 1)
    v = np.zeros_like(x)
    kes = []
    for blk in range(40):
        x, v = oracle.baoab(s, it, x, v, blk * 100, 100)
        if blk >= 5:
            kes.append(oracle.kinetic(s, v))
    kT = 1.380649e-23 * 6.02214076e23 / 1000 * 300.0
    assert np.mean(kes) == pytest.approx(1.5 * n * kT, rel=0.12)


# ------------------------------------------------------- cell-list oracle
def _cells_cases():
    hps = configs.hps_slab(n_chains=12, chain_len=30, box=(8.0, 8.0, 30.0), slab_z=6.0, seed=4)
    go = configs.aicg2_chain(56, seed=2)
    x_go = go[2] + np.random.default_rng(2).normal(scale=0.02, size=go[2].shape)
    return [(hps[0], hps[2]), (go[0], x_go)]


def test_eval_cells_matches_brute_force(oracle):
    """The cell-list evaluation (used on systems too large for the O(N^2)
    oracle) finds the same pairs with the same formulas: forces and energies
    equal the brute-force oracle up to summation order, on a periodic slab
    (three pair kinds, a Debye-Hueckel group of charged beads only) and an
    open AICG2+ chain (every bonded kind)."""
    for s, x in _cells_cases():
        Fb, Eb = oracle.evaluate(s, x)
        Fc, Ec = oracle.evaluate_cells(s, x, nthreads=4)
        scale = np.abs(Fb).max()
        assert np.abs(Fc - Fb).max() <= 1e-12 * scale
        assert np.abs(Ec - Eb).max() <= 1e-12 * np.abs(Eb).sum()
        rows = np.random.default_rng(0).choice(s.n, 50, replace=False)
        Fr, _ = oracle.evaluate_cells(s, x, rows=rows, nthreads=2)
        assert np.array_equal(Fr, Fc[rows])


def test_eval_cells_counts_each_pair_once_across_images(oracle):
    """Two beads one cutoff-fraction apart across the periodic boundary:
    the minimum image interacts, exactly once."""
    s = SystemSpec(masses=np.full(2, 100.0), box=(10.0, 10.0, 10.0))
    s.add(ForceSpec(abi.EXCLUDED_VOLUME, "ExcludedVolume", use_periodic=True,
                    particle_params=np.array([[0.2], [0.2]]), cutoff=0.8, globals=np.array([1.0])))
    x = np.array([[0.1, 5.0, 5.0], [9.7, 5.0, 5.0]])
    Fb, Eb = oracle.evaluate(s, x)
    Fc, Ec = oracle.evaluate_cells(s, x)
    r = 0.4
    assert Eb.sum() == pytest.approx(0.4 ** 12 * (r ** -12 - 0.8 ** -12))
    assert np.allclose(Fc, Fb, rtol=1e-13) and np.allclose(Ec, Eb, rtol=1e-13)


def test_baoab_cells_matches_baoab(oracle):
    """oracle_baoab_cells (the CPU baseline's step) against oracle_baoab: the
    same trajectory up to force summation order over 20 steps."""
    s, it, x = configs.hps_slab(n_chains=12, chain_len=30, box=(8.0, 8.0, 30.0), slab_z=6.0, seed=6)
    v = np.random.default_rng(6).normal(scale=0.2, size=x.shape)
    x1, v1 = oracle.baoab(s, it, x, v, 3, 20)
    x2, v2 = oracle.baoab_cells(s, it, x, v, 3, 20, 4)
    assert np.abs(x2 - x1).max() <= 1e-9
    assert np.abs(v2 - v1).max() <= 1e-7 * np.abs(v1).max()
    assert np.abs(x1 - x).max() > 1e-3   # it moved


# ------------------------------------------- membrane and capped potentials
def test_capped_go_contact_cap_is_the_tangent_line(oracle):
    """Inside its cap (r / r0 <= c) the capped contact is the tangent line of
    the 12-10 curve at r = c r0: value and slope match there, the energy is
    linear below it (zero second difference), and beyond the cap it is the
    plain 12-10 contact (minimum -k at r0)."""
    k, r0, c = 2.0, 0.6, 0.8
    def E(r):
        s, x = _two_body(abi.CAPPED_GO_CONTACT, [k, r0, c], r)
        return _energy(oracle, s, x)
    def go(r):
        s, x = _two_body(abi.GO_CONTACT, [k, r0], r)
        return _energy(oracle, s, x)
    rc, h = c * r0, 1e-6
    assert E(r0) == pytest.approx(-k, rel=1e-12)
    assert E(0.9 * r0) == pytest.approx(go(0.9 * r0), rel=1e-12)
    assert E(rc) == pytest.approx(go(rc), rel=1e-10)
    slope_go = (go(rc + h) - go(rc - h)) / (2 * h)
    slope_cap = (E(rc - h) - E(rc - 3 * h)) / (2 * h)
    assert slope_cap == pytest.approx(slope_go, rel=1e-5)
    a, b, m = E(0.2 * r0), E(0.4 * r0), E(0.3 * r0)
    assert m == pytest.approx(0.5 * (a + b), rel=1e-12)   # linear below the cap
    assert E(0.1 * r0) < 1e3 * k                          # no r^-12 wall inside the cap


@pytest.mark.parametrize("n", [1, 2, 3])
def test_cosine_dihedral_closed_form(oracle, n):
    """k (1 - cos(n (phi - t0))): zero at t0, 2k half a period away, period 2 pi / n."""
    t0 = 0.4
    s = SystemSpec(masses=np.ones(4))
    s.add(ForceSpec(abi.COSINE_DIHEDRAL, "c", indices=np.array([[0, 1, 2, 3]]), params=np.array([[1.5, t0, n]])))
    assert _energy(oracle, s, _dihedral_positions(t0)) == pytest.approx(0.0, abs=1e-12)
    assert _energy(oracle, s, _dihedral_positions(t0 + math.pi / n)) == pytest.approx(3.0, rel=1e-12)
    p = -1.1
    assert _energy(oracle, s, _dihedral_positions(p)) == pytest.approx(
        _energy(oracle, s, _dihedral_positions(p + 2 * math.pi / n)), rel=1e-10, abs=1e-12)


def test_wca_closed_form_and_mixing(oracle):
    """WCA: epsilon mixed geometrically, sigma arithmetically; E(sigma) = eps,
    zero from 2^(1/6) sigma on."""
    pp = [[0.5, 1.0], [0.7, 4.0]]          # sigma, eps: sigma 0.6, eps sqrt(4) = 2
    s = _pair_system(abi.WCA, pp, 0.7 * 2 ** (1 / 6))
    assert _energy(oracle, s, _at(0.6)) == pytest.approx(2.0, rel=1e-12)
    rstar = 0.6 * 2 ** (1 / 6)
    assert _energy(oracle, s, _at(rstar * (1 - 1e-9))) == pytest.approx(0.0, abs=1e-6)
    assert _energy(oracle, s, _at(rstar * 1.0001)) == 0.0


def test_uniform_wca_and_lj_tables_act_between_their_categories(oracle):
    """Uniform (table) WCA between categories 0 and 1 only: E(sigma) = eps for
    an A-B pair, nothing for A-A.  Uniform LJ attraction: -eps inside r*, the
    LJ tail outside, zero at the switched cutoff."""
    sg, ep = 0.5, 1.3
    s = SystemSpec(masses=np.ones(3))
    s.add(ForceSpec(abi.UNIFORM_WCA, "u", globals=np.array([sg, ep]), category=np.array([0, 1, 0], np.int32),
                    n_categories=2, allow=np.array([[0, 1], [1, 0]], np.uint8), cutoff=sg * 2 ** (1 / 6)))
    x = np.array([[1.0, 1.0, 1.0], [1.0, 1.0 + sg, 1.0], [1.0, 1.0, 1.0 + 0.45]])   # 0-1 at sigma, 0-2 at 0.45
    assert _energy(oracle, s, x) == pytest.approx(ep, rel=1e-12)
    s2 = SystemSpec(masses=np.ones(2))
    rc = 2.5 * sg
    s2.add(ForceSpec(abi.UNIFORM_LJ_ATTRACTIVE, "l", globals=np.array([sg, ep]), cutoff=rc, switch_distance=0.9 * rc))
    assert _energy(oracle, s2, _at(0.9 * sg)) == pytest.approx(-ep, rel=1e-12)
    r = 1.0
    assert _energy(oracle, s2, _at(r)) == pytest.approx(4 * ep * ((sg / r) ** 12 - (sg / r) ** 6), rel=1e-12)
    assert _energy(oracle, s2, _at(rc * (1 - 1e-7))) == pytest.approx(0.0, abs=1e-9)


def test_isolf_closed_form(oracle):
    """iSoLF: -eps up to b = 2^(1/6) sigma, -eps cos^2(pi (r - b) / (2 omega))
    out to b + omega, zero beyond; eps mixed geometrically."""
    pp = [[0.6, 1.0, 0.9], [0.6, 9.0, 1.1]]   # sigma 0.6, eps 3, omega 1.0
    b = 0.6 * 2 ** (1 / 6)
    s = _pair_system(abi.ISOLF_ATTRACTIVE, pp, b + 1.2)
    assert _energy(oracle, s, _at(0.9 * b)) == pytest.approx(-3.0, rel=1e-12)
    assert _energy(oracle, s, _at(b + 0.5)) == pytest.approx(-1.5, rel=1e-12)      # cos^2(pi/4)
    assert _energy(oracle, s, _at(b + 0.999999)) == pytest.approx(0.0, abs=1e-9)
    assert _energy(oracle, s, _at(b + 1.1)) == 0.0


def test_finite_differences_membrane_and_capped(oracle):
    s, it, x = configs.lipid_patch(n_lipids=64, box=(6.0, 6.0, 6.0), seed=1)
    x = x + np.random.default_rng(3).normal(scale=0.03, size=x.shape)
    kinds = {f.kind for f in s.forces}
    assert {abi.WCA, abi.ISOLF_ATTRACTIVE, abi.UNIFORM_WCA, abi.UNIFORM_LJ_ATTRACTIVE} <= kinds
    _fd_check(oracle, s, x, range(0, s.n, 13), h=1e-7, tol=1e-5)
    s, it, x = configs.capped_go_chain(50, seed=2, capping_ratio=1.0)
    x = x + np.random.default_rng(4).normal(scale=0.02, size=x.shape)
    _fd_check(oracle, s, x, range(0, 50, 3))


@pytest.mark.parametrize("which", ["hps", "go", "exv_gas"])
def test_fast_cpu_step_matches_oracle(oracle, which):
    """The CPU step timed as the benchmark's baseline (Verlet lists with a
    skin, both directions per row, no pow) against the oracle: forces equal
    up to summation order, and the same BAOAB trajectory over 30 steps
    (with list rebuilds inside) within FP64 rounding."""
    if which == "hps":
        s, it, x = configs.hps_slab(n_chains=12, chain_len=30, box=(8.0, 8.0, 30.0), slab_z=6.0, seed=6)
    elif which == "go":
        s, it, x = configs.go_chain(70, seed=4)
        x = x + np.random.default_rng(4).normal(scale=0.02, size=x.shape)
    else:
        rng = np.random.default_rng(5)
        n = 300
        x = rng.uniform(0, 5.0, size=(n, 3))
        s = SystemSpec(masses=np.full(n, 80.0), box=(5.0, 5.0, 5.0))
        s.add(ForceSpec(abi.EXCLUDED_VOLUME, "ExcludedVolume", use_periodic=True,
                        particle_params=np.full((n, 1), 0.12), globals=np.array([1.0]), cutoff=0.5))
        from opencafemol_amd.model import IntegratorSpec
        it = IntegratorSpec(300.0, 0.5, 0.005, 2)
    Fb, _ = oracle.evaluate(s, x)
    Ff = oracle.fast_forces(s, x, skin=0.3, nthreads=4)
    assert np.abs(Ff - Fb).max() <= 1e-9 * np.abs(Fb).max()
    v = np.random.default_rng(1).normal(scale=0.1, size=x.shape)
    x1, v1 = oracle.baoab(s, it, x, v, 7, 30)
    x2, v2, builds = oracle.baoab_fast(s, it, x, v, 7, 30, nthreads=4, skin=0.05)
    assert builds >= 2          # the small skin forces rebuilds inside the run
    assert np.abs(x2 - x1).max() <= 1e-9
    assert np.abs(v2 - v1).max() <= 1e-7 * np.abs(v1).max()
