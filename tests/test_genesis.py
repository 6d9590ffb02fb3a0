"""GENESIS-format front end (.inp / .top / .itp / .gro -> generators ->
SystemGenerator) on the CPU.

Pins, in order of strength:
* the reference's own 1SRL sample (when present), read independently here
  by whitespace-splitting the files, against the C++ fixed-column reader,
  and its native-structure energies, which the potentials' forms fix
  (a Go contact at r = r0 and a Gaussian term at its centre each give -k);
* the sh3 protein written in GENESIS units from the numbers the TOML reader
  produces, which must give the TOML reader's potentials group by group;
* the readers' error behaviour (ReadGenesisInput.cpp, ReadGenesisForceFieldGenerator.cpp).
"""
import os
import re

import numpy as np
import pytest

from opencafemol_amd import abi, aicg2, host_abi
from tests import sh3_system
from tests.golden import sh3_genesis, sh3_toml

REF_GENESIS = "/root/reference/sample/genesis_input"
KJ = 4.184


def _by_kind(spec):
    out = {}
    for f in spec.forces:
        out.setdefault(f.kind, []).append(f)
    return out


def _rows(forces):
    """{indices: params} over forces of one kind; FLA/FLD keyed with their table."""
    out = {}
    for f in forces:
        for i, p in zip(f.indices, f.params):
            key = tuple(int(v) for v in i)
            if f.globals is not None and len(f.globals):
                key = key + (tuple(np.round(f.globals, 9)),)
            assert key not in out
            out[key] = np.asarray(p, dtype=np.float64)
    return out


def _sections(path, base):
    """[ section ] -> whitespace-split rows, following #include, for the tests'
    own reading of a topology."""
    out, cur = {}, None
    for raw in open(os.path.join(base, path)):
        line = raw.split(";")[0].rstrip()
        m = re.match(r'^#include "(\S+)"', raw)
        if m:
            for k, v in _sections(m.group(1), base).items():
                out.setdefault(k, []).extend(v)
            continue
        m = re.match(r"^\s*\[\s*(\S.*?)\s*\]\s*$", line)
        if m:
            cur = m.group(1)
            out.setdefault(cur, [])
        elif line.strip():
            out[cur].append(line.split())
    return out


# --------------------------------------------------- the reference's sample
@pytest.mark.skipif(not os.path.isdir(REF_GENESIS), reason="reference samples not present")
def test_reference_sample_against_whitespace_reading():
    spec, integ, x, v = host_abi.load_genesis(os.path.join(REF_GENESIS, "1SRL_cg.inp"))
    top = _sections("1SRL_cg.top", REF_GENESIS)
    atoms = top["atoms"]
    n = len(atoms)
    types = [a[1] for a in atoms]
    assert spec.n == n == 56
    np.testing.assert_array_equal(spec.masses, [float(a[7]) for a in atoms])
    gro = open(os.path.join(REF_GENESIS, "1SRL_cg.gro")).read().splitlines()
    np.testing.assert_array_equal(x, [[float(t) for t in l.split()[4:7]] for l in gro[2:2 + n]])
    assert v is None
    assert (integ.temperature, integ.friction, integ.dt, integ.seed) == (300.0, 0.01, 0.010, 12345)

    k = _by_kind(spec)
    bonds = _rows(k[abi.HARMONIC_BOND])
    assert bonds.keys() == {(int(r[0]) - 1, int(r[1]) - 1) for r in top["bonds"]}
    for r in top["bonds"]:
        np.testing.assert_array_equal(bonds[(int(r[0]) - 1, int(r[1]) - 1)], [float(r[3]), 0.5 * float(r[4])])
    gb = _rows(k[abi.GAUSSIAN_BOND])
    a21 = [r for r in top["angles"] if r[3] == "21"]
    assert len(gb) == len(a21) == 54
    for r in a21:
        np.testing.assert_array_equal(gb[(int(r[0]) - 1, int(r[2]) - 1)], [-float(r[5]), float(r[4]), float(r[6])])
    gd = _rows(k[abi.GAUSSIAN_DIHEDRAL])
    d41 = [r for r in top["dihedrals"] if r[4] == "41"]
    assert len(gd) == len(d41) == 53
    for r in d41:
        got = gd[tuple(int(t) - 1 for t in r[:4])]
        np.testing.assert_allclose(got, [-float(r[6]), np.radians(float(r[5])), float(r[7])], rtol=1e-15)
    go = _rows(k[abi.GO_CONTACT])
    assert len(go) == len(top["pairs"]) == 215
    for r in top["pairs"]:
        np.testing.assert_array_equal(go[(int(r[0]) - 1, int(r[1]) - 1)], [float(r[4]), float(r[3])])

    # FLA: every function-22 angle once, in the table of its middle bead's type
    fla = _rows(k[abi.FLEXIBLE_LOCAL_ANGLE])
    a22 = [tuple(int(t) - 1 for t in r[:3]) for r in top["angles"] if r[3] == "22"]
    assert len(fla) == len(a22) == 54
    for tri in a22:
        key = tri + (tuple(np.round(aicg2.fla_globals(types[tri[1]]), 9)),)
        np.testing.assert_array_equal(fla[key], [KJ])

    # FLD: each function-52 dihedral in every class its middle types select;
    # X-GLY and X-PRO dihedrals land in two classes
    fld = _rows(k[abi.FLEXIBLE_LOCAL_DIHEDRAL])
    expect = set()
    for r in top["dihedrals"]:
        if r[4] != "52":
            continue
        q = tuple(int(t) - 1 for t in r[:4])
        tj, tk = types[q[1]], types[q[2]]
        classes = [(tj, "R3")]
        if tk == "GLY":
            classes.append(("R1", "GLY"))
        if tk == "PRO":
            classes.append(("GLY", "PRO") if tj == "GLY" else ("R2", "PRO"))
        for c in classes:
            expect.add(q + (tuple(np.round(aicg2.fld_globals(c), 9)),))
    assert set(fld) == expect and len(expect) == 53 + 6

    (exv,) = k[abi.EXCLUDED_VOLUME]
    # rmin is read from columns 31-40 (ReadGenesisForceFieldGenerator.cpp:249);
    # the sample's rmin field runs one column further, so its last digit is
    # dropped: 0.3943 is read as 0.394
    raw = [l for l in open(os.path.join(REF_GENESIS, "param", "atom_types.itp")) if l.startswith("  ALA  6")][0]
    assert raw.split()[5] == "0.3943" and raw[30:40] == "     0.394"
    atomtypes = {}
    for line in open(os.path.join(REF_GENESIS, "param", "atom_types.itp")):
        m = re.match(r"^\s*\[\s*(\S.*?)\s*\]", line)
        if m:
            section = m.group(1)
        elif section == "atomtypes" and line.strip() and not line.lstrip().startswith(";"):
            atomtypes[line.split()[0]] = float(line[30:40])
    np.testing.assert_array_equal(exv.particle_params[:, 0], [atomtypes[t] * 0.5 for t in types])
    rmin = {r[0]: float(r[5]) for r in top["atomtypes"]}
    np.testing.assert_allclose(exv.particle_params[:, 0], [rmin[t] * 0.5 for t in types], rtol=0, atol=5e-4)
    assert exv.globals[0] == 2.5104
    srt = np.sort(exv.particle_params[:, 0])
    assert exv.cutoff == pytest.approx(2.0 * (srt[-1] + srt[-2]), rel=1e-15)
    bond_edges = [(int(r[0]) - 1, int(r[1]) - 1) for r in top["bonds"]]
    contacts = [(int(r[0]) - 1, int(r[1]) - 1) for r in top["pairs"]]
    excl = set(sh3_system._pairs_within(bond_edges, n, 3)) | set(sh3_system._pairs_within(contacts, n, 1))
    assert {tuple(e) for e in exv.exclusions} == excl
    assert spec.group_ids == {"HarmonicBond": 0, "GaussianBond": 1, "FlexibleLocalAngle": 2, "GaussianDihedral": 3,
                              "FlexibleLocalDihedral": 4, "GoContact": 5, "ExcludedVolume": 6}


@pytest.mark.skipif(not os.path.isdir(REF_GENESIS), reason="reference samples not present")
def test_reference_sample_native_energies(oracle):
    """At the native structure every Go contact sits at its r0 and every
    Gaussian term at its centre (up to the 1e-4 nm rounding of the .gro), so
    each group's energy is minus the sum of its coefficients, and the bonds
    are at rest."""
    spec, integ, x, v = host_abi.load_genesis(os.path.join(REF_GENESIS, "1SRL_cg.inp"))
    top = _sections("1SRL_cg.top", REF_GENESIS)
    F, E = oracle.evaluate(spec, x)
    g = spec.group_ids
    assert E[g["GoContact"]] == pytest.approx(-sum(float(r[4]) for r in top["pairs"]), rel=1e-3)
    assert E[g["GaussianBond"]] == pytest.approx(-sum(float(r[5]) for r in top["angles"] if r[3] == "21"), rel=1e-3)
    assert E[g["GaussianDihedral"]] == pytest.approx(
        -sum(float(r[6]) for r in top["dihedrals"] if r[4] == "41"), rel=1e-3)
    assert 0 <= E[g["HarmonicBond"]] < 0.05


# ------------------------------------------------ sh3 through both readers
def test_sh3_genesis_matches_toml_reader(tmp_path, oracle):
    gen, ig, xg, vg = host_abi.load_genesis(sh3_genesis.write(str(tmp_path)))
    tom, it, xt, vt = host_abi.load_toml(sh3_toml.write(str(tmp_path / "sh3.toml")))
    assert gen.n == tom.n == 56 and vg is None
    np.testing.assert_array_equal(gen.masses, tom.masses)
    np.testing.assert_allclose(xg, xt, rtol=0, atol=5.1e-5)     # .gro keeps 4 decimals in nm
    kg, kt = _by_kind(gen), _by_kind(tom)
    for kind in (abi.GAUSSIAN_BOND, abi.GO_CONTACT, abi.GAUSSIAN_DIHEDRAL, abi.FLEXIBLE_LOCAL_ANGLE):
        rg, rt = _rows(kg[kind]), _rows(kt[kind])
        assert rg.keys() == rt.keys(), kind
        for key in rt:
            np.testing.assert_allclose(rg[key], rt[key], rtol=1e-8, err_msg=str(kind))
    # bonds: the GENESIS reader halves the file's coefficient, written here as 2k
    rg, rt = _rows(kg[abi.HARMONIC_BOND]), _rows(kt[abi.HARMONIC_BOND])
    assert rg.keys() == rt.keys()
    for key in rt:
        np.testing.assert_allclose(rg[key], rt[key], rtol=1e-10)
    # FLD: the TOML classes plus, for each X-GLY / X-PRO central pair, the
    # X-R3 class the GENESIS reader also selects (and R2-PRO dropped for GLY-PRO)
    fg, ft = set(_rows(kg[abi.FLEXIBLE_LOCAL_DIHEDRAL])), set(_rows(kt[abi.FLEXIBLE_LOCAL_DIHEDRAL]))
    res = sh3_genesis.residues()
    extra = {k for k in fg - ft}
    assert all(res[k[2]] in ("GLY", "PRO") and k[4] == tuple(np.round(aicg2.fld_globals((res[k[1]], "R3")), 9))
               for k in extra)
    assert len(extra) == sum(1 for k in ft if res[k[2]] in ("GLY", "PRO") and
                             k[4] != tuple(np.round(aicg2.fld_globals(("GLY", "PRO")), 9)))
    assert ft - fg == set()
    (eg,), (et,) = kg[abi.EXCLUDED_VOLUME], kt[abi.EXCLUDED_VOLUME]
    np.testing.assert_allclose(eg.particle_params, et.particle_params, rtol=1e-15)
    assert eg.globals[0] == pytest.approx(et.globals[0], rel=1e-15)
    assert eg.cutoff == pytest.approx(et.cutoff, rel=1e-15)
    assert {tuple(e) for e in eg.exclusions} == {tuple(e) for e in et.exclusions}
    # integrator: GENESIS gives ps and 1/ps directly
    assert (ig.temperature, ig.friction, ig.dt, ig.seed) == (300.0, 0.01, 0.01, 12345)
    # same energies for every group but FLD, at one perturbed configuration
    xp = xt + np.random.default_rng(0).normal(scale=0.01, size=xt.shape)
    _, Eg = oracle.evaluate(gen, xp)
    _, Et = oracle.evaluate(tom, xp)
    for name in ("GaussianBond", "GoContact", "FlexibleLocalAngle", "GaussianDihedral", "ExcludedVolume",
                 "HarmonicBond"):
        assert Eg[gen.group_ids[name]] == pytest.approx(Et[tom.group_ids[name]], rel=1e-7, abs=1e-9), name


def test_periodic_box_from_boundary_section(tmp_path):
    s, *_ = host_abi.load_genesis(sh3_genesis.write(str(tmp_path), boundary=(80.0, 90.0, 100.0)))
    assert s.box == pytest.approx((8.0, 9.0, 10.0))
    assert all(f.use_periodic for f in s.forces)


# -------------------------------------------------------- reader behaviour
def _edit(path, old, new, count=1):
    text = open(path).read()
    assert old in text, old
    open(path, "w").write(text.replace(old, new, count))


def _load_error(inp):
    with pytest.raises(abi.OcmError) as e:
        host_abi.load_genesis(inp)
    assert e.value.code == abi.OCM_ERR_INVALID
    return str(e.value)


@pytest.mark.parametrize("file,old,new,needle", [
    ("sh3.itp", "[ atoms ]", "[ atoms_x ]", "There is no atoms section"),
    ("param/atom_types.itp", "[ atomtypes ]", "[ atomtypes_x ]", "There is no `[ atomtypes ]` section"),
    ("sh3.itp", "[ moleculetype ]", "[ moleculetype_x ]", "There is no `[ moleculetype ]` section"),
    ("param/atom_types.itp", "    2.5104\n", "    2.5105\n", "uniform eps"),
    ("param/atom_types.itp", "  PHE  6", "  THR  6", "`THR` key was duplicated"),
    ("sh3.top", "; common parameters\n", "stray line\n", "does not belong to any section"),
    ("sh3.itp", "         1        23", "        57        23", "outside 1..56"),
    ("sh3.inp", "type = NOBC", "type = PBC", "box_size_x"),
    ("sh3.inp", "grocrdfile = sh3.gro", "grocrdfile = none.gro", "file open error"),
    ("sh3.inp", "timestep = 0.01", "timestep = fast", "not a number"),
    ("sh3.inp", "[DYNAMICS]", "[DYN]", "There is no [DYNAMICS] section"),
    ("sh3.gro", "          56\n", "          57\n", "the .gro file ends before column 21"),
])
def test_reader_errors(tmp_path, file, old, new, needle):
    inp = sh3_genesis.write(str(tmp_path))
    _edit(str(tmp_path / file), old, new)
    assert needle in _load_error(inp)


def test_gro_shorter_than_its_count(tmp_path):
    inp = sh3_genesis.write(str(tmp_path))
    _edit(str(tmp_path / "sh3.gro"), "   0.0   0.0   0.0\n", "")
    _edit(str(tmp_path / "sh3.gro"), "          56\n", "          57\n")
    assert "ends after 56 of 57 atoms" in _load_error(inp)


def test_nrexcl_must_agree(tmp_path):
    inp = sh3_genesis.write(str(tmp_path))
    _edit(str(tmp_path / "sh3.itp"), "sh3                   3\n", "sh3                   3\nsh3b                  2\n")
    assert "all nrexcl need to be same" in _load_error(inp)


def test_inp_and_top_syntax(tmp_path):
    """Comments, repeated keys (the first is kept), repeated sections (their
    lines concatenate), header spacing, and an #include with a trailing
    comment, which stays a plain line as the reference reads it."""
    inp = sh3_genesis.write(str(tmp_path))
    _edit(str(tmp_path / "sh3.inp"), "nsteps = 1000", "nsteps = 1000 # steps\nnsteps = 5\n  iseed=77  ")
    _edit(str(tmp_path / "sh3.inp"), "[ENSEMBLE]", "  [ ENSEMBLE ]  ")
    s, it, x, v = host_abi.load_genesis(inp)
    assert it.seed == 77          # the first `iseed` line wins over the later one
    # a second [ pairs ] section adds its lines to the first
    itp = str(tmp_path / "sh3.itp")
    text = open(itp).read()
    head, pairs = text.split("[ pairs ]\n")
    rows = pairs.strip().splitlines()
    open(itp, "w").write(head + "[ pairs ]\n" + "\n".join(rows[:100]) + "\n[pairs]\n" + "\n".join(rows[100:]) + "\n")
    s2, *_ = host_abi.load_genesis(inp)
    go = [f for f in s2.forces if f.kind == abi.GO_CONTACT][0]
    assert len(go.indices) == 215
    _edit(str(tmp_path / "sh3.top"), '#include "sh3.itp"', '#include "sh3.itp" ; molecule')
    # the unexpanded line joins [ atomtypes ], and the molecule never arrives
    assert "There is no atoms section" in _load_error(inp)


def test_without_iseed_the_seed_is_drawn(tmp_path):
    # iseed absent -> seed 0 (ReadGenesisInput.cpp:179-183), which the Langevin
    # integrator generator replaces with a random one, as OpenMM does for 0
    inp = sh3_genesis.write(str(tmp_path), iseed=None)
    seeds = {host_abi.load_genesis(inp)[1].seed for _ in range(3)}
    assert 0 not in seeds and len(seeds) == 3


def test_reference_sample_fixtures_load_and_match_the_readers():
    """The committed fixtures of the reference's samples (used by the GPU
    tests) rebuild the systems the readers give for the sample files."""
    from tests.golden import reference_inputs
    for name, (kind, path) in reference_inputs.SAMPLES.items():
        s, it, x, v = reference_inputs.load(name)
        assert x.shape == (s.n, 3) and len(s.forces) > 0
        if not os.path.exists(path):
            continue
        loader = host_abi.load_toml if kind == "toml" else host_abi.load_genesis
        r, rit, rx, rv = loader(path)
        assert np.array_equal(r.masses, s.masses) and np.array_equal(rx, x)
        assert (rit.temperature, rit.friction, rit.dt, rit.seed) == (it.temperature, it.friction, it.dt, it.seed)
        assert len(r.forces) == len(s.forces)
        for a, b in zip(r.forces, s.forces):
            assert (a.kind, a.group, a.use_periodic, a.cutoff, a.switch_distance) == \
                (b.kind, b.group, b.use_periodic, b.cutoff, b.switch_distance)
            for fld in reference_inputs.FIELDS:
                fa, fb = getattr(a, fld), getattr(b, fld)
                assert (fa is None) == (fb is None), fld
                if fa is not None:
                    assert np.array_equal(np.asarray(fa), np.asarray(fb)), fld
