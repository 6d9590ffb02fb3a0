"""C++ front end on the GPU: the ocm_run driver end to end, and forces of
a TOML-built system against the oracle."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from opencafemol_amd import abi, host_abi
from opencafemol_amd.context import Context
from tests.golden import sh3_toml
from tests.tolerances import assert_energies_close, assert_forces_close

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(host_abi.LIB_PATH), "..", "bin", "ocm_run")


@pytest.fixture(autouse=True)
def _announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()
    yield


def test_toml_system_forces_match_oracle(tmp_path, oracle):
    path = sh3_toml.write(str(tmp_path / "sh3.toml"))
    s, it, x, v = host_abi.load_toml(path)
    xt = x + np.random.default_rng(3).normal(scale=0.02, size=x.shape)
    with Context(s, it) as c:
        c.set_positions(xt)
        st = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY)
    Fo, Eo = oracle.evaluate(s, xt)
    assert_forces_close(st["forces"], Fo)
    assert_energies_close(st["energy_by_group"], Eo)


def test_membrane_toml_forces_match_oracle(tmp_path, oracle):
    """Per-bead WCA, iSoLF, table WCA and table LJ attraction, capped Go
    contacts and cosine dihedrals, read by the C++ front end, on the device;
    then 500 steps from the relaxed patch stay finite and bit-identical to a
    second run."""
    from tests import membrane_input
    p = tmp_path / "membrane.toml"
    p.write_text(membrane_input.toml_text(27, capping_ratio=0.95))
    s, it, x, v = host_abi.load_toml(str(p))
    xt = x + np.random.default_rng(4).normal(scale=0.02, size=x.shape)
    with Context(s, it) as c:
        c.set_positions(xt)
        st = c.get_state(abi.STATE_FORCES | abi.STATE_ENERGY)
    Fo, Eo = oracle.evaluate(s, xt)
    assert_forces_close(st["forces"], Fo)
    assert_energies_close(st["energy_by_group"], Eo)
    finals = []
    for _ in range(2):
        with Context(s, it) as c:
            c.set_positions(xt)
            c.minimize(tolerance=10.0, max_iterations=200)
            c.set_velocities_to_temperature(it.temperature, 9)
            assert c.step(500) == abi.OCM_OK
            finals.append(c.get_state(abi.STATE_POSITIONS)["positions"])
    assert np.isfinite(finals[0]).all()
    assert np.array_equal(finals[0], finals[1])


def test_cli_runs_sh3_and_writes_pdb_and_energy(tmp_path):
    path = sh3_toml.write(str(tmp_path / "sh3.toml"), total_step=500, save_step=100, with_velocity=True)
    r = subprocess.run([EXE, path], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    pdb = (tmp_path / "sh3.pdb").read_text()
    assert pdb.count("MODEL") == 5 and pdb.count("ATOM") == 5 * 56
    ene = (tmp_path / "sh3.ene").read_text().splitlines()
    assert ene[0] == "# unit of energy : kcal/mol"
    cols = ene[1].split()
    assert cols[:2] == ["#", "timestep"] and cols[-1] == "kinetic_energy"
    rows = [list(map(float, l.split())) for l in ene[2:]]
    assert len(rows) == 5 and all(np.isfinite(r).all() for r in rows)
    assert rows[0][0] == 0 and rows[-1][0] == 400


def test_cli_writes_dcd_for_periodic_input(tmp_path):
    from tests.test_frontend import HPS
    text = HPS.replace("total_step = 1000_000", "total_step = 300").replace("save_step = 1000", "save_step = 100")
    p = tmp_path / "hps.toml"
    p.write_text(text)
    r = subprocess.run([EXE, str(p)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    data = (tmp_path / "hps_position.dcd").read_bytes()
    assert struct.unpack("<i4s", data[:8]) == (84, b"CORD")
    nset, first, interval, nsteps = struct.unpack("<4i", data[8:24])
    assert (nset, interval, nsteps) == (3, 100, 300)
    natom = struct.unpack("<i", data[8 + 84 + 8 + 84 + 4:8 + 84 + 8 + 84 + 8])[0]
    assert natom == 4
    header = 4 + 84 + 4 + 4 + 84 + 4 + 4 + 4 + 4
    frame = (4 + 48 + 4) + 3 * (4 + 4 * natom + 4)
    assert len(data) == header + 3 * frame
    assert (tmp_path / "hps_velocity.dcd").stat().st_size == header + 3 * (3 * (4 + 4 * natom + 4))


def test_cli_runs_genesis_input(tmp_path, oracle):
    """ocm_run on a GENESIS .inp: the initial-state PDB, DCD frames with the
    timestep in ps, and an .ene whose step-0 row is the oracle's energies at
    the .gro positions with zero kinetic energy (this input mode sets no
    velocities, ReadGenesisInput.cpp:499-503)."""
    from tests.golden import sh3_genesis
    inp = sh3_genesis.write(str(tmp_path), nsteps=300, crdout_period=100, platform="CUDA")
    spec, integ, x, v = host_abi.load_genesis(inp)
    r = subprocess.run([EXE, inp], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    pdb = (tmp_path / "sh3_init.pdb").read_text().splitlines()
    assert pdb[0] == "MODEL     0" and pdb[-1] == "ENDMDL" and len(pdb) == 58
    for i, line in enumerate(pdb[1:-1]):
        assert line.startswith("ATOM  %5d   CA " % (i + 1)) and line.endswith("  1.00  0.00")
        np.testing.assert_allclose([float(line[30:38]), float(line[38:46]), float(line[46:54])], x[i] * 10,
                                   atol=5e-4)
    data = (tmp_path / "sh3_md_position.dcd").read_bytes()
    assert struct.unpack("<i4s", data[:8]) == (84, b"CORD")
    assert struct.unpack("<4i", data[8:24]) == (3, 0, 100, 300)
    assert struct.unpack("<f", data[44:48])[0] == pytest.approx(0.01)
    ene = (tmp_path / "sh3_md.ene").read_text().splitlines()
    names = ene[1].split()[2:-1]
    rows = [list(map(float, l.split())) for l in ene[2:]]
    assert [r[0] for r in rows] == [0, 100, 200]
    _, E = oracle.evaluate(spec, x)
    got = dict(zip(names, rows[0][1:-1]))
    for name, g in spec.group_ids.items():
        assert got[name] == pytest.approx(E[g] / 4.184, rel=1e-6, abs=2e-6), name
    assert rows[0][-1] == 0.0 and rows[-1][-1] > 0.0
