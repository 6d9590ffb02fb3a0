"""The reference's own sample inputs on the device, from the fixtures of
tests/golden/reference_inputs.py (the C++ front end's reading of
sample/toml_input/FUS_100_HPS.toml and sample/genesis_input/1SRL_cg.*).

- FUS_100_HPS: 16,300 HPS beads in a periodic 15.3 x 17.7 x 113.8 nm box.
  Its start is a packed one with overlapping beads (pair forces up to
  7e7 kJ/mol/nm), which is why the sample asks for energy minimization; the
  device minimizes it, then forces and energies are compared with the FP64
  oracle, and again after 1,000 steps.
- 1SRL (GENESIS AICG2+ input, 56 beads, open boundary): forces and energies
  at the .gro coordinates and after 2,000 steps.
"""
import sys

import numpy as np
import pytest

from opencafemol_amd import abi
from opencafemol_amd.context import Context
from tests.golden import reference_inputs
from tests.tolerances import assert_energies_close, assert_forces_close

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    sys.__stderr__.write(f"\n[gpu-case] {request.node.nodeid}\n")
    sys.__stderr__.flush()
    yield


def _compare(c, s, oracle, cells):
    st = c.get_state(abi.STATE_POSITIONS | abi.STATE_FORCES | abi.STATE_ENERGY)
    Fo, Eo = (oracle.evaluate_cells if cells else oracle.evaluate)(s, st["positions"])
    assert_forces_close(st["forces"], Fo)
    assert_energies_close(st["energy_by_group"], Eo)
    return st


def test_fus_100_hps_sample(oracle):
    s, it, x, v = reference_inputs.load("fus_100_hps")
    assert s.n == 16300 and s.box is not None
    with Context(s, it) as c:
        c.set_positions(x)
        c.minimize(tolerance=10.0, max_iterations=2000)
        _compare(c, s, oracle, True)
        c.set_velocities_to_temperature(it.temperature, it.seed)
        assert c.step(1000) == abi.OCM_OK
        st = _compare(c, s, oracle, True)
        assert st["step"] == 1000 and np.isfinite(st["positions"]).all()


def test_1srl_genesis_sample(oracle):
    s, it, x, v = reference_inputs.load("1srl_cg")
    assert s.n == 56 and s.box is None
    with Context(s, it) as c:
        c.set_positions(x)
        _compare(c, s, oracle, False)
        assert c.step(2000) == abi.OCM_OK
        st = _compare(c, s, oracle, False)
        assert st["step"] == 2000
