"""Golden trajectories (tests/golden/*.npz, written by make_golden.py): the
oracle must reproduce them; the device path is checked against them in
test_gpu_golden."""
from pathlib import Path

import numpy as np
import pytest

from tests.oracle_lib import OracleModel, rel_l2_per_species

GOLDEN = sorted((Path(__file__).parent / "golden").glob("*.npz"))


@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_oracle_reproduces_golden(path):
    g = np.load(path)
    traj, dt, ini = g["trajectory"], float(g["dt"]), str(g["ini"])
    o = OracleModel(ini)
    o.set_state(traj[0], 0.0)
    for k in range(1, len(traj)):
        assert o.step(dt)[0] == 0
        u, _ = o.get_state()
        assert max(rel_l2_per_species(o.compartments(), u, traj[k])) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_gpu_golden(dca, path):
    g = np.load(path)
    traj, dt, ini = g["trajectory"], float(g["dt"]), str(g["ini"])
    m = dca.Model(ini, "", 0)
    m.set_state(traj[0], 0.0)
    t = 0.0
    for k in range(1, len(traj)):
        st = m.step(t, dt)
        assert st.status == 0
        t += dt
        u, _ = m.get_state()
        assert max(rel_l2_per_species(m.compartments(), u, traj[k])) <= 1e-8, k
