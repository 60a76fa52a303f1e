"""Writes the golden trajectories in this folder with the CPU oracle.

Each fixture holds the seeded initial state and the oracle's state after
every step of a small run of one benchmark model.  The oracle's own pins
(tests/test_oracle_pins.py) are what make these trustworthy; the fixtures
freeze them so that oracle regressions and the device path are both checked
against fixed data.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from tests.conftest import _load  # noqa: E402
from tests.oracle_lib import OracleModel  # noqa: E402

wl = _load("dca_workloads", ROOT / "dune-copasi_amd" / "workloads.py")

TIGHT = dict(nl_rtol=1e-24, nl_atol=1e-28, lin_rtol=1e-13)
FIXTURES = {
    "gray_scott_q1_16": (lambda: wl.gray_scott(n=16, cell_type="cube", precond="ILU", **TIGHT), 6),
    "gray_scott_p1_16": (lambda: wl.gray_scott(n=16, cell_type="simplex", precond="ILU", **TIGHT), 6),
    "mitchell_schaefer_q1_6": (lambda: wl.mitchell_schaefer(n=6, cell_type="cube", **TIGHT), 5),
    "mitchell_schaefer_p1_9": (lambda: wl.mitchell_schaefer(n=9, cell_type="simplex", **TIGHT), 5),
    "two_compartments_q1_20": (lambda: wl.two_compartments(n=20, cell_type="cube", **TIGHT), 5),
}


def run(name):
    make, steps = FIXTURES[name]
    w = make()
    o = OracleModel(w.ini)
    u0 = w.initial_state(o)
    o.set_state(u0, 0.0)
    traj = [u0]
    for _ in range(steps):
        rc, _, _ = o.step(w.dt)
        assert rc == 0
        traj.append(o.get_state()[0])
    return w, np.array(traj)


if __name__ == "__main__":
    for name in FIXTURES:
        w, traj = run(name)
        np.savez_compressed(HERE / f"{name}.npz", trajectory=traj, dt=w.dt, ini=w.ini)
        print(name, traj.shape)
