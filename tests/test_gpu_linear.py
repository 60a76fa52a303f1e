"""The linear defect-correction path (model.is_linear = true): one residual,
one Jacobian and one linear solve per stage instead of Newton
(make_step_operator.hh:210-245), on the device against the oracle.

Models: the reference's test/exp.ini (u' = -2u, whose implicit Euler and
Alexander2 steps have closed forms) and the heat kernel of its
test/gauss.ini, whose own u_error reduce key must stay below the reference's
bound of 0.5."""
import math

import numpy as np
import pytest

from tests.oracle_lib import OracleModel, rel_l2_per_species, rel_mass_l2_per_species
from tests.test_oracle_pins import ALPHA, _exp_ini, heat_kernel_ini

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cell_type", ["simplex", "cube"])
@pytest.mark.parametrize("rk", ["ImplicitEuler", "Alexander2"])
def test_linear_path_closed_form_on_device(dca, cell_type, rk):
    m = dca.Model(_exp_ini(2, cell_type, rk, D=0.3), "model.is_linear = true", 0)
    m.set_state(np.ones(m.n_dof), 0.0)
    dt, n = 0.1, 6
    t = 0.0
    for _ in range(n):
        st = m.step(t, dt)
        assert st.status == 0
        assert st.newton_iterations == (1 if rk == "ImplicitEuler" else 2)  # one correction per stage
        t += dt
    u, _ = m.get_state()
    z = -2.0 * dt
    R = 1.0 / (1.0 - z) if rk == "ImplicitEuler" else (1 + (1 - 2 * ALPHA) * z) / (1 - ALPHA * z) ** 2
    np.testing.assert_allclose(u, R ** n, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("cell_type,storage", [("simplex", "auto"), ("cube", "points"), ("cube", "matrix")])
def test_heat_kernel_linear_path_matches_oracle(dca, cell_type, storage):
    ini = heat_kernel_ini(cell_type)
    ov = f"model.time_step_operator.linear_solver.storage = {storage}\n"
    gpu = dca.Model(ini, ov, 0)
    orc = OracleModel(ini)
    new = dca.Model(ini, ov + "model.is_linear = false\n", 0)
    u0 = gpu.initial_state()  # initial.expression at time_begin = 1
    for m in (gpu, orc, new):
        m.set_state(u0, 1.0)
    vals = gpu.reduce()
    assert vals["u_error"] < 0.5
    t = 1.0
    for k in range(4):
        st = gpu.step(t, 0.05)
        assert st.status == 0 and st.newton_iterations == 2  # Alexander2, one correction per stage
        assert orc.step(0.05)[0] == 0
        assert new.step(t, 0.05).status == 0
        t += 0.05
        ug, tg = gpu.get_state()
        uo, _ = orc.get_state()
        un, _ = new.get_state()
        assert max(rel_l2_per_species(gpu.compartments(), ug, uo)) <= 1e-8, k
        assert max(rel_mass_l2_per_species(orc, ug, uo)) <= 1e-8, k
        # linear path and Newton solve the same linear stage equations
        assert max(rel_l2_per_species(gpu.compartments(), ug, un)) <= 1e-8, k
    assert abs(tg - 1.2) < 1e-12
    # the reference's own assertions on this model: max, min and L2 error
    vals = gpu.reduce()  # raises if an error expression fires
    ref, flags, rc = orc.reduce()
    assert rc == 0
    assert 0.0 < vals["u_error"] < 0.5
    for key in ("u_max", "u_min", "u_error"):
        assert vals[key] == pytest.approx(ref[key], rel=1e-8, abs=1e-12)
    assert vals["u_max"] < 1.0 / (4 * math.pi * 0.005)
