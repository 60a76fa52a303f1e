"""The reference's other one-step schemes (model/make_step_operator.hh:
408-443) on the device against the oracle: ExplicitEuler, Heun, Shu3,
RungeKutta4, FractionalStepTheta and Alexander3 (ImplicitEuler and
Alexander2 are covered by the parity and evolve tests).

The oracle's tables are pinned by the stability functions and orders of
accuracy in tests/test_oracle_pins.py; here every step of a nonlinear
Gray-Scott model must match the oracle's: same Newton iterations, BiCGSTAB
iterations within a few percent (summation order), states to the parity bar,
with the assembled matrix and with point-coefficient storage.
"""
import numpy as np
import pytest

from tests.oracle_lib import OracleModel, rel_l2_per_species

pytestmark = pytest.mark.gpu

SCHEMES = ["ExplicitEuler", "Heun", "Shu3", "RungeKutta4", "FractionalStepTheta", "Alexander3"]


@pytest.mark.parametrize("storage", ["matrix", "points"])
@pytest.mark.parametrize("rk", SCHEMES)
def test_scheme_steps_match_oracle(dca, wl, rk, storage):
    print(f"case: scheme {rk} storage={storage}", flush=True)
    w = wl.gray_scott(n=16, cell_type="cube", precond="Jacobi", nl_rtol=1e-20, nl_atol=1e-26, lin_rtol=1e-12)
    ov = f"model.time_step_operator.type = {rk}\n"
    gpu = dca.Model(w.ini, ov + f"model.time_step_operator.linear_solver.storage = {storage}\n", 0)
    orc = OracleModel(w.ini, ov)
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    t = 0.0
    for k in range(4):
        st = gpu.step(t, w.dt)
        rc, newton, linear = orc.step(w.dt)
        assert st.status == 0 and rc == 0, (k, st.status, rc)
        assert st.newton_iterations == newton, (k, st.newton_iterations, newton)
        assert abs(st.linear_iterations - linear) <= max(2, linear // 20), (k, st.linear_iterations, linear)
        t += w.dt
        ug, tg = gpu.get_state()
        uo, _ = orc.get_state()
        assert abs(tg - t) < 1e-12
        assert max(rel_l2_per_species(gpu.compartments(), ug, uo)) <= 1e-8, k


@pytest.mark.parametrize("rk", ["Shu3", "Alexander3"])
def test_scheme_steps_on_simplices(dca, wl, rk):
    w = wl.mitchell_schaefer(n=4, cell_type="simplex", precond="Jacobi", nl_rtol=1e-20, nl_atol=1e-26,
                             lin_rtol=1e-12)
    ov = f"model.time_step_operator.type = {rk}\n"
    gpu = dca.Model(w.ini, ov, 0)
    orc = OracleModel(w.ini, ov)
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    t = 0.0
    for k in range(3):
        st = gpu.step(t, w.dt)
        rc, newton, _ = orc.step(w.dt)
        assert st.status == 0 and rc == 0 and st.newton_iterations == newton, k
        t += w.dt
        ug, _ = gpu.get_state()
        uo, _ = orc.get_state()
        assert max(rel_l2_per_species(gpu.compartments(), ug, uo)) <= 1e-8, k
