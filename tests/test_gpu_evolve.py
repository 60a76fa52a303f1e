"""The adaptive time loop (dcg_evolve) on the device against the oracle's.

The oracle's loop is pinned on the CPU (tests/test_oracle_pins.py: closed
form and snapping, retry from the unchanged state, minimum step).  Here the
device loop must take the same decisions: the same accepted step sizes, the
same number of failed attempts, and the same state to the parity bar.
"""
import numpy as np
import pytest

from tests.oracle_lib import OracleModel, rel_l2_per_species
from tests.test_oracle_pins import ALPHA, _exp_ini, _gs_fail_ini, _gs_state

pytestmark = pytest.mark.gpu


def test_evolve_closed_form_on_device(dca):
    print("case: evolve exp.ini Alexander2", flush=True)
    m = dca.Model(_exp_ini(2, "cube", "Alexander2"), "", 0)
    m.set_state(np.ones(m.n_dof), 0.0)
    st, acc = m.evolve(t_end=1.7, dt=0.05, increase_factor=1.1, decrease_factor=0.5, min_step=0.0, max_step=0.0,
                       snap_to_end_time=1)
    assert st.status == 0 and st.failed_steps == 0
    u, t = m.get_state()
    assert abs(t - 1.7) <= 1e-14 and abs(acc.sum() - 1.7) <= 1e-13
    z = -2.0 * acc
    R = (1.0 + (1.0 - 2.0 * ALPHA) * z) / (1.0 - ALPHA * z) ** 2
    assert np.abs(u - np.prod(R)).max() <= 1e-12


def _ini(nl_rtol, lin_rtol="1e-12"):
    return _gs_fail_ini(3).replace(
        "nonlinear_solver.convergence_condition.relative_tolerance = 1e-16",
        f"nonlinear_solver.convergence_condition.relative_tolerance = {nl_rtol}").replace(
        "linear_solver.convergence_condition.relative_tolerance = 1e-12",
        f"linear_solver.convergence_condition.relative_tolerance = {lin_rtol}")


def _oracle_evolve(ini, u0, dt, inc, max_step):
    o = OracleModel(ini)
    o.set_state(u0, 0.0)
    rc, acc, failed, dt_next = o.evolve(40.0, dt, increase=inc, min_step=1e-3, max_step=max_step)
    return o, rc, acc, failed, dt_next


# (Newton tolerance on the squared defect, first step, growth, maximum step).
# Whether a step converges within three Newton iterations depends on how far
# each adaptive linear solve got, so a step size close to the edge can go
# either way under rounding.  These runs are chosen so that every decision
# survives halving and doubling the Newton tolerance and a ten times looser
# linear tolerance in the oracle (checked below before the device runs):
# decisions the device must then reproduce exactly.
CASES = [("1e-12", 10.0, 2.0, 0.0), ("1e-10", 6.0, 2.0, 6.0)]


@pytest.mark.parametrize("nl_rtol,dt,inc,max_step", CASES)
def test_evolve_matches_oracle_with_failures(dca, nl_rtol, dt, inc, max_step):
    print(f"case: evolve with retries, rtol={nl_rtol} dt={dt} inc={inc} max_step={max_step}", flush=True)
    ini = _ini(nl_rtol)
    u0 = _gs_state(OracleModel(ini))
    orc, rc, acc_o, failed_o, dt_o = _oracle_evolve(ini, u0, dt, inc, max_step)
    assert rc == 0 and failed_o >= 3
    e = float(nl_rtol)
    for other in (_ini(repr(0.5 * e)), _ini(repr(2.0 * e)), _ini(nl_rtol, "1e-11")):
        _, rc2, acc2, failed2, _ = _oracle_evolve(other, u0, dt, inc, max_step)
        assert rc2 == 0 and failed2 == failed_o and acc2.size == acc_o.size
        assert np.allclose(acc2, acc_o, rtol=1e-12, atol=0)
    m = dca.Model(ini, "", 0)
    m.set_state(u0, 0.0)
    st, acc = m.evolve(t_end=40.0, dt=dt, increase_factor=inc, decrease_factor=0.5, min_step=1e-3,
                       max_step=max_step, snap_to_end_time=1)
    assert st.status == 0
    assert st.failed_steps == failed_o
    assert acc.size == acc_o.size and np.allclose(acc, acc_o, rtol=1e-12, atol=0)
    assert abs(st.dt - dt_o) <= 1e-12 * dt_o
    if max_step:
        assert acc.max() <= max_step * (1 + 1e-12)
    ug, tg = m.get_state()
    uo, to = orc.get_state()
    assert abs(tg - 40.0) <= 1e-12 and abs(to - 40.0) <= 1e-12
    assert max(rel_l2_per_species(m.compartments(), ug, uo)) <= 1e-8


def test_evolve_minimum_step_leaves_state(dca):
    print("case: evolve below the minimum step", flush=True)
    ini = _gs_fail_ini()
    m = dca.Model(ini, "", 0)
    u0 = _gs_state(OracleModel(ini))
    m.set_state(u0, 0.0)
    st, acc = m.evolve(t_end=40.0, dt=16.0, increase_factor=1.5, decrease_factor=0.5, min_step=4.0, max_step=0.0,
                       snap_to_end_time=1)
    assert st.status == 5 and st.failed_steps == 3 and acc.size == 0
    u, t = m.get_state()
    assert t == 0.0 and np.array_equal(u, u0)


@pytest.mark.parametrize("precond", ["Jacobi", "ILU"])
def test_solver_iterations_match_oracle(dca, precond):
    """Same preconditioner as the reference's (istl SeqJac on the scalar
    Jacobian is the inverse of its diagonal; ILU0 over dense species blocks
    is the scalar ILU0): the device's Newton and BiCGSTAB iteration counts
    follow the oracle's step by step."""
    print(f"case: iteration counts, {precond}", flush=True)
    ini = _ini("1e-10").replace("linear_solver.preconditioner.type = Jacobi",
                                f"linear_solver.preconditioner.type = {precond}")
    o = OracleModel(ini)
    u0 = _gs_state(o)
    o.set_state(u0, 0.0)
    # the assembled matrix, as the reference solves (point-coefficient storage
    # is the same operator with its sums in another order)
    m = dca.Model(ini, "model.time_step_operator.linear_solver.storage = matrix\n", 0)
    for dt in (0.5, 1.0, 1.25, 0.75):
        u, t = o.get_state()
        m.set_state(u, t)
        st = m.step(t, dt)
        rc, newton, linear = o.step(dt)
        assert (st.status, st.newton_iterations) == (rc, newton) == (0, newton)
        assert abs(st.linear_iterations - linear) <= 1, (dt, st.linear_iterations, linear)


@pytest.mark.parametrize("case", ["gs_q1_2d_64", "ms_q1_3d_9", "two_q1_2d_24", "gs_p1_2d_33"])
def test_ilu_sweeps_match_oracle(dca, wl, case):
    """Block ILU0 through the one-workgroup sweeps (per-row entry tables, the
    last levels' results in an LDS ring, the factorisation's precomputed
    update pairs): the same Newton iterations, BiCGSTAB iterations within one
    and the oracle's state to the parity bar, on 2D and 3D box grids, two
    compartments (rows of two block sizes) and the Kuhn split."""
    print(f"case: ILU0 sweeps {case}", flush=True)
    make = {
        "gs_q1_2d_64": lambda: wl.gray_scott(n=64, cell_type="cube", precond="ILU", nl_rtol=1e-12, lin_rtol=1e-12),
        "ms_q1_3d_9": lambda: wl.mitchell_schaefer(n=9, cell_type="cube", precond="ILU", nl_rtol=1e-12,
                                                   lin_rtol=1e-12),
        "two_q1_2d_24": lambda: wl.two_compartments(n=24, cell_type="cube", precond="ILU", nl_rtol=1e-20),
        "gs_p1_2d_33": lambda: wl.gray_scott(n=33, cell_type="simplex", precond="ILU", nl_rtol=1e-12,
                                             lin_rtol=1e-12),
    }[case]
    w = make()
    o = OracleModel(w.ini)
    m = dca.Model(w.ini, "model.time_step_operator.linear_solver.storage = matrix\n", 0)
    u0 = w.initial_state(m)
    o.set_state(u0, 0.0)
    m.set_state(u0, 0.0)
    t = 0.0
    for _ in range(2):
        st = m.step(t, w.dt)
        rc, newton, linear = o.step(w.dt)
        assert (st.status, rc) == (0, 0)
        assert st.newton_iterations == newton
        assert abs(st.linear_iterations - linear) <= max(1, linear // 50), (st.linear_iterations, linear)
        t += w.dt
    u, _ = m.get_state()
    uo, _ = o.get_state()
    assert max(rel_l2_per_species(m.compartments(), u, uo)) < 1e-8
    o.close()
