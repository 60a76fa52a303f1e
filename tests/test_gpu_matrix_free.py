"""The matrix-free Jacobian (linear_solver.matrix_free = true) against the oracle.

The apply y = J(x) v is checked against the oracle's assembled Jacobian times
v (1e-12 relative: summation order only), and matrix-free time steps against
the oracle's steps with the same operator and the identity preconditioner
(the reference's Richardson with weight 1): same Newton iterations, BiCGSTAB
iterations within a few percent, states within the parity bar.
"""
import numpy as np
import pytest

from tests.oracle_lib import OracleModel, rel_l2_per_species

pytestmark = pytest.mark.gpu

MF = ("model.time_step_operator.linear_solver.matrix_free = true\n"
      "model.time_step_operator.linear_solver.preconditioner.type = Richardson\n")

APPLY_CASES = [
    ("gs_q1", lambda wl: wl.gray_scott(n=12, cell_type="cube")),
    ("gs_p1", lambda wl: wl.gray_scott(n=12, cell_type="simplex")),
    ("ms_q1", lambda wl: wl.mitchell_schaefer(n=6, cell_type="cube")),
    ("ms_p1", lambda wl: wl.mitchell_schaefer(n=5, cell_type="simplex")),
    # blocks above 4 x 4 (one lane per block row)
    ("req_2d_6sp_1c", lambda wl: wl.random_request(11, side2=(10, 12), steps_range=(3, 3))),
    ("req_3d_8sp_1c", lambda wl: wl.random_request(18, side3=(5, 6), steps_range=(3, 3))),
]


def _perturbed_state(w, model, seed=1):
    u = w.initial_state(model)
    rng = np.random.default_rng(seed)
    return u + 0.05 * rng.standard_normal(u.size)


@pytest.mark.parametrize("name,make", APPLY_CASES, ids=[c[0] for c in APPLY_CASES])
@pytest.mark.parametrize("cm,ca", [(1.0, 0.0), (0.0, 0.7), (1.0, 0.3)])
def test_jacobian_apply_matches_oracle_matrix(dca, wl, name, make, cm, ca):
    print(f"case: apply {name} cm={cm} ca={ca}", flush=True)
    w = make(wl)
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    x = _perturbed_state(w, gpu)
    v = np.random.default_rng(7).standard_normal(gpu.n_dof)
    y = gpu.jacobian_apply(x, 0.3, cm, ca, v)
    _, vals = orc.assemble(x, 0.3, cm, ca)
    ref = orc.csr(vals) @ v
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    # the same operator as the device's own assembled Jacobian
    _, J = gpu.assemble(x, 0.3, cm, ca)
    assert np.abs(y - gpu.spmv(J, v)).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("make", [lambda wl: wl.two_compartments(n=12, cell_type="cube"),
                                  lambda wl: wl.two_compartments(n=10, cell_type="simplex"),
                                  lambda wl: wl.random_request(4, side3=(5, 6), steps_range=(3, 3))],
                         ids=["two_q1", "two_p1", "req_3d_8sp_2c"])
def test_jacobian_apply_with_transmission_terms(dca, wl, make):
    """Interface blocks (both sides, both directions) are applied by the facet
    kernels on top of the tiles' volume blocks."""
    w = make(wl)
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    x = _perturbed_state(w, gpu)
    v = np.random.default_rng(9).standard_normal(gpu.n_dof)
    for cm, ca in [(0.0, 0.7), (1.0, 0.3)]:
        y = gpu.jacobian_apply(x, 0.2, cm, ca, v)
        _, vals = orc.assemble(x, 0.2, cm, ca)
        ref = orc.csr(vals) @ v
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


STEP_CASES = [
    ("gs_q1", lambda wl: wl.gray_scott(n=24, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26, lin_rtol=1e-12), 6),
    ("ms_q1", lambda wl: wl.mitchell_schaefer(n=8, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26,
                                              lin_rtol=1e-12), 4),
    ("req_2d_6sp_1c", lambda wl: wl.random_request(11, side2=(10, 12), nl_rtol=1e-20, nl_atol=1e-28,
                                                   precision="fp64"), 3),
    ("two_q1", lambda wl: wl.two_compartments(n=16, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26,
                                              lin_rtol=1e-12), 3),
]


@pytest.mark.parametrize("name,make,steps", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_matrix_free_steps_match_oracle(dca, wl, name, make, steps):
    print(f"case: matrix-free steps {name}", flush=True)
    w = make(wl)
    gpu = dca.Model(w.ini, MF, 0)
    orc = OracleModel(w.ini, MF)
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    t = 0.0
    for k in range(steps):
        st = gpu.step(t, w.dt)
        rc, newton, linear = orc.step(w.dt)
        assert st.status == 0 and rc == 0, (k, st.status, rc)
        assert st.newton_iterations == newton, (k, st.newton_iterations, newton)
        # unpreconditioned BiCGSTAB to 1e-12 is sensitive to summation order:
        # the counts agree to a few percent, the states to the parity bar
        assert abs(st.linear_iterations - linear) <= max(2, linear // 20), (k, st.linear_iterations, linear)
        t += w.dt
        ug, _ = gpu.get_state()
        uo, _ = orc.get_state()
        assert max(rel_l2_per_species(gpu.compartments(), ug, uo)) <= 1e-8, k


def test_matrix_free_holds_no_matrix(dca, wl):
    w = wl.gray_scott(n=64, cell_type="cube")
    assembled = dca.Model(w.ini, "", 0)
    free = dca.Model(w.ini, MF, 0)
    B = assembled.block_size
    matrix = 8 * assembled.n_blocks * B * B
    assert free.device_bytes() <= assembled.device_bytes() - matrix
    with pytest.raises(dca.DcgError, match="no assembled Jacobian"):
        free.assemble(w.initial_state(free), 0.0, 1.0, 0.1)
