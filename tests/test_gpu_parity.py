"""Device path vs the CPU oracle, through the C ABI.

Assembly (residual and Jacobian), SpMV, the Krylov solve and whole time-step
sequences are compared on the same seeded inputs.  FP64 tolerances:
assembly 1e-12 relative (summation order only), solutions 1e-8 relative L2
per species (the parity bar), linear solves to their tolerance.
"""
import numpy as np
import pytest

from tests.oracle_lib import OracleModel, bcsr_to_dense_csr, rel_l2_per_species, rel_mass_l2_per_species

pytestmark = pytest.mark.gpu

def _fallback_model(kind, cell_type):
    """Gray-Scott variants that need the coloured assembly kernel (kind "" is
    the plain model, for the tiled kernel)."""
    extra_u = {"adv": "velocity.x.expression = 0.02\nvelocity.y.expression = -0.01*v\n"
                      "velocity.jacobian.v.y.expression = -0.01\n",
               "nld": "",
               "num": "",
               "": ""}[kind]
    du = ("cross_diffusion.u.expression = 1e-3*(1+u^2)\ncross_diffusion.u.jacobian.u.expression = 2e-3*u\n"
          if kind == "nld" else "cross_diffusion.u.expression = 2e-3\n")
    jt = "[model]\njacobian.type = numerical\n" if kind == "num" else ""
    return f"""[grid]
dimension = 2
cells = 10 10
cell_type = {cell_type}
[compartments]
domain.expression = 1
{jt}[model.scalar_field.u]
compartment = domain
storage.expression = 1
{du}{extra_u}reaction.expression = -u*v^2 + 0.04*(1-u)
reaction.jacobian.u.expression = -v^2 - 0.04
reaction.jacobian.v.expression = -2*u*v
[model.scalar_field.v]
compartment = domain
storage.expression = 1
cross_diffusion.v.expression = 1e-3
reaction.expression = u*v^2 - 0.1*v
reaction.jacobian.u.expression = v^2
reaction.jacobian.v.expression = 2*u*v - 0.1
[model.time_step_operator]
type = ImplicitEuler
linear_solver.preconditioner.type = Jacobi
linear_solver.convergence_condition.relative_tolerance = 1e-12
linear_solver.convergence_condition.iteration_range = 1 2000
nonlinear_solver.convergence_condition.relative_tolerance = 1e-20
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-26
"""


class _Inline:
    def __init__(self, ini, dt=0.5):
        self.ini, self.dt = ini, dt

    def initial_state(self, model):
        ids = model.compartment_nodes(0)
        out = np.empty(2 * ids.size)
        out[0::2] = 1.0 - 0.5 * (ids % 7) / 7.0
        out[1::2] = 0.25 * (ids % 5) / 5.0
        return out


FALLBACK = [(f"{k}_{c}", (lambda k=k, c=c: _Inline(_fallback_model(k, c))))
            for k in ("adv", "nld", "num") for c in ("cube", "simplex")]


@pytest.mark.parametrize("name,make", FALLBACK, ids=[f[0] for f in FALLBACK])
def test_coloured_kernel_models(dca, name, make):
    w = make()
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    x = _perturbed_state(w, gpu)
    r_g, J_g = gpu.assemble(x, 0.1, 0.8, 0.6)
    r_o, v_o = orc.assemble(x, 0.1, 0.8, 0.6)
    assert np.abs(r_g - r_o).max() <= 1e-12 * np.abs(r_o).max()
    A_g = bcsr_to_dense_csr(gpu, J_g).toarray()
    A_o = orc.csr(v_o).toarray()
    tol = 1e-6 if name.startswith("num") else 1e-12  # forward differences round differently
    assert np.abs(A_g - A_o).max() <= tol * np.abs(A_o).max()
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    t = 0.0
    for k in range(4):
        assert gpu.step(t, w.dt).status == 0
        assert orc.step(w.dt)[0] == 0
        t += w.dt
        err = max(rel_l2_per_species(gpu.compartments(), gpu.get_state()[0], orc.get_state()[0]))
        assert err <= 1e-8, (k, err)


def test_interpolation_context_on_device(dca):
    """A reaction rate from a parser_context interpolation table (a jump in
    the middle, constant beyond the ends): device assembly and steps = oracle."""
    ini = _fallback_model("", "cube").replace(
        "reaction.expression = -u*v^2 + 0.04*(1-u)\nreaction.jacobian.u.expression = -v^2 - 0.04\n",
        "reaction.expression = -u*v^2 + k(position_x)*(1-u)\nreaction.jacobian.u.expression = -v^2 - k(position_x)\n")
    ini = ini.replace("[compartments]", "[parser_context.k]\ntype = interpolation\ndomain = 0.2 0.45 0.45 0.7\n"
                      "range = 0.02 0.06 0.01 0.08\n[compartments]")
    assert "k(position_x)" in ini
    w = _Inline(ini)
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    x = _perturbed_state(w, gpu)
    r_g, J_g = gpu.assemble(x, 0.1, 0.8, 0.6)
    r_o, v_o = orc.assemble(x, 0.1, 0.8, 0.6)
    assert np.abs(r_g - r_o).max() <= 1e-12 * np.abs(r_o).max()
    assert np.abs(bcsr_to_dense_csr(gpu, J_g).toarray() - orc.csr(v_o).toarray()).max() <= 1e-12 * np.abs(v_o).max()
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    t = 0.0
    for k in range(3):
        assert gpu.step(t, w.dt).status == 0 and orc.step(w.dt)[0] == 0
        t += w.dt
        assert max(rel_l2_per_species(gpu.compartments(), gpu.get_state()[0], orc.get_state()[0])) <= 1e-8, k


CASES = [
    ("gs_q1", lambda wl: wl.gray_scott(n=12, cell_type="cube")),
    ("gs_p1", lambda wl: wl.gray_scott(n=12, cell_type="simplex")),
    ("ms_q1", lambda wl: wl.mitchell_schaefer(n=6, cell_type="cube")),
    ("ms_p1", lambda wl: wl.mitchell_schaefer(n=5, cell_type="simplex")),
    ("two_q1", lambda wl: wl.two_compartments(n=20, cell_type="cube")),
    ("two_p1", lambda wl: wl.two_compartments(n=16, cell_type="simplex")),
    # random mass-action requests with large blocks (block rows split over lanes)
    ("req_3d_8sp_2c", lambda wl: wl.random_request(4, side3=(5, 6), steps_range=(3, 3))),
    ("req_2d_6sp_1c", lambda wl: wl.random_request(11, side2=(10, 12), steps_range=(3, 3))),
]


def _perturbed_state(w, model, seed=1):
    u = w.initial_state(model)
    rng = np.random.default_rng(seed)
    return u + 0.05 * rng.standard_normal(u.size)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("cm,ca", [(1.0, 0.0), (0.0, 0.7), (1.0, 0.3)])
def test_assembly_matches_oracle(dca, wl, name, make, cm, ca):
    w = make(wl)
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    assert gpu.n_dof == orc.n_dof
    x = _perturbed_state(w, gpu)
    r_g, J_g = gpu.assemble(x, 0.3, cm, ca)
    r_o, v_o = orc.assemble(x, 0.3, cm, ca)
    scale = np.abs(r_o).max() + 1e-300
    assert np.abs(r_g - r_o).max() <= 1e-12 * scale
    A_g = bcsr_to_dense_csr(gpu, J_g).toarray()
    A_o = orc.csr(v_o).toarray()
    jscale = np.abs(A_o).max()
    assert np.abs(A_g - A_o).max() <= 1e-12 * jscale


@pytest.mark.parametrize("name,make", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_spmv_matches_scipy(dca, wl, name, make):
    w = make(wl)
    gpu = dca.Model(w.ini, "", 0)
    rng = np.random.default_rng(3)
    J = rng.standard_normal((gpu.n_blocks, gpu.block_size, gpu.block_size))
    x = rng.standard_normal(gpu.n_dof)
    y = gpu.spmv(J, x)
    A = bcsr_to_dense_csr(gpu, J)
    ref = A @ x
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("precond", ["Jacobi", "ILU"])
def test_linear_solve(dca, wl, precond):
    w = wl.gray_scott(n=16, cell_type="cube", precond=precond)
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    x0 = _perturbed_state(w, gpu)
    _, J = gpu.assemble(x0, 0.0, 1.0, 1.0)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(gpu.n_dof)
    x, rc, its = gpu.linear_solve(J, b, rtol=1e-12, maxit=2000)
    assert rc == 0 and its > 0
    A = bcsr_to_dense_csr(gpu, J)
    assert np.linalg.norm(A @ x - b) <= 1e-11 * np.linalg.norm(b)


STEP_CASES = [
    ("gs_q1_ilu", lambda wl: wl.gray_scott(n=24, cell_type="cube", precond="ILU", nl_rtol=1e-20, nl_atol=1e-26,
                                           lin_rtol=1e-12), 8),
    ("gs_p1_jac", lambda wl: wl.gray_scott(n=24, cell_type="simplex", precond="Jacobi", nl_rtol=1e-20,
                                           nl_atol=1e-26, lin_rtol=1e-12), 8),
    ("ms_q1", lambda wl: wl.mitchell_schaefer(n=8, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26,
                                              lin_rtol=1e-12), 6),
    ("ms_p1", lambda wl: wl.mitchell_schaefer(n=6, cell_type="simplex", nl_rtol=1e-20, nl_atol=1e-26,
                                              lin_rtol=1e-12), 6),
    ("two_q1", lambda wl: wl.two_compartments(n=24, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26,
                                              lin_rtol=1e-12), 6),
    ("req_3d_8sp_2c", lambda wl: wl.random_request(4, side3=(5, 6), nl_rtol=1e-20, nl_atol=1e-28,
                                                   precision="fp64"), 3),
    ("req_2d_6sp_1c", lambda wl: wl.random_request(11, side2=(10, 12), nl_rtol=1e-20, nl_atol=1e-28,
                                                   precision="fp64"), 3),
]


@pytest.mark.parametrize("name,make,steps", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_time_steps_match_oracle(dca, wl, name, make, steps):
    w = make(wl)
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    t = 0.0
    for k in range(steps):
        st = gpu.step(t, w.dt)
        assert st.status == 0, (k, st.status)
        rc, _, _ = orc.step(w.dt)
        assert rc == 0
        t += w.dt
        ug, tg = gpu.get_state()
        uo, _ = orc.get_state()
        err = max(rel_l2_per_species(gpu.compartments(), ug, uo))
        assert err <= 1e-8, (k, err)
        assert abs(tg - t) < 1e-12
    # the parity bar also in the mass-matrix norm: per species, relative
    errm = max(rel_mass_l2_per_species(orc, ug, uo, t))
    assert errm <= 1e-8, errm


@pytest.mark.parametrize("dim,cell_type,n", [(2, "cube", 33), (2, "simplex", 27), (3, "cube", 10)])
def test_transmission_residual_is_flux_times_interface_measure(dca, dim, cell_type, n):
    """The device's facet residual against the closed form that pins the
    oracle's: per species, the sum of the transmission residual at a constant
    state is +-k_t (D - A) |Gamma_h| (local_operator.hh:918-927,956-963)."""
    from tests.test_oracle_pins import _BALL, _DISK, _interface_measure, _transmission_ini

    kt = 0.5
    gpu = dca.Model(_transmission_ini(dim, cell_type, n, kt, *(_DISK if dim == 2 else _BALL)), "", 0)
    gamma = _interface_measure(dim, cell_type, n)
    cs = gpu.compartments()
    x = np.empty(gpu.n_dof)
    x[cs[0]["dof_offset"]: cs[0]["dof_offset"] + cs[0]["n_nodes"]] = 0.3
    x[cs[1]["dof_offset"]: cs[1]["dof_offset"] + cs[1]["n_nodes"]] = 0.8
    r, _ = gpu.assemble(x, 0.0, 0.0, 1.0, jacobian=False)
    s_in = r[cs[0]["dof_offset"]: cs[0]["dof_offset"] + cs[0]["n_nodes"]].sum()
    s_out = r[cs[1]["dof_offset"]: cs[1]["dof_offset"] + cs[1]["n_nodes"]].sum()
    flux = kt * (0.3 - 0.8)
    assert s_in == pytest.approx(flux * gamma, rel=1e-12)
    assert s_out == pytest.approx(-flux * gamma, rel=1e-12)
