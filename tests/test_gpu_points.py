"""The point-coefficient Jacobian (linear_solver.storage = points) against the
assembled Jacobian and the oracle.

With points the Krylov operator keeps, per cell and quadrature point, the
coefficient block of the mass and reaction terms (the stage weights folded
in) and applies it, with the uniform diffusion, to each direction: the
operator of the assembled Jacobian with its sums in another order.  Checks:

* y = J(x) v against the oracle's assembled Jacobian times v and against the
  device's own assembled Jacobian (1e-12 relative: summation order only), on
  box and simplex cells, 2D and 3D, blocks of 2 to 8 species, two
  compartments with and without interface terms;
* time steps against the oracle (the parity bar) with the same Newton
  iterations and BiCGSTAB iterations within a few percent (summation order),
  for Jacobi and vertex-block Jacobi, FP64 and FP32 Krylov inside FP64
  refinement;
* the model holds no matrix, and the kernel-level entry points still take one.
"""
import numpy as np
import pytest

from tests.oracle_lib import OracleModel, bcsr_to_dense_csr, rel_l2_per_species

pytestmark = pytest.mark.gpu

# (point coefficients need a preconditioner built from diagonal blocks)
PTS = ("model.time_step_operator.linear_solver.storage = points\n"
       "model.time_step_operator.linear_solver.preconditioner.type = Jacobi\n")


def _no_interface(w):
    """The two-compartment workload without its transmission (outflow) keys."""
    w.ini = "\n".join(line for line in w.ini.splitlines() if "outflow." not in line) + "\n"
    return w


APPLY_CASES = [
    ("gs_q1", lambda wl: wl.gray_scott(n=12, cell_type="cube")),
    ("gs_p1", lambda wl: wl.gray_scott(n=12, cell_type="simplex")),
    ("ms_q1", lambda wl: wl.mitchell_schaefer(n=6, cell_type="cube")),
    ("ms_q1_odd", lambda wl: wl.mitchell_schaefer(n=11, cell_type="cube")),
    ("ms_p1", lambda wl: wl.mitchell_schaefer(n=5, cell_type="simplex")),
    ("two_q1_no_interface", lambda wl: _no_interface(wl.two_compartments(n=14, cell_type="cube"))),
    # interface (transmission) terms applied by the facet kernels
    ("two_q1", lambda wl: wl.two_compartments(n=14, cell_type="cube")),
    ("two_p1", lambda wl: wl.two_compartments(n=10, cell_type="simplex")),
    ("req_3d_8sp_2c", lambda wl: wl.random_request(4, side3=(5, 6), steps_range=(3, 3))),
    # blocks above 4 x 4 (corner values read from LDS)
    ("req_2d_6sp_1c", lambda wl: wl.random_request(11, side2=(10, 12), steps_range=(3, 3))),
    ("req_3d_8sp_1c", lambda wl: wl.random_request(18, side3=(5, 6), steps_range=(3, 3))),
]


def _perturbed_state(w, model, seed=1):
    u = w.initial_state(model)
    rng = np.random.default_rng(seed)
    return u + 0.05 * rng.standard_normal(u.size)


@pytest.mark.parametrize("name,make", APPLY_CASES, ids=[c[0] for c in APPLY_CASES])
@pytest.mark.parametrize("cm,ca", [(1.0, 0.0), (0.0, 0.7), (1.0, 0.3)])
def test_point_apply_matches_assembled(dca, wl, name, make, cm, ca):
    print(f"case: point apply {name} cm={cm} ca={ca}", flush=True)
    w = make(wl)
    gpu = dca.Model(w.ini, PTS, 0)
    orc = OracleModel(w.ini)
    x = _perturbed_state(w, gpu)
    v = np.random.default_rng(7).standard_normal(gpu.n_dof)
    y = gpu.jacobian_apply(x, 0.3, cm, ca, v)
    _, vals = orc.assemble(x, 0.3, cm, ca)
    ref = orc.csr(vals) @ v
    scale = np.abs(ref).max()
    assert np.abs(y - ref).max() <= 1e-12 * scale
    # the device's own assembled Jacobian (built on request for this entry point)
    _, J = gpu.assemble(x, 0.3, cm, ca)
    assert np.abs(y - bcsr_to_dense_csr(gpu, J) @ v).max() <= 1e-12 * scale


TILE_CASES = [
    ("gs_q1_2d", lambda wl: wl.gray_scott(n=37, cell_type="cube")),
    ("ms_q1_3d", lambda wl: wl.mitchell_schaefer(n=11, cell_type="cube")),
    ("two_q1_2d", lambda wl: wl.two_compartments(n=23, cell_type="cube")),
]


@pytest.mark.parametrize("name,make", TILE_CASES, ids=[c[0] for c in TILE_CASES])
@pytest.mark.parametrize("tile", [str(k) for k in range(10)])
def test_point_apply_every_tile_shape(dca, wl, name, make, tile):
    """The point apply forced to each candidate tile shape (sides that no
    shape divides, so edge tiles are partial) against the oracle's assembled
    Jacobian; with "auto" the set-up autotuner reports one timed entry per
    candidate and exactly one chosen shape per compartment."""
    print(f"case: point apply {name} point_tile={tile}", flush=True)
    w = make(wl)
    gpu = dca.Model(w.ini, PTS + f"model.time_step_operator.linear_solver.point_tile = {tile}\n", 0)
    orc = OracleModel(w.ini)
    x = _perturbed_state(w, gpu)
    v = np.random.default_rng(11).standard_normal(gpu.n_dof)
    y = gpu.jacobian_apply(x, 0.3, 1.0, 0.4, v)
    _, vals = orc.assemble(x, 0.3, 1.0, 0.4)
    ref = orc.csr(vals) @ v
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    # a forced shape times nothing and is the one in use in every compartment
    forced = gpu.point_tuning()
    assert forced and all(us == 0.0 for _, _, us, _ in forced)
    for c in sorted({c for c, _, _, _ in forced}):
        flags = [ch for c2, _, _, ch in forced if c2 == c]
        assert flags.count(True) == 1
        if int(tile) < len(flags):
            assert flags[int(tile)]
    if tile == "0":
        auto = dca.Model(w.ini, PTS, 0)
        tuning = auto.point_tuning()
        comps = sorted({c for c, _, _, _ in tuning})
        assert comps and all(sum(ch for c2, _, _, ch in tuning if c2 == c) == 1 for c in comps)
        assert all(us > 0 for _, _, us, _ in tuning)
        y2 = auto.jacobian_apply(x, 0.3, 1.0, 0.4, v)
        assert np.abs(y2 - ref).max() <= 1e-12 * np.abs(ref).max()


def _step_pair(dca, w, overrides, steps, oracle_overrides=None):
    pts = dca.Model(w.ini, PTS + overrides, 0)
    mat = dca.Model(w.ini, overrides, 0)
    orc = OracleModel(w.ini, overrides if oracle_overrides is None else oracle_overrides)
    u0 = w.initial_state(pts)
    for m in (pts, mat):
        m.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    t = 0.0
    for k in range(steps):
        sp = pts.step(t, w.dt)
        sm = mat.step(t, w.dt)
        rc, newton, linear = orc.step(w.dt)
        assert sp.status == 0 and sm.status == 0 and rc == 0, (k, sp.status, sm.status, rc)
        assert sp.newton_iterations == sm.newton_iterations, (k, sp.newton_iterations, sm.newton_iterations)
        # summation order moves a tight BiCGSTAB by an iteration or two
        assert abs(sp.linear_iterations - sm.linear_iterations) <= max(2, sm.linear_iterations // 20), \
            (k, sp.linear_iterations, sm.linear_iterations)
        t += w.dt
        up, _ = pts.get_state()
        uo, _ = orc.get_state()
        assert max(rel_l2_per_species(pts.compartments(), up, uo)) <= 1e-8, k
    return pts


STEP_CASES = [
    ("gs_q1", lambda wl: wl.gray_scott(n=24, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26, lin_rtol=1e-12), 4),
    ("ms_q1", lambda wl: wl.mitchell_schaefer(n=8, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26,
                                              lin_rtol=1e-12), 3),
    ("gs_p1", lambda wl: wl.gray_scott(n=20, cell_type="simplex", nl_rtol=1e-20, nl_atol=1e-26, lin_rtol=1e-12), 3),
    ("two_q1_no_interface", lambda wl: _no_interface(wl.two_compartments(n=16, cell_type="cube", nl_rtol=1e-20,
                                                                         nl_atol=1e-26, lin_rtol=1e-12)), 3),
    ("two_q1", lambda wl: wl.two_compartments(n=16, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26,
                                              lin_rtol=1e-12), 3),
    ("req_3d_8sp_1c", lambda wl: wl.random_request(18, side3=(6, 7), nl_rtol=1e-20, nl_atol=1e-28,
                                                   precision="fp64"), 3),
]


@pytest.mark.parametrize("name,make,steps", STEP_CASES, ids=[c[0] for c in STEP_CASES])
@pytest.mark.parametrize("precond", ["Jacobi", "VertexBlockJacobi"])
def test_point_steps_match_oracle(dca, wl, name, make, steps, precond):
    print(f"case: point steps {name} {precond}", flush=True)
    w = make(wl)
    pre = "model.time_step_operator.linear_solver.preconditioner.type = {}\n"
    # the oracle has no vertex-block Jacobi: its Jacobi steps reach the same
    # states at these tolerances; iterations are compared with the device's
    # assembled-matrix path
    _step_pair(dca, w, pre.format(precond), steps, pre.format("Jacobi"))


@pytest.mark.parametrize("make", [
    lambda wl: wl.mitchell_schaefer(n=10, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26, lin_rtol=1e-12),
    lambda wl: wl.two_compartments(n=16, cell_type="cube", nl_rtol=1e-20, nl_atol=1e-26, lin_rtol=1e-12),
], ids=["ms_q1", "two_q1"])
@pytest.mark.parametrize("floor", [0.0, 1e-4, 1e-3])
def test_point_steps_mixed_precision(dca, wl, make, floor):
    """FP32 point-coefficient applies (with FP32 facet applies for interface
    terms) inside FP64 refinement: the states meet the parity bar against the
    oracle's FP64 steps, also when each Newton correction is solved only to
    the tolerance floor (Newton's FP64 defect test decides convergence)."""
    w = make(wl)
    mix = ("model.time_step_operator.linear_solver.precision = mixed\n"
           f"model.time_step_operator.linear_solver.refinement.tolerance_floor = {floor}\n")
    gpu = dca.Model(w.ini, PTS + mix, 0)
    orc = OracleModel(w.ini)
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    gpu.profile_enable(True)
    t = 0.0
    for k in range(3):
        st = gpu.step(t, w.dt)
        rc, _, _ = orc.step(w.dt)
        assert st.status == 0 and rc == 0 and st.refinement_steps > 0
        t += w.dt
        ug, _ = gpu.get_state()
        uo, _ = orc.get_state()
        assert max(rel_l2_per_species(gpu.compartments(), ug, uo)) <= 1e-8, k
    assert gpu.profile("point_apply32")[0] > 0 and gpu.profile("spmv32")[0] == 0


def test_point_model_holds_no_matrix(dca, wl):
    w = wl.mitchell_schaefer(n=16, cell_type="cube")
    mat = dca.Model(w.ini, "", 0)
    pts = dca.Model(w.ini, PTS, 0)
    B = mat.block_size
    matrix = 8 * mat.n_blocks * B * B
    pts.set_state(w.initial_state(pts), 0.0)
    pts.profile_enable(True)
    assert pts.step(0.0, w.dt).status == 0
    assert pts.profile("point_setup")[0] > 0 and pts.profile("point_apply64")[0] > 0
    assert pts.profile("spmv64")[0] == 0 and pts.profile("assembly_jacobian")[0] == 0
    # the kernel-level entry points build a matrix on request (so none was held)
    x = _perturbed_state(w, pts)
    before = pts.device_bytes()
    _, J = pts.assemble(x, 0.0, 1.0, 0.1)
    assert pts.device_bytes() - before >= matrix
    # and a given matrix is solved with the matrix
    b = np.random.default_rng(3).standard_normal(pts.n_dof)
    sol, rc, _ = pts.linear_solve(J, b, rtol=1e-12, maxit=500)
    assert rc == 0
    A = bcsr_to_dense_csr(pts, J)
    assert np.linalg.norm(A @ sol - b) <= 1e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("seed", [1, 2, 3, 5, 7, 8, 9, 12, 21, 33])
def test_random_requests_with_default_storage(dca, wl, seed):
    """Random mass-action requests of the serving stream (1-2 compartments,
    2-8 species, transmission between compartments) with the default storage:
    point coefficients where `auto` takes them, the matrix elsewhere; two
    steps against the oracle to the parity bar."""
    w = wl.random_request(seed, precision="fp64", nl_rtol=1e-20, nl_atol=1e-28, side2=(10, 14), side3=(5, 7),
                          steps_range=(2, 2))
    print(f"case: random request {seed}", flush=True)
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    t = 0.0
    for k in range(2):
        st = gpu.step(t, w.dt)
        rc, newton, _ = orc.step(w.dt)
        assert st.status == 0 and rc == 0 and st.newton_iterations == newton, (k, st.status, rc)
        t += w.dt
        ug, _ = gpu.get_state()
        uo, _ = orc.get_state()
        assert max(rel_l2_per_species(gpu.compartments(), ug, uo)) <= 1e-8, k


def test_readback_batches_change_nothing_but_launches(dca, wl):
    """BiCGSTAB runs its iterations in batches between read-backs; applies
    launched past convergence return at once and are not counted as applies.
    Read-backs after every iteration (readback_iterations = 1) reach bit-equal
    states with equal iterations, and both count exactly two applies per
    BiCGSTAB iteration.  (Both use the same tile shape: the shape sets the
    order of the partial dot sums, so shapes timed independently could differ
    in the last bits.)"""
    w = wl.mitchell_schaefer(n=10, cell_type="cube")
    runs = []
    for extra in ("", "model.time_step_operator.linear_solver.readback_iterations = 1\n"):
        m = dca.Model(w.ini, PTS + extra + "model.time_step_operator.linear_solver.point_tile = 1\n", 0)
        m.set_state(w.initial_state(m), 0.0)
        m.profile_enable(True)
        t, lin, newton = 0.0, 0, 0
        for _ in range(3):
            st = m.step(t, w.dt)
            assert st.status == 0
            lin += st.linear_iterations
            newton += st.newton_iterations
            t += w.dt
        assert m.profile("point_apply64")[0] == 2 * lin
        runs.append((m.get_state()[0], lin, newton))
    assert runs[0][1:] == runs[1][1:]
    assert np.array_equal(runs[0][0], runs[1][0])
