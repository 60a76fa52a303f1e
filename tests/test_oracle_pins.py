"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors for this path (its test/data is Git-LFS
stubs), so the oracle is checked against:
  * closed forms of the reference's own test/exp.ini (u' = -2u): implicit
    Euler gives exactly (1 + 2 dt)^-n at every node, Alexander2, the explicit
    schemes and the fractional-step theta scheme their stability functions,
    on P1 and Q1 meshes in 1D/2D/3D; every scheme reaches its order of
    accuracy, and the L-stable ones damp a huge step to zero;
  * textbook element matrices (P1 triangle/tetrahedron and Q1 square mass and
    stiffness), assembled by numpy on 1-cell and 2x2-cell meshes;
  * invariants: stiffness rows sum to zero, mass entries sum to |Omega|,
    species mass is conserved under pure diffusion;
  * central finite differences of the oracle's residual for its Jacobian
    (reaction, transmission and discontinuous-gate models);
  * the exact discrete propagator of implicit Euler and Alexander2 for a
    linear two-species ODE (matrix functions in numpy), which pins the
    Runge-Kutta coefficients and the signs of the coupling Jacobian;
  * the root property: the state after a step makes the stage residual vanish.
"""
import ctypes
import itertools
import math

import numpy as np
import pytest

from tests.oracle_lib import OracleModel, load

ALPHA = 1.0 - math.sqrt(2.0) / 2.0


def _exp_ini(dim, cell_type, rk, n=3, D=0.0):
    cells = " ".join([str(n)] * dim)
    diff = f"cross_diffusion.u.expression = {D}\n" if D else ""
    return f"""[grid]
dimension = {dim}
cells = {cells}
cell_type = {cell_type}
[compartments]
domain.expression = 1
[parser_context.grow_rate]
type = constant
value = -2.
[model.scalar_field.u]
compartment = domain
storage.expression = 1
reaction.expression = grow_rate*u
reaction.jacobian.u.expression = grow_rate
{diff}[model.time_step_operator]
type = {rk}
linear_solver.convergence_condition.relative_tolerance = 1e-14
nonlinear_solver.convergence_condition.relative_tolerance = 1e-24
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-30
"""


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("cell_type", ["simplex", "cube"])
@pytest.mark.parametrize("rk", ["ImplicitEuler", "Alexander2"])
@pytest.mark.parametrize("D", [0.0, 0.3])
def test_exp_ini_closed_form(dim, cell_type, rk, D):
    o = OracleModel(_exp_ini(dim, cell_type, rk, D=D))
    o.set_state(np.ones(o.n_dof), 0.0)
    dt, n = 0.1, 7
    for _ in range(n):
        rc, _, _ = o.step(dt)
        assert rc == 0
    u, t = o.get_state()
    z = -2.0 * dt
    R = 1.0 / (1.0 - z) if rk == "ImplicitEuler" else (1 + (1 - 2 * ALPHA) * z) / (1 - ALPHA * z) ** 2
    assert abs(t - n * dt) < 1e-12
    np.testing.assert_allclose(u, R ** n, rtol=1e-12, atol=1e-14)


# Stability functions of the explicit schemes (the Taylor polynomial of
# exp(z) of the scheme's order) and of the fractional-step theta scheme
# (product of its three theta-steps).
_TH = 1.0 - math.sqrt(2.0) / 2.0
_FS_A = (1.0 - 2.0 * _TH) / (1.0 - _TH)
_R = {
    "ExplicitEuler": lambda z: 1 + z,
    "Heun": lambda z: 1 + z + z * z / 2,
    "Shu3": lambda z: 1 + z + z * z / 2 + z ** 3 / 6,
    "RungeKutta4": lambda z: 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24,
    "FractionalStepTheta": lambda z: ((1 + (1 - _FS_A) * _TH * z) / (1 - _FS_A * _TH * z)) ** 2
    * (1 + _FS_A * (1 - 2 * _TH) * z) / (1 - (1 - _FS_A) * (1 - 2 * _TH) * z),
}


@pytest.mark.parametrize("rk", sorted(_R))
@pytest.mark.parametrize("dim,cell_type", [(1, "cube"), (2, "simplex"), (3, "cube")])
def test_exp_ini_stability_functions(rk, dim, cell_type):
    """exp.ini (u' = -2u, every node alone): n steps multiply by R(z)^n.  The
    explicit schemes run without diffusion (with it, dt = 0.1 on this mesh is
    beyond their stability limit and rounding noise grows)."""
    o = OracleModel(_exp_ini(dim, cell_type, rk, D=0.3 if rk == "FractionalStepTheta" else 0.0))
    o.set_state(np.ones(o.n_dof), 0.0)
    dt, n = 0.1, 6
    for _ in range(n):
        rc, _, _ = o.step(dt)
        assert rc == 0
    u, t = o.get_state()
    assert abs(t - n * dt) < 1e-12
    np.testing.assert_allclose(u, _R[rk](-2.0 * dt) ** n, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("rk,order", [("ExplicitEuler", 1), ("ImplicitEuler", 1), ("Heun", 2), ("Alexander2", 2),
                                      ("FractionalStepTheta", 2), ("Shu3", 3), ("Alexander3", 3),
                                      ("RungeKutta4", 4)])
def test_schemes_reach_their_order(rk, order):
    """The error at t = 0.4 against exp(-2t) falls by 2^order when dt halves:
    a wrong coefficient (order conditions) or a wrong Alexander3 alpha shows
    here.  Alexander3 is also stiffly accurate: one huge step gives ~0."""
    errs = []
    for dt in (0.05, 0.025, 0.0125):
        o = OracleModel(_exp_ini(2, "cube", rk))
        o.set_state(np.ones(o.n_dof), 0.0)
        for _ in range(int(round(0.4 / dt))):
            assert o.step(dt)[0] == 0
        u, t = o.get_state()
        errs.append(abs(u[0] - math.exp(-2.0 * t)))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(abs(r - order) < 0.2 for r in rates), (rk, errs, rates)
    if rk in ("Alexander2", "Alexander3", "ImplicitEuler"):
        o = OracleModel(_exp_ini(2, "cube", rk))
        o.set_state(np.ones(o.n_dof), 0.0)
        assert o.step(1e6)[0] == 0
        assert np.abs(o.get_state()[0]).max() < 1e-5


def test_interpolation_context_values():
    """parser_context type = interpolation in the oracle: a reaction f(u) on
    a constant state c gives residual rows -f(c) times the rows of the mass
    matrix, so f(c) is read back at values below, on, between (both sides of
    the jump at the duplicated 0.3) and above the points; the values are the
    reference's rule (context.cc:72-97) worked by hand."""
    base = """[grid]
dimension = 2
cells = 3 3
cell_type = cube
[parser_context.f]
type = interpolation
domain = 0.1 0.3 0.3 0.8
range = 0.5 -0.25 2.0 1.0
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
storage.expression = 1
reaction.expression = f(u)
reaction.jacobian.u.expression = 0
"""
    o = OracleModel(base)
    mass, _ = o.assemble(np.ones(o.n_dof), 0.0, 1.0, 0.0, jacobian=False)
    # (not at 0.3 itself: the points' values of a constant state round to
    # either side of the jump)
    cases = {-1.0: 0.5, 0.1: 0.5, 0.2: 0.125, 0.29: -0.2125, 0.31: 1.98, 0.55: 1.5, 0.8: 1.0, 5.0: 1.0}
    for c, fc in cases.items():
        r, _ = o.assemble(np.full(o.n_dof, c), 0.0, 0.0, 1.0, jacobian=False)
        np.testing.assert_allclose(r, -fc * mass, rtol=1e-9, atol=1e-15)


def _diffusion_ini(dim, cell_type, cells, D=1.0, ext=None):
    ext = ext or " ".join(["1"] * dim)
    return f"""[grid]
dimension = {dim}
cells = {cells}
extensions = {ext}
cell_type = {cell_type}
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
storage.expression = 1
cross_diffusion.u.expression = {D}
[model.time_step_operator]
type = ImplicitEuler
linear_solver.convergence_condition.relative_tolerance = 1e-14
nonlinear_solver.convergence_condition.relative_tolerance = 1e-24
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-30
"""


def _matrices(o):
    x = np.zeros(o.n_dof)
    _, vm = o.assemble(x, 0.0, 1.0, 0.0)
    _, vk = o.assemble(x, 0.0, 0.0, 1.0)
    return o.csr(vm).toarray(), o.csr(vk).toarray()


def test_q1_square_textbook_matrices():
    h = 0.5
    o = OracleModel(_diffusion_ini(2, "cube", "1 1", ext="0.5 0.5"))
    M, K = _matrices(o)
    m1 = h / 6 * np.array([[2.0, 1.0], [1.0, 2.0]])
    k1 = 1 / h * np.array([[1.0, -1.0], [-1.0, 1.0]])
    # vertex order x fastest: index = i + 2 j  ->  kron(y-matrix, x-matrix)
    np.testing.assert_allclose(M, np.kron(m1, m1), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(K, np.kron(m1, k1) + np.kron(k1, m1), rtol=1e-13, atol=1e-15)


def test_p1_triangles_textbook_matrices():
    # unit square split along the (0,0)-(1,1) diagonal: triangles {0,1,3}, {0,2,3}
    o = OracleModel(_diffusion_ini(2, "simplex", "1 1"))
    M, K = _matrices(o)
    Me = 1.0 / 24 * np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]], dtype=float)  # area 1/2
    Mref = np.zeros((4, 4))
    Kref = np.zeros((4, 4))
    # right angle at vertex 1 in {0,1,3}; at vertex 2 in {0,2,3}
    for tri, right in (([0, 1, 3], 1), ([0, 2, 3], 2)):
        Ke = np.zeros((3, 3))
        # stiffness of a right isosceles triangle with unit legs
        others = [v for v in tri if v != right]
        li = {v: i for i, v in enumerate(tri)}
        r = li[right]
        a, b = li[others[0]], li[others[1]]
        Ke[r, r] = 1.0
        Ke[a, a] = Ke[b, b] = 0.5
        Ke[r, a] = Ke[a, r] = Ke[r, b] = Ke[b, r] = -0.5
        for i, vi in enumerate(tri):
            for j, vj in enumerate(tri):
                Mref[vi, vj] += Me[i, j]
                Kref[vi, vj] += Ke[i, j]
    np.testing.assert_allclose(M, Mref, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(K, Kref, rtol=1e-13, atol=1e-15)


def test_p1_tetrahedra_mass_matrix():
    # each Kuhn tetrahedron of the unit cube has volume 1/6 and mass V/20 (2 on the diagonal)
    o = OracleModel(_diffusion_ini(3, "simplex", "1 1 1"))
    M, K = _matrices(o)
    Mref = np.zeros((8, 8))
    Me = (1.0 / 6) / 20 * (np.ones((4, 4)) + np.eye(4))
    import itertools
    for p in itertools.permutations(range(3)):
        verts, bits = [0], 0
        for ax in p:
            bits |= 1 << ax
            verts.append(bits)
        for i, vi in enumerate(verts):
            for j, vj in enumerate(verts):
                Mref[vi, vj] += Me[i, j]
    np.testing.assert_allclose(M, Mref, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(K.sum(axis=1), 0.0, atol=1e-13)


@pytest.mark.parametrize("dim,cell_type,cells", [(2, "simplex", "5 3"), (2, "cube", "4 6"), (3, "simplex", "3 2 2"),
                                                 (3, "cube", "3 3 2"), (1, "cube", "7")])
def test_mass_and_stiffness_invariants(dim, cell_type, cells):
    ext = " ".join(["0.7"] * dim)
    o = OracleModel(_diffusion_ini(dim, cell_type, cells, ext=ext))
    M, K = _matrices(o)
    np.testing.assert_allclose(M.sum(), 0.7 ** dim, rtol=1e-13)
    np.testing.assert_allclose(K.sum(axis=1), 0.0, atol=1e-12)
    np.testing.assert_allclose(M, M.T, atol=1e-16)
    np.testing.assert_allclose(K, K.T, atol=1e-13)
    assert np.linalg.eigvalsh(M).min() > 0


@pytest.mark.parametrize("cell_type", ["simplex", "cube"])
def test_pure_diffusion_conserves_mass(wl, cell_type):
    o = OracleModel(_diffusion_ini(2, cell_type, "9 7", D=0.05))
    rng = np.random.default_rng(0)
    u0 = rng.uniform(0, 1, o.n_dof)
    o.set_state(u0, 0.0)
    m0 = o.mass(u0)
    for _ in range(5):
        assert o.step(0.3)[0] == 0
    u, _ = o.get_state()
    np.testing.assert_allclose(o.mass(u), m0, rtol=1e-11)
    assert np.std(u) < np.std(u0)  # and it does diffuse


def _fd_check(o, x, t, cm, ca, cols, h=1e-6, rtol=2e-6):
    _, v = o.assemble(x, t, cm, ca)
    A = o.csr(v)
    for j in cols:
        e = np.zeros_like(x)
        e[j] = h
        rp, _ = o.assemble(x + e, t, cm, ca, jacobian=False)
        rm, _ = o.assemble(x - e, t, cm, ca, jacobian=False)
        fd = (rp - rm) / (2 * h)
        col = A[:, j].toarray().ravel()
        scale = max(np.abs(col).max(), np.abs(fd).max(), 1e-30)
        assert np.abs(fd - col).max() <= rtol * scale, (j, np.abs(fd - col).max(), scale)


@pytest.mark.parametrize("make", [
    lambda wl: wl.gray_scott(n=5, cell_type="simplex"),
    lambda wl: wl.gray_scott(n=4, cell_type="cube"),
    lambda wl: wl.mitchell_schaefer(n=3, cell_type="cube"),
    lambda wl: wl.two_compartments(n=10, cell_type="cube"),
    lambda wl: wl.two_compartments(n=8, cell_type="simplex"),
], ids=["gs_p1", "gs_q1", "ms_q1", "two_q1", "two_p1"])
def test_jacobian_matches_finite_differences(wl, make):
    w = make(wl)
    o = OracleModel(w.ini)
    x = w.initial_state(o)
    rng = np.random.default_rng(2)
    x = x + 0.05 * rng.standard_normal(x.size)
    cols = rng.choice(o.n_dof, size=min(40, o.n_dof), replace=False)
    _fd_check(o, x, 0.2, 1.0, 0.7, cols)


def test_numerical_jacobian_agrees_with_analytical(wl):
    w = wl.gray_scott(n=4, cell_type="cube")
    o_a = OracleModel(w.ini)
    o_n = OracleModel(w.ini, "model.jacobian.type=numerical")
    x = w.initial_state(o_a) + 0.01
    _, va = o_a.assemble(x, 0.0, 1.0, 1.0)
    _, vn = o_n.assemble(x, 0.0, 1.0, 1.0)
    assert np.abs(va - vn).max() <= 1e-6 * np.abs(va).max()


def _linear_ode_ini(rk):
    # u' = -0.3 u + 0.8 v ; v' = -0.5 u - 0.2 v (rotation with decay), no diffusion
    return f"""[grid]
dimension = 2
cells = 2 2
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
storage.expression = 1
reaction.expression = -0.3*u + 0.8*v
reaction.jacobian.u.expression = -0.3
reaction.jacobian.v.expression = 0.8
[model.scalar_field.v]
compartment = domain
storage.expression = 1
reaction.expression = -0.5*u - 0.2*v
reaction.jacobian.u.expression = -0.5
reaction.jacobian.v.expression = -0.2
[model.time_step_operator]
type = {rk}
linear_solver.convergence_condition.relative_tolerance = 1e-14
nonlinear_solver.convergence_condition.relative_tolerance = 1e-26
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-30
"""


@pytest.mark.parametrize("rk", ["ImplicitEuler", "Alexander2"])
def test_linear_ode_discrete_propagator(rk):
    import scipy.linalg

    A = np.array([[-0.3, 0.8], [-0.5, -0.2]])
    dt, n = 0.25, 8
    Z = dt * A
    I = np.eye(2)
    if rk == "ImplicitEuler":
        R = np.linalg.inv(I - Z)
    else:
        Ri = np.linalg.inv(I - ALPHA * Z)
        R = Ri @ Ri @ (I + (1 - 2 * ALPHA) * Z)
    o = OracleModel(_linear_ode_ini(rk))
    u0 = np.array([1.0, 0.5])
    o.set_state(np.tile(u0, o.n_dof // 2), 0.0)
    for _ in range(n):
        assert o.step(dt)[0] == 0
    u, _ = o.get_state()
    ref = np.linalg.matrix_power(R, n) @ u0
    np.testing.assert_allclose(u.reshape(-1, 2), np.tile(ref, (o.n_dof // 2, 1)), rtol=1e-12, atol=1e-14)
    # and the scheme approximates the exact flow to its order
    exact = scipy.linalg.expm(n * dt * A) @ u0
    err = np.abs(ref - exact).max()
    assert err < (0.1 if rk == "ImplicitEuler" else 5e-3)


@pytest.mark.parametrize("make", [lambda wl: wl.gray_scott(n=6, cell_type="cube", nl_rtol=1e-24, nl_atol=1e-28,
                                                           lin_rtol=1e-13),
                                  lambda wl: wl.two_compartments(n=8, cell_type="cube", nl_rtol=1e-24, nl_atol=1e-28,
                                                                 lin_rtol=1e-13)], ids=["gs", "two"])
def test_step_solves_the_stage_equation(wl, make):
    # implicit Euler: M(u1) - M(u0) + dt A(u1) = 0 at the returned state
    w = make(wl)
    o = OracleModel(w.ini)
    u0 = w.initial_state(o)
    o.set_state(u0, 0.0)
    assert o.step(w.dt)[0] == 0
    u1, _ = o.get_state()
    r1, _ = o.assemble(u1, w.dt, 1.0, w.dt, jacobian=False)
    r0, _ = o.assemble(u0, 0.0, 1.0, 0.0, jacobian=False)
    ref = np.abs(r0).max()
    assert np.abs(r1 - r0).max() <= 1e-11 * ref


def _eval(expr, names=(), values=(), context=""):
    L = load()
    L.orc_eval.restype = ctypes.c_int
    L.orc_eval.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
                           ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    n = len(names)
    nm = (ctypes.c_char_p * max(n, 1))(*[s.encode() for s in names])
    vv = (ctypes.c_double * max(n, 1))(*values)
    out = ctypes.c_double()
    rc = L.orc_eval(context.encode(), expr.encode(), nm, vv, n, ctypes.byref(out))
    if rc < 0:
        raise ValueError(L.orc_last_error().decode())
    return out.value, rc


EXPRESSIONS = [
    ("-x^2", lambda x, y: -(x ** 2)),
    ("x^y^0.5", lambda x, y: x ** (y ** 0.5)),
    ("2^-y", lambda x, y: 2 ** -y),
    ("x*y - y/x + 3", lambda x, y: x * y - y / x + 3),
    ("x < y ? x : y", lambda x, y: x if x < y else y),
    ("(x > 1) && (y <= 3)", lambda x, y: float(x > 1 and y <= 3)),
    ("x == 2 || y != 3", lambda x, y: float(x == 2 or y != 3)),
    ("not (x > y)", lambda x, y: float(not (x > y))),
    ("min(x, y, 0.5) + max(x, -y)", lambda x, y: min(x, y, 0.5) + max(x, -y)),
    ("exp(-x) * log(y) + sqrt(x*y) - abs(-x)", lambda x, y: math.exp(-x) * math.log(y) + math.sqrt(x * y) - x),
    ("sin(x)^2 + cos(x)^2", lambda x, y: 1.0),
    ("atan2(y, x) + tanh(x) - pow(x, y)", lambda x, y: math.atan2(y, x) + math.tanh(x) - x ** y),
    ("1e-3 * .5e2 + 2.5E+1", lambda x, y: 1e-3 * 50 + 25),
    ("if(x > y, 1, 2) + floor(y / x)", lambda x, y: (1 if x > y else 2) + math.floor(y / x)),
]


@pytest.mark.parametrize("text,ref", EXPRESSIONS, ids=[e[0] for e in EXPRESSIONS])
def test_oracle_expressions(text, ref):
    for x, y in [(2.0, 3.0), (0.7, 0.2), (1.5, 4.5)]:
        v, rc = _eval(text, ("x", "y"), (x, y))
        assert rc == 0
        assert v == pytest.approx(ref(x, y), rel=1e-14, abs=1e-14)


def test_oracle_context_functions_and_zero_literal():
    ctx = "[parser_context.f]\ntype = function\nexpression = a, b: a*10 + b\n[parser_context.k]\ntype = constant\nvalue = 4\n"
    v, rc = _eval("f(x, k) + f(1, 2)", ("x",), (3.0,), ctx)
    assert (v, rc) == (34.0 + 12.0, 0)
    for zero in ["0", "0.0", "-0", "", "  0.  "]:
        assert _eval(zero)[1] == 1  # term absent, as the reference treats it
    assert _eval("no_value")[0] == np.finfo(float).max


# ---------------------------------------------------------------- time loop
def _gs_fail_ini(nl_max=3):
    """Gray-Scott with three Newton iterations allowed: steps of 2 and more
    fail, steps of 1 and less converge."""
    return f"""[grid]
dimension = 2
cells = 6 6
cell_type = cube
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
storage.expression = 1
cross_diffusion.u.expression = 2e-3
reaction.expression = -u*v^2 + 0.04*(1-u)
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
nonlinear_solver.convergence_condition.relative_tolerance = 1e-16
nonlinear_solver.convergence_condition.iteration_range = 0 {nl_max}
"""


def _gs_state(o):
    n = o.n_dof // 2
    u = np.empty(o.n_dof)
    k = np.arange(n)
    u[0::2] = 0.5 + 0.4 * np.sin(0.7 * k)
    u[1::2] = 0.3 + 0.25 * np.cos(1.3 * k)
    return u


@pytest.mark.parametrize("rk", ["ImplicitEuler", "Alexander2"])
def test_evolve_closed_form_and_snap(rk):
    """exp.ini under the adaptive loop: the accepted steps tile [0, T]
    exactly, grow geometrically until two steps before T, finish with equal
    snapped steps, and the state is the product of the scheme's stability
    function over exactly those steps."""
    o = OracleModel(_exp_ini(2, "cube", rk))
    o.set_state(np.ones(o.n_dof), 0.0)
    T, dt0, inc = 1.7, 0.05, 1.1
    rc, acc, failed, _ = o.evolve(T, dt0, increase=inc)
    assert rc == 0 and failed == 0
    u, t = o.get_state()
    assert abs(t - T) <= 1e-14 and abs(acc.sum() - T) <= 1e-13
    # growth phase: each accepted step is 1.1 times the one before it
    k = 0
    while k + 1 < acc.size and abs(acc[k + 1] - inc * acc[k]) <= 1e-14 * acc[k + 1]:
        k += 1
    assert k >= 5
    # snap phase: the remaining steps split the remainder into equal parts
    # (each snapped step recomputes n = ceil(remaining / (1.1 dt)))
    assert acc[k + 1:].size >= 1
    # stability function of the scheme over the accepted steps
    z = -2.0 * acc
    if rk == "ImplicitEuler":
        R = 1.0 / (1.0 - z)
    else:
        R = (1.0 + (1.0 - 2.0 * ALPHA) * z) / (1.0 - ALPHA * z) ** 2
    assert np.abs(u - np.prod(R)).max() <= 1e-12


def test_evolve_retries_failed_steps_from_the_same_state():
    """Steps too large for one Newton iteration fail and are retried at half
    the size from the unchanged state: replaying the accepted sizes with
    plain steps gives the same state bit for bit."""
    o = OracleModel(_gs_fail_ini())
    u0 = _gs_state(o)
    o.set_state(u0, 0.0)
    rc, acc, failed, _ = o.evolve(40.0, 16.0, increase=1.5, min_step=1e-3)
    assert rc == 0 and failed >= 4
    u, t = o.get_state()
    assert abs(t - 40.0) <= 1e-12
    r = OracleModel(_gs_fail_ini())
    r.set_state(u0, 0.0)
    for h in acc:
        rc, _, _ = r.step(h)
        assert rc == 0
    ur, tr = r.get_state()
    assert np.array_equal(u, ur) and tr == t


def test_evolve_stops_at_the_minimum_step():
    o = OracleModel(_gs_fail_ini())
    u0 = _gs_state(o)
    o.set_state(u0, 0.0)
    rc, acc, failed, _ = o.evolve(40.0, 16.0, min_step=4.0)
    # 16, 8 and 4 are tried and fail; 2 is below the minimum
    assert rc == 5 and acc.size == 0 and failed == 3
    u, t = o.get_state()
    assert t == 0.0 and np.array_equal(u, u0)


# ---------------------------------------------------------------- transform / reduce
def _nodes_xyz(o, dim):
    """Vertex coordinates of the oracle's compartment-0 nodes on the unit box
    with 3 cells per axis (vertex id = i + 4 j + 16 k)."""
    ids = o.compartment_nodes(0)
    x = (ids % 4) / 3.0
    y = ((ids // 4) % 4) / 3.0 if dim > 1 else 0 * x
    z = (ids // 16) / 3.0 if dim > 2 else 0 * x
    return x, y, z


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("cell_type", ["simplex", "cube"])
def test_reduce_rule_integrates_degree_four_exactly(dim, cell_type):
    """The order-4 rule: integration_factor times every monomial of degree
    <= 4 sums to its integral over the unit box (1 / prod(p_k + 1)); a plain
    evaluation counts the points (no integration factor), and the initial
    value enters twice (fold start and final combination)."""
    o = OracleModel(_exp_ini(dim, cell_type, "ImplicitEuler"))
    o.set_state(np.ones(o.n_dof), 0.0)
    exps = [e for e in itertools.product(range(5), repeat=dim) if sum(e) <= 4]
    ax = "xyz"
    keys = []
    for n, e in enumerate(exps):
        mono = "*".join(f"position_{ax[k]}^{p}" for k, p in enumerate(e) if p) or "1"
        keys.append(f"[m{n}]\nevaluation.expression = {mono}*integration_factor\n")
    keys.append("[count]\nevaluation.expression = 1\ninitial.value = 5\n")
    vals, flags, rc = o.reduce("".join(keys))
    assert rc == 0
    for n, e in enumerate(exps):
        exact = 1.0 / np.prod([p + 1 for p in e])
        assert abs(vals[f"m{n}"] - exact) <= 1e-13, (e, vals[f"m{n}"], exact)
    cells = 3 ** dim
    per_elem = {1: 3, 2: 6, 3: 11}[dim] if cell_type == "simplex" and dim > 1 else 3 ** dim
    elems = cells * (math.factorial(dim) if cell_type == "simplex" and dim > 1 else 1)
    if dim == 1:
        per_elem = 3
    assert vals["count"] == elems * per_elem + 10


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("cell_type", ["simplex", "cube"])
def test_reduce_keys_of_the_reference_tests(dim, cell_type):
    """Keys as test/exp.ini and test/gauss.ini write them, on u = 1 + x + 2y
    (exact in P1 and Q1): max / min over the points with the (init, val)
    reduction, the L2 norm of u - (1 + x + 2y) through a transformation,
    gradients, and error / warn expressions."""
    o = OracleModel(_exp_ini(dim, cell_type, "ImplicitEuler"))
    x, y, z = _nodes_xyz(o, dim)
    o.set_state(1.0 + x + 2.0 * y, 0.0)
    cfg = """[u_max]
evaluation.expression = u
reduction.expression = init, val: max(init, val)
initial.value = -1e100
error.expression = arg: arg > 3.5
[u_min]
evaluation.expression = u
reduction.expression = init, val: min(init, val)
initial.value = 1e100
warn.expression = arg: arg < 1.5
[u_error]
evaluation.expression = (u - (1 + position_x + 2*position_y))^2 * integration_factor
transformation.expression = arg: sqrt(arg)
[grad]
evaluation.expression = (grad_u_x + 10*grad_u_y) * integration_factor
[mass]
evaluation.expression = u * integration_factor
"""
    vals, flags, rc = o.reduce(cfg)
    # the points are interior: max/min of 1 + x + 2y over them, below the corner values
    assert 1.0 < vals["u_min"] < 1.2 and 3.8 < vals["u_max"] < 4.0
    assert flags["u_max"] == 2 and rc == 1          # 3.5 < max fires the error
    assert flags["u_min"] == 1                       # min < 1.5 fires the warning
    assert vals["u_error"] <= 1e-14
    assert abs(vals["grad"] - 21.0) <= 1e-12
    assert abs(vals["mass"] - 2.5) <= 1e-13          # int 1 + x + 2y over the unit box
    # the model's own model.reduce keys are read when no config is given
    o2 = OracleModel(_exp_ini(dim, cell_type, "ImplicitEuler") + "[model.reduce]\nm.evaluation.expression = u*integration_factor\n")
    o2.set_state(np.full(o2.n_dof, 2.0), 0.0)
    v2, _, _ = o2.reduce()
    assert abs(v2["m"] - 2.0) <= 1e-13


# ---------------------------------------------------------------- transmission and outflow residual
# The facet terms (local_operator.hh:918-927,956-963) are pinned by what the
# weak form fixes independently of the code: with T = k_t (D_in - A_out) on the
# inside species and -T on the outside one (the sign convention of
# test/two_disks.ini:24-35), the residual summed over a species' nodes is
# T * |Gamma_h| for a constant state (the hat functions are a partition of
# unity on every facet), where |Gamma_h| is the measure of the facets between
# cells marked inside and outside, counted here from the cell-centre rule on
# its own; along a straight interface the nodal residual is the 1D P1 mass
# matrix applied to the interface values; the total of both species is
# conserved by a time step, and D falls while D > A.
def _transmission_ini(dim, cell_type, n, kt, comp_in, comp_out, rk="ImplicitEuler"):
    cells = " ".join([str(n)] * dim)
    return f"""[grid]
dimension = {dim}
cells = {cells}
extensions = {" ".join(["1"] * dim)}
cell_type = {cell_type}
[compartments]
inner.expression = {comp_in}
outer.expression = {comp_out}
[model.scalar_field.D]
compartment = inner
storage.expression = 1
outflow.outer.expression = {kt}*(D - A)
outflow.outer.jacobian.D.expression = {kt}
outflow.outer.jacobian.A.expression = {-kt}
[model.scalar_field.A]
compartment = outer
storage.expression = 1
outflow.inner.expression = -{kt}*(D - A)
outflow.inner.jacobian.D.expression = {-kt}
outflow.inner.jacobian.A.expression = {kt}
[model.time_step_operator]
type = {rk}
linear_solver.convergence_condition.relative_tolerance = 1e-14
nonlinear_solver.convergence_condition.relative_tolerance = 1e-26
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-30
"""


_DISK = ("(position_x-0.5)^2 + (position_y-0.5)^2 < 0.09", "(position_x-0.5)^2 + (position_y-0.5)^2 >= 0.09")
_BALL = ("(position_x-0.5)^2 + (position_y-0.5)^2 + (position_z-0.5)^2 < 0.09",
         "(position_x-0.5)^2 + (position_y-0.5)^2 + (position_z-0.5)^2 >= 0.09")


def _interface_measure(dim, cell_type, n):
    """|Gamma_h|: measure of the facets between cells whose centres lie inside
    and outside the disk / ball of radius 0.3 about the domain centre."""
    h = 1.0 / n
    inside = lambda c: float(np.sum((np.asarray(c) - 0.5) ** 2)) < 0.09
    if cell_type == "cube":
        idx = np.indices([n] * dim).reshape(dim, -1).T
        mark = {tuple(i): inside((i + 0.5) * h) for i in idx}
        total = 0.0
        for i in idx:
            for ax in range(dim):
                j = i.copy()
                j[ax] += 1
                if j[ax] < n and mark[tuple(i)] != mark[tuple(j)]:
                    total += h ** (dim - 1)
        return total
    assert dim == 2
    # Kuhn triangles of each cell: (0,0)-(1,0)-(1,1) and (0,0)-(0,1)-(1,1)
    edges = {}
    for i in range(n):
        for j in range(n):
            for tri in (((0, 0), (1, 0), (1, 1)), ((0, 0), (0, 1), (1, 1))):
                pts = [(i + a, j + b) for a, b in tri]
                c = np.mean(np.array(pts, dtype=float), axis=0) * h
                m = inside(c)
                for a, b in ((0, 1), (1, 2), (0, 2)):
                    e = tuple(sorted((pts[a], pts[b])))
                    edges.setdefault(e, []).append(m)
    total = 0.0
    for (p, q), marks in edges.items():
        if len(marks) == 2 and marks[0] != marks[1]:
            total += h * math.hypot(q[0] - p[0], q[1] - p[1])
    return total


def _inner_measure(dim, cell_type, n):
    """Measure of the cells whose centres lie in the disk / ball of radius 0.3."""
    h = 1.0 / n
    centres = [np.array(c) + 0.5 for c in np.ndindex(*([n] * dim))]
    if cell_type == "cube":
        return sum(h ** dim for c in centres if np.sum((c * h - 0.5) ** 2) < 0.09)
    total = 0.0
    for i, j in np.ndindex(n, n):
        for c in ((i + 2 / 3, j + 1 / 3), (i + 1 / 3, j + 2 / 3)):
            if np.sum((np.array(c) * h - 0.5) ** 2) < 0.09:
                total += 0.5 * h * h
    return total


def _vertex_xyz(o):
    """Coordinates of every grid vertex (rows = vertex ids)."""
    L = load()
    n = L.orc_n_vertices(o._h)
    xyz = np.zeros((n, 3))
    L.orc_vertex_coords(o._h, ctypes.c_void_p(xyz.ctypes.data))
    return xyz


def _species_sums(o, r):
    sums = []
    for c in o.compartments():
        b, off, nn = c["n_species"], c["dof_offset"], c["n_nodes"]
        sums.append(r[off: off + b * nn].reshape(nn, b).sum(axis=0))
    return np.concatenate(sums)


@pytest.mark.parametrize("dim,cell_type,n", [(2, "cube", 20), (2, "cube", 33), (2, "simplex", 20), (2, "simplex", 27),
                                             (3, "cube", 10)])
def test_transmission_residual_is_flux_times_interface_measure(dim, cell_type, n):
    kt = 0.5
    comp = _DISK if dim == 2 else _BALL
    o = OracleModel(_transmission_ini(dim, cell_type, n, kt, *comp))
    gamma = _interface_measure(dim, cell_type, n)
    assert gamma > 0.5  # about the circumference 1.885 (2D) / area 1.131 (3D) of the exact interface
    for d_val, a_val in [(1.0, 0.0), (0.3, 0.8), (2.0, 2.0)]:
        x = np.empty(o.n_dof)
        cs = o.compartments()
        x[cs[0]["dof_offset"]: cs[0]["dof_offset"] + cs[0]["n_nodes"]] = d_val
        x[cs[1]["dof_offset"]: cs[1]["dof_offset"] + cs[1]["n_nodes"]] = a_val
        r, _ = o.assemble(x, 0.0, 0.0, 1.0, jacobian=False)
        s = _species_sums(o, r)
        flux = kt * (d_val - a_val)
        np.testing.assert_allclose(s[0], flux * gamma, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(s[1], -flux * gamma, rtol=1e-12, atol=1e-14)
        # the mass operator carries no facet term: its sums are the values
        # times the measure of each compartment's cells
        rm, _ = o.assemble(x, 0.0, 1.0, 0.0, jacobian=False)
        vol_in = _inner_measure(dim, cell_type, n)
        np.testing.assert_allclose(_species_sums(o, rm), [d_val * vol_in, a_val * (1.0 - vol_in)], rtol=1e-12)


@pytest.mark.parametrize("cell_type", ["cube", "simplex"])
def test_transmission_residual_along_a_straight_interface(cell_type):
    # inner = x < 0.5 on an 8 x 8 grid: the interface is the line x = 0.5; with
    # D = 1 + y and A = y^2 the nodal residual of D at (0.5, y_j) is the
    # integral of k_t (D_h - A_h) phi_j over the line, D_h and A_h the linear
    # interpolants on each segment (a degree-2 integrand, so any rule of
    # order 2 gives it exactly; it is evaluated here with 2-point Gauss)
    n, kt = 8, 0.5
    o = OracleModel(_transmission_ini(2, cell_type, n, kt, "position_x < 0.5", "position_x >= 0.5"))
    h = 1.0 / n
    cs = o.compartments()
    xyz = _vertex_xyz(o)
    x = np.zeros(o.n_dof)
    for c, comp in enumerate(cs):
        ids = o.compartment_nodes(c)
        y = xyz[ids, 1]
        x[comp["dof_offset"]: comp["dof_offset"] + comp["n_nodes"]] = (1.0 + y) if c == 0 else y ** 2
    r, _ = o.assemble(x, 0.0, 0.0, 1.0, jacobian=False)
    g = np.array([0.5 - 0.5 / math.sqrt(3.0), 0.5 + 0.5 / math.sqrt(3.0)])
    for c, comp in enumerate(cs):
        ids = o.compartment_nodes(c)
        rc = r[comp["dof_offset"]: comp["dof_offset"] + comp["n_nodes"]]
        on = np.abs(xyz[ids, 0] - 0.5) < 1e-12
        np.testing.assert_allclose(rc[~on], 0.0, atol=1e-15)
        sign = 1.0 if c == 0 else -1.0
        for node, yj in zip(np.nonzero(on)[0], xyz[ids[on], 1]):
            ref = 0.0
            for lo in (yj - h, yj):
                if lo < -1e-12 or lo + h > 1 + 1e-12:
                    continue
                for gq in g:
                    yq = lo + gq * h
                    phi = 1.0 - abs(yq - yj) / h
                    # D and A are interpolated linearly on the segment
                    d_h = 1.0 + yq
                    a_h = np.interp(yq, [lo, lo + h], [lo ** 2, (lo + h) ** 2])
                    ref += 0.5 * h * kt * (d_h - a_h) * phi
            assert rc[node] == pytest.approx(sign * ref, rel=1e-12, abs=1e-15)


@pytest.mark.parametrize("cell_type", ["cube", "simplex"])
def test_transmission_conserves_and_relaxes(cell_type):
    kt = 0.5
    o = OracleModel(_transmission_ini(2, cell_type, 16, kt, *_DISK))
    cs = o.compartments()
    x = np.empty(o.n_dof)
    x[cs[0]["dof_offset"]: cs[0]["dof_offset"] + cs[0]["n_nodes"]] = 1.0
    x[cs[1]["dof_offset"]: cs[1]["dof_offset"] + cs[1]["n_nodes"]] = 0.0
    o.set_state(x, 0.0)
    m0 = o.mass(x)
    total0 = m0.sum()
    prev = m0
    for _ in range(6):
        assert o.step(0.2)[0] == 0
        u, _ = o.get_state()
        m = o.mass(u)
        assert m.sum() == pytest.approx(total0, rel=1e-12)
        assert m[0] < prev[0] and m[1] > prev[1]  # D flows out while D > A
        prev = m


def _outflow_ini(dim, cell_type, n, k):
    cells = " ".join([str(n)] * dim)
    return f"""[grid]
dimension = {dim}
cells = {cells}
extensions = {" ".join(["1"] * dim)}
cell_type = {cell_type}
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
storage.expression = 1
outflow.domain.expression = {k}*u
outflow.domain.jacobian.u.expression = {k}
[model.time_step_operator]
type = ImplicitEuler
linear_solver.convergence_condition.relative_tolerance = 1e-14
nonlinear_solver.convergence_condition.relative_tolerance = 1e-26
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-30
"""


@pytest.mark.parametrize("dim,cell_type", [(2, "cube"), (2, "simplex"), (3, "cube"), (3, "simplex")])
def test_boundary_outflow_is_flux_times_boundary_measure(dim, cell_type):
    # outflow.<own compartment> on the domain boundary (local_operator.hh:850-852):
    # sum of the residual = k u |dOmega| for a constant u, zero at interior nodes,
    # and a corner node of the unit square gets k u h (two half edges)
    k, n, u = 0.7, 5, 1.3
    o = OracleModel(_outflow_ini(dim, cell_type, n, k))
    x = np.full(o.n_dof, u)
    r, _ = o.assemble(x, 0.0, 0.0, 1.0, jacobian=False)
    assert r.sum() == pytest.approx(k * u * 2 * dim, rel=1e-12)
    xyz = _vertex_xyz(o)[o.compartment_nodes(0), :dim]
    interior = np.all((xyz > 1e-12) & (xyz < 1 - 1e-12), axis=1)
    np.testing.assert_allclose(r[interior], 0.0, atol=1e-15)  # round-off of the basis on far facets
    if dim == 2:
        corner = np.all(np.abs(xyz) < 1e-12, axis=1)
        assert r[corner][0] == pytest.approx(k * u / n, rel=1e-12)
        # outflow of u = 1 decays the total as d/dt int u = -k int_dOmega u
        o.set_state(np.ones(o.n_dof), 0.0)
        assert o.step(0.1)[0] == 0
        u1, _ = o.get_state()
        r1, _ = o.assemble(u1, 0.1, 0.0, 1.0, jacobian=False)
        assert o.mass(u1).sum() == pytest.approx(o.mass(np.ones(o.n_dof)).sum() - 0.1 * r1.sum(), rel=1e-12)


# ---------------------------------------------------------------- linear defect correction
# model.is_linear = true replaces Newton by one residual, one Jacobian and one
# linear solve per stage (make_step_operator.hh:210-245).  For a linear model
# that is the exact stage solution up to the linear tolerance, so the closed
# forms of test/exp.ini hold unchanged, and the heat kernel of test/gauss.ini
# keeps the reference's own error bound (its u_error key, limit 0.5).
@pytest.mark.parametrize("cell_type", ["simplex", "cube"])
@pytest.mark.parametrize("rk", ["ImplicitEuler", "Alexander2"])
def test_linear_defect_correction_closed_form(cell_type, rk):
    o = OracleModel(_exp_ini(2, cell_type, rk, D=0.3), "model.is_linear = true")
    o.set_state(np.ones(o.n_dof), 0.0)
    dt, n = 0.1, 6
    for _ in range(n):
        rc, newton, _ = o.step(dt)
        assert rc == 0
        assert newton == (1 if rk == "ImplicitEuler" else 2)  # one correction per stage
    u, _ = o.get_state()
    z = -2.0 * dt
    R = 1.0 / (1.0 - z) if rk == "ImplicitEuler" else (1 + (1 - 2 * ALPHA) * z) / (1 - ALPHA * z) ** 2
    np.testing.assert_allclose(u, R ** n, rtol=1e-12, atol=1e-14)


def heat_kernel_ini(cell_type="simplex", level=5):
    """The model of the reference's test/gauss.ini (heat kernel of a point
    source from t = 1, diffusion 0.005, on [-1, 1]^2, model.is_linear = true,
    with its max/min/error reduce keys), with a tight linear tolerance."""
    return f"""[grid]
origin = -1 -1
extensions = 2 2
refinement_level = {level}
cell_type = {cell_type}
[compartments.domain]
type = expression
expression = 1
[parser_context.diffusion]
type = constant
value = 0.005
[parser_context.gauss]
type = function
expression = x, y, z, t: exp(-(x^2+y^2+z^2)/(4*t*diffusion)) / (4*3.14159265359*t*diffusion)
[model]
is_linear = true
[model.scalar_field.u]
compartment = domain
storage.expression = 1
cross_diffusion.u.expression = diffusion
initial.expression = gauss(position_x, position_y, position_z, time)
[model.time_step_operator]
time_begin = 1
time_end = 1.2
linear_solver.convergence_condition.relative_tolerance = 1e-13
linear_solver.convergence_condition.iteration_range = 1 5000
nonlinear_solver.convergence_condition.relative_tolerance = 1e-26
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-30
[model.reduce]
u_max.evaluation.expression = u
u_max.reduction.expression = init, val: max(init, val)
u_max.initial.value = -1e100
u_max.error.expression = arg: arg > 1/(4*3.14159265359*diffusion)
u_min.evaluation.expression = u
u_min.reduction.expression = init, val: min(init, val)
u_min.initial.value = 1e100
u_min.warn.expression = arg: arg < 0
u_min.error.expression = arg: arg < -1e-2
u_error.evaluation.expression = (u - gauss(position_x, position_y, position_z, time))^2 * integration_factor
u_error.transformation.expression = arg: sqrt(arg)
u_error.error.expression = arg: arg > 0.50
"""


def _heat_initial(o, level=5):
    n = 2 ** level
    xyz = _vertex_xyz(o)[o.compartment_nodes(0)]
    assert xyz[:, 0].min() == -1.0 and xyz[:, 0].max() == 1.0 and o.n_dof == (n + 1) ** 2
    r2 = (xyz ** 2).sum(axis=1)
    return np.exp(-r2 / (4 * 1.0 * 0.005)) / (4 * 3.14159265359 * 1.0 * 0.005)


@pytest.mark.parametrize("cell_type", ["simplex", "cube"])
def test_heat_kernel_linear_path_keeps_the_reference_bound(cell_type):
    o_lin = OracleModel(heat_kernel_ini(cell_type))
    o_new = OracleModel(heat_kernel_ini(cell_type), "model.is_linear = false")
    u0 = _heat_initial(o_lin)
    for o in (o_lin, o_new):
        o.set_state(u0, 1.0)
    vals, flags, rc = o_lin.reduce()
    assert rc == 0 and vals["u_error"] < 0.5
    for _ in range(4):
        rc, newton, _ = o_lin.step(0.05)
        assert rc == 0 and newton == 2  # Alexander2: one correction per stage
        assert o_new.step(0.05)[0] == 0
    ul, t = o_lin.get_state()
    un, _ = o_new.get_state()
    assert t == pytest.approx(1.2)
    # the linear path and Newton solve the same linear stage equations
    assert np.linalg.norm(ul - un) <= 1e-9 * np.linalg.norm(un)
    vals, flags, rc = o_lin.reduce()
    assert rc == 0 and 0.0 < vals["u_error"] < 0.5 and flags["u_max"] == 0
    # the heat kernel decays: its peak 1/(4 pi D t) falls from t = 1 to 1.2
    assert vals["u_max"] < 1.0 / (4 * math.pi * 0.005)
