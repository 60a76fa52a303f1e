"""Host side of the product, no GPU needed.

* libdcopasi.so loads and exports every function include/dcopasi.h declares;
* the host-only model (device < 0) builds the grid, compartments and BCSR
  pattern: block counts equal the stencil counts of the benchmark meshes,
  compartment node sets and the sparsity equal the oracle's (which is built by
  independent code), and bad input fails with the reference's messages;
* the expression compiler, exercised through `initial.expression`, agrees
  with Python's arithmetic;
* the generated assembly source compiles with hipRTC for gfx950.
"""
import ctypes
import math
import re

import numpy as np
import pytest

from tests.oracle_lib import OracleModel


def test_library_exports_every_declared_function(dca):
    lib = dca.load_library()
    header = dca.HEADER.read_text()
    names = re.findall(r"^\s*(?:int|void|const char\*)\s+(dcg_\w+)\s*\(", header, flags=re.M)
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("cell_type,blocks", [("cube", 148225), ("simplex", 115457)])
def test_gray_scott_128_pattern_counts(dca, wl, cell_type, blocks):
    m = dca.Model(wl.gray_scott(n=128, cell_type=cell_type).ini, "", -1)
    assert m.n_dof == 33282 and m.n_slots == 16641 and m.block_size == 2
    assert m.n_blocks == blocks


def test_mitchell_schaefer_96_pattern_counts(dca, wl):
    m = dca.Model(wl.mitchell_schaefer(n=96, cell_type="cube").ini, "", -1)
    assert m.n_dof == 1825346 and m.n_blocks == 24137569


def test_two_compartment_counts(dca, wl):
    m = dca.Model(wl.two_compartments(n=1024, cell_type="cube").ini, "", -1)
    comps = m.compartments()
    assert [c["n_nodes"] for c in comps] == [297745, 755336]
    assert m.n_dof == 4 * (297745 + 755336)
    rp, col = m.pattern()
    # inner rows couple to inner columns only in the volume part: 2,672,329 blocks;
    # the outer volume part has 6,778,368; the rest are interface couplings
    inner = comps[0]["n_nodes"]
    rows = np.repeat(np.arange(m.n_slots), np.diff(rp))
    same = (rows < inner) == (col < inner)
    assert int(np.sum(same & (rows < inner))) == 2672329
    assert int(np.sum(same & (rows >= inner))) == 6778368
    assert int(np.sum(~same)) > 0


CASES = [
    lambda wl: wl.gray_scott(n=7, cell_type="cube"),
    lambda wl: wl.gray_scott(n=6, cell_type="simplex"),
    lambda wl: wl.mitchell_schaefer(n=3, cell_type="simplex"),
    lambda wl: wl.mitchell_schaefer(n=4, cell_type="cube"),
    lambda wl: wl.two_compartments(n=14, cell_type="cube"),
    lambda wl: wl.two_compartments(n=12, cell_type="simplex"),
]


@pytest.mark.parametrize("make", CASES, ids=["gs_q1", "gs_p1", "ms_p1", "ms_q1", "two_q1", "two_p1"])
def test_layout_matches_oracle(dca, wl, make):
    w = make(wl)
    m = dca.Model(w.ini, "", -1)
    o = OracleModel(w.ini)
    assert m.n_dof == o.n_dof
    for c in range(m.n_compartments):
        assert np.array_equal(m.compartment_nodes(c), o.compartment_nodes(c))
    assert [s[0] for s in m.species()] == [s[0] for s in o.species()]
    # expand the BCSR pattern to scalar entries and compare with the oracle's CSR
    rp, col = m.pattern()
    first, cnt = m.slot_dofs()
    rows = np.repeat(np.arange(m.n_slots), np.diff(rp))
    ent = set()
    for r, c in zip(rows, col):
        for i in range(cnt[r]):
            for j in range(cnt[c]):
                ent.add((first[r] + i, first[c] + j))
    orp, ocol = o.pattern()
    orows = np.repeat(np.arange(o.n_dof), np.diff(orp))
    assert ent == set(zip(orows.tolist(), ocol.tolist()))


PY_EXPR = [
    ("-position_x^2 + 3*position_y", lambda x, y: -x * x + 3 * y),
    ("position_x < position_y ? exp(position_x) : log(1+position_y)",
     lambda x, y: math.exp(x) if x < y else math.log(1 + y)),
    ("max(position_x, position_y, 0.25) * min(2, position_x + 1)", lambda x, y: max(x, y, 0.25) * min(2, x + 1)),
    ("bump(position_x - 0.5, position_y - 0.5) + k", lambda x, y: 0.5 * math.exp(-100 * ((x - .5) ** 2 + (y - .5) ** 2)) + 4),
    ("(position_x > 0.3) && (position_y <= 0.6) || not (position_x > 0.1)",
     lambda x, y: float((x > 0.3 and y <= 0.6) or not (x > 0.1))),
    ("2^-position_x^2 + abs(position_x - position_y) + sqrt(position_y)", lambda x, y: 2 ** -(x * x) + abs(x - y) + math.sqrt(y)),
    ("atan2(position_y, position_x + 1) + tanh(position_x) + sin(position_y)^2",
     lambda x, y: math.atan2(y, x + 1) + math.tanh(x) + math.sin(y) ** 2),
]


@pytest.mark.parametrize("text,ref", PY_EXPR, ids=[p[0][:30] for p in PY_EXPR])
def test_expression_compiler_through_initial_conditions(dca, text, ref):
    ini = f"""[grid]
dimension = 2
cells = 5 4
[parser_context.bump]
type = function
expression = a, b: 0.5*exp(-100*(a^2 + b^2))
[parser_context.k]
type = constant
value = 4
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
initial.expression = {text}
"""
    m = dca.Model(ini, "", -1)
    u = m.initial_state()
    ids = m.compartment_nodes(0)
    xs = (ids % 6) / 5.0
    ys = (ids // 6) / 4.0
    ref_v = np.array([ref(x, y) for x, y in zip(xs, ys)])
    np.testing.assert_allclose(u, ref_v, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("ini,msg", [
    ("[grid]\ndimension = 2\n[compartments]\nd.expression = 1\n", "Basis has dimension 0"),
    ("[grid]\ndimension = 2\n[compartments]\nd.expression = 1\n[model.scalar_field.u]\ncompartment = d\n"
     "[model]\njacobian.type = foo\n", "must be either 'analytical', 'numerical' or 'symbolic'"),
    ("[grid]\ndimension = 2\n[compartments]\nd.expression = 1\n[model.scalar_field.u]\ncompartment = d\n"
     "reaction.expression = u*(\n", "expression"),
    ("[grid]\ndimension = 2\n[compartments]\nd.expression = 1\n[model.scalar_field.u]\ncompartment = e\n",
     "unknown compartment"),
    # istl's SeqJac sweeps and relaxation: only the defaults (one sweep, weight 1) are the diagonal inverse
    ("[grid]\ndimension = 2\n[compartments]\nd.expression = 1\n[model.scalar_field.u]\ncompartment = d\n"
     "[model.time_step_operator]\nlinear_solver.preconditioner.type = Jacobi\n"
     "linear_solver.preconditioner.relaxation = 0.8\n", "Jacobi with relaxation != 1"),
    ("[grid]\ndimension = 2\n[compartments]\nd.expression = 1\n[model.scalar_field.u]\ncompartment = d\n"
     "[model.time_step_operator]\nlinear_solver.preconditioner.type = AMG\n", "preconditioner 'AMG'"),
    # a matrix-free operator has no matrix for Jacobi / ILU (preconditioner.hh:95-117)
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\n"
     "[model.time_step_operator]\nlinear_solver.matrix_free = true\nlinear_solver.preconditioner.type = Jacobi\n",
     "needs an assembled matrix"),
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\n"
     "[model.time_step_operator]\nlinear_solver.matrix_free = true\nlinear_solver.precision = mixed\n",
     "mixed precision with the matrix-free Jacobian"),
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\nvelocity.x.expression = 0.1\n"
     "[model.time_step_operator]\nlinear_solver.matrix_free = true\n",
     "with the tiled kernel"),
    # point-coefficient storage: a preconditioner from diagonal blocks,
    # compartments of the tiled kernel
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\n"
     "[model.time_step_operator]\nlinear_solver.storage = points\nlinear_solver.preconditioner.type = ILU\n",
     "storage = points with an ILU0 preconditioner"),
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\nvelocity.x.expression = 0.1\n"
     "[model.time_step_operator]\nlinear_solver.storage = points\nlinear_solver.preconditioner.type = Jacobi\n",
     "storage = points with compartments outside the tiled kernel"),
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\n"
     "[model.time_step_operator]\nlinear_solver.storage = csr\n",
     "linear_solver.storage must be"),
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\n"
     "[model.time_step_operator]\nlinear_solver.readback_iterations = -1\n",
     "linear_solver.readback_iterations must be >= 0"),
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\n"
     "[model.time_step_operator]\nlinear_solver.point_tile = widest\n",
     "linear_solver.point_tile must be 'auto' or a candidate index"),
    ("[grid]\ndimension = 2\ncells = 4 4\ncell_type = cube\n[compartments]\nd.expression = 1\n"
     "[model.scalar_field.u]\ncompartment = d\nstorage.expression = 1\n"
     "[model.time_step_operator]\nlinear_solver.refinement.tolerance_floor = 1\n",
     "linear_solver.refinement.tolerance_floor must be in [0, 1)"),
])
def test_errors_are_reported(dca, ini, msg):
    with pytest.raises(dca.DcgError, match=re.escape(msg)):
        dca.Model(ini, "", -1)


@pytest.mark.parametrize("rk", ["ExplicitEuler", "ImplicitEuler", "Heun", "Shu3", "RungeKutta4", "Alexander2",
                                "FractionalStepTheta", "Alexander3"])
def test_time_stepping_types(dca, wl, rk):
    """Every type of model/make_step_operator.hh:422-439 builds; others are an IOError."""
    w = wl.gray_scott(n=4, cell_type="cube")
    dca.Model(w.ini, f"model.time_step_operator.type = {rk}\n", -1).close()


def test_unknown_time_stepping_type(dca, wl):
    w = wl.gray_scott(n=4, cell_type="cube")
    with pytest.raises(dca.DcgError, match="Not known time stepping type 'Euler'"):
        dca.Model(w.ini, "model.time_step_operator.type = Euler\n", -1)


_INTERP = """[grid]
dimension = 2
cells = 8 4
[parser_context.f]
type = interpolation
domain = 0.1 0.3 0.3 0.8
range = 0.5 -0.25 2.0 1.0
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
initial.expression = f(position_x) + 10*f(position_y - 0.2)
"""


def _interp_ref(x):
    """The reference's interpolation (context.cc:72-97) by hand: first domain
    point >= x; the end values outside."""
    d, r = [0.1, 0.3, 0.3, 0.8], [0.5, -0.25, 2.0, 1.0]
    i = next((k for k, v in enumerate(d) if v >= x), len(d))
    if i == 0:
        return r[0]
    if i == len(d):
        return r[-1]
    return r[i - 1] + (x - d[i - 1]) / (d[i] - d[i - 1]) * (r[i] - r[i - 1])


def test_interpolation_context(dca):
    """parser_context type = interpolation through initial conditions: nodes
    below, on and between the points and beyond the last (the duplicated 0.3
    is a jump); np.interp agrees away from the jump."""
    m = dca.Model(_INTERP, "", -1)
    u = m.initial_state()
    ids = m.compartment_nodes(0)
    xs, ys = (ids % 9) / 8.0, (ids // 9) / 4.0
    ref = np.array([_interp_ref(x) + 10 * _interp_ref(y - 0.2) for x, y in zip(xs, ys)])
    np.testing.assert_allclose(u, ref, rtol=1e-15, atol=1e-15)
    away = np.abs(xs - 0.3) > 1e-9
    np.testing.assert_allclose(u[away] - 10 * np.array([_interp_ref(y - 0.2) for y in ys[away]]),
                               np.interp(xs[away], [0.1, 0.3, 0.8], [0.5, -0.25, 1.0]) * (xs[away] < 0.3)
                               + np.interp(xs[away], [0.3, 0.8], [2.0, 1.0]) * (xs[away] > 0.3), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("dom,rng,msg", [("0 2 1", "1 2 3", "domain must be sorted"),
                                         ("0 1", "1 2 3", "at least two points and be the same size"),
                                         ("0", "1", "at least two points and be the same size")])
def test_interpolation_context_errors(dca, dom, rng, msg):
    ini = _INTERP.replace("domain = 0.1 0.3 0.3 0.8", f"domain = {dom}").replace("range = 0.5 -0.25 2.0 1.0",
                                                                                 f"range = {rng}")
    with pytest.raises(dca.DcgError, match=msg):
        dca.Model(ini, "", -1)


def test_overrides_apply_last(dca, wl):
    w = wl.gray_scott(n=8, cell_type="cube")
    m = dca.Model(w.ini, "grid.cells=4 4\n", -1)
    assert m.n_slots == 25


def _hiprtc_compile(src):
    L = ctypes.CDLL("/opt/rocm/lib/libhiprtc.so")
    prog = ctypes.c_void_p()
    assert L.hiprtcCreateProgram(ctypes.byref(prog), src.encode(), b"t.hip", 0, None, None) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17"]
    arr = (ctypes.c_char_p * len(opts))(*opts)
    rc = L.hiprtcCompileProgram(prog, len(opts), arr)
    n = ctypes.c_size_t()
    L.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
    log = ctypes.create_string_buffer(n.value + 1)
    L.hiprtcGetProgramLog(prog, log)
    return rc, log.value.decode(errors="replace")


@pytest.mark.parametrize("make", CASES[:2] + CASES[4:5], ids=["gs_q1", "gs_p1", "two_q1"])
def test_generated_kernels_compile_for_gfx950(dca, wl, make):
    m = dca.Model(make(wl).ini, "", -1)
    src = m.generated_source()
    # tile-eligible compartments get only the tiled kernel (the coloured one
    # would be dead code and dominates compile time)
    assert "dcg_tile_0" in src and "dcg_volume_0" not in src
    rc, log = _hiprtc_compile(src)
    assert rc == 0, log


def test_large_block_kernel_source_compiles(dca, wl):
    """8 species in 3D: the block rows are split over lanes (and the
    matrix-free apply keeps one block row per lane)."""
    m = dca.Model(wl.random_request(18, side3=(5, 6)).ini, "", -1)
    rc, log = _hiprtc_compile(m.generated_source())
    assert rc == 0, log


def test_coloured_kernel_source_compiles(dca):
    """A model with advection needs the coloured kernel; it is generated and compiles."""
    ini = """[grid]
dimension = 2
cells = 8 8
cell_type = cube
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
storage.expression = 1
cross_diffusion.u.expression = 1e-3
velocity.x.expression = 0.02
reaction.expression = -u
reaction.jacobian.u.expression = -1
[model.time_step_operator]
type = ImplicitEuler
"""
    src = dca.Model(ini, "", -1).generated_source()
    assert "dcg_volume_0" in src and "dcg_tile_0" not in src
    rc, log = _hiprtc_compile(src)
    assert rc == 0, log


def test_runtime_parameter_source_is_shared(dca, wl):
    """With run-time parameters, models that differ only in mesh size and
    parameter values generate the same source (one hipRTC module for a whole
    batch), and that source holds the batched tile kernel and compiles."""
    ws = wl.gray_scott_batch(count=4, seed=1, side_min=8, side_max=64)
    srcs = []
    for w in ws:
        m = dca.Model(w.ini, "", -1, runtime_parameters=True)
        assert m.param_names() == ["Du", "Dv", "F", "k"]
        p = m.params()
        assert 0.01 <= p[2] <= 0.08 and 0.04 <= p[3] <= 0.07 and 1.5 <= p[0] / p[1] <= 3.0
        srcs.append(m.generated_source())
    assert all(s == srcs[0] for s in srcs)
    assert "dcg_tile_0_batch" in srcs[0] and "#define DCG_GEO 1" in srcs[0]
    rc, log = _hiprtc_compile(srcs[0])
    assert rc == 0, log
    # folded models embed their constants and spacing instead
    folded = dca.Model(ws[0].ini, "", -1).generated_source()
    assert "#define DCG_GEO 0" in folded and "_batch" not in folded


def test_gray_scott_batch_workload(wl):
    ws = wl.gray_scott_batch(count=256)
    sides = np.array([w.cells[0] for w in ws])
    assert sides.min() >= 32 and sides.max() <= 512
    # log-uniform: about half the instances below the geometric mean 128
    assert 0.35 < np.mean(sides < 128) < 0.65
    assert len({w.ini for w in ws}) == 256


def test_footprint_estimate_scales_with_mesh(dca, wl):
    small = dca.estimate_device_bytes(wl.gray_scott(n=32, precision="mixed").ini)
    big = dca.estimate_device_bytes(wl.gray_scott(n=128, precision="mixed").ini)
    assert small > 0 and 12 < big / small < 20
    fp64 = dca.estimate_device_bytes(wl.gray_scott(n=128, precision="fp64").ini)
    assert fp64 < big


@pytest.mark.parametrize("make", [lambda wl: wl.gray_scott(n=6, cell_type="cube"),
                                  lambda wl: wl.mitchell_schaefer(n=4, cell_type="simplex"),
                                  lambda wl: wl.two_compartments(n=8, cell_type="cube")],
                         ids=["gs_q1", "ms_p1", "two_q1"])
def test_reduce_source_compiles_for_gfx950(dca, wl, make):
    m = dca.Model(make(wl).ini, "", -1)
    s0 = m.species()[0][0]
    cfg = (f"[mx]\nevaluation.expression = {s0}\nreduction.expression = a, b: max(a, b)\ninitial.value = -1e100\n"
           f"[l2]\nevaluation.expression = {s0}^2*integration_factor\ntransformation.expression = x: sqrt(x)\n"
           "[t]\nevaluation.expression = time*position_x\n")
    src = m.reduce_source(cfg)
    assert "dcg_reduce_0" in src
    rc, log = _hiprtc_compile(src)
    assert rc == 0, log


def test_reduce_configuration_errors(dca, wl):
    m = dca.Model(wl.gray_scott(n=6, cell_type="cube").ini, "", -1)
    with pytest.raises(dca.DcgError, match="exactly 2"):
        m.reduce_source("[k]\nevaluation.expression = u\nreduction.expression = a: a\n")
    with pytest.raises(dca.DcgError, match="exactly 1"):
        m.reduce_source("[k]\nevaluation.expression = u\ntransformation.expression = a, b: a\n")


# ---------------------------------------------------------------- grids read from gmsh files
@pytest.fixture(scope="module")
def gmsh_files(tmp_path_factory):
    from tests import gmsh_meshes
    d = tmp_path_factory.mktemp("gmsh")
    gmsh_meshes.square(d / "square.msh", 8)
    gmsh_meshes.cube(d / "cube.msh", 3)
    return d


@pytest.mark.parametrize("mesh,dim", [("square.msh", 2), ("cube.msh", 3)])
@pytest.mark.parametrize("model", ["gray_scott", "two_halves"])
def test_gmsh_layout_matches_oracle(dca, gmsh_files, mesh, dim, model):
    """Cells of the grid dimension, nodes numbered by increasing node number,
    compartments by physical entity (gmsh_id): the same DOF layout as the
    oracle's independent reader; the generated cell kernels compile."""
    from tests import unstructured_models
    from tests.oracle_lib import OracleModel
    ini = getattr(unstructured_models, model)(gmsh_files / mesh, dim)
    m = dca.Model(ini, "", -1)
    o = OracleModel(ini)
    assert m.n_dof == o.n_dof
    for c, (a, b) in enumerate(zip(m.compartments(), o.compartments())):
        assert (a["n_nodes"], a["dof_offset"], a["n_species"]) == (b["n_nodes"], b["dof_offset"], b["n_species"])
        assert np.array_equal(m.compartment_nodes(c), o.compartment_nodes(c))
    src = m.generated_source()
    assert "dcg_tile" not in src and "DCG_UNSTRUCTURED 1" in src
    rc, log = _hiprtc_compile(src)
    assert rc == 0, log


@pytest.mark.parametrize("mesh,dim,levels", [("square.msh", 2, 1), ("square.msh", 2, 2), ("cube.msh", 3, 1),
                                             ("cube.msh", 3, 2)])
def test_gmsh_refinement(dca, gmsh_files, mesh, dim, levels):
    """grid.refinement_level on a file grid: red refinement keeps the
    measure of every compartment, adds vertices, and gives the same layout as
    the oracle's own refinement."""
    from tests import unstructured_models
    from tests.oracle_lib import OracleModel
    base = unstructured_models.two_halves(gmsh_files / mesh, dim)
    ini = base + f"[grid]\nrefinement_level = {levels}\n"
    coarse = dca.Model(base, "", -1)
    m = dca.Model(ini, "", -1)
    o = OracleModel(ini)
    assert m.n_dof == o.n_dof
    for c, (a, b) in enumerate(zip(m.compartments(), o.compartments())):
        assert (a["n_nodes"], a["dof_offset"], a["n_species"]) == (b["n_nodes"], b["dof_offset"], b["n_species"])
        assert np.array_equal(m.compartment_nodes(c), o.compartment_nodes(c))
    # more vertices, and each compartment keeps its measure (the integral of 1)
    oc = OracleModel(base)
    for c in range(len(m.compartments())):
        assert len(coarse.compartment_nodes(c)) < len(m.compartment_nodes(c))
    np.testing.assert_allclose(o.mass(np.ones(o.n_dof)), oc.mass(np.ones(oc.n_dof)), rtol=1e-12)
    assert abs(sum(o.mass(np.ones(o.n_dof))) / 2 - 1.0 / 2) < 1e-12  # each species sums over its half; halves fill the unit box


def test_gmsh_errors(dca, gmsh_files, tmp_path):
    from tests import unstructured_models
    ini = unstructured_models.gray_scott(gmsh_files / "square.msh")
    with pytest.raises(dca.DcgError, match="refinement_level must not be negative"):
        dca.Model(ini + "[grid]\nrefinement_level = -1\n", "", -1)
    with pytest.raises(dca.DcgError, match="cannot open"):
        dca.Model(unstructured_models.gray_scott(tmp_path / "missing.msh"), "", -1)
    quad = tmp_path / "quad.msh"
    quad.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n4\n1 0 0 0\n2 1 0 0\n3 1 1 0\n4 0 1 0\n$EndNodes\n"
                    "$Elements\n1\n1 3 2 1 1 1 2 3 4\n$EndElements\n")
    with pytest.raises(dca.DcgError, match="simplices only"):
        dca.Model(unstructured_models.gray_scott(quad), "", -1)


def test_committed_traffic_covers_the_apply_variants():
    """The committed counter summary of the 96^3 FP64 workload gives each
    measured point-apply variant its calibrated HBM bytes per launch; the
    figure is at least the algorithmic bytes of one apply (every cell's 8
    points x 3 coefficients and class mask, then the direction, the result
    and one dot operand over 2 x 97^3 values, 8 bytes each, and a 4-byte
    slot per vertex) and not far above them."""
    import importlib.util as iu
    spec = iu.spec_from_file_location("bench_for_test", str(__import__("pathlib").Path(__file__).parents[1] / "bench.py"))
    bench = iu.module_from_spec(spec)
    spec.loader.exec_module(bench)
    algorithmic = 96 ** 3 * (8.0 * 24 + 1) + 3 * 8.0 * 2 * 97 ** 3 + 4.0 * 97 ** 3
    assert round(algorithmic) == 218213044
    seen = 0
    for k in range(9):
        t = bench.committed_traffic("ms96", "fp64", f"dcg_tile_0_p{k}")
        if t is None:
            continue
        seen += 1
        assert algorithmic <= t <= 1.5 * algorithmic, (k, t, algorithmic)
    assert seen >= 2
    assert bench.committed_traffic("ms96", "fp64", "dcg_tile_0_pf4") is None
    assert bench.committed_traffic("gs128", "fp64", "dcg_tile_0_p4") is None
