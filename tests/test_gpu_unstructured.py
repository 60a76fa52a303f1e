"""Grids read from gmsh files (unstructured triangles and tetrahedra, also refined) on the device
against the oracle: residual and Jacobian assembly to 1e-12, time-step
sequences to the parity bar, mass sums, and VTK output of simplex cells."""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from tests import gmsh_meshes, unstructured_models
from tests.oracle_lib import OracleModel, bcsr_to_dense_csr, rel_l2_per_species, rel_mass_l2_per_species

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes(tmp_path_factory):
    d = tmp_path_factory.mktemp("gmsh")
    gmsh_meshes.square(d / "square.msh", 10)
    gmsh_meshes.cube(d / "cube.msh", 4)
    return d


CASES = [("gs_tri", "gray_scott", "square.msh", 2), ("two_tri", "two_halves", "square.msh", 2),
         ("gs_tet", "gray_scott", "cube.msh", 3), ("two_tet", "two_halves", "cube.msh", 3),
         # grid.refinement_level on the file grids (red refinement)
         ("two_tri_r2", "two_halves", "square.msh", 2), ("gs_tet_r1", "gray_scott", "cube.msh", 3)]


def _ini(meshes, model, mesh, dim, name=""):
    ini = getattr(unstructured_models, model)(meshes / mesh, dim)
    if "_r" in name:
        ini += f"[grid]\nrefinement_level = {name.rsplit('_r', 1)[1]}\n"
    return ini


@pytest.mark.parametrize("name,model,mesh,dim", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("cm,ca", [(1.0, 0.0), (0.0, 0.7), (1.0, 0.3)])
def test_gmsh_assembly_matches_oracle(dca, meshes, name, model, mesh, dim, cm, ca):
    print(f"case: gmsh assembly {name} cm={cm} ca={ca}", flush=True)
    ini = _ini(meshes, model, mesh, dim, name)
    gpu = dca.Model(ini, "", 0)
    orc = OracleModel(ini)
    x = gpu.initial_state() + 0.05 * np.random.default_rng(5).standard_normal(gpu.n_dof)
    r_g, J_g = gpu.assemble(x, 0.3, cm, ca)
    r_o, v_o = orc.assemble(x, 0.3, cm, ca)
    assert np.abs(r_g - r_o).max() <= 1e-12 * (np.abs(r_o).max() + 1e-300)
    A_g = bcsr_to_dense_csr(gpu, J_g).toarray()
    A_o = orc.csr(v_o).toarray()
    assert np.abs(A_g - A_o).max() <= 1e-12 * np.abs(A_o).max()


@pytest.mark.parametrize("name,model,mesh,dim", CASES, ids=[c[0] for c in CASES])
def test_gmsh_time_steps_match_oracle(dca, meshes, name, model, mesh, dim):
    print(f"case: gmsh steps {name}", flush=True)
    ini = _ini(meshes, model, mesh, dim, name)
    gpu = dca.Model(ini, "", 0)
    orc = OracleModel(ini)
    u0 = gpu.initial_state()
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    dt = 0.5 if model == "gray_scott" else 0.1  # the models' time_step_initial
    t = 0.0
    for k in range(4):
        st = gpu.step(t, dt)
        rc, _, _ = orc.step(dt)
        assert st.status == 0 and rc == 0
        t += dt
        ug, _ = gpu.get_state()
        uo, _ = orc.get_state()
        assert max(rel_l2_per_species(gpu.compartments(), ug, uo)) <= 1e-8, k
        # jittered cells: the mass-matrix norm weighs nodes by their cells' measure
        assert max(rel_mass_l2_per_species(orc, ug, uo, t)) <= 1e-8, k
    assert np.allclose(gpu.mass(), orc.mass(ug), rtol=1e-12, atol=1e-15)


def _array(el):
    raw = base64.b64decode(el.text.strip())
    n = struct.unpack("<I", raw[:4])[0]
    dt = {"Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}[el.get("type")]
    return np.frombuffer(raw[4:4 + n], dtype=dt)


def test_gmsh_vtk_cells(dca, meshes, tmp_path):
    ini = _ini(meshes, "two_halves", "cube.msh", 3)
    m = dca.Model(ini, "", 0)
    m.set_state(m.initial_state(), 0.0)
    m.write_vtk(tmp_path / "t", append=False)
    total = 0
    for comp in m.compartments():
        root = ET.parse(tmp_path / "t" / f"t-{comp['name']}-00000.vtu").getroot()
        piece = root.find("UnstructuredGrid/Piece")
        arr = {a.get("Name"): _array(a) for a in piece.iter("DataArray")}
        assert np.all(arr["types"] == 10) and int(piece.get("NumberOfPoints")) == comp["n_nodes"]
        total += arr["types"].size
    assert total == 6 * 4 ** 3


@pytest.mark.parametrize("name,model,mesh,dim", CASES, ids=[c[0] for c in CASES])
def test_gmsh_reduce_matches_oracle(dca, meshes, name, model, mesh, dim):
    from tests.test_gpu_reduce import _config
    ini = _ini(meshes, model, mesh, dim, name)
    gpu = dca.Model(ini, "", 0)
    orc = OracleModel(ini)
    u = gpu.initial_state() + 0.05 * np.random.default_rng(3).standard_normal(gpu.n_dof)
    gpu.set_state(u, 0.4)
    orc.set_state(u, 0.4)
    cfg = _config([s[0] for s in gpu.species()], dim) + "[area]\nevaluation.expression = integration_factor\n"
    got = gpu.reduce(cfg)
    ref, _, rc = orc.reduce(cfg)
    assert rc == 0 and list(got) == list(ref)
    assert abs(got["area"] - 1.0) <= 1e-12  # the meshes fill the unit box
    for k, v in ref.items():
        tol = 1e-14 * abs(v) if k.startswith(("max", "min")) else 1e-12 * max(1.0, abs(v))
        assert abs(got[k] - v) <= tol, (k, got[k], v)
