"""VTK output (dcg_write_vtk) read back: the files hold the state, mesh and series.

The reference writes its state through dune-grid's VTKWriter; there is no
written file among its test data to compare with, so the files are checked
against what they must contain: per compartment, every node's coordinates
and species values as get_state returns them, cells whose corners are the
compartment's nodes with the right VTK cell type and positive measure, and a
.pvd series that lists every step written since the last restart.
"""
import base64
import re
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _array(el):
    raw = base64.b64decode(el.text.strip())
    n = struct.unpack("<I", raw[:4])[0]
    dt = {"Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}[el.get("type")]
    a = np.frombuffer(raw[4:4 + n], dtype=dt)
    k = int(el.get("NumberOfComponents", "1"))
    return a.reshape(-1, k) if k > 1 else a


def _read(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arrays = {a.get("Name"): _array(a) for a in piece.iter("DataArray")}
    return int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells")), arrays


@pytest.mark.parametrize("make,dim,cell,nodes", [
    (lambda wl: wl.gray_scott(n=6, cell_type="cube"), 2, 9, 4),
    (lambda wl: wl.gray_scott(n=5, cell_type="simplex"), 2, 5, 3),
    (lambda wl: wl.mitchell_schaefer(n=3, cell_type="cube"), 3, 12, 8),
    (lambda wl: wl.two_compartments(n=12, cell_type="cube"), 2, 9, 4),
], ids=["gs_q1", "gs_p1", "ms_q1", "two_q1"])
def test_vtk_files_hold_the_state(dca, wl, tmp_path, make, dim, cell, nodes):
    w = make(wl)
    m = dca.Model(w.ini, "", 0)
    m.set_state(w.initial_state(m), 0.0)
    out = tmp_path / "run"
    m.write_vtk(out, append=False)
    u1 = w.initial_state(m) + np.random.default_rng(2).random(m.n_dof)
    m.set_state(u1, 0.5)
    m.write_vtk(out)
    u, t = m.get_state()
    for c, comp in enumerate(m.compartments()):
        base = out / f"run-{comp['name']}"
        npts, ncell, arr = _read(f"{base}-00001.vtu")
        b, off = comp["n_species"], comp["dof_offset"]
        assert npts == comp["n_nodes"]
        # coordinates of the compartment's vertices, in node order
        ids = m.compartment_nodes(c)
        n = [int(v) for v in re.search(r"cells\s*=\s*([\d ]+)", w.ini).group(1).split()]
        ijk = [ids % (n[0] + 1), (ids // (n[0] + 1)) % (n[1] + 1)] + ([ids // ((n[0] + 1) * (n[1] + 1))] if dim > 2 else [])
        for k in range(dim):
            assert np.allclose(arr["Coordinates"][:, k], ijk[k] / n[k], rtol=0, atol=1e-14)
        for i in range(b):
            sp = [s[0] for s in m.species() if s[1] == c and s[2] == i][0]
            assert np.array_equal(arr[sp], u[off + i: off + b * npts: b])
        assert np.all(arr["types"] == cell) and ncell == arr["types"].size
        conn = arr["connectivity"]
        assert np.array_equal(arr["offsets"], nodes * np.arange(1, ncell + 1))
        assert conn.min() >= 0 and conn.max() < npts
        # every cell has positive measure in the given corner order
        X = arr["Coordinates"][conn.reshape(ncell, nodes)][:, :, :dim]
        if cell == 5:
            e1, e2 = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]
            assert np.all(np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]) > 0)
        elif cell == 9:
            e1, e2 = X[:, 1] - X[:, 0], X[:, 3] - X[:, 0]
            assert np.all(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0] > 0)  # counter-clockwise quads
        series = ET.parse(f"{base}.pvd").getroot().findall("Collection/DataSet")
        assert [d.get("file") for d in series] == [f"run-{comp['name']}-00000.vtu", f"run-{comp['name']}-00001.vtu"]
        assert float(series[1].get("timestep")) == t
    # a restart without append begins a new series
    m.write_vtk(out, append=False)
    comp = m.compartments()[0]
    series = ET.parse(out / f"run-{comp['name']}.pvd").getroot().findall("Collection/DataSet")
    assert len(series) == 1
