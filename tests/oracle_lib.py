"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: used by the tests, the smoke entry and the benchmark's
CPU-baseline leg only.  The oracle is a plain C++ restatement of the
reference's diffusion-reaction step (see oracle/src/oracle.cc) with scalar
CSR matrices, sequential ILU0/BiCGSTAB and OpenMP-coloured assembly.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "_build" / "liboracle.so"
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        raise FileNotFoundError(f"{LIB} missing: run `make -C oracle`")
    L = ctypes.CDLL(str(LIB))
    P, D, I, Lg = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_long
    sig = {
        "orc_create": (P, [ctypes.c_char_p, ctypes.c_char_p]),
        "orc_destroy": (None, [P]),
        "orc_last_error": (ctypes.c_char_p, []),
        "orc_set_threads": (None, [I]),
        "orc_set_krylov_threads": (None, [I]),
        "orc_reduce": (I, [P, ctypes.c_char_p, ctypes.c_char_p, I, P, P, I, ctypes.POINTER(I)]),
        "orc_max_threads": (I, []),
        "orc_n_dof": (Lg, [P]),
        "orc_n_species": (I, [P]),
        "orc_n_compartments": (I, [P]),
        "orc_n_vertices": (Lg, [P]),
        "orc_n_elements": (Lg, [P]),
        "orc_n_facets": (Lg, [P]),
        "orc_species": (I, [P, I, ctypes.c_char_p, I, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "orc_comp_nodes": (Lg, [P, I, P]),
        "orc_comp_dof_offset": (Lg, [P, I]),
        "orc_comp_species_count": (I, [P, I]),
        "orc_vertex_coords": (I, [P, P]),
        "orc_set_state": (I, [P, P, D]),
        "orc_get_state": (I, [P, P, P]),
        "orc_step": (I, [P, D, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "orc_evolve": (I, [P, D, D, D, D, D, D, I, P, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(D)]),
        "orc_last_assembly_seconds": (D, [P]),
        "orc_last_solve_seconds": (D, [P]),
        "orc_nnz": (Lg, [P]),
        "orc_pattern": (I, [P, P, P]),
        "orc_assemble": (I, [P, P, D, D, D, P, P]),
        "orc_solve": (I, [P, P, P, P, D, I, ctypes.POINTER(I)]),
        "orc_mass": (I, [P, P, P]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    _lib = L
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class OracleModel:
    def __init__(self, ini: str, overrides: str = ""):
        L = load()
        self._L = L
        self._h = L.orc_create(ini.encode(), overrides.encode())
        if not self._h:
            raise RuntimeError(L.orc_last_error().decode())
        self.n_dof = L.orc_n_dof(self._h)
        self.n_species = L.orc_n_species(self._h)
        self.n_compartments = L.orc_n_compartments(self._h)
        self.time = 0.0

    def close(self):
        if self._h:
            self._L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def compartments(self):
        out = []
        for c in range(self.n_compartments):
            n = self._L.orc_comp_nodes(self._h, c, None)
            out.append({"n_nodes": n, "dof_offset": self._L.orc_comp_dof_offset(self._h, c),
                        "n_species": self._L.orc_comp_species_count(self._h, c)})
        return out

    def compartment_nodes(self, c):
        n = self._L.orc_comp_nodes(self._h, c, None)
        ids = np.zeros(n, dtype=np.int64)
        self._L.orc_comp_nodes(self._h, c, _p(ids))
        return ids

    def species(self):
        out = []
        for s in range(self.n_species):
            buf = ctypes.create_string_buffer(256)
            c, l = ctypes.c_int(), ctypes.c_int()
            self._L.orc_species(self._h, s, buf, 256, ctypes.byref(c), ctypes.byref(l))
            out.append((buf.value.decode(), c.value, l.value))
        return out

    def set_state(self, u, time=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._L.orc_set_state(self._h, _p(u), time)

    def get_state(self):
        u = np.zeros(self.n_dof)
        t = ctypes.c_double()
        self._L.orc_get_state(self._h, _p(u), ctypes.byref(t))
        return u, t.value

    def evolve(self, t_end, dt, decrease=0.5, increase=1.1, min_step=0.0, max_step=0.0, snap=True):
        """Adaptive time loop; returns (status, accepted step sizes, failed attempts, next dt)."""
        acc = np.zeros(100000)
        na, nf, dn = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        rc = self._L.orc_evolve(self._h, t_end, dt, decrease, increase, min_step, max_step, 1 if snap else 0,
                                _p(acc), acc.size, ctypes.byref(na), ctypes.byref(nf), ctypes.byref(dn))
        if rc < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return rc, acc[:na.value].copy(), nf.value, dn.value

    def reduce(self, config=""):
        """Transform/reduce on the current state: ({key: value}, {key: flag}, rc)."""
        names = ctypes.create_string_buffer(1 << 14)
        vals = np.zeros(256)
        flags = np.zeros(256, dtype=np.int32)
        n = ctypes.c_int()
        rc = self._L.orc_reduce(self._h, config.encode() if config else None, names, len(names), _p(vals), _p(flags),
                                256, ctypes.byref(n))
        if rc < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        keys = names.value.decode().split("\n") if n.value else []
        return ({k: vals[i] for i, k in enumerate(keys)}, {k: int(flags[i]) for i, k in enumerate(keys)}, rc)

    def step(self, dt):
        ni, li = ctypes.c_int(), ctypes.c_int()
        rc = self._L.orc_step(self._h, dt, ctypes.byref(ni), ctypes.byref(li))
        if rc < 0:
            raise RuntimeError(self._L.orc_last_error().decode())
        return rc, ni.value, li.value

    def pattern(self):
        nnz = self._L.orc_nnz(self._h)
        rp = np.zeros(self.n_dof + 1, dtype=np.int64)
        col = np.zeros(nnz, dtype=np.int32)
        self._L.orc_pattern(self._h, _p(rp), _p(col))
        return rp, col

    def assemble(self, x, t, cm, ca, jacobian=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(self.n_dof)
        vals = np.zeros(self._L.orc_nnz(self._h)) if jacobian else None
        rc = self._L.orc_assemble(self._h, _p(x), t, cm, ca, _p(r), _p(vals) if vals is not None else None)
        if rc:
            raise RuntimeError(self._L.orc_last_error().decode())
        return r, vals

    def mass(self, x=None):
        if x is None:
            x, _ = self.get_state()
        out = np.zeros(self.n_species)
        self._L.orc_mass(self._h, _p(np.ascontiguousarray(x, dtype=np.float64)), _p(out))
        return out

    def csr(self, vals):
        import scipy.sparse as sp
        rp, col = self.pattern()
        return sp.csr_matrix((vals, col, rp), shape=(self.n_dof, self.n_dof))


def bcsr_to_dense_csr(model, J):
    """Expand a product BCSR Jacobian (n_blocks, B, B) to a scipy CSR matrix over
    packed DOFs (padded species dropped)."""
    import scipy.sparse as sp
    rp, col = model.pattern()
    first, cnt = model.slot_dofs()
    rows, cols, vals = [], [], []
    B = model.block_size
    nslot = model.n_slots
    blk_row = np.repeat(np.arange(nslot), np.diff(rp))
    for i in range(B):
        for j in range(B):
            ok = (i < cnt[blk_row]) & (j < cnt[col])
            rows.append(first[blk_row[ok]] + i)
            cols.append(first[col[ok]] + j)
            vals.append(J[ok, i, j])
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.concatenate(vals)
    return sp.csr_matrix((vals, (rows, cols)), shape=(model.n_dof, model.n_dof))


def rel_l2_per_species(model_info, a, b, weights=None):
    """max over species of ||a-b|| / ||b|| (Euclidean over nodes; per species)."""
    out = []
    for comp in model_info:
        n, off, bsp = comp["n_nodes"], comp["dof_offset"], comp["n_species"]
        for i in range(bsp):
            sa = a[off + i: off + bsp * n: bsp]
            sb = b[off + i: off + bsp * n: bsp]
            den = np.linalg.norm(sb)
            out.append(np.linalg.norm(sa - sb) / (den if den > 0 else 1.0))
    return out


def rel_mass_l2_per_species(oracle_model, a, b, t=0.0):
    """max over species of ||a-b||_M / ||b||_M, M the oracle's FP64 mass
    matrix restricted to the species (the mass operator's Jacobian at b, which
    is the consistent mass matrix times the storage coefficient)."""
    _, vals = oracle_model.assemble(b, t, 1.0, 0.0)
    M = oracle_model.csr(vals)
    out = []
    for comp in oracle_model.compartments():
        n, off, bsp = comp["n_nodes"], comp["dof_offset"], comp["n_species"]
        for i in range(bsp):
            idx = off + i + bsp * np.arange(n)
            # x^T M x of a vector that is zero off the species is the species
            # block's quadratic form (no submatrix is formed)
            e = np.zeros(len(a))
            e[idx] = a[idx] - b[idx]
            bs = np.zeros(len(b))
            bs[idx] = b[idx]
            den = float(bs @ (M @ bs))
            num = float(e @ (M @ e))
            out.append(np.sqrt(max(num, 0.0) / den) if den > 0 else np.sqrt(max(num, 0.0)))
    return out
