"""ctypes binding of libdcopasi.so (the C ABI declared in include/dcopasi.h).

The library is the product: grid/compartment/BCSR set-up and the hipRTC
assembly compiler on the host, the HIP kernels of assembly, SpMV, BiCGSTAB,
Jacobi and ILU0 on the GPU.  This module only moves numpy arrays across
the boundary; it has no numerical code of its own and no fallback: if the
library is missing, loading fails with an error naming the build step.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "lib" / "libdcopasi.so"
HEADER = _HERE.parent / "include" / "dcopasi.h"

_lib = None


class DcgError(RuntimeError):
    """A call through the C ABI returned a negative status."""


class StepStats(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("stages", ctypes.c_int),
        ("newton_iterations", ctypes.c_int),
        ("linear_iterations", ctypes.c_int),
        ("refinement_steps", ctypes.c_int),
        ("residual_norm", ctypes.c_double),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("finished", ctypes.c_int),
        ("steps_done", ctypes.c_int),
        ("steps", ctypes.c_int),
        ("newton_iterations", ctypes.c_long),
        ("linear_iterations", ctypes.c_long),
        ("n_dof", ctypes.c_size_t),
        ("device_bytes", ctypes.c_size_t),
        ("t_arrival", ctypes.c_double),
        ("t_ready", ctypes.c_double),
        ("t_admitted", ctypes.c_double),
        ("t_finished", ctypes.c_double),
    ]


class EvolveOptions(ctypes.Structure):
    _fields_ = [("t_end", ctypes.c_double), ("dt_initial", ctypes.c_double), ("decrease_factor", ctypes.c_double),
                ("increase_factor", ctypes.c_double), ("min_step", ctypes.c_double), ("max_step", ctypes.c_double),
                ("snap_to_end_time", ctypes.c_int)]


class EvolveStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("steps", ctypes.c_int), ("failed_steps", ctypes.c_int),
                ("time", ctypes.c_double), ("dt", ctypes.c_double), ("newton_iterations", ctypes.c_long),
                ("linear_iterations", ctypes.c_long)]


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_S = ctypes.c_size_t
_SIGS = {
    "dcg_model_create": (_I, [ctypes.c_char_p, ctypes.c_char_p, _I, ctypes.POINTER(_P)]),
    "dcg_model_destroy": (None, [_P]),
    "dcg_state_size": (_I, [_P, ctypes.POINTER(_S), ctypes.POINTER(_S), ctypes.POINTER(_S)]),
    "dcg_set_state": (_I, [_P, _P, _S, _D]),
    "dcg_get_state": (_I, [_P, _P, _S, ctypes.POINTER(_D)]),
    "dcg_initial_state": (_I, [_P, _P, _S]),
    "dcg_step": (_I, [_P, _D, _D, ctypes.POINTER(StepStats)]),
    "dcg_mass": (_I, [_P, _P]),
    "dcg_last_error": (ctypes.c_char_p, []),
    "dcg_species_info": (_I, [_P, _S, ctypes.c_char_p, _S, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "dcg_compartment_info": (_I, [_P, _S, ctypes.c_char_p, _S, ctypes.POINTER(_S), ctypes.POINTER(_S),
                                  ctypes.POINTER(_I)]),
    "dcg_compartment_nodes": (_I, [_P, _S, _P]),
    "dcg_pattern_info": (_I, [_P, ctypes.POINTER(_S), ctypes.POINTER(_S), ctypes.POINTER(_I)]),
    "dcg_pattern": (_I, [_P, _P, _P]),
    "dcg_generated_source": (_I, [_P, ctypes.c_char_p, _S, ctypes.POINTER(_S)]),
    "dcg_assemble": (_I, [_P, _P, _D, _D, _D, _P, _P]),
    "dcg_spmv": (_I, [_P, _P, _P, _P]),
    "dcg_jacobian_apply": (_I, [_P, _P, _D, _D, _D, _P, _P]),
    "dcg_reduce": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(_S)]),
    "dcg_write_vtk": (_I, [_P, ctypes.c_char_p, _I]),
    "dcg_reduce_source": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _S, ctypes.POINTER(_S)]),
    "dcg_reduce_result": (_I, [_P, _S, ctypes.c_char_p, _S, ctypes.POINTER(_D), ctypes.POINTER(_I)]),
    "dcg_linear_solve": (_I, [_P, _P, _P, _P, _D, _I, ctypes.POINTER(_I)]),
    "dcg_profile_enable": (_I, [_P, _I]),
    "dcg_profile_read": (_I, [_P, _I, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "dcg_spmv_tuning": (_I, [_P, ctypes.POINTER(_I), _P, _P, ctypes.POINTER(_I)]),
    "dcg_point_tuning": (_I, [_P, ctypes.POINTER(_I), _P, _P, _P, _P]),
    "dcg_probe_copy": (_I, [_P, _S, ctypes.POINTER(_D)]),
    "dcg_model_create_ex": (_I, [ctypes.c_char_p, ctypes.c_char_p, _I, ctypes.c_uint, ctypes.POINTER(_P)]),
    "dcg_param_info": (_I, [_P, _S, ctypes.c_char_p, _S, ctypes.POINTER(_S)]),
    "dcg_set_params": (_I, [_P, _P, _S]),
    "dcg_get_params": (_I, [_P, _P, _S]),
    "dcg_device_bytes": (_I, [_P, ctypes.POINTER(_S)]),
    "dcg_estimate_device_bytes": (_I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_S)]),
    "dcg_device_memory": (_I, [_I, ctypes.POINTER(_S), ctypes.POINTER(_S)]),
    "dcg_batch_create": (_I, [ctypes.c_char_p, _I, ctypes.POINTER(_P)]),
    "dcg_batch_destroy": (None, [_P]),
    "dcg_batch_add": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_I)]),
    "dcg_batch_remove": (_I, [_P, _I]),
    "dcg_batch_ids": (_I, [_P, _P, _S, ctypes.POINTER(_S)]),
    "dcg_batch_member": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "dcg_batch_set_time_step": (_I, [_P, _I, _D]),
    "dcg_batch_step": (_I, [_P, _P, _S]),
    "dcg_batch_device_bytes": (_I, [_P, ctypes.POINTER(_S)]),
    "dcg_batch_profile_enable": (_I, [_P, _I]),
    "dcg_batch_profile_read": (_I, [_P, _I, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "dcg_evolve_defaults": (_I, [_P, ctypes.POINTER(EvolveOptions)]),
    "dcg_evolve": (_I, [_P, ctypes.POINTER(EvolveOptions), ctypes.POINTER(EvolveStats), _P, _P, _P, _S]),
    "dcg_server_create": (_I, [ctypes.c_char_p, _I, ctypes.POINTER(_P)]),
    "dcg_server_destroy": (None, [_P]),
    "dcg_server_submit": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _I, ctypes.POINTER(_I)]),
    "dcg_server_tick": (_I, [_P, ctypes.POINTER(_I)]),
    "dcg_server_result": (_I, [_P, _I, _P, _S, ctypes.POINTER(RequestInfo)]),
    "dcg_server_forget": (_I, [_P, _I]),
    "dcg_server_stats": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I),
                              ctypes.POINTER(_S), ctypes.POINTER(_S), ctypes.POINTER(_D)]),
}

RUNTIME_PARAMETERS = 1

PROFILE_KINDS = {"spmv64": 0, "spmv32": 1, "assembly_jacobian": 2, "assembly_residual": 3, "precond_setup": 4,
                 "jacobian_apply": 5, "point_setup": 6, "point_apply64": 7, "point_apply32": 8}


def load_library(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Load libdcopasi.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(make -C dune-copasi_amd/csrc)")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _check(rc: int) -> int:
    if rc < 0:
        raise DcgError(load_library().dcg_last_error().decode())
    return rc


class Model:
    """One model handle (dcg_model_t)."""

    def __init__(self, ini: str, overrides: str = "", device: int = 0, runtime_parameters: bool = False,
                 _handle=None):
        lib = load_library()
        if _handle is not None:
            h = _handle
        else:
            h = ctypes.c_void_p()
            if runtime_parameters:
                _check(lib.dcg_model_create_ex(ini.encode(), overrides.encode(), device, RUNTIME_PARAMETERS,
                                               ctypes.byref(h)))
            else:
                _check(lib.dcg_model_create(ini.encode(), overrides.encode(), device, ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self.device = device
        nd, ns, nc = _S(), _S(), _S()
        _check(lib.dcg_state_size(h, ctypes.byref(nd), ctypes.byref(ns), ctypes.byref(nc)))
        self.n_dof, self.n_species, self.n_compartments = nd.value, ns.value, nc.value
        sl, nb, bs = _S(), _S(), _I()
        _check(lib.dcg_pattern_info(h, ctypes.byref(sl), ctypes.byref(nb), ctypes.byref(bs)))
        self.n_slots, self.n_blocks, self.block_size = sl.value, nb.value, bs.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dcg_model_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- layout
    def species(self):
        out = []
        for s in range(self.n_species):
            buf = ctypes.create_string_buffer(256)
            c, l = _I(), _I()
            _check(self._lib.dcg_species_info(self._h, s, buf, 256, ctypes.byref(c), ctypes.byref(l)))
            out.append((buf.value.decode(), c.value, l.value))
        return out

    def compartments(self):
        out = []
        for c in range(self.n_compartments):
            buf = ctypes.create_string_buffer(256)
            nn, off, nsp = _S(), _S(), _I()
            _check(self._lib.dcg_compartment_info(self._h, c, buf, 256, ctypes.byref(nn), ctypes.byref(off),
                                                  ctypes.byref(nsp)))
            out.append({"name": buf.value.decode(), "n_nodes": nn.value, "dof_offset": off.value,
                        "n_species": nsp.value})
        return out

    def compartment_nodes(self, c: int) -> np.ndarray:
        n = self.compartments()[c]["n_nodes"]
        ids = np.zeros(n, dtype=np.int64)
        _check(self._lib.dcg_compartment_nodes(self._h, c, _ptr(ids)))
        return ids

    def pattern(self):
        rp = np.zeros(self.n_slots + 1, dtype=np.int32)
        col = np.zeros(self.n_blocks, dtype=np.int32)
        _check(self._lib.dcg_pattern(self._h, _ptr(rp), _ptr(col)))
        return rp, col

    def generated_source(self) -> str:
        need = _S()
        _check(self._lib.dcg_generated_source(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(self._lib.dcg_generated_source(self._h, buf, need.value, None))
        return buf.value.decode()

    # ---- state and stepping
    def initial_state(self) -> np.ndarray:
        u = np.zeros(self.n_dof)
        _check(self._lib.dcg_initial_state(self._h, _ptr(u), self.n_dof))
        return u

    def set_state(self, u: np.ndarray, time: float = 0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.size == self.n_dof
        _check(self._lib.dcg_set_state(self._h, _ptr(u), self.n_dof, time))

    def get_state(self):
        u = np.zeros(self.n_dof)
        t = _D()
        _check(self._lib.dcg_get_state(self._h, _ptr(u), self.n_dof, ctypes.byref(t)))
        return u, t.value

    def step(self, t: float, dt: float) -> StepStats:
        st = StepStats()
        _check(self._lib.dcg_step(self._h, t, dt, ctypes.byref(st)))
        return st

    def reduce(self, config: str = "", check: bool = True):
        """Transform/reduce over the current state (the model's `model.reduce`
        keys, or `config` whose top-level sections are the keys): {key: value}.
        With check, a fired error expression raises DcgError (the reference's
        ReductionError); flags() of the last call gives 0/1 (warn)/2 (error)."""
        n = _S()
        rc = _check(self._lib.dcg_reduce(self._h, config.encode() if config else None, ctypes.byref(n)))
        out, flags = {}, {}
        for i in range(n.value):
            buf = ctypes.create_string_buffer(256)
            v, f = _D(), _I()
            _check(self._lib.dcg_reduce_result(self._h, i, buf, 256, ctypes.byref(v), ctypes.byref(f)))
            out[buf.value.decode()] = v.value
            flags[buf.value.decode()] = f.value
        self.reduce_flags = flags
        if check and rc == 1:
            bad = [k for k, f in flags.items() if f == 2]
            raise DcgError(f"reduction error expression fired for {bad}: {[out[k] for k in bad]}")
        return out

    def write_vtk(self, path, append: bool = True):
        """VTK files of the current state (one .vtu per compartment, .pvd series)."""
        _check(self._lib.dcg_write_vtk(self._h, str(path).encode(), 1 if append else 0))

    def reduce_source(self, config: str = "") -> str:
        need = _S()
        cfg = config.encode() if config else None
        _check(self._lib.dcg_reduce_source(self._h, cfg, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(self._lib.dcg_reduce_source(self._h, cfg, buf, need.value, None))
        return buf.value.decode()

    def mass(self) -> np.ndarray:
        m = np.zeros(self.n_species)
        _check(self._lib.dcg_mass(self._h, _ptr(m)))
        return m

    # ---- kernel-level entry points
    def assemble(self, x: np.ndarray, t: float, mass_weight: float, stiffness_weight: float, jacobian: bool = True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(self.n_dof)
        B = self.block_size
        J = np.zeros(self.n_blocks * B * B) if jacobian else None
        _check(self._lib.dcg_assemble(self._h, _ptr(x), t, mass_weight, stiffness_weight, _ptr(r),
                                      _ptr(J) if J is not None else None))
        return r, (J.reshape(self.n_blocks, B, B) if J is not None else None)

    def jacobian_apply(self, x: np.ndarray, t: float, mass_weight: float, stiffness_weight: float,
                       v: np.ndarray) -> np.ndarray:
        """y = J(x) v without assembling J (the matrix-free Jacobian apply)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        y = np.zeros(self.n_dof)
        _check(self._lib.dcg_jacobian_apply(self._h, _ptr(x), t, mass_weight, stiffness_weight, _ptr(v), _ptr(y)))
        return y

    def spmv(self, J: np.ndarray, x: np.ndarray) -> np.ndarray:
        J = np.ascontiguousarray(J, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n_dof)
        _check(self._lib.dcg_spmv(self._h, _ptr(J), _ptr(x), _ptr(y)))
        return y

    def linear_solve(self, J: np.ndarray, b: np.ndarray, rtol: float = 1e-10, maxit: int = 1000):
        J = np.ascontiguousarray(J, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n_dof)
        its = _I()
        rc = _check(self._lib.dcg_linear_solve(self._h, _ptr(J), _ptr(b), _ptr(x), rtol, maxit, ctypes.byref(its)))
        return x, rc, its.value

    # ---- measurement
    def profile_enable(self, on: bool = True):
        _check(self._lib.dcg_profile_enable(self._h, 1 if on else 0))

    def profile(self, kind: str):
        """(launches, total milliseconds, total algorithmic bytes) of one kernel class."""
        l, ms, b = ctypes.c_long(), _D(), _D()
        _check(self._lib.dcg_profile_read(self._h, PROFILE_KINDS[kind], ctypes.byref(l), ctypes.byref(ms),
                                          ctypes.byref(b)))
        return l.value, ms.value, b.value

    def probe_copy(self, nbytes: int) -> float:
        """Milliseconds of a plain device copy of nbytes (probe_copy_k)."""
        ms = _D()
        _check(self._lib.dcg_probe_copy(self._h, nbytes, ctypes.byref(ms)))
        return ms.value

    def spmv_tuning(self):
        """{variant name: microseconds} of the set-up autotuner and the chosen variant."""
        n, ch = _I(), _I()
        var = np.zeros(16, dtype=np.int32)
        us = np.zeros(16)
        _check(self._lib.dcg_spmv_tuning(self._h, ctypes.byref(n), _ptr(var), _ptr(us), ctypes.byref(ch)))
        names = {0: "lds_entries", 1: "rows_g8", 2: "rows_g16", 3: "rows_g32", 4: "rows_g64", 5: "lds_blocks"}
        out = {}
        for v, t in zip(var[:n.value], us[:n.value]):
            out[names[int(v) % 100] + ("_fp32" if v >= 100 else "")] = float(t)
        return out, names[ch.value]

    def point_tuning(self):
        """[(compartment, (tx, ty, tz), microseconds, chosen)] of the point-apply tile autotuner;
        point_tuning_waves() gives each candidate's register bound (waves per SIMD)."""
        return [(c, t[:3], us, ch) for c, t, us, ch in self._point_tuning_raw()]

    def point_tuning_waves(self):
        """Waves per SIMD each point-apply candidate's registers are bounded for (1: none)."""
        return [t[3] for _, t, _, _ in self._point_tuning_raw()]

    def _point_tuning_raw(self):
        n = _I()
        comp = np.zeros(64, dtype=np.int32)
        tile = np.zeros(256, dtype=np.int32)
        us = np.zeros(64)
        ch = np.zeros(64, dtype=np.int32)
        _check(self._lib.dcg_point_tuning(self._h, ctypes.byref(n), _ptr(comp), _ptr(tile), _ptr(us), _ptr(ch)))
        return [(int(comp[k]), tuple(int(x) for x in tile[4 * k:4 * k + 4]), float(us[k]), bool(ch[k]))
                for k in range(n.value)]

    # ---- time loop
    def evolve_defaults(self) -> EvolveOptions:
        o = EvolveOptions()
        _check(self._lib.dcg_evolve_defaults(self._h, ctypes.byref(o)))
        return o

    def evolve(self, t_end=None, dt=None, **kw):
        """Adaptive time loop from the current state (options default to the INI keys);
        returns (EvolveStats, accepted step sizes)."""
        o = self.evolve_defaults()
        if t_end is not None:
            o.t_end = t_end
        if dt is not None:
            o.dt_initial = dt
        for k, v in kw.items():
            setattr(o, k, v)
        st = EvolveStats()
        acc = np.zeros(100000)
        rc = self._lib.dcg_evolve(self._h, ctypes.byref(o), ctypes.byref(st), None, None, _ptr(acc), acc.size)
        _check(rc)
        return st, acc[:st.steps].copy()

    # ---- run-time parameters and footprint
    def param_names(self):
        n = _S()
        _check(self._lib.dcg_param_info(self._h, 0, None, 0, ctypes.byref(n)))
        out = []
        for i in range(n.value):
            buf = ctypes.create_string_buffer(256)
            _check(self._lib.dcg_param_info(self._h, i, buf, 256, None))
            out.append(buf.value.decode())
        return out

    def params(self) -> np.ndarray:
        v = np.zeros(len(self.param_names()))
        _check(self._lib.dcg_get_params(self._h, _ptr(v), v.size))
        return v

    def set_params(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        _check(self._lib.dcg_set_params(self._h, _ptr(v), v.size))

    def device_bytes(self) -> int:
        b = _S()
        _check(self._lib.dcg_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    # ---- helpers on the layout
    def slot_dofs(self):
        """For every slot: (first packed dof, species count)."""
        first = np.zeros(self.n_slots, dtype=np.int64)
        cnt = np.zeros(self.n_slots, dtype=np.int64)
        s = 0
        for c in self.compartments():
            n, off, b = c["n_nodes"], c["dof_offset"], c["n_species"]
            first[s:s + n] = off + np.arange(n) * b
            cnt[s:s + n] = b
            s += n
        return first, cnt


def estimate_device_bytes(ini: str, overrides: str = "") -> int:
    """HBM bytes a model of this INI would hold (computed on the host)."""
    b = _S()
    _check(load_library().dcg_estimate_device_bytes(ini.encode(), overrides.encode(), ctypes.byref(b)))
    return b.value


def device_memory(device: int = 0):
    """(free, total) bytes of a device."""
    f, t = _S(), _S()
    _check(load_library().dcg_device_memory(device, ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


class Batch:
    """Many independent models advanced together (dcg_batch_t)."""

    def __init__(self, options: str = "", device: int = 0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.dcg_batch_create(options.encode(), device, ctypes.byref(h)))
        self._h = h
        self._lib = lib
        self.device = device

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dcg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add(self, ini: str, overrides: str = "") -> int:
        i = _I()
        _check(self._lib.dcg_batch_add(self._h, ini.encode(), overrides.encode(), ctypes.byref(i)))
        return i.value

    def remove(self, member_id: int):
        _check(self._lib.dcg_batch_remove(self._h, member_id))

    def ids(self):
        n = _S()
        _check(self._lib.dcg_batch_ids(self._h, None, 0, ctypes.byref(n)))
        v = np.zeros(max(1, n.value), dtype=np.int32)
        _check(self._lib.dcg_batch_ids(self._h, _ptr(v), v.size, None))
        return [int(x) for x in v[:n.value]]

    def member(self, member_id: int) -> Model:
        """A Model view of a member (state, parameters, mass); the member stays in the batch."""
        h = ctypes.c_void_p()
        _check(self._lib.dcg_batch_member(self._h, member_id, ctypes.byref(h)))
        return Model("", "", self.device, _handle=h)

    def set_time_step(self, member_id: int, dt: float):
        _check(self._lib.dcg_batch_set_time_step(self._h, member_id, dt))

    def step(self):
        """One step of every member; list of StepStats in ids() order."""
        n = len(self.ids())
        st = (StepStats * max(1, n))()
        _check(self._lib.dcg_batch_step(self._h, ctypes.cast(st, ctypes.c_void_p), max(1, n)))
        return [st[i] for i in range(n)]

    def device_bytes(self) -> int:
        b = _S()
        _check(self._lib.dcg_batch_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def profile_enable(self, on: bool = True):
        _check(self._lib.dcg_batch_profile_enable(self._h, 1 if on else 0))

    def profile(self, kind: str):
        """(launches, total ms, total algorithmic bytes) of the batched SpMV ("spmv64" or "spmv32")."""
        l, ms, b = ctypes.c_long(), _D(), _D()
        _check(self._lib.dcg_batch_profile_read(self._h, {"spmv64": 0, "spmv32": 1}[kind], ctypes.byref(l),
                                                ctypes.byref(ms), ctypes.byref(b)))
        return l.value, ms.value, b.value


class Server:
    """Request loop over one GPU with continuous batching (dcg_server_t)."""

    def __init__(self, options: str = "", device: int = 0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.dcg_server_create(options.encode(), device, ctypes.byref(h)))
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dcg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, ini: str, steps: int, overrides: str = "") -> int:
        i = _I()
        _check(self._lib.dcg_server_submit(self._h, ini.encode(), overrides.encode(), steps, ctypes.byref(i)))
        return i.value

    def tick(self) -> int:
        """Admit, advance every resident request one step, retire; returns unfinished requests."""
        o = _I()
        _check(self._lib.dcg_server_tick(self._h, ctypes.byref(o)))
        return o.value

    def info(self, rid: int) -> RequestInfo:
        ri = RequestInfo()
        rc = self._lib.dcg_server_result(self._h, rid, None, 0, ctypes.byref(ri))
        if rc < 0:
            raise DcgError(self._lib.dcg_last_error().decode())
        return ri

    def result(self, rid: int):
        """(state, info) of a finished request, or (None, info) while it runs."""
        ri = self.info(rid)
        if not ri.finished:
            return None, ri
        if ri.status < 0:
            raise DcgError(self._lib.dcg_last_error().decode())
        u = np.zeros(ri.n_dof)
        _check(self._lib.dcg_server_result(self._h, rid, _ptr(u), u.size, ctypes.byref(ri)))
        return u, ri

    def forget(self, rid: int):
        _check(self._lib.dcg_server_forget(self._h, rid))

    def stats(self):
        q, r, res, fin = _I(), _I(), _I(), _I()
        used, budget, sec = _S(), _S(), _D()
        _check(self._lib.dcg_server_stats(self._h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(res),
                                          ctypes.byref(fin), ctypes.byref(used), ctypes.byref(budget),
                                          ctypes.byref(sec)))
        return {"queued": q.value, "ready": r.value, "resident": res.value, "finished": fin.value,
                "used_bytes": used.value, "budget_bytes": budget.value, "seconds": sec.value}
