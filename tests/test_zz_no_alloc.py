"""dcg_linear_solve_noalloc: the linear solve with no device allocation
(work space in the library's fixed per-device array), through the C ABI,
against dcg_linear_solve on the same inputs.  Two calls in a row share the
array, the second on a smaller model; a model whose work space exceeds the
array and a mixed-precision model are rejected with an error.  The caller's
device buffers are allocated here with hipMalloc from the HIP runtime the
library itself uses."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")  # the runtime already loaded with libdcopasi.so
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def to_device(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(8, a.nbytes)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, n):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(8, 8 * n)) == 0
        return p

    def to_host(self, p, n):
        out = np.empty(n)
        assert self.L.hipMemcpy(out.ctypes.data, p, 8 * n, D2H) == 0
        return out

    def free(self, *ps):
        for p in ps:
            self.L.hipFree(p)


def _solve_noalloc(dca, hip, model, J, b, rtol, maxit):
    lib = dca.load_library()
    fn = lib.dcg_linear_solve_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_int)]
    dJ, db, dx = hip.to_device(J), hip.to_device(b), hip.empty(model.n_dof)
    assert hip.L.hipDeviceSynchronize() == 0
    its = ctypes.c_int()
    rc = fn(model._h, dJ, db, dx, rtol, maxit, ctypes.byref(its))
    err = lib.dcg_last_error().decode() if rc < 0 else ""
    x = hip.to_host(dx, model.n_dof) if rc >= 0 else None
    hip.free(dJ, db, dx)
    return rc, x, its.value, err


def _case(wl, n, precond):
    w = wl.gray_scott(n=n, cell_type="cube", precond=precond)
    return w


def test_noalloc_solve_matches_linear_solve(dca, wl):
    hip = _Hip()
    rng = np.random.default_rng(11)
    # larger model first, then a smaller one (the array keeps the first call's
    # contents), then the larger one again with the other preconditioner
    for n, precond in [(64, "Jacobi"), (20, "ILU"), (48, "ILU"), (9, "Jacobi")]:
        print(f"case: no-alloc solve gray_scott n={n} {precond}", flush=True)
        w = _case(wl, n, precond)
        m = dca.Model(w.ini, "model.time_step_operator.linear_solver.storage = matrix\n", 0)
        x0 = w.initial_state(m) + 0.05 * rng.standard_normal(m.n_dof)
        _, J = m.assemble(x0, 0.0, 1.0, 1.0)
        b = rng.standard_normal(m.n_dof)
        x_ref, rc_ref, its_ref = m.linear_solve(J, b, rtol=1e-12, maxit=3000)
        assert rc_ref == 0
        rc, x, its, err = _solve_noalloc(dca, hip, m, J, b, 1e-12, 3000)
        assert rc == 0, err
        assert its == its_ref
        assert np.abs(x - x_ref).max() <= 1e-12 * np.abs(x_ref).max()
        m.close()


def test_noalloc_solve_rejects_what_it_cannot_hold(dca, wl):
    hip = _Hip()
    # 1024^2 Gray-Scott: about 2.1 M DOF, 11 work vectors alone exceed 64 MiB
    big = dca.Model(wl.gray_scott(n=1024, cell_type="cube", precond="Jacobi").ini, "", 0)
    J = np.zeros(big.n_blocks * big.block_size ** 2)
    rc, _, _, err = _solve_noalloc(dca, hip, big, J, np.zeros(big.n_dof), 1e-8, 10)
    assert rc < 0 and "work space" in err and "67108864" in err
    big.close()
    mixed = dca.Model(wl.gray_scott(n=12, cell_type="cube", precond="Jacobi", precision="mixed").ini,
                      "model.time_step_operator.linear_solver.storage = matrix\n", 0)
    J = np.zeros(mixed.n_blocks * mixed.block_size ** 2)
    rc, _, _, err = _solve_noalloc(dca, hip, mixed, J, np.zeros(mixed.n_dof), 1e-8, 10)
    assert rc < 0 and "fp64" in err
    mixed.close()
