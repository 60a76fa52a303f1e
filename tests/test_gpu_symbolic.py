"""Jacobians derived by the expression compiler (model.jacobian.type =
symbolic, no jacobian.<species> keys) on the device, against the oracle's
Jacobian of the same model given the partials as explicit keys (the
reference's user-supplied jacobian keys, local_equations.hh:551-579).  The
hand-written partials are pinned by tests/test_symbolic_keys.py; the cases
cover every derivative rule (powers with variable exponents, min/max, abs,
atan2, conditionals, fmod, the elementary functions), state-dependent
storage and diffusion, nonlinear transmission, and random mass-action
requests of the serving stream."""
import numpy as np
import pytest

from tests import symbolic_cases as sc
from tests.oracle_lib import OracleModel, bcsr_to_dense_csr, rel_l2_per_species

pytestmark = pytest.mark.gpu


def _uv_state(model, seed, u_range=(0.6, 1.4), v_range=(0.25, 0.75)):
    rng = np.random.default_rng(seed)
    u = np.empty(model.n_dof)
    u[0::2] = rng.uniform(*u_range, model.n_dof // 2)
    u[1::2] = rng.uniform(*v_range, model.n_dof // 2)
    return u


def _compare(dca, sym_ini, keys_ini, x, t=0.3, cm=1.0, ca=0.7, tol=1e-12):
    gpu = dca.Model(sym_ini, "", 0)
    orc = OracleModel(keys_ini)
    assert gpu.n_dof == orc.n_dof
    r_g, J_g = gpu.assemble(x, t, cm, ca)
    r_o, v_o = orc.assemble(x, t, cm, ca)
    assert np.abs(r_g - r_o).max() <= tol * np.abs(r_o).max()
    A_g = bcsr_to_dense_csr(gpu, J_g).toarray()
    A_o = orc.csr(v_o).toarray()
    assert np.abs(A_g - A_o).max() <= tol * np.abs(A_o).max()
    return gpu, orc


@pytest.mark.parametrize("cell_type", ["cube", "simplex"])
@pytest.mark.parametrize("with_diffusion", [False, True])
def test_symbolic_rules_match_explicit_keys(dca, cell_type, with_diffusion):
    sym = sc.rules_ini(cell_type, symbolic=True, with_diffusion=with_diffusion)
    keys = sc.rules_ini(cell_type, symbolic=False, with_diffusion=with_diffusion)
    probe = dca.Model(sym, "", -1)
    x = _uv_state(probe, 3)
    _compare(dca, sym, keys, x)
    # time steps of the model without its jumps (floor, fmod), from a state
    # on one side of its conditional (u > v), so Newton meets a smooth residual
    sym = sc.rules_ini(cell_type, symbolic=True, with_diffusion=with_diffusion, smooth=True)
    keys = sc.rules_ini(cell_type, symbolic=False, with_diffusion=with_diffusion, smooth=True)
    x = _uv_state(probe, 3, (0.9, 1.4), (0.25, 0.6))
    gpu, orc = _compare(dca, sym, keys, x)
    gpu.set_state(x, 0.0)
    orc.set_state(x, 0.0)
    t = 0.0
    for k in range(3):
        assert gpu.step(t, 0.01).status == 0
        assert orc.step(0.01)[0] == 0
        t += 0.01
        assert max(rel_l2_per_species(gpu.compartments(), gpu.get_state()[0], orc.get_state()[0])) <= 1e-8, k
    probe.close()


@pytest.mark.parametrize("cell_type", ["cube", "simplex"])
def test_symbolic_transmission_matches_explicit_keys(dca, cell_type):
    sym = sc.transmission_ini(cell_type, symbolic=True)
    keys = sc.transmission_ini(cell_type, symbolic=False)
    probe = dca.Model(sym, "", -1)
    x = np.random.default_rng(4).uniform(0.2, 1.5, probe.n_dof)
    probe.close()
    _compare(dca, sym, keys, x)


@pytest.mark.parametrize("seed", [4, 11, 17, 23])
def test_symbolic_random_requests_match_explicit_keys(dca, wl, seed):
    w = wl.random_request(seed, side2=(10, 12), side3=(4, 5), steps_range=(3, 3), precision="fp64")
    sym = sc.strip_jacobians(w.ini)
    assert ".jacobian." not in sym
    probe = dca.Model(w.ini, "", -1)
    x = probe.initial_state() + 0.05 * np.random.default_rng(seed).standard_normal(probe.n_dof)
    probe.close()
    _compare(dca, sym, w.ini, x)
