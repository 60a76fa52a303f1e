"""Transform/reduce (the reference's `model.reduce` keys) on the device against the oracle.

The oracle's reduce is pinned on the CPU (tests/test_oracle_pins.py: exact
integrals of every monomial of degree <= 4, point counts and the doubled
initial value, the keys of the reference's test INIs).  Here the device's
generated reduce kernels must give the oracle's values on the same state:
sums within 1e-12 relative (the device folds in a tree), maxima and minima
to the last bits, and the same error and warn flags.
"""
import re

import numpy as np
import pytest

from tests.oracle_lib import OracleModel

pytestmark = pytest.mark.gpu

CASES = [
    ("gs_q1", lambda wl: wl.gray_scott(n=12, cell_type="cube")),
    ("gs_p1", lambda wl: wl.gray_scott(n=12, cell_type="simplex")),
    ("ms_q1", lambda wl: wl.mitchell_schaefer(n=6, cell_type="cube")),
    ("ms_p1", lambda wl: wl.mitchell_schaefer(n=5, cell_type="simplex")),
    ("two_q1", lambda wl: wl.two_compartments(n=20, cell_type="cube")),
    ("req_3d_8sp_2c", lambda wl: wl.random_request(4, side3=(5, 6), steps_range=(3, 3))),
]


def _config(species, dim):
    ax = "xyz"[:dim]
    lines = []
    for s in species:
        lines += [f"[mass_{s}]", f"evaluation.expression = {s}*integration_factor"]
        grad = " + ".join(f"grad_{s}_{a}^2" for a in ax)
        lines += [f"[h1_{s}]", f"evaluation.expression = ({grad})*integration_factor",
                  "transformation.expression = arg: sqrt(arg)"]
    s0 = species[0]
    lines += ["[max0]", f"evaluation.expression = {s0}", "reduction.expression = init, val: max(init, val)",
              "initial.value = -1e100", "[min0]", f"evaluation.expression = {s0}",
              "reduction.expression = a, b: min(a, b)", "initial.value = 1e100",
              "[moment]", f"evaluation.expression = position_x*{s0}*time*integration_factor",
              "[count]", "evaluation.expression = 1", "initial.value = 3",
              "[all]", "evaluation.expression = " + " + ".join(species)]
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_reduce_matches_oracle(dca, wl, name, make):
    print(f"case: reduce {name}", flush=True)
    w = make(wl)
    gpu = dca.Model(w.ini, "", 0)
    orc = OracleModel(w.ini)
    u = w.initial_state(gpu) + 0.05 * np.random.default_rng(3).standard_normal(gpu.n_dof)
    gpu.set_state(u, 0.7)
    orc.set_state(u, 0.7)
    species = [s[0] for s in gpu.species()]
    dim = int(re.search(r"dimension\s*=\s*(\d)", w.ini).group(1))
    cfg = _config(species, dim)
    got = gpu.reduce(cfg)
    ref, flags, rc = orc.reduce(cfg)
    assert rc == 0 and list(got) == list(ref)
    for k, v in ref.items():
        if k.startswith(("max", "min")):
            assert abs(got[k] - v) <= 1e-14 * abs(v), (k, got[k], v)
        else:
            assert abs(got[k] - v) <= 1e-12 * max(1.0, abs(v)), (k, got[k], v)


def test_reduce_flags_and_model_keys(dca, wl):
    w = wl.gray_scott(n=10, cell_type="cube")
    ini = w.ini + ("\n[model.reduce]\nu_max.evaluation.expression = u\n"
                   "u_max.reduction.expression = init, val: max(init, val)\nu_max.initial.value = -1e100\n"
                   "u_max.error.expression = arg: arg > 0.5\n"
                   "v_min.evaluation.expression = v\nv_min.reduction.expression = a, b: min(a, b)\n"
                   "v_min.initial.value = 1e100\nv_min.warn.expression = arg: arg < 1\n")
    gpu = dca.Model(ini, "", 0)
    orc = OracleModel(ini)
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    orc.set_state(u0, 0.0)
    with pytest.raises(dca.DcgError, match="u_max"):
        gpu.reduce()
    got = gpu.reduce(check=False)
    ref, flags, rc = orc.reduce()
    assert rc == 1 and gpu.reduce_flags == flags == {"u_max": 2, "v_min": 1}
    assert got == pytest.approx(ref, rel=1e-14)


def test_reduce_reads_run_time_parameters(dca, wl):
    """Constants of parser_context stay run-time parameters: the reduce kernel
    reads them like the assembly does, and set_params changes the result."""
    w = wl.gray_scott(n=10, cell_type="cube", constants=True)
    gpu = dca.Model(w.ini, "", 0, runtime_parameters=True)
    u0 = w.initial_state(gpu)
    gpu.set_state(u0, 0.0)
    names = gpu.param_names()
    p = gpu.params()
    cfg = f"[scaled]\nevaluation.expression = {names[0]}*u*integration_factor\n"
    a = gpu.reduce(cfg)["scaled"]
    orc = OracleModel(w.ini)
    orc.set_state(u0, 0.0)
    assert a == pytest.approx(orc.reduce(cfg)[0]["scaled"], rel=1e-12)
    p2 = p.copy()
    p2[0] *= 3.0
    gpu.set_params(p2)
    assert gpu.reduce(cfg)["scaled"] == pytest.approx(3.0 * a, rel=1e-12)
