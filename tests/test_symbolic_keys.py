"""Pins of the hand-written partial derivatives in tests/symbolic_cases.py.

The GPU tests compare the device's derived Jacobians (model.jacobian.type =
symbolic) with the oracle's Jacobian of the same model given these partials
as explicit jacobian keys, so the partials themselves are checked here
against central differences of their expressions, evaluated by the oracle's
expression evaluator (itself pinned against Python in test_oracle_pins.py).
"""
import numpy as np
import pytest

from tests import symbolic_cases as sc
from tests.test_oracle_pins import _eval

CASES = ([("reaction", s, e, p) for s, (e, p) in sc.REACTIONS.items()] +
         [("storage", s, e, p) for s, (e, p) in sc.STORAGE.items()] +
         [("diffusion", s, e, p) for s, (e, p) in sc.DIFFUSION.items()])
OUT_CASES = [("outflow", s, e, p) for s, (e, p) in sc.OUTFLOW.items()]


def _points(rng, k=40):
    pts = []
    while len(pts) < k:
        u, v = rng.uniform(0.5, 1.5), rng.uniform(0.2, 0.8)
        # away from the kinks of min/max/abs/conditionals/floor/fmod
        kinks = [u - v, u - 0.6, v - 0.6, u - 1.05, u * u - v, 4 * u - round(4 * u),
                 (u * v) / 0.3 - round((u * v) / 0.3)]
        if min(abs(x) for x in kinks) > 2e-3:
            pts.append((u, v))
    return pts


def _fd(expr, names, point, i, h=1e-6):
    xp, xm = list(point), list(point)
    xp[i] += h
    xm[i] -= h
    return (_eval(expr, names, xp)[0] - _eval(expr, names, xm)[0]) / (2 * h)


@pytest.mark.parametrize("kind,species,expr,partials", CASES, ids=[f"{c[0]}_{c[1]}" for c in CASES])
def test_hand_written_partials_match_central_differences(kind, species, expr, partials):
    rng = np.random.default_rng(7)
    names = ("u", "v")
    for pt in _points(rng):
        for w, dexpr in partials.items():
            i = names.index(w)
            ref = _fd(expr, names, pt, i)
            val, rc = _eval(dexpr, names, pt)
            assert rc == 0
            assert val == pytest.approx(ref, rel=2e-6, abs=2e-7), (kind, species, w, pt)
        for w in set(names) - set(partials):
            assert abs(_fd(expr, names, pt, names.index(w))) < 1e-7  # no dependence


@pytest.mark.parametrize("kind,species,expr,partials", OUT_CASES, ids=[f"{c[0]}_{c[1]}" for c in OUT_CASES])
def test_hand_written_outflow_partials(kind, species, expr, partials):
    rng = np.random.default_rng(8)
    names = ("D", "A")
    for _ in range(30):
        pt = (rng.uniform(0.1, 2.0), rng.uniform(0.1, 2.0))
        for w, dexpr in partials.items():
            i = names.index(w)
            ref = _fd(expr, names, pt, i)
            val, rc = _eval(dexpr, names, pt)
            assert val == pytest.approx(ref, rel=2e-6, abs=2e-7), (species, w, pt)


def test_symbolic_models_build_on_the_host(dca):
    """The symbolic forms of every case build on the host (no device) from
    INI text that holds no jacobian.<species> key."""
    for ini in (sc.rules_ini("cube"), sc.rules_ini("simplex", with_diffusion=True), sc.transmission_ini("cube")):
        m = dca.Model(ini, "", -1)
        assert m.n_dof > 0
        assert ".jacobian." not in ini
        m.close()
