"""Expressions that exercise every derivative rule of the expression compiler
(model.jacobian.type = symbolic), with their partial derivatives written out
by hand.  tests/test_symbolic_keys.py pins the hand-written partials against
central differences (through the oracle's own evaluator) and sympy; the GPU
tests hand them to the oracle as explicit jacobian keys and compare its
Jacobian with the device's derived one.  States are kept away from kinks:
u in [0.5, 1.5], v in [0.2, 0.8], so min/max/abs/conditionals pick one branch
in a neighbourhood of every point."""

# name -> (expression, {species: partial})
REACTIONS = {
    "u": ("u^v + min(u, v, 0.6) + abs(u - 1.05)*atan2(v, u) - exp(-u*v)/(1 + v) + (u*v) % 0.3",
          {"u": "v*u^(v - 1) + ((u <= v) && (u <= 0.6) ? 1 : 0) + abs(u - 1.05)/(u - 1.05)*atan2(v, u)"
                " - abs(u - 1.05)*v/(u^2 + v^2) + v*exp(-u*v)/(1 + v) + v",
           "v": "u^v*log(u) + ((v < u) && (v <= 0.6) ? 1 : 0) + abs(u - 1.05)*u/(u^2 + v^2)"
                " + exp(-u*v)*(u*(1 + v) + 1)/(1 + v)^2 + u"}),
    "v": ("(u > v ? sqrt(u*v) : log(u + v)) + max(u^2, v)*tanh(v) + sin(u)*cos(v) + pow(v, 1.5)/u"
          " - 0.1*tan(0.3*u) + asin(0.5*v) - acos(0.4*u*v) + atan(u - v) + sinh(0.2*u)*cosh(0.3*v)"
          " + log10(u + 2) - log2(v + 1) + 3*floor(4*u)",
          {"u": "(u > v ? v/(2*sqrt(u*v)) : 1/(u + v)) + (u^2 >= v ? 2*u : 0)*tanh(v) + cos(u)*cos(v)"
                " - pow(v, 1.5)/u^2 - 0.03*(1 + tan(0.3*u)^2) + 0.4*v/sqrt(1 - (0.4*u*v)^2)"
                " + 1/(1 + (u - v)^2) + 0.2*cosh(0.2*u)*cosh(0.3*v) + 1/((u + 2)*log(10))",
           "v": "(u > v ? u/(2*sqrt(u*v)) : 1/(u + v)) + (u^2 >= v ? 0 : 1)*tanh(v)"
                " + max(u^2, v)*(1 - tanh(v)^2) - sin(u)*sin(v) + 1.5*sqrt(v)/u + 0.5/sqrt(1 - 0.25*v^2)"
                " + 0.4*u/sqrt(1 - (0.4*u*v)^2) - 1/(1 + (u - v)^2) + 0.3*sinh(0.2*u)*sinh(0.3*v)"
                " - 1/((v + 1)*log(2))"}),
}

STORAGE = {"u": ("1 + 0.1*u^2", {"u": "0.2*u"})}

DIFFUSION = {"u": ("1e-3*(1 + u^2 + 0.5*v)", {"u": "2e-3*u", "v": "0.5e-3"})}

# transmission between compartments (inside species D, outside species A)
OUTFLOW = {"D": ("0.5*(D^2 - A) + 0.1*D*exp(-A)", {"D": "D + 0.1*exp(-A)", "A": "-0.5 - 0.1*D*exp(-A)"}),
           "A": ("-0.5*(D^2 - A) - 0.1*D*exp(-A)", {"D": "-D - 0.1*exp(-A)", "A": "0.5 + 0.1*D*exp(-A)"})}


# the discontinuous terms (floor, fmod) and their partials: dropped for time
# steps, where Newton needs a residual without jumps
_JUMPS = {"u": (" + (u*v) % 0.3", {"u": " + v", "v": " + u"}), "v": (" + 3*floor(4*u)", {})}


def _reaction(s, smooth):
    r, rj = REACTIONS[s]
    if not smooth:
        return r, rj
    term, parts = _JUMPS[s]
    assert r.endswith(term)
    rj = {w: (e[: -len(parts[w])] if w in parts else e) for w, e in rj.items()}
    for w, tail in parts.items():
        assert REACTIONS[s][1][w].endswith(tail)
    return r[: -len(term)], rj


def rules_ini(cell_type="cube", n=6, symbolic=True, with_diffusion=False, smooth=False):
    """Two species u, v with REACTIONS (and STORAGE on u); with_diffusion adds
    the state-dependent DIFFUSION of u.  symbolic: no jacobian keys and
    model.jacobian.type = symbolic; else the hand-written keys.  smooth drops
    the floor and fmod terms (for time steps)."""
    lines = [f"[grid]\ndimension = 2\ncells = {n} {n}\ncell_type = {cell_type}\n[compartments]\ndomain.expression = 1\n"]
    if symbolic:
        lines.append("[model]\njacobian.type = symbolic\n")
    for s in ("u", "v"):
        sec = [f"[model.scalar_field.{s}]", "compartment = domain"]
        st, stj = STORAGE.get(s, ("1", {}))
        sec.append(f"storage.expression = {st}")
        if not symbolic:
            sec += [f"storage.jacobian.{w}.expression = {e}" for w, e in stj.items()]
        if with_diffusion and s in DIFFUSION:
            d, dj = DIFFUSION[s]
            sec.append(f"cross_diffusion.{s}.expression = {d}")
            if not symbolic:
                sec += [f"cross_diffusion.{s}.jacobian.{w}.expression = {e}" for w, e in dj.items()]
        else:
            sec.append(f"cross_diffusion.{s}.expression = {'2e-3' if s == 'u' else '1e-3'}")
        r, rj = _reaction(s, smooth)
        sec.append(f"reaction.expression = {r}")
        if not symbolic:
            sec += [f"reaction.jacobian.{w}.expression = {e}" for w, e in rj.items()]
        lines.append("\n".join(sec) + "\n")
    lines.append(_SOLVER)
    return "".join(lines)


def transmission_ini(cell_type="cube", n=10, symbolic=True):
    lines = [f"[grid]\ndimension = 2\ncells = {n} {n}\ncell_type = {cell_type}\n[compartments]\n"
             "inner.expression = (position_x-0.5)^2 + (position_y-0.5)^2 < 0.09\n"
             "outer.expression = (position_x-0.5)^2 + (position_y-0.5)^2 >= 0.09\n"]
    if symbolic:
        lines.append("[model]\njacobian.type = symbolic\n")
    for s, comp, other in (("D", "inner", "outer"), ("A", "outer", "inner")):
        e, ej = OUTFLOW[s]
        sec = [f"[model.scalar_field.{s}]", f"compartment = {comp}", "storage.expression = 1",
               f"cross_diffusion.{s}.expression = 1e-3", f"outflow.{other}.expression = {e}"]
        if not symbolic:
            sec += [f"outflow.{other}.jacobian.{w}.expression = {x}" for w, x in ej.items()]
        lines.append("\n".join(sec) + "\n")
    lines.append(_SOLVER)
    return "".join(lines)


def strip_jacobians(ini):
    """The INI without any jacobian.* key, computing its Jacobian symbolically."""
    keep = [l for l in ini.splitlines() if ".jacobian." not in l]
    return "[model]\njacobian.type = symbolic\n" + "\n".join(keep) + "\n"


_SOLVER = """[model.time_step_operator]
type = ImplicitEuler
linear_solver.preconditioner.type = Jacobi
linear_solver.convergence_condition.relative_tolerance = 1e-12
linear_solver.convergence_condition.iteration_range = 1 3000
nonlinear_solver.convergence_condition.relative_tolerance = 1e-20
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-26
"""
