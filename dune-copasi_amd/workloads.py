"""Benchmark and parity workloads as reference-style INI text plus seeded
nodal initial states.

Each workload is the model of one benchmark configuration written with the
reference's own keys (model.scalar_field.*, compartments.*, grid.*,
model.time_step_operator.*).  Two keys are additions of this build:
`grid.cell_type = cube` (Q1 cells; the reference builds simplices) and
`model.time_step_operator.linear_solver.precision = mixed`.

Random initial data come from a counter-based SplitMix64 stream keyed by
(seed, stream, global vertex id), so the same vertex gets the same value on
any layout, device or process.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

_M64 = (1 << 64) - 1


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser applied to uint64 counters (Steele et al. 2014)."""
    z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, stream: int, ids: np.ndarray) -> np.ndarray:
    key = (np.uint64((seed * 0x100000001B3) & _M64) ^ np.uint64((stream << 40) & _M64))
    z = splitmix64(np.asarray(ids, dtype=np.uint64) ^ key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass
class Workload:
    name: str
    ini: str
    dt: float
    steps: int
    dim: int
    cells: tuple
    init: callable  # (vertex coordinates per compartment, vertex ids) -> packed state
    stages: int = 1

    def initial_state(self, model) -> np.ndarray:
        """Packed initial state for a model handle (product or oracle wrapper)."""
        return self.init(model)


def vertex_coords(ids: np.ndarray, cells, ext=None, origin=None) -> np.ndarray:
    d = len(cells)
    ext = ext or (1.0,) * d
    origin = origin or (0.0,) * d
    ids = np.asarray(ids, dtype=np.int64)
    out = np.zeros((ids.size, 3))
    rem = ids.copy()
    for k in range(d):
        n = cells[k] + 1
        idx = rem % n
        rem //= n
        out[:, k] = origin[k] + ext[k] * idx / cells[k]
    return out


def _solver_block(rk, precond, nl_rtol, nl_atol, lin_rtol, precision="fp64", maxit=2000, dt=None):
    pre = ("linear_solver.preconditioner.type = BlockJacobi\n"
           "linear_solver.preconditioner.diagonal_solver.type = Loop\n"
           "linear_solver.preconditioner.diagonal_solver.preconditioner.type = ILU\n"
           "linear_solver.preconditioner.diagonal_solver.convergence_condition.iteration_range = 1 1\n"
           if precond == "ILU" else f"linear_solver.preconditioner.type = {precond}\n")
    return (f"[model.time_step_operator]\n"
            f"type = {rk}\n"
            + (f"time_step_initial = {dt!r}\n" if dt is not None else "") +
            f"linear_solver.type = BiCGSTAB\n"
            f"{pre}"
            f"linear_solver.convergence_condition.relative_tolerance = {lin_rtol:g}\n"
            f"linear_solver.convergence_condition.iteration_range = 1 {maxit}\n"
            f"linear_solver.precision = {precision}\n"
            f"nonlinear_solver.convergence_condition.relative_tolerance = {nl_rtol:g}\n"
            f"nonlinear_solver.convergence_condition.absolute_tolerance = {nl_atol:g}\n"
            f"nonlinear_solver.convergence_condition.iteration_range = 0 40\n")


# --------------------------------------------------------------------- Gray-Scott
def gray_scott(n=128, cell_type="cube", seed=0, F=0.04, k=0.06, Du=2e-5, Dv=1e-5, precond="ILU",
               nl_rtol=1e-10, nl_atol=0.0, lin_rtol=1e-10, precision="fp64", steps=200, dt=1.0,
               constants=False) -> Workload:
    """Gray-Scott on [0,1]^2, zero flux, implicit Euler (the ILU0 benchmark case at n = 128).

    constants=True writes F, k, Du, Dv as parser_context constants (the form a
    parameter sweep uses); the model is the same."""
    if constants:
        ctx = (f"[parser_context]\nF.type = constant\nF.value = {F!r}\nk.type = constant\nk.value = {k!r}\n"
               f"Du.type = constant\nDu.value = {Du!r}\nDv.type = constant\nDv.value = {Dv!r}\n\n")
        sF, sk, sDu, sDv = "F", "k", "Du", "Dv"
    else:
        ctx = ""
        sF, sk, sDu, sDv = f"{F:g}", f"{k:g}", f"{Du:g}", f"{Dv:g}"
    ini = f"""[grid]
dimension = 2
cells = {n} {n}
extensions = 1 1
cell_type = {cell_type}

{ctx}[compartments]
domain.expression = 1

[model.scalar_field.u]
compartment = domain
storage.expression = 1
cross_diffusion.u.expression = {sDu}
reaction.expression = -u*v^2 + {sF}*(1-u)
reaction.jacobian.u.expression = -v^2 - {sF}
reaction.jacobian.v.expression = -2*u*v

[model.scalar_field.v]
compartment = domain
storage.expression = 1
cross_diffusion.v.expression = {sDv}
reaction.expression = u*v^2 - ({sF}+{sk})*v
reaction.jacobian.u.expression = v^2
reaction.jacobian.v.expression = 2*u*v - ({sF}+{sk})

""" + _solver_block("ImplicitEuler", precond, nl_rtol, nl_atol, lin_rtol, precision, dt=dt)

    def init(model):
        ids = model.compartment_nodes(0)
        x = vertex_coords(ids, (n, n))
        inside = (x[:, 0] >= 0.4) & (x[:, 0] <= 0.6) & (x[:, 1] >= 0.4) & (x[:, 1] <= 0.6)
        u = np.where(inside, 0.5, 1.0) + 0.02 * (uniform01(seed, 0, ids) - 0.5)
        v = np.where(inside, 0.25, 0.0) + 0.02 * (uniform01(seed, 1, ids) - 0.5)
        out = np.empty(2 * ids.size)
        out[0::2] = u
        out[1::2] = v
        return out

    return Workload(f"gray_scott_{cell_type}_{n}", ini, dt, steps, 2, (n, n), init)


def gray_scott_batch(count=256, seed=0, side_min=32, side_max=512, precision="mixed", precond="Jacobi",
                     nl_rtol=1e-10, nl_atol=0.0, lin_rtol=1e-10, steps=100, dt=1.0, cell_type="cube"):
    """Independent Gray-Scott instances (the batch benchmark at count = 256):
    sides log-uniform in [side_min, side_max], F ~ U[0.01, 0.08],
    k ~ U[0.04, 0.07], D_u / D_v ~ U[1.5, 3] with D_v = 1e-5, initial state as
    in gray_scott seeded per instance, implicit Euler."""
    ids = np.arange(count)
    side = np.exp(np.log(side_min) + (np.log(side_max) - np.log(side_min)) * uniform01(seed, 200, ids))
    side = np.clip(np.rint(side).astype(int), side_min, side_max)
    F = 0.01 + 0.07 * uniform01(seed, 201, ids)
    k = 0.04 + 0.03 * uniform01(seed, 202, ids)
    ratio = 1.5 + 1.5 * uniform01(seed, 203, ids)
    out = []
    for i in range(count):
        out.append(gray_scott(n=int(side[i]), cell_type=cell_type, seed=seed * 100003 + i, F=float(F[i]),
                              k=float(k[i]), Du=float(ratio[i]) * 1e-5, Dv=1e-5, precond=precond, nl_rtol=nl_rtol,
                              nl_atol=nl_atol, lin_rtol=lin_rtol, precision=precision, steps=steps, dt=dt, constants=True))
    return out


# ------------------------------------------------------------- Mitchell-Schaefer
def mitchell_schaefer(n=96, cell_type="cube", seed=0, D=1e-3, tau_in=0.3, tau_out=6.0, tau_open=120.0,
                      tau_close=150.0, u_gate=0.13, precond="Jacobi", nl_rtol=1e-12, nl_atol=0.0,
                      lin_rtol=1e-10, precision="fp64", steps=500, dt=0.1) -> Workload:
    """Two-species excitable model on [0,1]^3, Alexander2 (the headline benchmark at n = 96)."""
    ini = f"""[grid]
dimension = 3
cells = {n} {n} {n}
extensions = 1 1 1
cell_type = {cell_type}

[parser_context]
tau_in.type = constant
tau_in.value = {tau_in:g}
tau_out.type = constant
tau_out.value = {tau_out:g}
tau_open.type = constant
tau_open.value = {tau_open:g}
tau_close.type = constant
tau_close.value = {tau_close:g}
u_gate.type = constant
u_gate.value = {u_gate:g}

[compartments]
domain.expression = 1

[model.scalar_field.u]
compartment = domain
storage.expression = 1
cross_diffusion.u.expression = {D:g}
reaction.expression = h*u^2*(1-u)/tau_in - u/tau_out
reaction.jacobian.u.expression = h*(2*u - 3*u^2)/tau_in - 1/tau_out
reaction.jacobian.h.expression = u^2*(1-u)/tau_in

[model.scalar_field.h]
compartment = domain
storage.expression = 1
reaction.expression = (u < u_gate) ? (1-h)/tau_open : -h/tau_close
reaction.jacobian.h.expression = (u < u_gate) ? -1/tau_open : -1/tau_close

""" + _solver_block("Alexander2", precond, nl_rtol, nl_atol, lin_rtol, precision, dt=dt)

    rng = np.random.Generator(np.random.PCG64(seed))
    centre = rng.uniform(0.1, 0.9, size=3)

    def init(model):
        ids = model.compartment_nodes(0)
        x = vertex_coords(ids, (n, n, n))
        r = np.linalg.norm(x - centre[None, :], axis=1)
        out = np.empty(2 * ids.size)
        out[0::2] = np.where(r <= 0.1, 1.0, 0.0)
        out[1::2] = 1.0
        return out

    return Workload(f"mitchell_schaefer_{cell_type}_{n}", ini, dt, steps, 3, (n, n, n), init, stages=2)


# ------------------------------------------------------------ two compartments
def two_compartments(n=1024, cell_type="cube", seed=0, kt=0.5, precond="Jacobi", nl_rtol=1e-12, nl_atol=0.0,
                     lin_rtol=1e-10, precision="fp64", steps=400, dt=0.05) -> Workload:
    """Inner disk r < 0.3 and its complement, 4 species each, A+B<->C, C->D inside,
    D->A outside, transmission k_t (D_in - A_out) (the two-compartment benchmark at n = 1024)."""
    ds = np.exp(np.log(1e-4) + (np.log(1e-2) - np.log(1e-4)) * uniform01(seed, 99, np.arange(8)))
    names_in = ["A_in", "B_in", "C_in", "D_in"]
    names_out = ["A_out", "B_out", "C_out", "D_out"]
    R = {
        "A_in": ("-(1*A_in*B_in - 0.1*C_in)", {"A_in": "-B_in", "B_in": "-A_in", "C_in": "0.1"}),
        "B_in": ("-(1*A_in*B_in - 0.1*C_in)", {"A_in": "-B_in", "B_in": "-A_in", "C_in": "0.1"}),
        "C_in": ("(1*A_in*B_in - 0.1*C_in) - 0.05*C_in", {"A_in": "B_in", "B_in": "A_in", "C_in": "-0.15"}),
        "D_in": ("0.05*C_in", {"C_in": "0.05"}),
        "A_out": ("0.02*D_out", {"D_out": "0.02"}),
        "B_out": (None, {}),
        "C_out": (None, {}),
        "D_out": ("-0.02*D_out", {"D_out": "-0.02"}),
    }
    parts = [f"""[grid]
dimension = 2
cells = {n} {n}
extensions = 1 1
cell_type = {cell_type}

[compartments]
inner.expression = (position_x-0.5)^2 + (position_y-0.5)^2 < 0.09
outer.expression = (position_x-0.5)^2 + (position_y-0.5)^2 >= 0.09
"""]
    for i, s in enumerate(names_in + names_out):
        comp = "inner" if s.endswith("_in") else "outer"
        sec = [f"\n[model.scalar_field.{s}]", f"compartment = {comp}", "storage.expression = 1",
               f"cross_diffusion.{s}.expression = {ds[i]:.17g}"]
        expr, jac = R[s]
        if expr:
            sec.append(f"reaction.expression = {expr}")
            for w, e in jac.items():
                sec.append(f"reaction.jacobian.{w}.expression = {e}")
        if s == "D_in":
            sec += [f"outflow.outer.expression = {kt:g}*(D_in - A_out)",
                    f"outflow.outer.jacobian.D_in.expression = {kt:g}",
                    f"outflow.outer.jacobian.A_out.expression = {-kt:g}"]
        if s == "A_out":
            sec += [f"outflow.inner.expression = -{kt:g}*(D_in - A_out)",
                    f"outflow.inner.jacobian.D_in.expression = {-kt:g}",
                    f"outflow.inner.jacobian.A_out.expression = {kt:g}"]
        parts.append("\n".join(sec) + "\n")
    parts.append("\n" + _solver_block("ImplicitEuler", precond, nl_rtol, nl_atol, lin_rtol, precision, dt=dt))
    ini = "".join(parts)

    def init(model):
        out = np.empty(model.n_dof)
        for c, comp in enumerate(model.compartments()):
            ids = model.compartment_nodes(c)
            b, off = comp["n_species"], comp["dof_offset"]
            for i in range(b):
                out[off + i: off + b * ids.size: b] = uniform01(seed, 10 + 4 * c + i, ids)
        return out

    return Workload(f"two_compartments_{cell_type}_{n}", ini, dt, steps, 2, (n, n), init)


# ------------------------------------------------------------- serving requests
def random_request(seed: int, precision="mixed", nl_rtol=1e-20, nl_atol=0.0, side2=(64, 512), side3=(32, 96),
                   steps_range=(50, 500), dt=0.05) -> Workload:
    """One seeded random model of the serving benchmark's request stream:
    1 or 2 compartments (inner disk / ball of radius 0.3 and its complement)
    on a 2D Q1 mesh with side log-uniform in side2 or a 3D Q1 mesh with side
    log-uniform in side3; 2-8 species; a random mass-action network of 2-12
    reactions with rates log-uniform in [1e-3, 1]; diffusion log-uniform in
    [1e-4, 1e-2]; with two compartments one transmission flux
    k_t (a_in - b_out) across the interface; implicit Euler or Alexander2;
    a step count uniform in steps_range.  Rates, diffusion and k_t are
    parser_context constants; Jacobians are written out (mass action)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dim = int(rng.integers(2, 4))
    lo, hi = side2 if dim == 2 else side3
    n = int(np.clip(np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi)))), lo, hi))
    ncomp = int(rng.integers(1, 3))
    nsp = int(rng.integers(2, 9))
    if ncomp == 2:
        nin = int(rng.integers(1, nsp))
        comps = ["inner"] * nin + ["outer"] * (nsp - nin)
    else:
        comps = ["domain"] * nsp
    names = [f"s{i}" for i in range(nsp)]
    nreac = int(rng.integers(2, 13))
    logu = lambda a, b: float(np.exp(rng.uniform(np.log(a), np.log(b))))
    consts = {}
    terms = {s: [] for s in names}      # (sign, rate expr, reactants)
    for r in range(nreac):
        c = comps[int(rng.integers(0, nsp))]
        pool = [s for s, cc in zip(names, comps) if cc == c]
        nr = int(rng.integers(1, min(2, len(pool)) + 1))
        reac = list(rng.choice(pool, size=nr, replace=False))
        npr = int(rng.integers(0, min(2, len(pool)) + 1))
        prod = list(rng.choice(pool, size=npr, replace=False)) if npr else []
        k = f"k{r}"
        consts[k] = logu(1e-3, 1.0)
        for s in reac:
            terms[s].append((-1, k, reac))
        for s in prod:
            terms[s].append((+1, k, reac))
    for i, s in enumerate(names):
        consts[f"D{i}"] = logu(1e-4, 1e-2)
    trans = None
    if ncomp == 2:
        a = names[int(rng.integers(0, nin))]
        b = names[int(rng.integers(nin, nsp))]
        consts["kt"] = logu(1e-3, 1.0)
        trans = (a, b)
    rk = "ImplicitEuler" if rng.integers(0, 2) == 0 else "Alexander2"
    steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
    ic = {s: (float(rng.uniform(0.2, 0.8)), float(rng.uniform(0.0, 0.2)), rng.uniform(1.0, 6.0, size=3))
          for s in names}

    axes = ["position_x", "position_y", "position_z"][:dim]
    cells = " ".join([str(n)] * dim)
    ext = " ".join(["1"] * dim)
    r2 = " + ".join(f"({ax}-0.5)^2" for ax in axes)
    lines = [f"[grid]\ndimension = {dim}\ncells = {cells}\nextensions = {ext}\ncell_type = cube\n",
             "[parser_context]"]
    for k2, v in consts.items():
        lines.append(f"{k2}.type = constant\n{k2}.value = {v!r}")
    lines.append("\n[compartments]")
    if ncomp == 2:
        lines.append(f"inner.expression = {r2} < 0.09\nouter.expression = {r2} >= 0.09")
    else:
        lines.append("domain.expression = 1")

    def mono(reac, skip=None):
        f = [x for x in reac if x != skip]
        return "*".join(f) if f else "1"

    for i, s in enumerate(names):
        sec = [f"\n[model.scalar_field.{s}]", f"compartment = {comps[i]}", "storage.expression = 1",
               f"cross_diffusion.{s}.expression = D{i}"]
        a0, a1, w = ic[s]
        arg = " + ".join(f"{w[j]:.6f}*{ax}" for j, ax in enumerate(axes))
        sec.append(f"initial.expression = {a0:.6f} + {a1:.6f}*sin({arg})")
        if terms[s]:
            expr = " ".join(f"{'+' if sg > 0 else '-'} {k}*{mono(reac)}" for sg, k, reac in terms[s])
            sec.append(f"reaction.expression = {expr}")
            jac = {}
            for sg, k, reac in terms[s]:
                for x in reac:
                    jac.setdefault(x, []).append(f"{'+' if sg > 0 else '-'} {k}*{mono(reac, x)}")
            for x, parts in jac.items():
                sec.append(f"reaction.jacobian.{x}.expression = {' '.join(parts)}")
        if trans and s == trans[0]:
            a, b = trans
            sec += [f"outflow.outer.expression = kt*({a} - {b})", f"outflow.outer.jacobian.{a}.expression = kt",
                    f"outflow.outer.jacobian.{b}.expression = -kt"]
        if trans and s == trans[1]:
            a, b = trans
            sec += [f"outflow.inner.expression = -kt*({a} - {b})", f"outflow.inner.jacobian.{a}.expression = -kt",
                    f"outflow.inner.jacobian.{b}.expression = kt"]
        lines.append("\n".join(sec))
    ini = "\n".join(lines) + "\n\n" + _solver_block(rk, "Jacobi", nl_rtol, nl_atol, 1e-10, precision, dt=dt)

    def init(model):
        out = np.zeros(model.n_dof)
        model_init = getattr(model, "initial_state", None)
        return model_init() if model_init is not None else out

    w = Workload(f"request_{seed}_{dim}d_{n}_{nsp}sp_{ncomp}c_{rk}", ini, dt, steps, dim, (n,) * dim, init,
                 stages=1 if rk == "ImplicitEuler" else 2)
    return w


# ------------------------------------------------------------- closed forms
def exponential_decay(dim=2, n=4, cell_type="simplex", rk="ImplicitEuler", rate=-2.0, D=0.0) -> Workload:
    """u' = rate * u with storage 1 (test/exp.ini); uniform u stays uniform."""
    cells = " ".join([str(n)] * dim)
    diff = f"cross_diffusion.u.expression = {D:g}\n" if D else ""
    ini = f"""[grid]
dimension = {dim}
cells = {cells}
cell_type = {cell_type}

[compartments]
domain.expression = 1

[parser_context.grow_rate]
type = constant
value = {rate:g}

[model.scalar_field.u]
compartment = domain
initial.expression = 1
storage.expression = 1
reaction.expression = grow_rate*u
reaction.jacobian.u.expression = grow_rate
{diff}
""" + _solver_block(rk, "ILU", 1e-20, 1e-28, 1e-13)

    def init(model):
        return np.ones(model.n_dof)

    return Workload(f"exp_{dim}d_{cell_type}_{rk}", ini, 0.1, 10, dim, (n,) * dim, init,
                    stages=1 if rk == "ImplicitEuler" else 2)
