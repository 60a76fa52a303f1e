"""INI models on the gmsh test meshes (tests/gmsh_meshes.py)."""


def gray_scott(path, dim=2):
    return f"""[grid]
dimension = {dim}
path = {path}
[compartments]
domain.expression = 1
[model.scalar_field.u]
compartment = domain
initial.expression = 1 - 0.5*(position_x > 0.3)*(position_x < 0.6)
storage.expression = 1
cross_diffusion.u.expression = 2e-3
reaction.expression = -u*v^2 + 0.04*(1-u)
reaction.jacobian.u.expression = -v^2 - 0.04
reaction.jacobian.v.expression = -2*u*v
[model.scalar_field.v]
compartment = domain
initial.expression = 0.25*(position_y > 0.4)*(position_y < 0.7)
storage.expression = 1
cross_diffusion.v.expression = 1e-3
reaction.expression = u*v^2 - 0.1*v
reaction.jacobian.u.expression = v^2
reaction.jacobian.v.expression = 2*u*v - 0.1
[model.time_step_operator]
type = ImplicitEuler
time_step_initial = 0.5
linear_solver.preconditioner.type = Jacobi
linear_solver.convergence_condition.relative_tolerance = 1e-12
nonlinear_solver.convergence_condition.relative_tolerance = 1e-20
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-26
"""


def two_halves(path, dim=2):
    """Compartments by physical entity, species a (left) and b (right) with a
    transmission flux k (a - b) across x = 1/2, as test/two_disks.ini writes it."""
    return f"""[grid]
dimension = {dim}
path = {path}
[compartments]
left.expression = (gmsh_id == 1)
right.expression = (gmsh_id == 2)
[parser_context.k]
type = constant
value = 0.7
[model.scalar_field.a]
compartment = left
initial.expression = 1 + position_y
storage.expression = 1
cross_diffusion.a.expression = 0.05
reaction.expression = -0.1*a
reaction.jacobian.a.expression = -0.1
outflow.right.expression = k*(a - b)
outflow.right.jacobian.a.expression = k
outflow.right.jacobian.b.expression = -k
[model.scalar_field.b]
compartment = right
initial.expression = 0.2*position_x
storage.expression = 1
cross_diffusion.b.expression = 0.02
outflow.left.expression = k*(b - a)
outflow.left.jacobian.b.expression = k
outflow.left.jacobian.a.expression = -k
[model.time_step_operator]
type = Alexander2
time_step_initial = 0.1
linear_solver.preconditioner.type = ILU
linear_solver.convergence_condition.relative_tolerance = 1e-12
nonlinear_solver.convergence_condition.relative_tolerance = 1e-20
nonlinear_solver.convergence_condition.absolute_tolerance = 1e-26
"""
