/*
 * C ABI of the MI355X (gfx950) diffusion-reaction time step.
 *
 * This is the drop-in boundary for DuneCopasi's per-time-step path: a model
 * read from the reference's own INI keys, advanced one time step at a time
 * exactly where the reference calls `one_step.apply(in, out)`.  Every
 * function returns 0 on success, a positive status for a numerical failure
 * (the reference's PDELab::ErrorCondition) or a negative value on error; the
 * error text is then in dcg_last_error().  No function throws.
 *
 * Ownership: a handle owns all device memory of its model (mesh tables,
 * BCSR matrix, state and work vectors, generated kernels).  Host buffers are
 * the caller's and are copied.  A handle is used by one host thread at a
 * time; different handles are independent and may run concurrently.
 *
 * State layout (matches the reference's node-grouped "EntityGrouping" basis,
 * dune/copasi/model/diffusion_reaction/model_single_compartment_traits.hh:26-27):
 * compartments in INI order, inside a compartment its vertices in increasing
 * vertex order, inside a vertex the compartment's species in INI order.
 */
#ifndef DCOPASI_H
#define DCOPASI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dcg_model* dcg_model_t;

/* statistics of one dcg_step call */
typedef struct dcg_step_stats {
  int status;            /* 0 ok, 1 Newton not converged, 2 Krylov not converged, 3 non-finite, 4 breakdown */
  int stages;            /* Runge-Kutta stages completed */
  int newton_iterations; /* summed over stages */
  int linear_iterations; /* BiCGSTAB iterations summed over all solves */
  int refinement_steps;  /* FP64 refinement steps of the mixed-precision solve */
  double residual_norm;  /* last Newton residual in the configured norm (two_norm2 for l_2) */
} dcg_step_stats;

/* Build a model from INI text plus "key=value" override lines.
 * Replaces DiffusionReaction::make_model + Model::make_state +
 * Model::make_step_operator (dune/copasi/model/diffusion_reaction/factory.hh:46-136,
 * dune/copasi/model/model.hh:24-101,
 * dune/copasi/model/diffusion_reaction/model_single_compartment.impl.hh:195-232,
 * dune/copasi/model/make_step_operator.hh:159-444).
 * device >= 0 selects a GPU; device < 0 builds the host side only (grid,
 * compartments, BCSR pattern, generated kernel source) without touching a GPU. */
int dcg_model_create(const char* ini_text, const char* overrides, int device, dcg_model_t* out);

/* Release the handle and its device memory. */
void dcg_model_destroy(dcg_model_t model);

/* Sizes of the state vector (dune/copasi/model/model.hh:36-42 State::coefficients). */
int dcg_state_size(dcg_model_t model, size_t* n_dof, size_t* n_species, size_t* n_compartments);

/* Copy a host state in / out (Model::interpolate and State::coefficients,
 * dune/copasi/model/model.hh:60-75). */
int dcg_set_state(dcg_model_t model, const double* state, size_t n, double time);
int dcg_get_state(dcg_model_t model, double* state, size_t n, double* time);

/* Evaluate the `initial.expression` keys at the DOF vertices
 * (dune/copasi/model/make_initial.hh:24-90); no device needed. */
int dcg_initial_state(dcg_model_t model, double* state, size_t n);

/* One time step t -> t+dt from the current state: PDELab::OneStep::apply as
 * called by TimeStepper::do_step (dune/copasi/common/stepper.hh:95-125).
 * On a positive status the stored state is unchanged. */
int dcg_step(dcg_model_t model, double t, double dt, dcg_step_stats* stats);

/* FP64 integral of every species over its compartment, species in INI order
 * (the mass sums of dune/copasi/model/diffusion_reaction/reduce.hh:160-185). */
int dcg_mass(dcg_model_t model, double* per_species);

/* Message of the last failed call on this thread (format_exception text,
 * dune/copasi/common/exceptions.hh:21-33). */
const char* dcg_last_error(void);

/* ---- layout queries (no device needed) ---- */

/* Name, compartment and index inside the compartment of species s. */
int dcg_species_info(dcg_model_t model, size_t s, char* name, size_t name_len, int* compartment, int* local_index);

/* Name, vertex count, first DOF and species count of compartment c. */
int dcg_compartment_info(dcg_model_t model, size_t c, char* name, size_t name_len, size_t* n_nodes,
                         size_t* dof_offset, int* n_species);

/* Global vertex ids of compartment c's nodes, in DOF order. */
int dcg_compartment_nodes(dcg_model_t model, size_t c, long* vertex_ids);

/* Jacobian sparsity (local_operator.hh:276-399, make_step_operator.hh:295-401)
 * as BCSR over slots (one slot = one vertex of one compartment, slots in DOF
 * order) with block_size x block_size row-major blocks. */
int dcg_pattern_info(dcg_model_t model, size_t* n_slots, size_t* n_blocks, int* block_size);
int dcg_pattern(dcg_model_t model, int* row_ptr, int* col);

/* HIP source generated for the model's assembly kernels; *needed gets the
 * size including the terminating zero. */
int dcg_generated_source(dcg_model_t model, char* buf, size_t len, size_t* needed);

/* ---- kernel-level entry points (device required) ---- */

/* residual = mass_weight * R^M(x) + stiffness_weight * R^A(x) at time t and,
 * when jacobian != NULL, its Jacobian as BCSR values (n_blocks * block_size^2).
 * The local-operator residual/Jacobian of local_operator.hh:417-1452 combined
 * as the instationary assembler does (make_step_operator.hh:403-405). */
int dcg_assemble(dcg_model_t model, const double* x, double t, double mass_weight, double stiffness_weight,
                 double* residual, double* jacobian);

/* y = J(x) v without assembling J: the matrix-free Jacobian apply
 * (localAssembleJacobianVolumeApply, local_operator.hh:510-524, through the
 * matrix-free instationary assembler of make_step_operator.hh:291-293), the
 * operator BiCGSTAB uses with linear_solver.matrix_free = true, transmission
 * and outflow blocks included (local_operator.hh:1354-1452).  Models whose
 * compartments all take the tiled kernel (structured grids, analytical
 * Jacobians, no advection or gradient-dependent diffusion). */
int dcg_jacobian_apply(dcg_model_t model, const double* x, double t, double mass_weight, double stiffness_weight,
                       const double* v, double* y);

/* y = J x for BCSR values J on the model's pattern. */
int dcg_spmv(dcg_model_t model, const double* jacobian, const double* x, double* y);

/* Solve J x = b with the model's configured Krylov solver and preconditioner
 * (LinearSolver::apply, make_step_operator.hh:102-146). */
int dcg_linear_solve(dcg_model_t model, const double* jacobian, const double* b, double* x, double rtol, int maxit,
                     int* iterations);

/* dcg_linear_solve with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  jacobian (BCSR values on the
 * model's pattern), b and x (packed state layout) are DEVICE pointers the
 * caller owns, ready before the call (synchronise the stream that wrote them);
 * the call returns with x written.  Every other buffer of the FP64 BiCGSTAB
 * solve with the model's preconditioner lives in one fixed device array of
 * DCG_NOALLOC_WORKSPACE_BYTES per GPU, held until the process ends; a model
 * whose work space exceeds it, or that solves in mixed precision, is
 * rejected with an error.  Calls that overlap on one GPU share that array
 * and race: serialise them. */
#define DCG_NOALLOC_WORKSPACE_BYTES (64u << 20)
int dcg_linear_solve_noalloc(dcg_model_t model, const double* jacobian, const double* b, double* x, double rtol,
                             int maxit, int* iterations);

/* ---- measurement ----
 * HIP-event timing on the model's own stream, the device counterpart of the
 * reference's Perfetto counters (dune/copasi/common/stepper.hh:100-102,
 * dune/copasi/model/make_step_operator.hh:137-139).  kind: 0 FP64 SpMV,
 * 1 FP32 SpMV, 2 Jacobian assembly pass, 3 residual assembly pass,
 * 4 preconditioner set-up, 5 matrix-free Jacobian apply, 6 point-coefficient
 * set-up, 7 FP64 and 8 FP32 point-coefficient apply (linear_solver.storage =
 * points).  Enabling resets the counters; `bytes` sums the algorithmic bytes
 * of the timed launches.  Krylov operator launches that follow a BiCGSTAB
 * solve's convergence (the rest of a batch of iterations; they return at
 * once) are not counted, so averages are over working launches. */
int dcg_profile_enable(dcg_model_t model, int on);
int dcg_profile_read(dcg_model_t model, int kind, long* launches, double* milliseconds, double* bytes);

/* Copy `bytes` on the model's device with a plain streaming kernel
 * (probe_copy_k) and return its duration: the reference point that HBM
 * traffic counters and bandwidth figures are calibrated against. */
int dcg_probe_copy(dcg_model_t model, size_t bytes, double* milliseconds);

/* SpMV variants the set-up autotuner timed on this model's pattern: for each
 * candidate k < *count, variant[k] (0 LDS-staged entries, 5 LDS-staged whole
 * blocks, 1..4 row-split with 8/16/32/64 lanes per row; +100 for the FP32
 * matrix of the mixed-precision solve) and microseconds[k]; *chosen is the
 * FP64 variant used.  Arrays hold at least 16 entries. */
int dcg_spmv_tuning(dcg_model_t model, int* count, int* variant, double* microseconds, int* chosen);

/* Tile shapes the set-up autotuner timed for the point-coefficient apply
 * (linear_solver.storage = points): for each candidate k < *count, its
 * compartment comp[k], tile[4k..4k+2] (row vertices per workgroup along x,
 * y, z), tile[4k+3] (the waves per SIMD its registers are bounded for; 1:
 * none) and the microseconds of one apply; chosen[k] = 1 for the candidate in
 * use.  No candidates when the model has no point storage or a single shape;
 * a shape fixed by linear_solver.point_tile lists the candidates with zero
 * times.  Arrays hold at least 64 entries (tile: 256). */
int dcg_point_tuning(dcg_model_t model, int* count, int* comp, int* tile, double* microseconds, int* chosen);

/* ---------------------------------------------------------------- transform / reduce
 * The reference's `model.reduce` keys (dune/copasi/model/diffusion_reaction/
 * reduce.hh:40-285, evaluated by the CLI after every step,
 * src/dune_copasi.cc:420-423) on the model's current state: per key,
 * evaluation.expression at every order-4 quadrature point of every cell,
 * folded by reduction.expression ("a, b: ...", default +) from
 * initial.value, then transformation.expression, error.expression and
 * warn.expression.  config_ini: NULL or "" for the model's own `model.reduce`
 * keys, else INI text whose top-level sections are the keys.  Returns 0, or
 * 1 when an error expression fired (the reference's ReductionError); the
 * results stay readable either way.  flag: 0, 1 warn fired, 2 error fired. */
int dcg_reduce(dcg_model_t model, const char* config_ini, size_t* count);
int dcg_reduce_result(dcg_model_t model, size_t i, char* name, size_t name_len, double* value, int* flag);
/* HIP source of the reduce kernels for config_ini (as dcg_reduce reads it);
 * no device needed.  *needed gets the size including the terminating zero. */
int dcg_reduce_source(dcg_model_t model, const char* config_ini, char* buf, size_t len, size_t* needed);

/* VTK output of the current state (the reference's write_vtk,
 * model_single_compartment.impl.hh:250-319 / model_multi_compartment.impl.hh:
 * 219-300, used by the CLI for model.writer.vtk.path): per compartment
 * "<path>/<stem>-<compartment>-NNNNN.vtu" (stem: last component of path)
 * with one scalar vertex field per species, and the time series
 * "<path>/<stem>-<compartment>.pvd"; append = 0 restarts the series. */
int dcg_write_vtk(dcg_model_t model, const char* path, int append);

/* ---------------------------------------------------------------- time loop
 * The reference's SimpleAdaptiveStepper with TimeStepper::evolve and
 * snap_to_time (dune/copasi/common/stepper.hh:95-240,330-366), as its CLI
 * configures it (src/dune_copasi.cc:384-437): on a failed step the step size
 * is multiplied by decrease_factor and the step retried from the same state;
 * after a success the next size is multiplied by increase_factor (clamped
 * to max_step); the loop stops two steps before t_end and snaps to it with
 * equal steps.  min_step / max_step of 0 mean "no limit". */
typedef struct dcg_evolve_options {
  double t_end, dt_initial, decrease_factor, increase_factor, min_step, max_step;
  int snap_to_end_time;
} dcg_evolve_options;

typedef struct dcg_evolve_stats {
  int status; /* 0 ok, 5 step size left [min_step, max_step], else the failing step's status */
  int steps, failed_steps;
  double time, dt; /* reached time, suggested next step size */
  long newton_iterations, linear_iterations;
} dcg_evolve_stats;

/* options from the model's INI keys model.time_step_operator.{time_end,
 * time_step_initial, time_step_min, time_step_max, time_step_decrease_factor,
 * time_step_increase_factor} */
int dcg_evolve_defaults(dcg_model_t model, dcg_evolve_options* options);
/* advance from the model's current time; on_step (may be NULL) after every
 * accepted step; accepted_dt (may be NULL) receives up to capacity accepted
 * step sizes.  Returns stats.status (or -1 on error). */
int dcg_evolve(dcg_model_t model, const dcg_evolve_options* options, dcg_evolve_stats* stats,
               void (*on_step)(void* user, double time), void* user, double* accepted_dt, size_t capacity);

/* ---------------------------------------------------------------- run-time parameters
 * The reference re-parses `parser_context` constants whenever it builds a
 * model (dune/copasi/parser/context.cc:56-173); parameter sweeps and fits
 * rebuild it per value.  With DCG_RUNTIME_PARAMETERS the constants and the
 * mesh spacing are read by the generated kernels at run time instead of being
 * folded in, so one compiled module serves every parameter value and mesh
 * size, and dcg_set_params changes values in place. */
#define DCG_RUNTIME_PARAMETERS 1u
int dcg_model_create_ex(const char* ini_text, const char* overrides, int device, unsigned flags, dcg_model_t* out);
/* count = number of parameters; name of parameter i (sorted by name) when i < count */
int dcg_param_info(dcg_model_t model, size_t i, char* name, size_t name_len, size_t* count);
int dcg_set_params(dcg_model_t model, const double* values, size_t n);
int dcg_get_params(dcg_model_t model, double* values, size_t n);

/* Device memory held by a model, and the estimate for an INI before anything
 * is allocated (admission by HBM footprint). */
int dcg_device_bytes(dcg_model_t model, size_t* bytes);
int dcg_estimate_device_bytes(const char* ini_text, const char* overrides, size_t* bytes);
/* hipMemGetInfo of a device */
int dcg_device_memory(int device, size_t* free_bytes, size_t* total_bytes);

/* ---------------------------------------------------------------- batches
 * Many independent models advanced together: every member is built with
 * run-time parameters on the batch's device and stream and keeps its own HBM
 * buffers; dcg_batch_step advances every member by one step of its own
 * scheme and step size from its own time, with one launch per device
 * operation for all members (batched assembly per shared module, batched
 * BiCGSTAB with per-member scalars, Jacobi, per-member FP64 refinement
 * of the FP32 solve, hipGraph-captured iteration blocks).  The reference
 * has no counterpart: it advances one model per process
 * (dune/copasi/common/stepper.hh:95-175).
 * options_ini keys: precision = mixed|fp64,
 * preconditioner = Jacobi (diagonal, as istl's SeqJac) | VertexBlockJacobi
 * (inverse of each vertex's species block; no counterpart in the reference) | None,
 * refinement.inner_relative_tolerance, refinement.max_steps,
 * iteration_limit, graph_iterations (0: no graphs). */
typedef struct dcg_batch* dcg_batch_t;
int dcg_batch_create(const char* options_ini, int device, dcg_batch_t* out);
void dcg_batch_destroy(dcg_batch_t batch);
/* add a member; *id identifies it until removed */
int dcg_batch_add(dcg_batch_t batch, const char* ini_text, const char* overrides, int* id);
int dcg_batch_remove(dcg_batch_t batch, int id);
/* member ids in batch order (the order of dcg_batch_step's stats) */
int dcg_batch_ids(dcg_batch_t batch, int* ids, size_t capacity, size_t* count);
/* a model handle onto a member (state, parameters, mass, sizes); release it
 * with dcg_model_destroy, which leaves the member in the batch */
int dcg_batch_member(dcg_batch_t batch, int id, dcg_model_t* out);
int dcg_batch_set_time_step(dcg_batch_t batch, int id, double dt);
/* one step of every member; stats[k] for the k-th id (capacity >= members) */
int dcg_batch_step(dcg_batch_t batch, dcg_step_stats* stats, size_t capacity);
int dcg_batch_device_bytes(dcg_batch_t batch, size_t* bytes);
/* HIP-event timing of the batched SpMVs (kind 0: FP64, 1: FP32); while it is
 * on, BiCGSTAB iterations run as plain launches instead of graphs.  Enabling
 * resets the counters; bytes are the algorithmic bytes of the members that
 * were still iterating. */
int dcg_batch_profile_enable(dcg_batch_t batch, int on);
int dcg_batch_profile_read(dcg_batch_t batch, int kind, long* launches, double* milliseconds, double* bytes);

/* ---------------------------------------------------------------- serving
 * A request loop over one GPU: requests (an INI model and a number of time
 * steps from its initial state) are built and compiled on host worker
 * threads, admitted to a resident batch while their estimated HBM footprint
 * fits the budget and a member slot is free, advanced one step per
 * dcg_server_tick together with every other resident request (continuous
 * batching), and retired with their final state kept for dcg_server_result.
 * No counterpart in the reference, which solves one model per process
 * (src/dune_copasi.cc:229-290).
 * options_ini keys: the batch keys plus hbm_budget_bytes (0: 90% of free
 * memory), max_members, workers. */
typedef struct dcg_server* dcg_server_t;

typedef struct dcg_request_info {
  int status;      /* 0 ok, > 0 numerical failure (dcg_step_stats codes), < 0 set-up error */
  int finished;    /* 1 once the result is available */
  int steps_done, steps;
  long newton_iterations, linear_iterations;
  size_t n_dof, device_bytes;
  double t_arrival, t_ready, t_admitted, t_finished; /* seconds since the server started */
} dcg_request_info;

int dcg_server_create(const char* options_ini, int device, dcg_server_t* out);
void dcg_server_destroy(dcg_server_t server);
int dcg_server_submit(dcg_server_t server, const char* ini_text, const char* overrides, int steps, int* id);
/* admit, advance every resident request one step, retire; *open = requests
 * not finished.  One thread drives dcg_server_tick; dcg_server_submit,
 * dcg_server_result, dcg_server_forget and dcg_server_stats may be called on
 * other threads at the same time. */
int dcg_server_tick(dcg_server_t server, int* open);
/* info always; state (n == info.n_dof) once finished; returns 1 while unfinished */
int dcg_server_result(dcg_server_t server, int id, double* state, size_t n, dcg_request_info* info);
int dcg_server_forget(dcg_server_t server, int id);
/* queued (waiting for a worker), ready (built, waiting for memory), resident, finished; bytes */
int dcg_server_stats(dcg_server_t server, int* queued, int* ready, int* resident, int* finished, size_t* used_bytes,
                     size_t* budget_bytes, double* seconds);

#ifdef __cplusplus
}
#endif

#endif /* DCOPASI_H */
