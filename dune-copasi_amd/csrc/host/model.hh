// A diffusion-reaction model resident on one GPU: the drop-in for the
// reference's ModelSingleCompartment / ModelMultiCompartment step operator
// (model/diffusion_reaction/model_*_compartment.impl.hh, make_step_operator
// in model/make_step_operator.hh:159-444).  One Model owns its mesh tables,
// BCSR matrix, state, work vectors and generated kernels; `step` is
// PDELab::OneStep::apply for one time step of the configured Runge-Kutta
// scheme with Newton (or one defect correction for linear models) per stage.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../kernels/kernels.hh"
#include "codegen.hh"
#include "config.hh"
#include "device.hh"
#include "layout.hh"
#include "reduce.hh"
#include "rtc.hh"
#include "spec.hh"

namespace dcopasi {

struct StepStats {
  int status = 0;  // 0 ok, 1 Newton not converged, 2 Krylov not converged, 3 non-finite, 4 breakdown
  int stages = 0;
  int newton_iterations = 0;
  int linear_iterations = 0;
  int refinement_steps = 0;
  double residual_norm = 0.0;  // last Newton residual in the configured norm
};

struct RkTable {
  int stages = 1;
  double a[4][5] = {{0}}, b[4][5] = {{0}}, d[5] = {0};
};
RkTable rk_table(const std::string& type);

// Jacobi: istl's SeqJac as the reference applies it to its scalar Jacobian,
// one sweep from zero = the inverse of the diagonal; VertexBlockJacobi (no
// counterpart in the reference): the inverse of each vertex's species block;
// ILU0: incomplete LU over the species-blocked pattern.
enum class Preconditioner { None, Jacobi, VertexBlockJacobi, ILU0 };
enum class Precision { FP64, Mixed };

struct SolverConfig {
  std::string rk = "Alexander2";
  Preconditioner precond = Preconditioner::Jacobi;
  Precision precision = Precision::FP64;
  double lin_rtol = 1e-4;
  int lin_maxit = 500;
  double nl_rtol = 1e-4, nl_atol = 0.0;
  int nl_min = 0, nl_max = 40;
  int norm = 0;  // 0: two_norm2, 1: l_inf, 2: l_1
  bool fixed_linear_tolerance = false;  // nonlinear_solver.dx_inverse_fixed_tolerance
  double min_linear_reduction = 0.1;    // nonlinear_solver.dx_inverse_min_relative_tolerance
  double linearization_threshold = 0.0; // re-linearize when defect_k/defect_{k-1} exceeds it
  // linear_solver.matrix_free: BiCGSTAB applies the Jacobian at the
  // linearization point from the generated kernels, no matrix is assembled
  bool matrix_free = false;
  // linear_solver.storage: the Jacobian as point coefficients per cell (the
  // BiCGSTAB operator applies them) instead of an assembled BCSR matrix
  bool points = false;
  // linear_solver.readback_iterations: BiCGSTAB iterations launched between
  // read-backs of its scalars (0: as many as the same solve took in the last
  // time step, then 4); 1 launches none past convergence (for profilers)
  int readback_iterations = 0;
  // linear_solver.point_tile: the point apply's tile shape, "auto" (timed at
  // set-up) or the index of a candidate shape (compartments with fewer
  // candidates keep their first)
  int point_tile = -1;
  double ir_inner_rtol = 1e-3;  // FP32 Krylov tolerance inside refinement
  int ir_max = 50;
  // linear_solver.refinement.tolerance_floor: mixed precision solves each
  // Newton correction to max(requested, floor) (0: the requested tolerance);
  // Newton's own FP64 defect test still decides convergence
  double ir_floor = 0.0;
  std::vector<std::string> warnings;
};

// Creation options beyond the INI keys.
struct ModelOptions {
  // parser_context constants and the mesh spacing are read at run time, so
  // models that differ only in them share one generated module (batches,
  // parameter sweeps); set_params changes the constants without recompiling
  bool runtime_parameters = false;
  // member of a Batch: the batch tunes the SpMV and chooses the (Jacobi)
  // preconditioner; `stream` (owned by the batch) replaces the model's own
  bool batch_member = false;
  hipStream_t stream = nullptr;
};

class Batch;

class Model {
 public:
  Model(const std::string& ini, const std::string& overrides, int device, const ModelOptions& opt = {});
  ~Model();
  Model(const Model&) = delete;
  Model& operator=(const Model&) = delete;

  // sizes and layout
  size_t n_dof() const { return static_cast<size_t>(L_.ndof); }
  size_t n_species() const { return spec_.species.size(); }
  size_t n_compartments() const { return spec_.comp_names.size(); }
  const Layout& layout() const { return L_; }
  const ModelSpec& spec() const { return spec_; }
  const SolverConfig& solver() const { return cfg_; }
  const ParameterTree& config() const { return cfg_tree_; }
  const std::string& source() const { return gen_.source; }
  bool on_device() const { return device_ >= 0; }
  double time() const { return time_; }
  double time_step() const { return dt_initial_; }
  double time_end() const { return t_end_; }

  // two-phase construction (device < 0 first, e.g. on a worker thread):
  // compile the generated kernels into the on-disk cache, then allocate and
  // upload on `device`, on `stream` when given
  void precompile() const;
  void attach(int device, hipStream_t stream = nullptr);

  void set_state(const double* packed, double t);
  void get_state(double* packed, double* t);
  void initial_state(double* packed) const;  // from initial.expression keys
  int step(double t, double dt, StepStats* st);
  void mass(double* per_species);
  // transform/reduce over the current state: the `model.reduce` keys, or the
  // keys of `reduce_ini` (top-level sections) when it is not empty
  std::vector<ReduceResult> reduce(const std::string& reduce_ini);
  // the generated source of those reduce kernels (no device needed)
  std::string reduce_source(const std::string& reduce_ini) const;
  // VTK files of the current state under `path` (one .vtu per compartment and
  // call, a .pvd time series per compartment; append keeps the earlier
  // steps written to that path in the series)
  void write_vtk(const std::string& path, bool append);
  // run-time parameters (parser_context constants by name); only models
  // created with runtime_parameters have them
  const std::vector<std::string>& param_names() const { return spec_.params; }
  void set_params(const double* values, size_t n);
  void get_params(double* values, size_t n) const;
  // bytes of device memory the model holds (after set-up) or would hold
  size_t device_bytes() const;
  static size_t estimate_device_bytes(const Layout& L, int stages, bool mixed, bool ilu, bool matrix_free = false);

  // kernel-level entry points (parity tests)
  void assemble(const double* x_packed, double t, double cm, double ca, double* r_packed, double* jac);
  void spmv(const double* jac, const double* x_packed, double* y_packed);
  int linear_solve(const double* jac, const double* b_packed, double* x_packed, double rtol, int maxit, int* its);
  // the same FP64 solve on the caller's device buffers with no device
  // allocation: work space from the library's fixed per-device array
  int linear_solve_noalloc(const double* jac_dev, const double* b_dev, double* x_dev, double rtol, int maxit,
                           int* its);
  // y = J(x) v without assembling J (the matrix-free Jacobian apply), for
  // models whose compartments all use the tiled kernel (facet terms included)
  void jacobian_apply(const double* x_packed, double t, double cm, double ca, const double* v_packed,
                      double* y_packed);
  bool can_apply_jacobian() const;

  // kernel timing with HIP events on the model's stream (benchmarks)
  enum ProfKind {
    kSpmv64 = 0,
    kSpmv32 = 1,
    kAssemblyJac = 2,
    kAssemblyRes = 3,
    kPrecond = 4,
    kJacobianApply = 5,
    kPointSetup = 6,    // point coefficients and diagonal blocks (linear_solver.storage = points)
    kPointApply64 = 7,  // y = J v from the point coefficients, FP64
    kPointApply32 = 8,  // the same in FP32 inside the refinement loop
    kProfKinds = 9
  };
  void set_profiling(bool on);
  // launches, summed milliseconds and summed algorithmic bytes of one kind
  void profile(int kind, long* launches, double* ms, double* bytes);
  // device copy of `bytes` through a plain kernel; returns its milliseconds
  double probe_copy(long bytes);

  // point-apply autotuning result: (compartment, tile shape, microseconds) per candidate
  struct PointTuning {
    int comp;
    std::array<int, 3> tile;
    int waves;  // waves per SIMD the candidate's registers are bounded for (1: none)
    double us;
    bool chosen;
  };
  const std::vector<PointTuning>& point_tuning() const { return point_tuning_; }
  // SpMV autotuning result: (variant, microseconds) per candidate; chosen variant
  const std::vector<std::pair<int, double>>& spmv_tuning() const { return tuning_; }
  int spmv_variant() const { return plan_.variant; }

  // device-side pieces used by batching/benchmarks
  hipStream_t stream() const { return stream_; }
  void synchronize() const;
  long matrix_values() const { return static_cast<long>(L_.nnzb()) * L_.B * L_.B; }
  const RkTable& rk() const { return rk_; }
  void set_time(double t) { time_ = t; }
  // device buffers of the model for batched kernels
  k::MemberView view() const;

 private:
  friend class Batch;
  void build(const std::string& ini, const std::string& overrides);
  void setup_device();
  void require_device() const;
  // assembly: r = base + cm*R^M(x) + ca*R^A(x) (base may be null or r itself)
  // and/or val = d/dx of it, complete (padding and interface blocks included)
  // skip_tiled: leave the tiled compartments to a batched launch; phases:
  // 1 prepare r / val, 2 volume kernels, 4 facet kernels
  void launch_assembly(const double* x, double t, double cm, double ca, const double* base, double* r, double* val,
                       int mode, bool skip_tiled = false, int phases = 7);
  // y = J(lin) v from the tiled kernels (mode 4); stop: a done flag that
  // turns the launches into no-ops
  void launch_apply(const double* lin, double t, double cm, double ca, const double* v, double* y,
                    const int* stop);
  // compartments without a batched tile kernel, or facets: per-model launches
  bool needs_own_assembly() const;
  bool all_tile_eligible() const;
  // point-coefficient Jacobian: coefficients and diagonal blocks at x, and
  // y = J v from them (v, y FP64 or FP32 as fp32 says); stop as in launch_apply
  void launch_points_setup(const double* x, double t, double cm, double ca);
  // with part: also the partial dots (y.w1, y.w2) per tile into part; returns
  // the number of partial pairs written
  int launch_points_apply(const void* v, void* y, const int* stop, bool fp32, const void* w1 = nullptr,
                          const void* w2 = nullptr, double* part = nullptr);
  // set-up timing of the point apply's candidate tile shapes
  void tune_points();
  // the assembled matrix for the kernel-level entry points of a model that
  // solves with point coefficients
  void ensure_matrix();
  double residual_norm(const double* r);
  // linear solves on the current val_
  int krylov(const double* b, double* x, double rtol, int maxit, int* its, int* refinements);
  // device buffers one FP64 Krylov solve works in: the model's own, or a
  // carved-out part of the library's fixed work space (linear_solve_noalloc)
  struct KrylovWork {
    double* kv[9] = {};  // r rh p v ph s sh t tmp
    double* part = nullptr;
    double* spart = nullptr;  // the s-update's partial norms
    k::BicgScalars* sc = nullptr;
    const double* val = nullptr;  // the assembled operator
    double* dinv = nullptr;
    double* lu = nullptr;
    int* status = nullptr;
  };
  KrylovWork own_work();
  int bicgstab64(const double* b, double* x, double rtol, double abstol, int maxit, int* its, const KrylovWork& w);
  int bicgstab32(const double* b, float* x, double rtol, int maxit, int* its);
  void precond_setup(const KrylovWork& w);
  bool jacobi_kind() const {
    return cfg_.precond == Preconditioner::Jacobi || cfg_.precond == Preconditioner::VertexBlockJacobi;
  }
  void compile_and_load();
  void tune_spmv();
  int tune_spmv_plan(bool fp32);
  void prof_begin(int kind);
  void prof_end(int kind, double bytes);
  void prof_collect();
  // launches of `kind` recorded so far (not yet collected), and dropping the
  // records after the first `keep` of them from `mark` on: a Krylov solve
  // launches its iterations in batches, and the operator launches after its
  // convergence return at once, so they are not launches of the kernel's work
  size_t prof_mark(int kind) const;
  void prof_keep(int kind, size_t mark, long keep);

 public:
  // algorithmic bytes of one SpMV on this model's matrix
  double spmv_bytes(int value_bytes, int vector_bytes, int extra_vectors) const;

 private:
  double assembly_bytes(bool jacobian) const;
  std::vector<ReduceKey> reduce_keys(const std::string& reduce_ini) const;
  double apply_bytes() const;

 public:
  // algorithmic bytes of one point-coefficient apply / set-up
  double point_apply_bytes(int value_bytes, int extra_vectors = 0) const;
  double point_setup_bytes() const;
  // the Krylov operator is the point-coefficient one
  bool point_storage() const { return cfg_.points; }

 private:

  ParameterTree cfg_tree_;
  ModelSpec spec_;
  Layout L_;
  GeneratedSource gen_;
  SolverConfig cfg_;
  RkTable rk_;
  int device_ = -1;
  ModelOptions opt_;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = true;
  std::shared_ptr<RtcModule> module_;
  std::vector<hipFunction_t> vol_fn_;
  std::vector<hipFunction_t> tile_fn_;      // tiled residual (_r); set when the compartment is tiled
  std::vector<hipFunction_t> tile_jac_fn_;  // tiled Jacobian (_j)
  std::vector<hipFunction_t> tile_app_fn_;  // tiled matrix-free apply (_a)
  // point coefficients (_q), their FP64 (_p) and FP32 (_pf) apply
  std::vector<hipFunction_t> pc_set_fn_, pc_app_fn_, pc_appf_fn_;
  // the point apply per candidate tile shape, and the shape in use, per compartment
  std::vector<std::vector<hipFunction_t>> pc_app_fns_, pc_appf_fns_;
  std::vector<int> apply_shape_;
  std::vector<PointTuning> point_tuning_;
  std::vector<hipFunction_t> facet_fn_;
  std::vector<hipFunction_t> facetf_fn_;  // FP32 apply entries (point storage)
  double time_ = 0.0, dt_initial_ = 0.0, t_end_ = 0.0;
  std::map<std::string, std::vector<double>> vtk_steps_;  // times written per output path
  int last_krylov_its_ = 8;
  // iterations of the n-th Krylov solve of the last time step (the first
  // batch of the n-th solve of the next one), and the solve counter
  std::vector<int> its_hint_;
  int solve_index_ = 0;
  // solves per time step that keep a hint (later ones, and solves outside a
  // step, start from the last solve's count)
  static constexpr int kMaxSolveHints = 4096;

  // device data
  DeviceArray<int> row_ptr_, col_, own_start_, diag_pos_, slot_b_, chunk_row_;
  std::vector<int> chunk_host_;  // SpMV chunk row boundaries
  DeviceArray<unsigned> own_mask_;
  DeviceArray<long> slot_dof_;
  struct CompDev {
    DeviceArray<int> cells, cell_slots;
    DeviceArray<unsigned char> cell_mask;
    // tiled kernel: vertex -> slot, cell -> class mask over the whole grid, tile list
    DeviceArray<int> g2l, tiles;
    // the point apply's candidate tile shapes: tile list and count per shape
    std::vector<DeviceArray<int>> atiles;
    std::vector<int> natiles;
    DeviceArray<unsigned char> cell_mask_grid;
    int ntiles = 0;
    bool g2l_identity = false;  // the compartment's slots are the grid's vertices in order
    // unstructured cells: geometry words and block positions per cell
    DeviceArray<double> egeo;
    DeviceArray<int> cpos;
  };
  bool all_tiled_ = false;
  DeviceArray<int> extra_pos_;  // BCSR blocks coupling two compartments (zeroed before facets)
  std::vector<CompDev> comp_dev_;
  struct FacetDev {
    DeviceArray<int> fint;
    DeviceArray<double> fdbl;
    std::vector<std::pair<int, int>> color_ranges;  // [first, count) per colour
  };
  std::vector<FacetDev> facet_dev_;
  DeviceArray<double> weights_;  // lumped mass per slot (mass sums)
  DeviceArray<double> prm_;      // geometry words + run-time parameters
  DeviceArray<double> u_, hist_, r_, z_, xs_, packed_;
  // matrix-free: the Newton iterate the Jacobian is applied at, and its weights
  DeviceArray<double> lin_x_;
  double lin_t_ = 0.0, lin_cm_ = 0.0, lin_ca_ = 0.0;
  // point-coefficient Jacobian: per compartment the coefficients of every grid
  // cell ([class, point, coefficient][cell]), their FP32 copy when mixed, the
  // rows' diagonal blocks, an identity block-position map for the block
  // inverse, and the time and stiffness weight they were computed with
  std::vector<DeviceArray<double>> pc_;
  std::vector<DeviceArray<float>> pc32_;
  DeviceArray<double> pdiag_;
  DeviceArray<int> diag_id_;
  DeviceArray<double> pc_tca_;  // (t, ca) of the point coefficients, read by the applies
  bool use_points_ = false;  // the Krylov operator is the point-coefficient one
  std::vector<DeviceArray<double>> stage_;
  DeviceArray<double> val_, lu_, dinv_;
  DeviceArray<float> val32_;
  DeviceArray<double> kv_[9];  // r rh p v ph s sh t tmp
  DeviceArray<float> kf_[10];  // same in FP32 + x
  DeviceArray<double> part_, scalar_out_;
  size_t spart_off_ = 0;  // offset of the s-update's partial norms in part_
  DeviceArray<k::BicgScalars> sc_;
  DeviceArray<int> status_;
  k::SpmvPlan plan_, plan32_;
  // ILU level schedules (host copies of the level pointers, device row lists)
  std::vector<int> lo_ptr_, up_ptr_;
  DeviceArray<int> lo_rows_, up_rows_, lo_lev_, up_lev_;
  // every level fits one workgroup in a few passes: one-workgroup sweeps
  // over all levels instead of one launch per level
  bool ilu_sweep_ = false;
  static constexpr int kIluSweepWidth = 4096;
  // the one-workgroup sweeps' tables (entry lists in level order, LDS ring
  // slots, the factorisation's operand positions) and their device view
  DeviceArray<int> lo_ent_, lo_blk_, lo_src_, up_ent_, up_blk_, up_src_, lo_chunk_, up_chunk_;
  DeviceArray<int2> ilu_fpush_;
  DeviceArray<int4> ilu_frow_;
  DeviceArray<int> ilu_fub_, ilu_fm_;
  DeviceArray<double> ilu_packed_, ilu_inbox_;
  k::IluSweep ilu_tab_;
  static constexpr size_t kIluLdsBytes = 64 * 1024;  // dynamic LDS of the sweeps
  void build_ilu_sweep_tables(const std::vector<int>& lrows, const std::vector<int>& urows);
  bool ilu_ready_ = false;
  std::vector<std::pair<int, double>> tuning_;
  // profiling
  bool prof_ = false;
  struct Prof {
    std::vector<hipEvent_t> pool;
    std::vector<double> pending_bytes;
    size_t used = 0;
    long launches = 0;
    double ms = 0.0, bytes = 0.0;
  };
  Prof prof_data_[kProfKinds];
};

}  // namespace dcopasi
