// Many independent diffusion-reaction models advanced together on one GPU.
//
// Each member is a Model built with run-time parameters and geometry, so
// members that differ only in mesh size and parameter values share one
// generated module; a member keeps its own HBM buffers.  One batch step runs,
// for every member, the reference's one-step operator (RK stages with a
// Newton solve per stage, model/make_step_operator.hh:159-444) with the
// member's own time, step size, scheme and Newton control, but every device
// operation is one launch for all members:
//   * assembly: the generated batched tile kernel over (member, tile) pairs
//     per module and compartment (coloured kernels and interface facets of
//     members that need them are launched per member);
//   * Krylov: batched BiCGSTAB kernels over (member, block) and (member,
//     chunk) work lists with per-member scalars, Jacobi preconditioning
//     and, for mixed precision, per-member FP64 iterative refinement; blocks of
//     iterations are captured once in a hipGraph and replayed;
//   * vector updates and norms: batched with per-member skip flags.
// Members are grouped by block size for the solver kernels (the kernels are
// specialised per block size) and by generated module for assembly.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "model.hh"

namespace dcopasi {

struct BatchOptions {
  Precision precision = Precision::Mixed;
  Preconditioner precond = Preconditioner::Jacobi;  // Jacobi, VertexBlockJacobi or None
  double ir_inner_rtol = 1e-3;
  double ir_floor = 0.0;  // mixed: lowest relative tolerance a correction is solved to (0: as requested)
  int ir_max = 50;
  int lin_maxit = 500;
  int graph_iterations = 4;  // BiCGSTAB iterations per captured graph; 0: plain launches
  int point_tile = -1;       // point-apply tile shape: -1 timed per launch, else the candidate index
  // keys: precision = fp64|mixed, preconditioner = Jacobi|VertexBlockJacobi|None,
  // refinement.inner_relative_tolerance, refinement.tolerance_floor, refinement.max_steps,
  // iteration_limit, graph_iterations, point_tile (auto or an index)
  static BatchOptions from_ini(const std::string& text);
};

class Batch {
 public:
  Batch(int device, const BatchOptions& opt);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // builds a member on the batch's device and stream; returns its id
  int add(const std::string& ini, const std::string& overrides);
  // member options and the overrides that put the batch's linear solver in
  // a member's INI (for members built elsewhere, e.g. on worker threads)
  ModelOptions member_options() const;
  std::string member_overrides(const std::string& overrides) const;
  // adopts a host-built member (created with member_options(), device < 0):
  // attaches it to the batch's device and stream
  int add(std::unique_ptr<Model> m);
  void remove(int id);
  Model& member(int id);
  std::vector<int> ids() const;
  size_t size() const { return members_.size(); }
  size_t device_bytes() const;
  void set_time_step(int id, double dt);
  double time_step(int id);

  // one time step of every member from its own time with its own step size;
  // stats in member order (ids()).  A member whose step fails keeps its state.
  void step(std::vector<StepStats>* stats);

  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }
  const BatchOptions& options() const { return opt_; }
  // SpMV variant chosen for each block-size group: (B, FP64 variant, FP32 variant)
  std::vector<std::array<int, 3>> spmv_choice();
  // point-apply tile shape chosen for each launch of each point group
  std::vector<std::array<int, 3>> point_tiles();

  // Kernel timing: with profiling on, BiCGSTAB iterations run as plain
  // launches with HIP events around every batched SpMV (kind 0: FP64, 1:
  // FP32); launches, summed milliseconds and summed algorithmic bytes.
  void set_profiling(bool on);
  void profile(int kind, long* launches, double* ms, double* bytes) const;
  // Newton DOF of all members (state entries, no padding)
  size_t n_dof() const;

 private:
  struct Entry {
    int id;
    std::unique_ptr<Model> m;
    double dt;
    int sg = -1, si = -1;  // solver group and index in it
    int ag = -1, ai = -1;  // assembly group and index in it
    bool per_member_asm = false;  // coloured compartments or facets
  };
  struct SolverGroup {
    int B = 0;
    // members hold point coefficients (the Krylov operator is the batched
    // point apply) instead of assembled matrices
    bool points = false;
    std::vector<int> mem;  // indices into members_
    DeviceArray<k::MemberView> mv;
    DeviceArray<int2> vwork, vrange, vwork4, vrange4, crange;
    DeviceArray<int4> cwork;
    k::BatchWork w;
    DeviceArray<k::BicgScalars> sc;
    DeviceArray<k::IrState> ir;
    DeviceArray<int> act;
    DeviceArray<double> rtol, sums;
    DeviceArray<double> part;
    std::vector<int> h_act, dev_act;  // host flags, and what the device holds
    std::vector<double> h_rtol, dev_rtol;
    int variant64 = 0, variant32 = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // FP64, FP32 iteration blocks
    int est_its = 8;
    // most iterations of any member in the n-th solve of the last step (the
    // first block count of the n-th solve of this one), and the solve counter
    std::vector<int> hints;
    int call = 0;
    // point storage: one launch per generated module over its members'
    // (member, tile) items; item c of launch l writes partial pair off + c,
    // prange[j] = member j's [first, last) pairs
    struct PointLaunch {
      // the kernels and tile shape in use, chosen by timing among the
      // module's candidate tile shapes
      hipFunction_t f64 = nullptr, f32 = nullptr;
      std::shared_ptr<RtcModule> module;
      std::array<int, 3> tile{};
      int nt = 256;  // workgroup size
      std::vector<hipFunction_t> f64s, f32s;
      std::vector<std::array<int, 3>> tiles;
      std::vector<int> nts;
      DeviceArray<int2> work;
      int nwork = 0, off = 0;
    };
    std::vector<PointLaunch> plaunch;
    DeviceArray<k::PcMember> pcm;
    DeviceArray<int2> prange;
    int nitems = 0;
    k::BatchOperator op64, op32;
  };
  struct AsmGroup {
    std::shared_ptr<RtcModule> module;
    std::vector<hipFunction_t> fn;  // batched tile kernel per compartment (or null)
    // point storage: the batched point set-up per compartment replaces the
    // Jacobian pass
    std::vector<hipFunction_t> qfn;
    bool points = false;
    std::vector<int> mem;
    DeviceArray<k::AsmMember> am;
    std::vector<DeviceArray<int2>> work;
    std::vector<int> nwork;
    DeviceArray<double> coef;  // [call][member][4]
    DeviceArray<int> act;      // [2][member]: residual, Jacobian
    std::vector<int> h_act, dev_act;
    std::vector<double> h_coef;
  };
  static constexpr int kCalls = 5;  // hist(1,0), hist(2,0), hist(2,1), stage 1, stage 2
  static int hist_call(int s, int j) { return (s - 1) * s / 2 + j; }
  static int stage_call(int s) { return 3 + (s - 1); }

  void rebuild();
  int jacobi_kind() const;
  void destroy_graphs();
  void tune(SolverGroup& g);
  void capture(SolverGroup& g, bool fp32);
  // flags of members (index into members_): residual/solve activity, Jacobian
  void push_flags(const std::vector<char>& res, const std::vector<char>& jac);
  void assemble(int call, const std::vector<char>& on, int xsel, int bsel, int rsel, int mode);
  void copy(const std::vector<char>& on, int dst, int src);
  void norms(const std::vector<char>& on, int x, std::vector<double>& out);
  // solve val z = r per member (members with on), relative tolerances rtol
  void solve(const std::vector<char>& on, const std::vector<double>& rtol, std::vector<int>& status,
             std::vector<int>& its, std::vector<int>& refs);
  void iterate(SolverGroup& g, bool fp32);
  // the batched point apply of group gi (selectors as the SpMV takes them)
  void point_apply(int gi, bool fp32, int x, int y, int w1, int w2, double* part, int skip);

  int device_;
  BatchOptions opt_;
  bool prof_ = false;
  hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
  struct Prof {
    long launches = 0;
    double ms = 0.0, bytes = 0.0;
  } prof_data_[2];
  hipStream_t stream_ = nullptr;
  std::vector<Entry> members_;
  int next_id_ = 0;
  bool dirty_ = true;
  std::vector<SolverGroup> sgroups_;
  // point-apply candidate chosen per generated module, kept across rebuilds
  std::map<const RtcModule*, std::pair<std::weak_ptr<RtcModule>, size_t>> tuned_;
  std::vector<AsmGroup> agroups_;
  DeviceArray<int> status_;
  // (t, cm, ca, on) per member for each call of the current step
  std::vector<std::array<std::array<double, 4>, kCalls>> coef_;
};

}  // namespace dcopasi
