#include "model.hh"

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cmath>
#include <cstring>

namespace dcopasi {

// ------------------------------------------------------------------ RK tables
// Stage r solves sum_{j<=r} a_rj M(u_j) + dt b_rj A(u_j, t + d_j dt) = 0, the
// PDELab one-step form the reference configures (model/make_step_operator.hh:
// 408-443); coefficients are the textbook ones of each scheme.
RkTable rk_table(const std::string& type) {
  RkTable t;
  // rows of a and b for stage r (1-based) from lists; d from a list
  auto set = [&](int stages, std::vector<std::vector<double>> a, std::vector<std::vector<double>> b,
                 std::vector<double> d) {
    t.stages = stages;
    for (int r = 0; r < stages; ++r)
      for (int j = 0; j <= stages; ++j) {
        t.a[r][j] = a[r][j];
        t.b[r][j] = b[r][j];
      }
    for (int j = 0; j <= stages; ++j) t.d[j] = d[j];
  };
  if (type == "ExplicitEuler") {
    set(1, {{-1.0, 1.0}}, {{1.0, 0.0}}, {0.0, 1.0});
  } else if (type == "ImplicitEuler") {
    set(1, {{-1.0, 1.0}}, {{0.0, 1.0}}, {0.0, 1.0});
  } else if (type == "Heun") {
    // explicit trapezoidal rule in Shu-Osher form: u1 = u0 + dt L(u0),
    // u2 = (u0 + u1 + dt L(u1)) / 2
    set(2, {{-1.0, 1.0, 0.0}, {-0.5, -0.5, 1.0}}, {{1.0, 0.0, 0.0}, {0.0, 0.5, 0.0}}, {0.0, 1.0, 1.0});
  } else if (type == "Shu3") {
    // Shu and Osher (1988) third-order strong-stability-preserving scheme
    set(3, {{-1.0, 1.0, 0.0, 0.0}, {-0.75, -0.25, 1.0, 0.0}, {-1.0 / 3.0, 0.0, -2.0 / 3.0, 1.0}},
        {{1.0, 0.0, 0.0, 0.0}, {0.0, 0.25, 0.0, 0.0}, {0.0, 0.0, 2.0 / 3.0, 0.0}}, {0.0, 1.0, 0.5, 1.0});
  } else if (type == "RungeKutta4") {
    // the classical fourth-order scheme, stage values u_r = u0 + c_{r+1} dt k_r
    set(4,
        {{-1.0, 1.0, 0.0, 0.0, 0.0}, {-1.0, 0.0, 1.0, 0.0, 0.0}, {-1.0, 0.0, 0.0, 1.0, 0.0},
         {-1.0, 0.0, 0.0, 0.0, 1.0}},
        {{0.5, 0.0, 0.0, 0.0, 0.0}, {0.0, 0.5, 0.0, 0.0, 0.0}, {0.0, 0.0, 1.0, 0.0, 0.0},
         {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 0.0}},
        {0.0, 0.5, 0.5, 1.0, 1.0});
  } else if (type == "Alexander2") {
    // Alexander (1977): 2-stage, L-stable SDIRK, alpha = 1 - sqrt(2)/2
    const double al = 1.0 - std::sqrt(2.0) / 2.0;
    set(2, {{-1.0, 1.0, 0.0}, {-1.0, 0.0, 1.0}}, {{0.0, al, 0.0}, {0.0, 1.0 - al, al}}, {0.0, al, 1.0});
  } else if (type == "FractionalStepTheta") {
    // Glowinski's fractional-step theta scheme: theta = 1 - sqrt(2)/2, steps
    // theta, 1 - 2 theta, theta with implicit weights alpha, beta, alpha
    const double th = 1.0 - std::sqrt(2.0) / 2.0, thp = 1.0 - 2.0 * th;
    const double al = 2.0 - std::sqrt(2.0), be = 1.0 - al;
    set(3, {{-1.0, 1.0, 0.0, 0.0}, {0.0, -1.0, 1.0, 0.0}, {0.0, 0.0, -1.0, 1.0}},
        {{be * th, al * th, 0.0, 0.0}, {0.0, al * thp, be * thp, 0.0}, {0.0, 0.0, be * th, al * th}},
        {0.0, th, 1.0 - th, 1.0});
  } else if (type == "Alexander3") {
    // Alexander (1977): 3-stage, third-order, L-stable SDIRK; alpha is the
    // root of x^3 - 3x^2 + 3x/2 - 1/6 in (1/6, 1/2)
    double al = 0.4358665215;
    for (int it = 0; it < 8; ++it)
      al -= (((al - 3.0) * al + 1.5) * al - 1.0 / 6.0) / ((3.0 * al - 6.0) * al + 1.5);
    const double tau2 = (1.0 + al) / 2.0;
    const double b1 = -(6.0 * al * al - 16.0 * al + 1.0) / 4.0, b2 = (6.0 * al * al - 20.0 * al + 5.0) / 4.0;
    set(3, {{-1.0, 1.0, 0.0, 0.0}, {-1.0, 0.0, 1.0, 0.0}, {-1.0, 0.0, 0.0, 1.0}},
        {{0.0, al, 0.0, 0.0}, {0.0, tau2 - al, al, 0.0}, {0.0, b1, b2, al}}, {0.0, al, tau2, 1.0});
  } else {
    throw IOError("Not known time stepping type '" + type + "'");
  }
  return t;
}

// ------------------------------------------------------------------ set-up
Model::Model(const std::string& ini, const std::string& overrides, int device, const ModelOptions& opt)
    : device_(device), opt_(opt) {
  build(ini, overrides);
  if (device_ >= 0) {
    setup_device();
    compile_and_load();
    tune_points();
  }
}

void Model::precompile() const { (void)rtc_compile(gen_.source, "dcopasi_assembly"); }

void Model::attach(int device, hipStream_t stream) {
  if (device_ >= 0) throw Error("model is already on a device");
  if (device < 0) throw Error("attach needs a device");
  device_ = device;
  opt_.stream = stream;
  setup_device();
  compile_and_load();
  tune_points();
}

Model::~Model() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& p : prof_data_)
    for (auto e : p.pool) (void)hipEventDestroy(e);
  if (stream_ && own_stream_) (void)hipStreamDestroy(stream_);
}

void Model::build(const std::string& ini, const std::string& overrides) {
  cfg_tree_ = ParameterTree::from_ini(ini, overrides);
  const auto& c = cfg_tree_;
  const int dim = c.get("grid.dimension", 2);
  spec_ = read_model(c, dim, !opt_.runtime_parameters);
  Grid g;
  if (c.has_key("grid.path")) {
    // unstructured simplices from a gmsh file (make_multi_domain_grid.hh:56-63)
    g = read_gmsh(c.get("grid.path", std::string()), dim);
    const int levels = c.get("grid.refinement_level", 0);
    if (levels < 0) throw IOError("grid.refinement_level must not be negative");
    refine_simplices(g, levels);
  } else {
    const std::string cell = c.get("grid.cell_type", std::string("simplex"));
    if (cell != "simplex" && cell != "cube") throw IOError("grid.cell_type must be 'simplex' or 'cube'");
    g = make_grid(dim, c.get_list("grid.cells"), c.get_list("grid.extensions"), c.get_list("grid.origin"),
                  c.get("grid.refinement_level", 0), cell == "cube");
  }
  auto elem_comp = mark_elements(g, spec_.comp_expr);
  std::vector<Compartment> comps(spec_.comp_names.size());
  for (size_t i = 0; i < comps.size(); ++i) {
    comps[i].name = spec_.comp_names[i];
    comps[i].species = spec_.comp_species[i];
    comps[i].b = static_cast<int>(comps[i].species.size());
  }
  L_ = make_layout(g, elem_comp, std::move(comps), spec_.has_outflow_boundary, spec_.has_outflow_interface);

  // solver configuration, keys of model/make_step_operator.hh:183-288
  const std::string p = "model.time_step_operator.";
  cfg_.rk = c.get(p + "type", std::string("Alexander2"));
  rk_ = rk_table(cfg_.rk);
  const std::string lin = c.get(p + "linear_solver.type", std::string("BiCGSTAB"));
  if (lin != "BiCGSTAB")
    cfg_.warnings.push_back("linear_solver.type '" + lin + "' runs as BiCGSTAB on the device");
  std::string pre = c.get(p + "linear_solver.preconditioner.type", std::string("Jacobi"));
  if (pre == "BlockJacobi") {
    // block_jacobi.hh:46-81 with an ILU diagonal solver is ILU0 of the one block
    const std::string inner = c.get(p + "linear_solver.preconditioner.diagonal_solver.preconditioner.type", std::string("ILU"));
    pre = inner == "ILU" ? "ILU" : "Jacobi";
  }
  if (pre == "ILU" || pre == "ILU0") cfg_.precond = Preconditioner::ILU0;
  else if (pre == "Jacobi") cfg_.precond = Preconditioner::Jacobi;
  else if (pre == "VertexBlockJacobi") cfg_.precond = Preconditioner::VertexBlockJacobi;
  else if (pre == "None" || pre == "Richardson") cfg_.precond = Preconditioner::None;
  else if (pre == "SSOR" || pre == "SOR" || pre == "GaussSeidel") {
    cfg_.precond = Preconditioner::ILU0;
    cfg_.warnings.push_back("preconditioner '" + pre + "' runs as block ILU0 on the device");
  } else {
    throw NotImplemented("preconditioner '" + pre + "'");
  }
  if (cfg_.precond == Preconditioner::Jacobi) {
    // SeqJac's sweeps and relaxation (istl defaults 1 and 1.0): only one
    // sweep from zero with weight 1 is the plain inverse of the diagonal
    const double w = c.get(p + "linear_solver.preconditioner.relaxation", 1.0);
    const int sweeps = c.get(p + "linear_solver.preconditioner.iterations", 1);
    if (w != 1.0 || sweeps != 1) throw NotImplemented("Jacobi with relaxation != 1 or iterations != 1");
  }
  if (opt_.batch_member && cfg_.precond == Preconditioner::ILU0) {
    cfg_.precond = Preconditioner::VertexBlockJacobi;
    cfg_.warnings.push_back("batch members precondition with vertex block-Jacobi in place of ILU0");
  }
  const std::string prec = c.get(p + "linear_solver.precision", std::string("fp64"));
  if (prec == "fp64") cfg_.precision = Precision::FP64;
  else if (prec == "mixed") cfg_.precision = Precision::Mixed;
  else throw IOError("linear_solver.precision must be 'fp64' or 'mixed'");
  cfg_.matrix_free = c.get(p + "linear_solver.matrix_free", false);  // make_step_operator.hh:199
  if (cfg_.matrix_free) {
    // a matrix-free operator reaches only the preconditioners that need no
    // matrix (solver/istl/factory/preconditioner.hh:95-98); the Richardson
    // step with weight 1 is the identity
    const std::string mp = c.get(p + "linear_solver.preconditioner.type", std::string("Richardson"));
    if (mp != "Richardson" && mp != "None")
      throw IOError("preconditioner '" + mp + "' needs an assembled matrix (linear_solver.matrix_free = true)");
    if (c.get(p + "linear_solver.preconditioner.relaxation", 1.0) != 1.0)
      throw NotImplemented("Richardson with relaxation != 1");
    cfg_.precond = Preconditioner::None;
    if (cfg_.precision != Precision::FP64) throw NotImplemented("mixed precision with the matrix-free Jacobian");
    if (opt_.batch_member) throw NotImplemented("matrix-free batch members");
    if (!all_tile_eligible())
      throw NotImplemented("the matrix-free Jacobian covers structured compartments with the tiled kernel");
  }
  // linear_solver.storage: how the Krylov solver holds the Jacobian.
  // "matrix" assembles the BCSR matrix (as the reference does); "points"
  // keeps only the coefficient blocks of every cell's quadrature points and
  // applies them (with the uniform diffusion) to each Krylov direction, which
  // is the same operator in fewer bytes; "auto" takes points on box grids
  // with blocks of up to 4 species where the model allows it.  Points need
  // compartments on the tiled kernel and a preconditioner built from diagonal
  // blocks (Jacobi, vertex-block Jacobi or none); interface and outflow terms
  // are applied from the linearisation point by the facet kernels.
  {
    const std::string st = c.get(p + "linear_solver.storage", std::string("auto"));
    if (st != "matrix" && st != "points" && st != "auto")
      throw IOError("linear_solver.storage must be 'matrix', 'points' or 'auto'");
    std::string why;
    int with_species = 0;
    for (const auto& C : L_.comps) with_species += C.b > 0 ? 1 : 0;
    if (cfg_.matrix_free) why = "the matrix-free Jacobian";
    else if (opt_.batch_member && (with_species != 1 || !L_.facets.empty()))
      why = "batch members with interface or outflow terms or several compartments";
    else if (cfg_.precond == Preconditioner::ILU0) why = "an ILU0 preconditioner";
    else if (!all_tile_eligible()) why = "compartments outside the tiled kernel";
    if (st == "points" && !why.empty()) throw NotImplemented("linear_solver.storage = points with " + why);
    // auto: box cells, and blocks of up to 4 species (larger blocks keep their
    // corner sums in LDS, where the assembled SpMV is still faster)
    int bmax = 0;
    for (const auto& C : L_.comps) bmax = std::max(bmax, C.b);
    cfg_.points = why.empty() && (st == "points" || (st == "auto" && L_.grid.cube && bmax <= 4));
  }
  {
    const std::string pt = c.get(p + "linear_solver.point_tile", std::string("auto"));
    if (pt != "auto") {
      char* end = nullptr;
      const long k2 = std::strtol(pt.c_str(), &end, 10);
      if (pt.empty() || *end || k2 < 0 || k2 > 15)
        throw IOError("linear_solver.point_tile must be 'auto' or a candidate index in [0, 15]");
      cfg_.point_tile = static_cast<int>(k2);
    }
  }
  cfg_.readback_iterations = c.get(p + "linear_solver.readback_iterations", 0);
  if (cfg_.readback_iterations < 0) throw IOError("linear_solver.readback_iterations must be >= 0");
  cfg_.lin_rtol = c.get(p + "linear_solver.convergence_condition.relative_tolerance", 1e-4);
  auto lr = c.get_list(p + "linear_solver.convergence_condition.iteration_range");
  if (lr.size() == 2) cfg_.lin_maxit = static_cast<int>(lr[1]);
  cfg_.ir_inner_rtol = c.get(p + "linear_solver.refinement.inner_relative_tolerance", 1e-3);
  cfg_.ir_max = c.get(p + "linear_solver.refinement.max_steps", 50);
  cfg_.ir_floor = c.get(p + "linear_solver.refinement.tolerance_floor", 0.0);
  if (cfg_.ir_floor < 0.0 || cfg_.ir_floor >= 1.0)
    throw IOError("linear_solver.refinement.tolerance_floor must be in [0, 1)");
  cfg_.nl_rtol = c.get(p + "nonlinear_solver.convergence_condition.relative_tolerance", 1e-4);
  cfg_.nl_atol = c.get(p + "nonlinear_solver.convergence_condition.absolute_tolerance", 0.0);
  auto nr = c.get_list(p + "nonlinear_solver.convergence_condition.iteration_range");
  if (nr.size() == 2) {
    cfg_.nl_min = static_cast<int>(nr[0]);
    cfg_.nl_max = static_cast<int>(nr[1]);
  }
  cfg_.fixed_linear_tolerance = c.get(p + "nonlinear_solver.dx_inverse_fixed_tolerance", false);
  cfg_.min_linear_reduction = c.get(p + "nonlinear_solver.dx_inverse_min_relative_tolerance", 0.1);
  cfg_.linearization_threshold = c.get(p + "nonlinear_solver.linearization_threshold", 0.0);
  const std::string norm = c.get(p + "nonlinear_solver.norm", std::string("l_2"));
  if (norm == "l_2") cfg_.norm = 0;
  else if (norm == "l_inf") cfg_.norm = 1;
  else if (norm == "l_1") cfg_.norm = 2;
  else throw IOError("Not known nonlinear norm solver of type '" + norm + "'");
  time_ = c.get(p + "time_begin", 0.0);
  t_end_ = c.get(p + "time_end", 1.0);
  dt_initial_ = c.get(p + "time_step_initial", c.get(p + "time_step_max", 0.1));
  gen_ = generate_assembly(spec_, L_, opt_.runtime_parameters, cfg_.points);
}

bool Model::all_tile_eligible() const {
  if (L_.grid.unstructured) return false;
  for (size_t c = 0; c < L_.comps.size(); ++c)
    if (L_.comps[c].b && !tile_eligible(spec_, static_cast<int>(c))) return false;
  return true;
}

bool Model::can_apply_jacobian() const {
  for (size_t c = 0; c < L_.comps.size(); ++c)
    if (L_.comps[c].b && gen_.tile_kernel[c].empty()) return false;
  return true;
}

void Model::require_device() const {
  if (device_ < 0) throw DeviceError("model was created without a device");
}

void Model::synchronize() const {
  if (stream_) DCG_HIP(hipStreamSynchronize(stream_));
}

void Model::setup_device() {
  DCG_HIP(hipSetDevice(device_));
  if (opt_.stream) {
    stream_ = opt_.stream;
    own_stream_ = false;
  } else {
    DCG_HIP(hipStreamCreate(&stream_));
  }
  const long ns = L_.nslot;
  const int B = L_.B;
  row_ptr_.upload(L_.row_ptr, stream_);
  col_.upload(L_.col, stream_);
  own_start_.upload(L_.own_start, stream_);
  own_mask_.upload(L_.own_mask, stream_);
  diag_pos_.upload(L_.diag_pos, stream_);
  std::vector<int> sb(ns);
  std::vector<long> sd(ns);
  for (const auto& C : L_.comps)
    for (long l = 0; l < C.nslot; ++l) {
      sb[C.slot_off + l] = C.b;
      sd[C.slot_off + l] = C.dof_off + l * C.b;
    }
  slot_b_.upload(sb, stream_);
  slot_dof_.upload(sd, stream_);
  // SpMV chunks: rows grouped until their blocks x B fill the LDS stage
  std::vector<int> chunk{0};
  long acc = 0;
  for (long r = 0; r < ns; ++r) {
    long len = static_cast<long>(L_.row_ptr[r + 1] - L_.row_ptr[r]) * B;
    if (len > k::kSpmvStage) throw NotImplemented("BCSR row too long for the SpMV stage");
    if (acc + len > k::kSpmvStage) {
      chunk.push_back(static_cast<int>(r));
      acc = 0;
    }
    acc += len;
  }
  if (chunk.back() != ns) chunk.push_back(static_cast<int>(ns));
  chunk_row_.upload(chunk, stream_);
  chunk_host_ = chunk;
  plan_.nchunk = static_cast<int>(chunk.size()) - 1;
  plan_.chunk_row = chunk_row_.get();
  plan_.rp = row_ptr_.get();
  plan_.col = col_.get();
  plan_.nrows = static_cast<int>(ns);
  // row-split launches: about 8 rows per thread group pass, grid a multiple of 8
  {
    long wgs = std::max<long>(8, std::min<long>(4096, (ns + 63) / 64));
    wgs = (wgs + 7) / 8 * 8;
    plan_.row_wgs = static_cast<int>(wgs);
    plan_.rows_per_wg = static_cast<int>((ns + wgs - 1) / wgs);
  }
  // per-compartment assembly work lists
  comp_dev_.resize(L_.comps.size());
  all_tiled_ = true;
  const Grid& gr = L_.grid;
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    const auto& C = L_.comps[c];
    comp_dev_[c].cells.upload(C.cells, stream_);
    comp_dev_[c].cell_mask.upload(C.cell_mask, stream_);
    comp_dev_[c].cell_slots.upload(C.cell_slots, stream_);
    if (gr.unstructured && C.b) {
      comp_dev_[c].egeo.upload(cell_geometry(gr, C.cells), stream_);
      comp_dev_[c].cpos.upload(C.cell_pos, stream_);
    }
    if (!C.b) continue;
    if (gen_.tile_kernel[c].empty()) {
      all_tiled_ = false;
      continue;
    }
    std::vector<int> g2l(gr.nvert, -1);
    for (long l = 0; l < C.nslot; ++l) g2l[C.nodes[l]] = static_cast<int>(C.slot_off + l);
    std::vector<unsigned char> cm(gr.ncell, 0);
    for (size_t i = 0; i < C.cells.size(); ++i) cm[C.cells[i]] = C.cell_mask[i];
    // the tiles of shape td holding at least one of the compartment's vertices
    auto tile_list = [&](std::array<int, 3> td) {
      const long ntx = (gr.n[0] + 1 + td[0] - 1) / td[0];
      const long nty = gr.dim > 1 ? (gr.n[1] + 1 + td[1] - 1) / td[1] : 1;
      const long ntz = gr.dim > 2 ? (gr.n[2] + 1 + td[2] - 1) / td[2] : 1;
      std::vector<char> used(ntx * nty * ntz, 0);
      for (long l = 0; l < C.nslot; ++l) {
        long v = C.nodes[l];
        long i = v % (gr.n[0] + 1), rest = v / (gr.n[0] + 1);
        long j = gr.dim > 1 ? rest % (gr.n[1] + 1) : 0, k = gr.dim > 2 ? rest / (gr.n[1] + 1) : 0;
        used[(i / td[0]) + ntx * ((gr.dim > 1 ? j / td[1] : 0) + nty * (gr.dim > 2 ? k / td[2] : 0))] = 1;
      }
      std::vector<int> out;
      for (long t = 0; t < ntx * nty * ntz; ++t)
        if (used[t]) out.push_back(static_cast<int>(t));
      return out;
    };
    std::vector<int> tiles = tile_list(gen_.tile_dims[c]);
    const auto& shapes = gen_.apply_tiles[c];
    comp_dev_[c].atiles.resize(shapes.size());
    comp_dev_[c].natiles.assign(shapes.size(), 0);
    for (size_t k2 = 0; k2 < shapes.size(); ++k2) {
      auto at = tile_list(shapes[k2]);
      comp_dev_[c].natiles[k2] = static_cast<int>(at.size());
      comp_dev_[c].atiles[k2].upload(at, stream_);
    }
    comp_dev_[c].g2l.upload(g2l, stream_);
    {
      bool id = C.slot_off == 0 && C.nslot == gr.nvert;
      for (long l = 0; id && l < C.nslot; ++l) id = C.nodes[l] == l;
      comp_dev_[c].g2l_identity = id;
    }
    comp_dev_[c].cell_mask_grid.upload(cm, stream_);
    comp_dev_[c].tiles.upload(tiles, stream_);
    comp_dev_[c].ntiles = static_cast<int>(tiles.size());
  }
  {
    // blocks outside a row's own-compartment run couple two compartments
    std::vector<int> extra;
    for (long s = 0; s < ns; ++s) {
      int own_end = L_.own_start[s] + __builtin_popcount(L_.own_mask[s]);
      for (int p2 = L_.row_ptr[s]; p2 < L_.row_ptr[s + 1]; ++p2)
        if (p2 < L_.own_start[s] || p2 >= own_end) extra.push_back(p2);
    }
    extra_pos_.upload(extra, stream_);
  }
  // facet records per facet kernel, grouped by colour
  facet_dev_.resize(gen_.facet_kernels.size());
  for (size_t k = 0; k < gen_.facet_kernels.size(); ++k) {
    const auto& fk = gen_.facet_kernels[k];
    std::vector<int> fi;
    std::vector<double> fd;
    auto& dev = facet_dev_[k];
    int count = 0;
    for (size_t col = 0; col + 1 < L_.facet_color_ptr.size(); ++col) {
      int first = count;
      for (int f = L_.facet_color_ptr[col]; f < L_.facet_color_ptr[col + 1]; ++f) {
        const Facet& F = L_.facets[f];
        if (F.comp_in != fk.comp_in || F.comp_out != fk.comp_out || F.boundary != fk.boundary) continue;
        for (int v = 0; v < fk.nvf; ++v) fi.push_back(F.slot_in[v]);
        for (int v = 0; v < fk.nvf; ++v) fi.push_back(F.boundary ? F.slot_in[v] : F.slot_out[v]);
        auto pos = [&](int rs, int cs) -> int {
          if (rs < 0 || cs < 0) return 0;
          auto b = L_.col.begin() + L_.row_ptr[rs], e = L_.col.begin() + L_.row_ptr[rs + 1];
          auto it = std::lower_bound(b, e, cs);
          if (it == e || *it != cs) throw Error("internal: facet block missing from pattern");
          return static_cast<int>(it - L_.col.begin());
        };
        const auto& si = F.slot_in;
        const auto& so = F.boundary ? F.slot_in : F.slot_out;
        for (int a = 0; a < fk.nvf; ++a)
          for (int b = 0; b < fk.nvf; ++b) fi.push_back(pos(si[a], si[b]));
        for (int a = 0; a < fk.nvf; ++a)
          for (int b = 0; b < fk.nvf; ++b) fi.push_back(F.boundary ? 0 : pos(si[a], so[b]));
        for (int a = 0; a < fk.nvf; ++a)
          for (int b = 0; b < fk.nvf; ++b) fi.push_back(F.boundary ? 0 : pos(so[a], si[b]));
        for (int a = 0; a < fk.nvf; ++a)
          for (int b = 0; b < fk.nvf; ++b) fi.push_back(F.boundary ? 0 : pos(so[a], so[b]));
        for (int q = 0; q < fk.nqf; ++q) fd.push_back(F.qw[q]);
        for (int q = 0; q < fk.nqf; ++q)
          for (int k2 = 0; k2 < 3; ++k2) fd.push_back(F.qx[3 * q + k2]);
        for (int k2 = 0; k2 < 3; ++k2) fd.push_back(F.qn[k2]);
        fd.push_back(F.volume);
        ++count;
      }
      if (count > first) dev.color_ranges.push_back({first, count - first});
    }
    dev.fint.upload(fi, stream_);
    dev.fdbl.upload(fd, stream_);
  }
  // lumped mass weights (integral of each hat function over its compartment)
  {
    std::vector<double> w(ns, 0.0);
    Rule r = element_rule(L_.grid.dim, L_.grid.cube);
    double det = 1.0;
    for (int k2 = 0; k2 < L_.grid.dim; ++k2) det *= L_.grid.h[k2];
    std::vector<double> wint(L_.grid.nvpe, 0.0);
    for (size_t q = 0; q < r.w.size(); ++q) {
      double x[3] = {r.x[q][0], r.x[q][1], r.x[q][2]};
      // hat-function values on the reference element
      for (int a = 0; a < L_.grid.nvpe; ++a) {
        double phi;
        if (L_.grid.cube) {
          phi = 1.0;
          for (int k2 = 0; k2 < L_.grid.dim; ++k2) phi *= ((a >> k2) & 1) ? x[k2] : 1.0 - x[k2];
        } else {
          double s = 0;
          for (int k2 = 0; k2 < L_.grid.dim; ++k2) s += x[k2];
          phi = a == 0 ? 1.0 - s : x[a - 1];
        }
        wint[a] += r.w[q] * det * phi;
      }
    }
    if (L_.grid.unstructured)
      for (auto& v : wint) v /= det;  // reference-simplex integrals; each cell scales by its own |det J|
    for (const auto& C : L_.comps)
      for (size_t i = 0; i < C.cells.size(); ++i) {
        const double s = L_.grid.unstructured ? cell_det(L_.grid, C.cells[i]) : 1.0;
        for (int cls = 0; cls < L_.grid.nclass; ++cls)
          if ((C.cell_mask[i] >> cls) & 1)
            for (int a = 0; a < L_.grid.nvpe; ++a)
              w[C.cell_slots[i * L_.grid.ncorner + L_.grid.corner[cls][a]]] += s * wint[a];
      }
    weights_.upload(w, stream_);
  }
  const size_t nv = static_cast<size_t>(ns) * B;
  u_.resize(nv);
  hist_.resize(nv);
  r_.resize(nv);
  z_.resize(nv);
  xs_.resize(nv);
  packed_.resize(std::max<size_t>(1, static_cast<size_t>(L_.ndof)));
  stage_.resize(rk_.stages + 1);
  for (auto& s : stage_) s.resize(nv);
  use_points_ = cfg_.points;
  if (cfg_.matrix_free) {
    lin_x_.resize(nv);
  } else if (cfg_.points) {
    pc_.resize(L_.comps.size());
    if (cfg_.precision == Precision::Mixed) pc32_.resize(L_.comps.size());
    for (size_t c = 0; c < L_.comps.size(); ++c) {
      // every cell of the grid, word-major (cells outside the compartment
      // are never read)
      if (!gen_.point_words[c] || comp_dev_[c].ntiles <= 0) continue;
      const size_t words = static_cast<size_t>(gen_.point_words[c]) * static_cast<size_t>(L_.grid.ncell);
      pc_[c].resize(words);
      if (cfg_.precision == Precision::Mixed) pc32_[c].resize(words);
    }
    pdiag_.resize(static_cast<size_t>(ns) * B * B);
    pc_tca_.resize(2);
    if (!L_.facets.empty()) lin_x_.resize(nv);  // the facet terms apply from the linearisation point
    dinv_.resize(static_cast<size_t>(ns) * B * B);
    std::vector<int> id(ns);
    for (long i = 0; i < ns; ++i) id[i] = static_cast<int>(i);
    diag_id_.upload(id, stream_);
  } else {
    val_.resize(static_cast<size_t>(matrix_values()));
    dinv_.resize(static_cast<size_t>(ns) * B * B);
  }
  for (auto& v : kv_) v.resize(nv);
  if (cfg_.precision == Precision::Mixed) {
    if (!cfg_.points) val32_.resize(static_cast<size_t>(matrix_values()));
    for (auto& v : kf_) v.resize(nv);
  }
  const size_t nparts = std::max<size_t>(std::max<size_t>(static_cast<size_t>(plan_.nchunk), plan_.row_wgs) + 8,
                                         static_cast<size_t>((ns + 255) / 256) + 8);
  size_t ntiles_all = 0;
  for (const auto& cd : comp_dev_) {
    int most = std::max(0, cd.ntiles);
    for (int n2 : cd.natiles) most = std::max(most, n2);
    ntiles_all += static_cast<size_t>(most);
  }
  // operator partial dots first, then the s-update's partial norms (their own
  // region: they are finalised after the next operator launch)
  spart_off_ = std::max<size_t>(std::max<size_t>(2 * nparts, 2 * ntiles_all + 16), 8 * 1024 + 64);
  part_.resize(spart_off_ + static_cast<size_t>((L_.nslot + 255) / 256) + 64);
  scalar_out_.resize(16);
  sc_.resize(1);
  status_.resize(4);
  k::fill(u_.get(), 0.0, static_cast<long>(nv), stream_);
  k::fill(status_.get(), 0, 4, stream_);
  if (cfg_.precond == Preconditioner::ILU0) {
    // level schedules of the block lower / upper triangular factors
    std::vector<int> lev(ns, 0);
    int nlev = 0;
    for (long i = 0; i < ns; ++i) {
      int l = 0;
      for (int p2 = L_.row_ptr[i]; p2 < L_.diag_pos[i]; ++p2) l = std::max(l, lev[L_.col[p2]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
    std::vector<int> cnt(nlev + 1, 0), rows(ns);
    for (long i = 0; i < ns; ++i) cnt[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) cnt[l + 1] += cnt[l];
    lo_ptr_ = cnt;
    for (long i = 0; i < ns; ++i) rows[cnt[lev[i]]++] = static_cast<int>(i);
    lo_rows_.upload(rows, stream_);
    std::fill(lev.begin(), lev.end(), 0);
    nlev = 0;
    for (long i = ns - 1; i >= 0; --i) {
      int l = 0;
      for (int p2 = L_.diag_pos[i] + 1; p2 < L_.row_ptr[i + 1]; ++p2) l = std::max(l, lev[L_.col[p2]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
    std::vector<int> cu(nlev + 1, 0), rows_u(ns);
    for (long i = 0; i < ns; ++i) cu[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) cu[l + 1] += cu[l];
    up_ptr_ = cu;
    for (long i = 0; i < ns; ++i) rows_u[cu[lev[i]]++] = static_cast<int>(i);
    up_rows_.upload(rows_u, stream_);
    // narrow levels (structured grids: a few hundred rows per anti-diagonal)
    // sweep every level in one workgroup instead of one launch per level
    int widest = 0;
    for (size_t l = 0; l + 1 < lo_ptr_.size(); ++l) widest = std::max(widest, lo_ptr_[l + 1] - lo_ptr_[l]);
    for (size_t l = 0; l + 1 < up_ptr_.size(); ++l) widest = std::max(widest, up_ptr_[l + 1] - up_ptr_[l]);
    ilu_sweep_ = widest <= kIluSweepWidth;
    lo_lev_.upload(lo_ptr_, stream_);
    up_lev_.upload(up_ptr_, stream_);
    if (ilu_sweep_) build_ilu_sweep_tables(rows, rows_u);
    lu_.resize(static_cast<size_t>(matrix_values()));
  }
  {
    // geometry words (spacing, reciprocals, cell measure) + parameters
    std::vector<double> prm(kGeoWords, 0.0);
    double cell = 1.0;
    for (int k2 = 0; k2 < 3; ++k2) {
      const bool on = k2 < L_.grid.dim;
      prm[k2] = on ? L_.grid.h[k2] : 0.0;
      prm[3 + k2] = on ? 1.0 / L_.grid.h[k2] : 0.0;
      if (on) cell *= L_.grid.h[k2];
    }
    prm[6] = cell;
    prm.insert(prm.end(), spec_.param_values.begin(), spec_.param_values.end());
    prm_.upload(prm, stream_);
  }
  if (!opt_.batch_member && !cfg_.matrix_free && !cfg_.points) tune_spmv();
  synchronize();
}

// Times every SpMV variant valid for this block size on the model's own
// pattern (zero values, as in the solver: x, y and two dot operands) and keeps
// the fastest.
void Model::tune_spmv() {
  tuning_.clear();
  plan_.variant = tune_spmv_plan(false);
  plan32_ = plan_;
  if (cfg_.precision == Precision::Mixed) plan32_.variant = tune_spmv_plan(true);
}

int Model::tune_spmv_plan(bool fp32) {
  const long nv = L_.nslot * L_.B;
  k::fill(val_.get(), 0.0, matrix_values(), stream_);
  for (int v = 0; v < 4; ++v) k::fill(kv_[v].get(), 1.0, nv, stream_);
  if (fp32) {
    k::fill(val32_.get(), 0.0f, matrix_values(), stream_);
    for (int v = 0; v < 4; ++v) k::fill(kf_[v].get(), 1.0f, nv, stream_);
  }
  k::SpmvPlan pl = plan_;
  hipEvent_t e0, e1;
  DCG_HIP(hipEventCreate(&e0));
  DCG_HIP(hipEventCreate(&e1));
  double best = 1e30;
  int best_v = 0;
  auto run = [&]() {
    if (fp32)
      k::spmv(L_.B, pl, val32_.get(), kf_[0].get(), kf_[1].get(), kf_[2].get(), kf_[3].get(), part_.get(), nullptr, 0,
              stream_);
    else
      k::spmv(L_.B, pl, val_.get(), kv_[0].get(), kv_[1].get(), kv_[2].get(), kv_[3].get(), part_.get(), nullptr, 0,
              stream_);
  };
  for (int variant : {5, 0, 1, 2, 3, 4}) {
    const int G = (variant == 0 || variant == 5) ? 0 : (4 << variant);  // 8, 16, 32, 64
    if (G && G % L_.B != 0) continue;
    pl.variant = variant;
    for (int w = 0; w < 3; ++w) run();
    const int reps = 5;
    DCG_HIP(hipEventRecord(e0, stream_));
    for (int w = 0; w < reps; ++w) run();
    DCG_HIP(hipEventRecord(e1, stream_));
    DCG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DCG_HIP(hipEventElapsedTime(&ms, e0, e1));
    const double us = 1e3 * ms / reps;
    tuning_.push_back({variant + (fp32 ? 100 : 0), us});
    if (us < best) {
      best = us;
      best_v = variant;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best_v;
}

// Times the point apply with every candidate tile shape on the model's own
// grid, as the Krylov solver launches it (with its two partial dots), and
// keeps the fastest per compartment.  Coefficients and vectors are zero (the
// work does not depend on the values).  Before each timed launch a copy
// through a buffer larger than the chip's last-level cache evicts the
// coefficients, as the solver's vector updates do between two applies.
void Model::tune_points() {
  if (!cfg_.points || opt_.batch_member) return;
  point_tuning_.clear();
  if (cfg_.point_tile >= 0) {
    for (size_t c = 0; c < L_.comps.size(); ++c) {
      const int nk = static_cast<int>(pc_app_fns_[c].size());
      apply_shape_[c] = cfg_.point_tile < nk ? cfg_.point_tile : 0;
      for (int k2 = 0; k2 < nk; ++k2)
        point_tuning_.push_back({static_cast<int>(c), gen_.apply_tiles[c][k2], gen_.apply_waves[c][k2], 0.0,
                                 k2 == apply_shape_[c]});
    }
    return;
  }
  bool any = false;
  for (const auto& f : pc_app_fns_) any = any || f.size() > 1;
  if (!any) return;
  size_t pc_bytes = 0;
  for (auto& a : pc_)
    if (a.size()) {
      k::fill(a.get(), 0.0, static_cast<long>(a.size()), stream_);
      pc_bytes += a.size() * sizeof(double);
    }
  for (int i = 0; i < 4; ++i) k::fill(kv_[i].get(), 0.0, L_.nslot * L_.B, stream_);
  k::fill(pc_tca_.get(), 0.0, 2, stream_);
  // the eviction copy: only where the coefficients could stay cached
  constexpr long kFlushBytes = 384L << 20;
  DeviceArray<char> fa, fb;
  if (pc_bytes > (8u << 20)) {
    try {
      fa.resize(kFlushBytes);
      fb.resize(kFlushBytes);
      DCG_HIP(hipMemsetAsync(fa.get(), 0, kFlushBytes, stream_));
    } catch (const std::exception&) {
      (void)hipGetLastError();  // no room: time without the eviction
      fa = DeviceArray<char>();
      fb = DeviceArray<char>();
    }
  }
  hipEvent_t e0, e1;
  DCG_HIP(hipEventCreate(&e0));
  DCG_HIP(hipEventCreate(&e1));
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    const size_t nk = pc_app_fns_[c].size();
    if (nk < 2) continue;
    std::vector<double> ms(nk, 0.0);
    constexpr int kRounds = 4;
    for (int round = 0; round < kRounds; ++round)
      for (size_t k2 = 0; k2 < nk; ++k2) {
        apply_shape_[c] = static_cast<int>(k2);
        if (fa.size()) k::probe_copy(fa.get(), fb.get(), kFlushBytes, stream_);
        DCG_HIP(hipEventRecord(e0, stream_));
        launch_points_apply(kv_[0].get(), kv_[1].get(), nullptr, false, kv_[2].get(), kv_[3].get(), part_.get());
        DCG_HIP(hipEventRecord(e1, stream_));
        DCG_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        DCG_HIP(hipEventElapsedTime(&t, e0, e1));
        if (round) ms[k2] += t;  // the first round warms up
      }
    size_t best = 0;
    for (size_t k2 = 1; k2 < nk; ++k2)
      if (ms[k2] < ms[best]) best = k2;
    for (size_t k2 = 0; k2 < nk; ++k2)
      point_tuning_.push_back({static_cast<int>(c), gen_.apply_tiles[c][k2], gen_.apply_waves[c][k2],
                               1e3 * ms[k2] / (kRounds - 1), k2 == best});
    apply_shape_[c] = static_cast<int>(best);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  synchronize();
}

void Model::compile_and_load() {
  module_ = rtc_module(gen_.source, "dcopasi_assembly", device_);
  vol_fn_.assign(L_.comps.size(), nullptr);
  tile_fn_.assign(L_.comps.size(), nullptr);
  tile_jac_fn_.assign(L_.comps.size(), nullptr);
  tile_app_fn_.assign(L_.comps.size(), nullptr);
  pc_set_fn_.assign(L_.comps.size(), nullptr);
  pc_app_fn_.assign(L_.comps.size(), nullptr);
  pc_appf_fn_.assign(L_.comps.size(), nullptr);
  pc_app_fns_.assign(L_.comps.size(), {});
  pc_appf_fns_.assign(L_.comps.size(), {});
  apply_shape_.assign(L_.comps.size(), 0);
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    if (!gen_.volume_kernel[c].empty()) vol_fn_[c] = module_->function(gen_.volume_kernel[c]);
    if (!gen_.tile_kernel[c].empty()) {
      tile_fn_[c] = module_->function(gen_.tile_kernel[c] + "_r");
      tile_jac_fn_[c] = module_->function(gen_.tile_kernel[c] + "_j");
      tile_app_fn_[c] = module_->function(gen_.tile_kernel[c] + "_a");
      if (cfg_.points) {
        pc_set_fn_[c] = module_->function(gen_.tile_kernel[c] + "_q");
        for (size_t k2 = 0; k2 < gen_.apply_tiles[c].size(); ++k2) {
          pc_app_fns_[c].push_back(module_->function(gen_.tile_kernel[c] + "_p" + std::to_string(k2)));
          pc_appf_fns_[c].push_back(module_->function(gen_.tile_kernel[c] + "_pf" + std::to_string(k2)));
        }
        pc_app_fn_[c] = pc_app_fns_[c].front();
        pc_appf_fn_[c] = pc_appf_fns_[c].front();
      }
    }
  }
  facet_fn_.clear();
  facetf_fn_.clear();
  for (const auto& fk : gen_.facet_kernels) {
    facet_fn_.push_back(module_->function(fk.name));
    facetf_fn_.push_back(cfg_.points ? module_->function(fk.name + "_f") : nullptr);
  }
}

// ------------------------------------------------------------------ assembly
void Model::launch_assembly(const double* x, double t, double cm, double ca, const double* base, double* r,
                            double* val, int mode, bool skip_tiled, int phases) {
  const Grid& g = L_.grid;
  const long n = L_.nslot * L_.B;
  const double* base_eff = base;
  if ((mode & 1) && !all_tiled_) {
    // the coloured kernels accumulate: start from base
    if (phases & 1) {
      if (!base) k::fill(r, 0.0, n, stream_);
      else if (base != r)
        DCG_HIP(hipMemcpyAsync(r, base, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    }
    base_eff = r;
  }
  if ((mode & 2) && (phases & 1)) {
    if (!all_tiled_) {
      k::init_jacobian(val, matrix_values(), diag_pos_.get(), slot_b_.get(), static_cast<int>(L_.nslot), L_.B,
                       stream_);
    } else if (extra_pos_.size()) {
      k::zero_blocks(val, extra_pos_.get(), static_cast<long>(extra_pos_.size()), L_.B, stream_);
    }
  }
  int n0 = static_cast<int>(g.n[0]), n1 = g.dim > 1 ? static_cast<int>(g.n[1]) : 1;
  int n2 = g.dim > 2 ? static_cast<int>(g.n[2]) : 1;
  double ox = g.origin[0], oy = g.origin[1], oz = g.origin[2];
  double hx = g.h[0], hy = g.dim > 1 ? g.h[1] : 0.0, hz = g.dim > 2 ? g.h[2] : 0.0;
  const int* own_start = own_start_.get();
  const unsigned* own_mask = own_mask_.get();
  const double* prm = prm_.get();
  for (size_t c = 0; c < L_.comps.size() && (phases & 2); ++c) {
    if (!vol_fn_[c] && !tile_fn_[c]) continue;
    const auto& C = L_.comps[c];
    if (tile_fn_[c]) {
      if (skip_tiled) continue;
      const int* tiles = comp_dev_[c].tiles.get();
      const int* g2l = comp_dev_[c].g2l.get();
      const unsigned char* cmask = comp_dev_[c].cell_mask_grid.get();
      const int* stop = nullptr;
      // residual and Jacobian are separate entry points (each with its own
      // registers and LDS); both read only x, so their order is free
      for (int part : {1, 2}) {
        if (!(mode & part) || comp_dev_[c].ntiles <= 0) continue;
        int pmode = part;
        void* args[] = {&tiles, &g2l, &cmask, &x, &base_eff, &t, &cm, &ca, &r, &val, &own_start, &own_mask,
                        &n0, &n1, &n2, &ox, &oy, &oz, &hx, &hy, &hz, &pmode, &prm, &stop};
        DCG_HIP(hipModuleLaunchKernel(part == 1 ? tile_fn_[c] : tile_jac_fn_[c], comp_dev_[c].ntiles, 1, 1, 256, 1, 1,
                                      0, stream_, args, nullptr));
      }
      continue;
    }
    const int* cells = comp_dev_[c].cells.get();
    const unsigned char* cmask = comp_dev_[c].cell_mask.get();
    const int* cslots = comp_dev_[c].cell_slots.get();
    // unstructured cells find their blocks in cpos and carry their geometry
    const int* starts = g.unstructured ? comp_dev_[c].cpos.get() : own_start;
    const double* egeo = comp_dev_[c].egeo.get();
    for (size_t col = 0; col + 1 < C.color_ptr.size(); ++col) {
      int first = C.color_ptr[col], count = C.color_ptr[col + 1] - first;
      if (count <= 0) continue;
      void* args[] = {&cells, &cmask, &cslots, &first, &count, &x, &t, &cm, &ca, &r, &val, &starts,
                      &own_mask, &n0, &n1, &ox, &oy, &oz, &hx, &hy, &hz, &mode, &prm, &egeo};
      const unsigned block = 128, grid = static_cast<unsigned>((count + block - 1) / block);
      DCG_HIP(hipModuleLaunchKernel(vol_fn_[c], grid, 1, 1, block, 1, 1, 0, stream_, args, nullptr));
    }
  }
  if (ca == 0.0 || !(phases & 4)) return;  // transmission terms belong to the stiffness operator
  for (size_t k = 0; k < facet_fn_.size(); ++k) {
    const int* fint = facet_dev_[k].fint.get();
    const double* fdbl = facet_dev_[k].fdbl.get();
    for (auto [first, count] : facet_dev_[k].color_ranges) {
      const double* dir = nullptr;
      void* args[] = {&fint, &fdbl, &first, &count, &x, &t, &ca, &r, &val, &mode, &prm, &dir};
      const unsigned block = 64, grid = static_cast<unsigned>((count + block - 1) / block);
      DCG_HIP(hipModuleLaunchKernel(facet_fn_[k], grid, 1, 1, block, 1, 1, 0, stream_, args, nullptr));
    }
  }
}

// The matrix-free Jacobian apply (local_operator.hh:510-524 with the
// matrix-free instationary assembler, make_step_operator.hh:291-293): every
// tile recomputes its cells' point coefficients at `lin` and applies the
// block rows it owns to v, so y holds J(lin) v for all rows (padding species
// rows are the identity, as in the assembled Jacobian).
void Model::launch_apply(const double* lin, double t, double cm, double ca, const double* v, double* y,
                         const int* stop) {
  if (!can_apply_jacobian()) throw NotImplemented("Jacobian apply without the tiled kernel");
  const Grid& g = L_.grid;
  int n0 = static_cast<int>(g.n[0]), n1 = g.dim > 1 ? static_cast<int>(g.n[1]) : 1;
  int n2 = g.dim > 2 ? static_cast<int>(g.n[2]) : 1;
  double ox = g.origin[0], oy = g.origin[1], oz = g.origin[2];
  double hx = g.h[0], hy = g.dim > 1 ? g.h[1] : 0.0, hz = g.dim > 2 ? g.h[2] : 0.0;
  const int* own_start = own_start_.get();
  const unsigned* own_mask = own_mask_.get();
  const double* prm = prm_.get();
  double* val = nullptr;
  int mode = 4;
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    if (!tile_fn_[c] || comp_dev_[c].ntiles <= 0) continue;
    const int* tiles = comp_dev_[c].tiles.get();
    const int* g2l = comp_dev_[c].g2l.get();
    const unsigned char* cmask = comp_dev_[c].cell_mask_grid.get();
    void* args[] = {&tiles, &g2l, &cmask, &lin, &v, &t, &cm, &ca, &y, &val, &own_start, &own_mask,
                    &n0, &n1, &n2, &ox, &oy, &oz, &hx, &hy, &hz, &mode, &prm, &stop};
    DCG_HIP(hipModuleLaunchKernel(tile_app_fn_[c], comp_dev_[c].ntiles, 1, 1, 256, 1, 1, 0, stream_, args, nullptr));
  }
  // transmission and outflow blocks added on top, colour by colour
  // (localAssembleJacobianSkeletonApply / BoundaryApply, local_operator.hh:1354-1452)
  if (ca == 0.0) return;
  for (size_t k = 0; k < facet_fn_.size(); ++k) {
    const int* fint = facet_dev_[k].fint.get();
    const double* fdbl = facet_dev_[k].fdbl.get();
    for (auto [first, count] : facet_dev_[k].color_ranges) {
      double* y_out = y;
      void* fargs[] = {&fint, &fdbl, &first, &count, &lin, &t, &ca, &y_out, &val, &mode, &prm, &v};
      const unsigned block = 64, grid = static_cast<unsigned>((count + block - 1) / block);
      DCG_HIP(hipModuleLaunchKernel(facet_fn_[k], grid, 1, 1, block, 1, 1, 0, stream_, fargs, nullptr));
    }
  }
}

// Point coefficients at x (the cells' quadrature-point blocks of the mass and
// reaction Jacobian, weights cm and ca folded in) and the rows' diagonal
// blocks, compartment by compartment; the uniform diffusion is applied from
// its tables, so t and ca are kept for the applies.
void Model::launch_points_setup(const double* x, double t, double cm, double ca) {
  const Grid& g = L_.grid;
  int n0 = static_cast<int>(g.n[0]), n1 = g.dim > 1 ? static_cast<int>(g.n[1]) : 1;
  int n2 = g.dim > 2 ? static_cast<int>(g.n[2]) : 1;
  double ox = g.origin[0], oy = g.origin[1], oz = g.origin[2];
  double hx = g.h[0], hy = g.dim > 1 ? g.h[1] : 0.0, hz = g.dim > 2 ? g.h[2] : 0.0;
  const int* own_start = own_start_.get();
  const unsigned* own_mask = own_mask_.get();
  const double* prm = prm_.get();
  const double* base = nullptr;
  const int* stop = nullptr;
  double* diag = pdiag_.get();
  int mode = 8;
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    if (!pc_set_fn_[c] || comp_dev_[c].ntiles <= 0) continue;
    const int* tiles = comp_dev_[c].tiles.get();
    const int* g2l = comp_dev_[c].g2l.get();
    const unsigned char* cmask = comp_dev_[c].cell_mask_grid.get();
    double* pc = pc_[c].get();
    void* args[] = {&tiles, &g2l, &cmask, &x, &base, &t, &cm, &ca, &diag, &pc, &own_start, &own_mask,
                    &n0, &n1, &n2, &ox, &oy, &oz, &hx, &hy, &hz, &mode, &prm, &stop};
    DCG_HIP(hipModuleLaunchKernel(pc_set_fn_[c], comp_dev_[c].ntiles, 1, 1, 256, 1, 1, 0, stream_, args, nullptr));
  }
  k::fill(pc_tca_.get(), t, 1, stream_);
  k::fill(pc_tca_.get() + 1, ca, 1, stream_);
  if (facet_fn_.empty()) return;
  // interface and outflow terms: their diagonal blocks join the rows' ones,
  // and the applies evaluate them at the linearisation point (a copy of x)
  DCG_HIP(hipMemcpyAsync(lin_x_.get(), x, lin_x_.bytes(), hipMemcpyDeviceToDevice, stream_));
  lin_t_ = t;
  lin_ca_ = ca;
  if (ca == 0.0) return;  // they belong to the stiffness operator
  const double* lin = lin_x_.get();
  double* r = nullptr;
  const double* dir = nullptr;
  int fmode = 8;
  for (size_t k2 = 0; k2 < facet_fn_.size(); ++k2) {
    const int* fint = facet_dev_[k2].fint.get();
    const double* fdbl = facet_dev_[k2].fdbl.get();
    for (auto [first, count] : facet_dev_[k2].color_ranges) {
      void* fargs[] = {&fint, &fdbl, &first, &count, &lin, &t, &ca, &r, &diag, &fmode, &prm, &dir};
      const unsigned block = 64, grid = static_cast<unsigned>((count + block - 1) / block);
      DCG_HIP(hipModuleLaunchKernel(facet_fn_[k2], grid, 1, 1, block, 1, 1, 0, stream_, fargs, nullptr));
    }
  }
}

// y = J v from the stored point coefficients: every compartment's tiles write
// the rows they own (padding species rows are the identity) and, with part,
// one pair of partial dots each; interface and outflow terms are then added
// by the facet kernels from the linearisation point, and no partial dots are
// written (the returned count is 0: the caller forms the dots).
int Model::launch_points_apply(const void* v, void* y, const int* stop, bool fp32, const void* w1, const void* w2,
                               double* part) {
  const Grid& g = L_.grid;
  int n0 = static_cast<int>(g.n[0]), n1 = g.dim > 1 ? static_cast<int>(g.n[1]) : 1;
  int n2 = g.dim > 2 ? static_cast<int>(g.n[2]) : 1;
  const double* prm = prm_.get();
  const double* tca = pc_tca_.get();
  int nparts = 0;
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    if (!pc_app_fn_[c] || comp_dev_[c].ntiles <= 0) continue;
    const int* g2l = comp_dev_[c].g2l_identity ? nullptr : comp_dev_[c].g2l.get();
    const unsigned char* cmask = comp_dev_[c].cell_mask_grid.get();
    const void* pc = fp32 ? static_cast<const void*>(pc32_[c].get()) : static_cast<const void*>(pc_[c].get());
    double* pp = part && facet_fn_.empty() ? part + 2 * static_cast<size_t>(nparts) : nullptr;
    // the apply's own tile shape (chosen at set-up) and its tile list
    const int k2 = apply_shape_[c];
    const int* tiles = comp_dev_[c].atiles[k2].get();
    int ntile = comp_dev_[c].natiles[k2];
    void* args[] = {&tiles, &ntile, &g2l, &cmask, &pc, &v, &y, &tca, &n0, &n1, &n2, &prm, &stop, &w1, &w2, &pp};
    const unsigned grid = static_cast<unsigned>((ntile + 7) / 8 * 8);  // the kernel maps workgroups to XCD runs
    const unsigned nt = static_cast<unsigned>(
        apply_tile_threads(L_.grid.dim, gen_.apply_tiles[c][k2], L_.comps[c].b));
    DCG_HIP(hipModuleLaunchKernel(fp32 ? pc_appf_fns_[c][k2] : pc_app_fns_[c][k2], grid, 1, 1, nt, 1, 1, 0, stream_,
                                  args, nullptr));
    nparts += ntile;
  }
  if (facet_fn_.empty()) return nparts;
  if (lin_ca_ != 0.0) {
    const double* lin = lin_x_.get();
    double t2 = lin_t_, ca2 = lin_ca_;
    double* val = nullptr;
    int fmode = 4;
    for (size_t k2 = 0; k2 < facet_fn_.size(); ++k2) {
      const int* fint = facet_dev_[k2].fint.get();
      const double* fdbl = facet_dev_[k2].fdbl.get();
      for (auto [first, count] : facet_dev_[k2].color_ranges) {
        void* fargs[] = {&fint, &fdbl, &first, &count, &lin, &t2, &ca2, &y, &val, &fmode, &prm, &v};
        const unsigned block = 64, grid = static_cast<unsigned>((count + block - 1) / block);
        DCG_HIP(hipModuleLaunchKernel(fp32 ? facetf_fn_[k2] : facet_fn_[k2], grid, 1, 1, block, 1, 1, 0, stream_,
                                      fargs, nullptr));
      }
    }
  }
  return 0;
}

void Model::ensure_matrix() {
  if (val_.size()) return;
  val_.resize(static_cast<size_t>(matrix_values()));
  if (cfg_.precision == Precision::Mixed) val32_.resize(static_cast<size_t>(matrix_values()));
  if (!dinv_.size()) dinv_.resize(static_cast<size_t>(L_.nslot) * L_.B * L_.B);
  if (cfg_.precond == Preconditioner::ILU0 && !lu_.size()) lu_.resize(static_cast<size_t>(matrix_values()));
  plan32_ = plan_;
}

double Model::residual_norm(const double* r) {
  const long n = L_.nslot * L_.B;
  k::reduce(n, r, nullptr, cfg_.norm == 0 ? 0 : cfg_.norm == 1 ? 2 : 1, part_.get(), scalar_out_.get(), stream_);
  double v = 0;
  DCG_HIP(hipMemcpyAsync(&v, scalar_out_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  return v;  // two_norm2 for l_2 (make_step_operator.hh:274-276)
}

// ------------------------------------------------------------------ Krylov
Model::KrylovWork Model::own_work() {
  KrylovWork w;
  for (int i = 0; i < 9; ++i) w.kv[i] = kv_[i].get();
  w.part = part_.get();
  w.spart = part_.get() + spart_off_;
  w.sc = sc_.get();
  w.val = val_.get();
  w.dinv = dinv_.get();
  w.lu = lu_.get();
  w.status = status_.get();
  return w;
}

// Tables of the one-workgroup ILU0 sweeps from the level schedules (see
// k::IluSweep): per row position its off-diagonal blocks and the LDS ring
// slot of each block's column; the levels in chunks for the staged apply;
// the factorisation's update pairs per lower block (in the order a merge of
// the two sorted rows visits them) laid out as each row's inbox slots and
// pushes; and the LDS sizes that decide which kernels run (the plain sweeps
// where the stages do not fit).
void Model::build_ilu_sweep_tables(const std::vector<int>& lrows, const std::vector<int>& urows) {
  const long ns = L_.nslot;
  const int B = L_.B;
  const auto& rp = L_.row_ptr;
  const auto& col = L_.col;
  const auto& dp = L_.diag_pos;
  k::IluSweep t;
  t.nlo = static_cast<int>(lo_ptr_.size()) - 1;
  t.nup = static_cast<int>(up_ptr_.size()) - 1;
  t.lo_lev = lo_lev_.get();
  t.lo_rows = lo_rows_.get();
  t.up_lev = up_lev_.get();
  t.up_rows = up_rows_.get();
  int width = 1;
  for (size_t l = 0; l + 1 < lo_ptr_.size(); ++l) width = std::max(width, lo_ptr_[l + 1] - lo_ptr_[l]);
  for (size_t l = 0; l + 1 < up_ptr_.size(); ++l) width = std::max(width, up_ptr_[l + 1] - up_ptr_[l]);
  // level and position of every row in each sweep's order
  auto positions = [&](const std::vector<int>& ptr, const std::vector<int>& rows, std::vector<int>& lev,
                       std::vector<int>& pos) {
    lev.assign(ns, 0);
    pos.assign(ns, 0);
    for (size_t l = 0; l + 1 < ptr.size(); ++l)
      for (int r = ptr[l]; r < ptr[l + 1]; ++r) {
        lev[rows[r]] = static_cast<int>(l);
        pos[rows[r]] = r;
      }
  };
  std::vector<int> llev, lpos, ulev, upos;
  positions(lo_ptr_, lrows, llev, lpos);
  positions(up_ptr_, urows, ulev, upos);
  int dist = 1, maxrow = 1, maxlower = 0, maxupper = 0;
  for (long i = 0; i < ns; ++i) {
    for (int p2 = rp[i]; p2 < dp[i]; ++p2) dist = std::max(dist, llev[i] - llev[col[p2]]);
    for (int p2 = dp[i] + 1; p2 < rp[i + 1]; ++p2) dist = std::max(dist, ulev[i] - ulev[col[p2]]);
    maxrow = std::max(maxrow, rp[i + 1] - rp[i]);
    maxlower = std::max(maxlower, dp[i] - rp[i]);
    maxupper = std::max(maxupper, rp[i + 1] - dp[i] - 1);
  }
  std::vector<int> uptr(static_cast<size_t>(L_.nnzb()) + 1, 0);
  std::vector<int2> mq;
  int maxpairs = 0;
  for (long i = 0; i < ns; ++i) {
    const size_t row_first = mq.size();
    for (int p2 = rp[i]; p2 < dp[i]; ++p2) {
      const int k2 = col[p2];
      int q = dp[k2] + 1, qe = rp[k2 + 1], m = p2 + 1, me = rp[i + 1];
      while (q < qe && m < me) {
        if (col[q] == col[m]) {
          mq.push_back(make_int2(m, q));
          ++q;
          ++m;
        } else if (col[q] < col[m]) {
          ++q;
        } else {
          ++m;
        }
      }
      uptr[p2 + 1] = static_cast<int>(mq.size());
    }
    for (int p2 = dp[i]; p2 < rp[i + 1]; ++p2) uptr[p2 + 1] = static_cast<int>(mq.size());
    maxpairs = std::max(maxpairs, static_cast<int>(mq.size() - row_first));
  }
  // the staged apply: the largest chunk length whose stage and ring fit LDS
  auto entries = [&](int ring, const std::vector<int>& ptr, const std::vector<int>& rows,
                     const std::vector<int>& lev, const std::vector<int>& pos, bool lower, std::vector<int>& ent,
                     std::vector<int>& blk, std::vector<int>& src) {
    ent.assign(ns + 1, 0);
    blk.clear();
    src.clear();
    for (long r = 0; r < ns; ++r) {
      const int i = rows[r];
      const int b0 = lower ? rp[i] : dp[i] + 1, e0 = lower ? dp[i] : rp[i + 1];
      for (int p2 = b0; p2 < e0; ++p2) {
        const int k2 = col[p2];
        blk.push_back(p2);
        src.push_back((lev[k2] % ring) * width + (pos[k2] - ptr[lev[k2]]));
      }
      ent[r + 1] = static_cast<int>(blk.size());
    }
  };
  std::vector<int> lent, lblk, lsrc, uent, ublk, usrc;
  entries(1, lo_ptr_, lrows, llev, lpos, true, lent, lblk, lsrc);  // entry offsets (slots redone below)
  entries(1, up_ptr_, urows, ulev, upos, false, uent, ublk, usrc);
  auto chunks = [&](int clen, const std::vector<int>& ptr, const std::vector<int>& ent, std::vector<int>& ch,
                    int& me, int& mr) {
    const int nl = static_cast<int>(ptr.size()) - 1;
    ch.clear();
    for (int l = 0; l < nl; l += clen) ch.push_back(l);
    ch.push_back(nl);
    for (size_t c = 0; c + 1 < ch.size(); ++c) {
      const int r0 = ptr[ch[c]], r1 = ptr[ch[c + 1]];
      mr = std::max(mr, r1 - r0);
      me = std::max(me, ent[r1] - ent[r0]);
    }
  };
  std::vector<int> lch, uch;
  t.chunked = false;
  t.width = width;
  // the apply's stage may use all of a workgroup's LDS (160 KiB on gfx950)
  size_t apply_lds = kIluLdsBytes;
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device_) == hipSuccess && v > 0)
      apply_lds = std::max(apply_lds, std::min<size_t>(static_cast<size_t>(v), 160 * 1024));
  }
  // the apply's register path: every row's blocks in registers (at most 32
  // values), one row per thread per level, and the stage padded to K places
  // per row
  t.reg_k = 0;
  if (width <= k::kIluThreads)
    for (int K : {4, 8, 16})
      if (std::max(maxlower, maxupper) <= K && K * B * B <= 32) {
        t.reg_k = K;
        break;
      }
  for (int clen = 32; clen >= 1 && !t.chunked; --clen) {
    int me = 0, mr = 0;
    chunks(clen, lo_ptr_, lent, lch, me, mr);
    chunks(clen, up_ptr_, uent, uch, me, mr);
    k::IluSweep c2 = t;
    c2.stage_ent = std::max(me, 1);
    c2.stage_rows = std::max(mr, 1);
    c2.stage_levels = clen;
    c2.ring = std::max(dist + 1, clen);
    if (k::ilu_chunk_lds_bytes(c2, B, static_cast<int>(sizeof(double))) <= apply_lds) {
      t = c2;
      t.chunked = true;
    }
  }
  if (!t.chunked) t.reg_k = 0;
  if (t.chunked) {
    entries(t.ring, lo_ptr_, lrows, llev, lpos, true, lent, lblk, lsrc);
    entries(t.ring, up_ptr_, urows, ulev, upos, false, uent, ublk, usrc);
    // replay both sweeps on the host: every ring read must find its
    // column's row, written at an earlier level and not yet overwritten
    std::vector<int> held(static_cast<size_t>(t.ring) * width, -1);
    auto replay = [&](const std::vector<int>& ptr, const std::vector<int>& rows, const std::vector<int>& ent,
                      const std::vector<int>& blk, const std::vector<int>& src) {
      std::fill(held.begin(), held.end(), -1);
      for (size_t l = 0; l + 1 < ptr.size(); ++l) {
        for (int r = ptr[l]; r < ptr[l + 1]; ++r)
          for (int e = ent[r]; e < ent[r + 1]; ++e)
            if (src[e] < 0 || src[e] >= t.ring * width || held[src[e]] != col[blk[e]])
              throw Error("internal: ILU0 sweep ring table");
        for (int r = ptr[l]; r < ptr[l + 1]; ++r)
          held[(static_cast<int>(l) % t.ring) * width + (r - ptr[l])] = rows[r];
      }
    };
    replay(lo_ptr_, lrows, lent, lblk, lsrc);
    replay(up_ptr_, urows, uent, ublk, usrc);
    if (t.reg_k > 0) {
      // row r's place q at r * K + q; padding has block -1
      const int K = t.reg_k;
      auto pad = [&](const std::vector<int>& ent, std::vector<int>& blk, std::vector<int>& src) {
        std::vector<int> pb(static_cast<size_t>(ns) * K, -1), ps(static_cast<size_t>(ns) * K, -1);
        for (long r = 0; r < ns; ++r)
          for (int e = ent[r]; e < ent[r + 1]; ++e) {
            pb[static_cast<size_t>(r) * K + (e - ent[r])] = blk[e];
            ps[static_cast<size_t>(r) * K + (e - ent[r])] = src[e];
          }
        blk.swap(pb);
        src.swap(ps);
      };
      pad(lent, lblk, lsrc);
      pad(uent, ublk, usrc);
    }
    lo_ent_.upload(lent, stream_);
    lo_blk_.upload(lblk.empty() ? std::vector<int>{0} : lblk, stream_);
    lo_src_.upload(lsrc.empty() ? std::vector<int>{0} : lsrc, stream_);
    up_ent_.upload(uent, stream_);
    up_blk_.upload(ublk.empty() ? std::vector<int>{0} : ublk, stream_);
    up_src_.upload(usrc.empty() ? std::vector<int>{0} : usrc, stream_);
    lo_chunk_.upload(lch, stream_);
    up_chunk_.upload(uch, stream_);
    t.nlo_chunk = static_cast<int>(lch.size()) - 1;
    t.nup_chunk = static_cast<int>(uch.size()) - 1;
    t.lo_chunk = lo_chunk_.get();
    t.up_chunk = up_chunk_.get();
    t.lo_ent = lo_ent_.get();
    t.lo_blk = lo_blk_.get();
    t.lo_src = lo_src_.get();
    t.up_ent = up_ent_.get();
    t.up_blk = up_blk_.get();
    t.up_src = up_src_.get();
  }
  // the inbox factorisation: small blocks and rows; every operand of a row
  // at a position fixed by the row's place in level order (k::IluSweep)
  if (B <= 2 && maxlower <= k::kIluRegLower && maxpairs <= k::kIluRegPairs && maxrow <= k::kIluRegRow) {
    constexpr int ML = k::kIluRegLower, MU = k::kIluRegPairs, MP = k::kIluRegPush, IN = ML + MU;
    const int BB = B * B;
    const size_t N = static_cast<size_t>(ns);
    std::vector<int4> frow(ns);
    constexpr int FS = k::kIluFubStride;
    static_assert(FS >= ML + 1 && FS % 4 == 0 && MU % 4 == 0 && MP % 2 == 0, "16-byte table rows");
    std::vector<int> fub(N * FS, 0), fm(N * MU, 0);
    std::vector<std::vector<int2>> pushes(ns);  // by pushing row
    for (long r = 0; r < ns; ++r) {
      const int i = lrows[r];
      const int p0 = rp[i], nl = dp[i] - rp[i], u0 = uptr[p0], nu = uptr[dp[i]] - u0;
      frow[r] = make_int4(i, p0, rp[i + 1] - p0, nl);
      for (int a = 0; a <= ML; ++a) fub[r * FS + a] = a <= nl ? uptr[p0 + a] - u0 : nu;
      for (int a = 0; a < nl; ++a) {
        const int k2 = col[p0 + a];
        pushes[k2].push_back(make_int2(-1, static_cast<int>((r * IN + a) * BB)));
        for (int u = uptr[p0 + a] - u0; u < uptr[p0 + a + 1] - u0; ++u) {
          const int2 pq = mq[u0 + u];
          fm[r * MU + u] = pq.x - p0;
          pushes[k2].push_back(make_int2(pq.y - rp[k2], static_cast<int>((r * IN + ML + u) * BB)));
        }
      }
    }
    size_t maxpush = 0;
    for (const auto& v : pushes) maxpush = std::max(maxpush, v.size());
    const size_t rowbytes = static_cast<size_t>(k::kIluRegRow) * BB * sizeof(double);
    int nt = std::min(256, (width + 63) / 64 * 64);
    while (nt >= 64 && nt * rowbytes > kIluLdsBytes) nt -= 64;
    const size_t inbox_doubles = N * IN * BB;
    if (maxpush <= static_cast<size_t>(MP) && nt >= 64 && inbox_doubles < (1u << 31)) {
      std::vector<int2> fpush(N * MP, make_int2(0, -1));
      for (long i = 0; i < ns; ++i)
        for (size_t j = 0; j < pushes[i].size(); ++j) fpush[static_cast<size_t>(lpos[i]) * MP + j] = pushes[i][j];
      ilu_frow_.upload(frow, stream_);
      ilu_fub_.upload(fub, stream_);
      ilu_fm_.upload(fm, stream_);
      ilu_fpush_.upload(fpush, stream_);
      ilu_packed_.resize(N * k::kIluRegRow * BB);
      ilu_inbox_.resize(inbox_doubles);
      t.inbox_path = true;
      t.factor_threads = nt;
      t.nrows = static_cast<int>(ns);
      t.frow = ilu_frow_.get();
      t.fub = ilu_fub_.get();
      t.fm = ilu_fm_.get();
      t.fpush = ilu_fpush_.get();
      t.packed = ilu_packed_.get();
      t.inbox = ilu_inbox_.get();
    }
  }
  ilu_tab_ = t;
}

void Model::precond_setup(const KrylovWork& w) {
  const int B = L_.B;
  const int ns = static_cast<int>(L_.nslot);
  if (jacobi_kind() && use_points_) {
    k::block_inverse(B, ns, pdiag_.get(), diag_id_.get(), w.dinv, w.status, cfg_.precond == Preconditioner::Jacobi,
                     stream_);
  } else if (jacobi_kind()) {
    k::block_inverse(B, ns, w.val, diag_pos_.get(), w.dinv, w.status, cfg_.precond == Preconditioner::Jacobi, stream_);
  } else if (cfg_.precond == Preconditioner::ILU0) {
    DCG_HIP(hipMemcpyAsync(w.lu, w.val, matrix_values() * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    if (ilu_sweep_) {
      k::ilu_factor_sweep(B, ilu_tab_, row_ptr_.get(), col_.get(), diag_pos_.get(), w.lu, w.dinv, w.status, stream_);
      return;
    }
    for (size_t l = 0; l + 1 < lo_ptr_.size(); ++l)
      k::ilu_factor(B, lo_ptr_[l + 1] - lo_ptr_[l], lo_rows_.get() + lo_ptr_[l], row_ptr_.get(), col_.get(),
                    diag_pos_.get(), w.lu, w.dinv, w.status, stream_);
  }
}

int Model::bicgstab64(const double* b, double* x, double rtol, double abstol, int maxit, int* its,
                      const KrylovWork& w) {
  const int B = L_.B;
  const int ns = static_cast<int>(L_.nslot);
  const long n = static_cast<long>(ns) * B;
  double *r = w.kv[0], *rh = w.kv[1], *p = w.kv[2], *v = w.kv[3], *ph = w.kv[4], *s = w.kv[5], *sh = w.kv[6],
         *t = w.kv[7], *tmp = w.kv[8];
  k::BicgScalars* sc = w.sc;
  double* part = w.part;
  k::BicgScalars host{};
  host.tol = rtol;
  host.abstol = abstol;
  host.maxit = maxit;
  DCG_HIP(hipMemcpyAsync(sc, &host, sizeof host, hipMemcpyHostToDevice, stream_));
  const int nred = std::max(1, std::min(1024, static_cast<int>((n + 255) / 256)));
  k::bicg_init<double>(n, b, x, r, rh, p, v, part, nred, sc, stream_);
  const double* dinv = jacobi_kind() ? w.dinv : nullptr;
  // the reference's Jacobi inverts the scalar diagonal: B values per slot
  const bool point_jacobi = cfg_.precond == Preconditioner::Jacobi;
  const bool ilu = cfg_.precond == Preconditioner::ILU0;
  auto apply_ilu = [&](const double* in, double* out) {
    if (ilu_sweep_) {
      k::ilu_apply_sweep<double>(B, ilu_tab_, row_ptr_.get(), col_.get(), diag_pos_.get(), w.lu, w.dinv, in, tmp,
                                 out, sc, stream_);
      return;
    }
    for (size_t l = 0; l + 1 < lo_ptr_.size(); ++l)
      k::ilu_lower<double>(B, lo_ptr_[l + 1] - lo_ptr_[l], lo_rows_.get() + lo_ptr_[l], row_ptr_.get(), col_.get(),
                           diag_pos_.get(), w.lu, in, tmp, sc, stream_);
    for (size_t l = 0; l + 1 < up_ptr_.size(); ++l)
      k::ilu_upper<double>(B, up_ptr_[l + 1] - up_ptr_[l], up_rows_.get() + up_ptr_[l], row_ptr_.get(), col_.get(),
                           diag_pos_.get(), w.lu, w.dinv, tmp, out, sc, stream_);
  };
  // out = J in with the partial dots (out.w1, out.w2) finalised into `stage`
  const int* done_flag =
      reinterpret_cast<const int*>(reinterpret_cast<const char*>(sc) + offsetof(k::BicgScalars, done));
  // the partial dots' finaliser; stage 3 first finalises the s-update's
  // partial norms (stage 2, deferred past the second operator launch)
  const int nblk_s = k::bicg_s_parts(ns);
  auto finalise = [&](int nparts, int stage) {
    if (stage == 3) k::bicg_fin23(nblk_s, w.spart, nparts, part, sc, stream_);
    else k::bicg_fin(nparts, part, 2, stage, sc, stream_);
  };
  auto matvec = [&](const double* in, double* out, const double* w1, const double* w2, int stage) {
    if (cfg_.matrix_free) {
      prof_begin(kJacobianApply);
      launch_apply(lin_x_.get(), lin_t_, lin_cm_, lin_ca_, in, out, done_flag);
      prof_end(kJacobianApply, apply_bytes());
      k::dot_parts(n, out, w1, w2, part, nred, sc, stream_);
      finalise(nred, stage);
      return;
    }
    if (use_points_) {
      prof_begin(kPointApply64);
      const int np = launch_points_apply(in, out, done_flag, false, w1, w2, part);
      prof_end(kPointApply64, point_apply_bytes(8, 1));
      if (!np) k::dot_parts(n, out, w1, w2, part, nred, sc, stream_);
      finalise(np ? np : nred, stage);
      return;
    }
    prof_begin(kSpmv64);
    k::spmv(B, plan_, w.val, in, out, w1, w2, part, sc, stage, stream_);
    finalise(plan_.parts(), stage);
    prof_end(kSpmv64, spmv_bytes(8, 8, 1));
  };
  const int akind = cfg_.matrix_free ? kJacobianApply : use_points_ ? kPointApply64 : kSpmv64;
  auto iteration = [&]() {
    k::bicg_p<double>(B, ns, r, p, v, ph, dinv, point_jacobi, sc, stream_);
    if (ilu) apply_ilu(p, ph);
    matvec(ph, v, rh, nullptr, 1);
    k::bicg_s<double>(B, ns, r, v, s, sh, dinv, point_jacobi, sc, w.spart, stream_);
    if (ilu) apply_ilu(s, sh);
    matvec(sh, t, s, t, 3);
    const int nxr = k::bicg_xr<double>(B, ns, x, ph, sh, s, t, r, rh, sc, part, stream_);
    k::bicg_fin(nxr, part, 2, 4, sc, stream_);
  };
  // iterations run in batches between read-backs of the scalars: the first
  // as many as this solve took in the last time step (or the last solve)
  const int slot = solve_index_++;
  const int guess = slot < static_cast<int>(its_hint_.size()) ? its_hint_[slot] : last_krylov_its_;
  const size_t mark = prof_mark(akind);
  const int fixed = cfg_.readback_iterations;
  int done_its = 0, batch = fixed ? fixed : std::max(2, std::min(guess, 64));
  for (;;) {
    for (int it = 0; it < batch; ++it) iteration();
    done_its += batch;
    DCG_HIP(hipMemcpyAsync(&host, sc, sizeof host, hipMemcpyDeviceToHost, stream_));
    synchronize();
    if (host.done) break;
    if (done_its > maxit + 4) {
      host.done = 2;
      break;
    }
    batch = fixed ? fixed : 4;
  }
  // both operator applies of every iteration ran (a breakdown keeps them all)
  if (host.done == 1 || host.done == 2) prof_keep(akind, mark, 2L * host.iter);
  *its = host.iter;
  last_krylov_its_ = std::max(2, host.iter);
  if (slot < kMaxSolveHints) {
    if (slot >= static_cast<int>(its_hint_.size())) its_hint_.resize(slot + 1, 0);
    its_hint_[slot] = last_krylov_its_;
  }
  return host.done == 1 ? 0 : host.done;
}

int Model::bicgstab32(const double* b, float* x, double rtol, int maxit, int* its) {
  const int B = L_.B;
  const int ns = static_cast<int>(L_.nslot);
  const long n = static_cast<long>(ns) * B;
  float *r = kf_[0].get(), *rh = kf_[1].get(), *p = kf_[2].get(), *v = kf_[3].get(), *ph = kf_[4].get(),
        *s = kf_[5].get(), *sh = kf_[6].get(), *t = kf_[7].get(), *tmp = kf_[8].get();
  k::BicgScalars host{};
  host.tol = rtol;
  host.abstol = 0.0;
  host.maxit = maxit;
  DCG_HIP(hipMemcpyAsync(sc_.get(), &host, sizeof host, hipMemcpyHostToDevice, stream_));
  const int nred = std::max(1, std::min(1024, static_cast<int>((n + 255) / 256)));
  k::bicg_init<float>(n, b, x, r, rh, p, v, part_.get(), nred, sc_.get(), stream_);
  const double* dinv = jacobi_kind() ? dinv_.get() : nullptr;
  const bool point_jacobi = cfg_.precond == Preconditioner::Jacobi;
  const bool ilu = cfg_.precond == Preconditioner::ILU0;
  auto apply_ilu = [&](const float* in, float* out) {
    if (ilu_sweep_) {
      k::ilu_apply_sweep<float>(B, ilu_tab_, row_ptr_.get(), col_.get(), diag_pos_.get(), lu_.get(), dinv_.get(), in,
                                tmp, out, sc_.get(), stream_);
      return;
    }
    for (size_t l = 0; l + 1 < lo_ptr_.size(); ++l)
      k::ilu_lower<float>(B, lo_ptr_[l + 1] - lo_ptr_[l], lo_rows_.get() + lo_ptr_[l], row_ptr_.get(), col_.get(),
                          diag_pos_.get(), lu_.get(), in, tmp, sc_.get(), stream_);
    for (size_t l = 0; l + 1 < up_ptr_.size(); ++l)
      k::ilu_upper<float>(B, up_ptr_[l + 1] - up_ptr_[l], up_rows_.get() + up_ptr_[l], row_ptr_.get(), col_.get(),
                          diag_pos_.get(), lu_.get(), dinv_.get(), tmp, out, sc_.get(), stream_);
  };
  const int* done_flag =
      reinterpret_cast<const int*>(reinterpret_cast<const char*>(sc_.get()) + offsetof(k::BicgScalars, done));
  double* spart = part_.get() + spart_off_;
  const int nblk_s = k::bicg_s_parts(ns);
  auto finalise = [&](int nparts, int stage) {
    if (stage == 3) k::bicg_fin23(nblk_s, spart, nparts, part_.get(), sc_.get(), stream_);
    else k::bicg_fin(nparts, part_.get(), 2, stage, sc_.get(), stream_);
  };
  auto matvec = [&](const float* in, float* out, const float* w1, const float* w2, int stage) {
    if (use_points_) {
      prof_begin(kPointApply32);
      const int np = launch_points_apply(in, out, done_flag, true, w1, w2, part_.get());
      prof_end(kPointApply32, point_apply_bytes(4, 1));
      if (!np) k::dot_parts(n, out, w1, w2, part_.get(), nred, sc_.get(), stream_);
      finalise(np ? np : nred, stage);
      return;
    }
    prof_begin(kSpmv32);
    k::spmv(B, plan32_, val32_.get(), in, out, w1, w2, part_.get(), sc_.get(), stage, stream_);
    finalise(plan32_.parts(), stage);
    prof_end(kSpmv32, spmv_bytes(4, 4, 1));
  };
  const int akind = use_points_ ? kPointApply32 : kSpmv32;
  auto iteration = [&]() {
    k::bicg_p<float>(B, ns, r, p, v, ph, dinv, point_jacobi, sc_.get(), stream_);
    if (ilu) apply_ilu(p, ph);
    matvec(ph, v, rh, nullptr, 1);
    k::bicg_s<float>(B, ns, r, v, s, sh, dinv, point_jacobi, sc_.get(), spart, stream_);
    if (ilu) apply_ilu(s, sh);
    matvec(sh, t, s, t, 3);
    const int nxr = k::bicg_xr<float>(B, ns, x, ph, sh, s, t, r, rh, sc_.get(), part_.get(), stream_);
    k::bicg_fin(nxr, part_.get(), 2, 4, sc_.get(), stream_);
  };
  // iterations run in batches between read-backs of the scalars: the first
  // as many as this solve took in the last time step (or the last solve)
  const int slot = solve_index_++;
  const int guess = slot < static_cast<int>(its_hint_.size()) ? its_hint_[slot] : last_krylov_its_;
  const size_t mark = prof_mark(akind);
  const int fixed = cfg_.readback_iterations;
  int done_its = 0, batch = fixed ? fixed : std::max(2, std::min(guess, 64));
  for (;;) {
    for (int it = 0; it < batch; ++it) iteration();
    done_its += batch;
    DCG_HIP(hipMemcpyAsync(&host, sc_.get(), sizeof host, hipMemcpyDeviceToHost, stream_));
    synchronize();
    if (host.done) break;
    if (done_its > maxit + 4) {
      host.done = 2;
      break;
    }
    batch = fixed ? fixed : 4;
  }
  // both operator applies of every iteration ran (a breakdown keeps them all)
  if (host.done == 1 || host.done == 2) prof_keep(akind, mark, 2L * host.iter);
  *its = host.iter;
  last_krylov_its_ = std::max(2, host.iter);
  if (slot < kMaxSolveHints) {
    if (slot >= static_cast<int>(its_hint_.size())) its_hint_.resize(slot + 1, 0);
    its_hint_[slot] = last_krylov_its_;
  }
  return host.done == 1 ? 0 : host.done;
}

// Solve val_ * x = b to relative tolerance rtol.  FP64: one preconditioned
// BiCGSTAB.  Mixed: FP32 BiCGSTAB on an FP32 copy of the matrix inside FP64
// iterative refinement; the refinement residual b - A x is FP64.
int Model::krylov(const double* b, double* x, double rtol, int maxit, int* its, int* refinements) {
  const KrylovWork w = own_work();
  prof_begin(kPrecond);
  precond_setup(w);
  prof_end(kPrecond, 0.0);
  *its = 0;
  *refinements = 0;
  const long n = L_.nslot * L_.B;
  if (cfg_.precision == Precision::FP64) return bicgstab64(b, x, rtol, 0.0, maxit, its, w);
  if (use_points_) {
    for (size_t c = 0; c < pc_.size(); ++c)
      if (pc_[c].size()) k::to_float(static_cast<long>(pc_[c].size()), pc_[c].get(), pc32_[c].get(), stream_);
  } else {
    k::to_float(matrix_values(), val_.get(), val32_.get(), stream_);
  }
  double* res = kv_[8].get();  // FP64 refinement residual
  double* ax = kv_[7].get();
  float* d = kf_[9].get();
  k::fill(x, 0.0, n, stream_);
  DCG_HIP(hipMemcpyAsync(res, b, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  k::reduce(n, b, nullptr, 0, part_.get(), scalar_out_.get(), stream_);
  double bn2 = 0;
  DCG_HIP(hipMemcpyAsync(&bn2, scalar_out_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  const double target = std::max(rtol, cfg_.ir_floor) * std::sqrt(bn2);
  if (bn2 == 0.0) return 0;
  double rn = std::sqrt(bn2);
  for (int k2 = 0; k2 < cfg_.ir_max; ++k2) {
    // inner tolerance relative to the current refinement residual, never
    // tighter than what reaches the outer target
    double inner = std::max(cfg_.ir_inner_rtol, 0.5 * target / rn);
    int it = 0;
    int rc = bicgstab32(res, d, std::min(inner, 0.5), maxit, &it);
    *its += it;
    *refinements += 1;
    if (rc == 3) return 3;
    // x += d (FP64), res = b - A x (FP64)
    k::to_double(n, d, ax, stream_);
    k::axpy(n, 1.0, ax, x, stream_);
    if (use_points_) launch_points_apply(x, ax, nullptr, false);
    else k::spmv(L_.B, plan_, val_.get(), x, ax, nullptr, nullptr, nullptr, nullptr, 0, stream_);
    DCG_HIP(hipMemcpyAsync(res, b, n * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    k::axpy(n, -1.0, ax, res, stream_);
    k::reduce(n, res, nullptr, 0, part_.get(), scalar_out_.get(), stream_);
    double r2 = 0;
    DCG_HIP(hipMemcpyAsync(&r2, scalar_out_.get(), sizeof(double), hipMemcpyDeviceToHost, stream_));
    synchronize();
    rn = std::sqrt(r2);
    if (!std::isfinite(rn)) return 3;
    if (rn <= target) return 0;
  }
  return 2;
}

// ------------------------------------------------------------------ time step
int Model::step(double t, double dt, StepStats* st) {
  require_device();
  StepStats local;
  StepStats& S = st ? *st : local;
  S = StepStats{};
  solve_index_ = 0;
  const long n = L_.nslot * L_.B;
  const size_t bytes = static_cast<size_t>(n) * sizeof(double);
  DCG_HIP(hipMemcpyAsync(stage_[0].get(), u_.get(), bytes, hipMemcpyDeviceToDevice, stream_));
  for (int s = 1; s <= rk_.stages; ++s) {
    S.stages = s;
    // history of earlier stages: sum_{j<s} a_sj M(u_j) + dt b_sj A(u_j)
    k::fill(hist_.get(), 0.0, n, stream_);
    for (int j = 0; j < s; ++j) {
      const double am = rk_.a[s - 1][j], ab = dt * rk_.b[s - 1][j];
      if (am == 0.0 && ab == 0.0) continue;
      launch_assembly(stage_[j].get(), t + rk_.d[j] * dt, am, ab, hist_.get(), hist_.get(), nullptr, 1);
    }
    const double cm = rk_.a[s - 1][s], ca = dt * rk_.b[s - 1][s], ts = t + rk_.d[s] * dt;
    double* x = xs_.get();
    DCG_HIP(hipMemcpyAsync(x, stage_[s - 1].get(), bytes, hipMemcpyDeviceToDevice, stream_));
    auto residual = [&]() {
      prof_begin(kAssemblyRes);
      launch_assembly(x, ts, cm, ca, hist_.get(), r_.get(), nullptr, 1);
      prof_end(kAssemblyRes, assembly_bytes(false));
    };
    auto jacobian = [&]() {
      if (cfg_.matrix_free) {
        // the operator keeps its linearization point (the Newton iterate)
        DCG_HIP(hipMemcpyAsync(lin_x_.get(), x, bytes, hipMemcpyDeviceToDevice, stream_));
        lin_t_ = ts;
        lin_cm_ = cm;
        lin_ca_ = ca;
        return;
      }
      if (use_points_) {
        prof_begin(kPointSetup);
        launch_points_setup(x, ts, cm, ca);
        prof_end(kPointSetup, point_setup_bytes());
        return;
      }
      prof_begin(kAssemblyJac);
      launch_assembly(x, ts, cm, ca, nullptr, nullptr, val_.get(), 2);
      prof_end(kAssemblyJac, assembly_bytes(true));
    };
    auto solve = [&](double rtol) -> int {
      int its = 0, refs = 0;
      int rc = krylov(r_.get(), z_.get(), rtol, cfg_.lin_maxit, &its, &refs);
      S.linear_iterations += its;
      S.refinement_steps += refs;
      return rc;
    };
    if (spec_.is_linear) {
      // linear defect correction (make_step_operator.hh:210-245)
      residual();
      jacobian();
      int rc = solve(cfg_.lin_rtol);
      if (rc) {
        S.status = rc == 3 ? 3 : rc == 4 ? 4 : 2;
        return S.status;
      }
      k::axpy(n, -1.0, z_.get(), x, stream_);
      S.newton_iterations += 1;
    } else {
      // Newton with the control of the reference's configuration
      // (make_step_operator.hh:246-288): stop when the defect (in the chosen
      // norm) falls below max(absolute, relative * defect_0); unless the
      // linear tolerance is fixed, each correction is solved to
      //   max(stop / (10 d_k), min(min_reduction, d_k^2 / d_{k-1}^2)),
      // and the Jacobian is re-assembled unless the defect rate stays below
      // linearization_threshold.
      residual();
      double nrm0 = residual_norm(r_.get()), nrm = nrm0, prev = nrm0;
      const double stop = std::max(cfg_.nl_atol, cfg_.nl_rtol * nrm0);
      int it = 0;
      for (;;) {
        S.residual_norm = nrm;
        if (!std::isfinite(nrm)) {
          S.status = 3;
          return 3;
        }
        const bool conv = nrm <= stop || nrm == 0.0;
        if (conv && it >= cfg_.nl_min) break;
        if (it >= cfg_.nl_max) {
          S.status = 1;
          return 1;
        }
        const bool relinearize = it == 0 || cfg_.linearization_threshold <= 0.0 ||
                                 nrm / prev > cfg_.linearization_threshold;
        if (relinearize) jacobian();
        double rtol = cfg_.lin_rtol;
        if (!cfg_.fixed_linear_tolerance) {
          const double ratio = it == 0 ? 1.0 : (nrm * nrm) / (prev * prev);
          rtol = std::max(stop / (10.0 * nrm), std::min(cfg_.min_linear_reduction, ratio));
        }
        int rc = solve(rtol);
        if (rc) {
          S.status = rc == 3 ? 3 : rc == 4 ? 4 : 2;
          return S.status;
        }
        k::axpy(n, -1.0, z_.get(), x, stream_);
        ++it;
        residual();
        prev = nrm;
        nrm = residual_norm(r_.get());
      }
      S.newton_iterations += it;
    }
    DCG_HIP(hipMemcpyAsync(stage_[s].get(), x, bytes, hipMemcpyDeviceToDevice, stream_));
  }
  DCG_HIP(hipMemcpyAsync(u_.get(), stage_[rk_.stages].get(), bytes, hipMemcpyDeviceToDevice, stream_));
  time_ = t + dt;
  synchronize();
  if (prof_) prof_collect();
  return 0;
}

// ------------------------------------------------------------------ state I/O
void Model::set_state(const double* packed, double t) {
  require_device();
  if (L_.ndof > 0)
    DCG_HIP(hipMemcpyAsync(packed_.get(), packed, L_.ndof * sizeof(double), hipMemcpyHostToDevice, stream_));
  k::unpack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), packed_.get(), u_.get(), stream_);
  time_ = t;
  synchronize();
}

void Model::get_state(double* packed, double* t) {
  require_device();
  k::pack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), u_.get(), packed_.get(), stream_);
  if (L_.ndof > 0)
    DCG_HIP(hipMemcpyAsync(packed, packed_.get(), L_.ndof * sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  if (t) *t = time_;
}

void Model::initial_state(double* packed) const {
  std::map<std::string, int> slots = {{"position_x", 0}, {"position_y", 1}, {"position_z", 2}, {"time", 3}};
  for (const auto& sp : spec_.species) slots[sp.name] = 4;  // species values are not defined yet: 0
  for (size_t s = 0; s < spec_.species.size(); ++s) {
    const auto& sp = spec_.species[s];
    const auto& C = L_.comps[sp.comp];
    expr::Bound f;
    bool have = spec_.initial[s] != nullptr;
    if (have) f = expr::bind(spec_.initial[s], slots);
    for (long l = 0; l < C.nslot; ++l) {
      double v[5] = {0, 0, 0, time_, 0};
      L_.grid.vertex_pos(C.nodes[l], v);
      packed[C.dof_off + l * C.b + sp.local] = have ? f(v) : 0.0;
    }
  }
}

void Model::mass(double* out) {
  require_device();
  for (const auto& C : L_.comps) {
    if (!C.b || !C.nslot) continue;
    k::weighted_species_sum(C.slot_off, C.nslot, L_.B, weights_.get(), u_.get(), part_.get(), scalar_out_.get(),
                            stream_);
    double sums[8];
    DCG_HIP(hipMemcpyAsync(sums, scalar_out_.get(), sizeof sums, hipMemcpyDeviceToHost, stream_));
    synchronize();
    for (int i = 0; i < C.b; ++i) out[C.species[i]] = sums[i];
  }
}

// ------------------------------------------------------------------ profiling
void Model::set_profiling(bool on) {
  require_device();
  synchronize();
  prof_collect();
  prof_ = on;
  for (auto& p : prof_data_) {
    p.launches = 0;
    p.ms = 0.0;
    p.bytes = 0.0;
  }
}

void Model::prof_begin(int kind) {
  if (!prof_) return;
  Prof& p = prof_data_[kind];
  if (p.used + 2 > p.pool.size()) {
    hipEvent_t a, b;
    DCG_HIP(hipEventCreate(&a));
    DCG_HIP(hipEventCreate(&b));
    p.pool.push_back(a);
    p.pool.push_back(b);
  }
  DCG_HIP(hipEventRecord(p.pool[p.used], stream_));
}

void Model::prof_end(int kind, double bytes) {
  if (!prof_) return;
  Prof& p = prof_data_[kind];
  DCG_HIP(hipEventRecord(p.pool[p.used + 1], stream_));
  p.used += 2;
  p.pending_bytes.push_back(bytes);
}

size_t Model::prof_mark(int kind) const { return prof_ ? prof_data_[kind].used / 2 : 0; }

void Model::prof_keep(int kind, size_t mark, long keep) {
  if (!prof_) return;
  Prof& p = prof_data_[kind];
  const size_t n = mark + static_cast<size_t>(std::max(0L, keep));
  if (p.used / 2 <= n) return;
  p.used = 2 * n;
  p.pending_bytes.resize(n);
}

void Model::prof_collect() {
  for (auto& p : prof_data_) {
    for (size_t i = 0; i + 1 < p.used; i += 2) {
      float ms = 0.f;
      DCG_HIP(hipEventElapsedTime(&ms, p.pool[i], p.pool[i + 1]));
      p.ms += ms;
      p.launches += 1;
      p.bytes += p.pending_bytes[i / 2];
    }
    p.used = 0;
    p.pending_bytes.clear();
  }
}

void Model::profile(int kind, long* launches, double* ms, double* bytes) {
  require_device();
  if (kind < 0 || kind >= kProfKinds) throw Error("profile kind out of range");
  synchronize();
  prof_collect();
  *launches = prof_data_[kind].launches;
  *ms = prof_data_[kind].ms;
  *bytes = prof_data_[kind].bytes;
}

double Model::probe_copy(long bytes) {
  require_device();
  bytes = bytes / 32 * 32;
  DeviceArray<char> a, b;
  a.resize(static_cast<size_t>(bytes));
  b.resize(static_cast<size_t>(bytes));
  DCG_HIP(hipMemsetAsync(a.get(), 0, static_cast<size_t>(bytes), stream_));
  hipEvent_t e0, e1;
  DCG_HIP(hipEventCreate(&e0));
  DCG_HIP(hipEventCreate(&e1));
  k::probe_copy(a.get(), b.get(), bytes, stream_);
  DCG_HIP(hipEventRecord(e0, stream_));
  k::probe_copy(a.get(), b.get(), bytes, stream_);
  DCG_HIP(hipEventRecord(e1, stream_));
  DCG_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  DCG_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  synchronize();
  return ms;
}

// Algorithmic bytes of one BCSR SpMV: values, column indices, row pointer,
// x read once, y written once, plus fused dot-product operands.
double Model::spmv_bytes(int vb, int xb, int extra_vectors) const {
  const double nb = static_cast<double>(L_.nnzb()), N = static_cast<double>(L_.nslot), B = L_.B;
  return vb * B * B * nb + 4.0 * nb + 4.0 * (N + 1) + xb * B * N * (2.0 + extra_vectors);
}

// Algorithmic bytes of one assembly pass: state read once, residual or
// Jacobian values written once, plus the cell tables (cell id, class mask,
// corner slots) and the per-row stencil tables the Jacobian scatter reads.
double Model::assembly_bytes(bool jacobian) const {
  const double N = static_cast<double>(L_.nslot), B = L_.B;
  double cells = 0;
  for (const auto& C : L_.comps) cells += static_cast<double>(C.cells.size());
  double b = 8.0 * B * N + cells * (4.0 + 1.0 + 4.0 * L_.grid.ncorner);
  if (jacobian) b += 8.0 * B * B * static_cast<double>(L_.nnzb()) + 8.0 * N;
  else b += 8.0 * B * N * 2.0;  // residual: history read + residual written
  return b;
}

// Algorithmic bytes of one matrix-free apply: the linearization point and v
// read, y written, and the same per-cell tables as the assembly.
double Model::apply_bytes() const {
  const double N = static_cast<double>(L_.nslot), B = L_.B;
  double cells = 0;
  for (const auto& C : L_.comps) cells += static_cast<double>(C.cells.size());
  return 8.0 * B * N * 3.0 + cells * (4.0 + 1.0 + 4.0 * L_.grid.ncorner);
}

// Algorithmic bytes of one point-coefficient apply: the coefficients of every
// cell of the compartments once, v read and y written once, extra_vectors
// more vectors read (the BiCGSTAB applies read one dot operand; their other
// operand is absent or y itself), the cell class masks and vertex -> slot
// maps the tiles read.
double Model::point_apply_bytes(int value_bytes, int extra_vectors) const {
  const double N = static_cast<double>(L_.nslot), B = L_.B;
  double b = (2.0 + extra_vectors) * value_bytes * B * N;
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    if (!gen_.point_words[c]) continue;
    const double cells = static_cast<double>(L_.comps[c].cells.size());
    b += cells * (value_bytes * gen_.point_words[c] + 1.0) + 4.0 * L_.comps[c].nslot;
  }
  return b;
}

// Algorithmic bytes of one point-coefficient set-up: the state read once, the
// coefficients and diagonal blocks written once, the same tile tables.
double Model::point_setup_bytes() const {
  const double N = static_cast<double>(L_.nslot), B = L_.B;
  double b = 8.0 * B * N + 8.0 * B * B * N;
  for (size_t c = 0; c < L_.comps.size(); ++c) {
    if (!gen_.point_words[c]) continue;
    const double cells = static_cast<double>(L_.comps[c].cells.size());
    b += cells * (8.0 * gen_.point_words[c] + 1.0) + 4.0 * L_.comps[c].nslot;
  }
  return b;
}

// ------------------------------------------------------------------ test entry points
void Model::assemble(const double* x_packed, double t, double cm, double ca, double* r_packed, double* jac) {
  require_device();
  if (jac && cfg_.matrix_free) throw Error("a matrix-free model holds no assembled Jacobian");
  if (jac) ensure_matrix();
  const long n = L_.nslot * L_.B;
  DCG_HIP(hipMemcpyAsync(packed_.get(), x_packed, L_.ndof * sizeof(double), hipMemcpyHostToDevice, stream_));
  k::unpack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), packed_.get(), xs_.get(), stream_);
  (void)n;
  launch_assembly(xs_.get(), t, cm, ca, nullptr, r_.get(), jac ? val_.get() : nullptr, jac ? 3 : 1);
  k::pack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), r_.get(), packed_.get(), stream_);
  DCG_HIP(hipMemcpyAsync(r_packed, packed_.get(), L_.ndof * sizeof(double), hipMemcpyDeviceToHost, stream_));
  if (jac) DCG_HIP(hipMemcpyAsync(jac, val_.get(), val_.bytes(), hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void Model::jacobian_apply(const double* x_packed, double t, double cm, double ca, const double* v_packed,
                           double* y_packed) {
  require_device();
  DCG_HIP(hipMemcpyAsync(packed_.get(), x_packed, L_.ndof * sizeof(double), hipMemcpyHostToDevice, stream_));
  k::unpack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), packed_.get(), xs_.get(), stream_);
  DCG_HIP(hipMemcpyAsync(packed_.get(), v_packed, L_.ndof * sizeof(double), hipMemcpyHostToDevice, stream_));
  k::unpack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), packed_.get(), z_.get(), stream_);
  if (cfg_.points) {
    // the point-coefficient operator the solver uses
    launch_points_setup(xs_.get(), t, cm, ca);
    launch_points_apply(z_.get(), r_.get(), nullptr, false);
  } else {
    launch_apply(xs_.get(), t, cm, ca, z_.get(), r_.get(), nullptr);
  }
  k::pack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), r_.get(), packed_.get(), stream_);
  DCG_HIP(hipMemcpyAsync(y_packed, packed_.get(), L_.ndof * sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
}

void Model::spmv(const double* jac, const double* x_packed, double* y_packed) {
  require_device();
  if (cfg_.matrix_free) throw Error("a matrix-free model holds no assembled Jacobian");
  ensure_matrix();
  DCG_HIP(hipMemcpyAsync(val_.get(), jac, val_.bytes(), hipMemcpyHostToDevice, stream_));
  DCG_HIP(hipMemcpyAsync(packed_.get(), x_packed, L_.ndof * sizeof(double), hipMemcpyHostToDevice, stream_));
  k::unpack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), packed_.get(), xs_.get(), stream_);
  k::spmv(L_.B, plan_, val_.get(), xs_.get(), r_.get(), nullptr, nullptr, nullptr, nullptr, 0, stream_);
  k::pack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), r_.get(), packed_.get(), stream_);
  DCG_HIP(hipMemcpyAsync(y_packed, packed_.get(), L_.ndof * sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
}

int Model::linear_solve(const double* jac, const double* b_packed, double* x_packed, double rtol, int maxit,
                        int* its) {
  require_device();
  if (cfg_.matrix_free) throw Error("a matrix-free model holds no assembled Jacobian");
  ensure_matrix();
  DCG_HIP(hipMemcpyAsync(val_.get(), jac, val_.bytes(), hipMemcpyHostToDevice, stream_));
  DCG_HIP(hipMemcpyAsync(packed_.get(), b_packed, L_.ndof * sizeof(double), hipMemcpyHostToDevice, stream_));
  k::unpack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), packed_.get(), r_.get(), stream_);
  int refs = 0;
  // the given matrix is the operator, whatever the model's storage; a solve
  // outside a time step takes no place in the per-step iteration hints
  struct Restore {
    Model* m;
    bool points;
    int index;
    ~Restore() {
      m->use_points_ = points;
      m->solve_index_ = index;
    }
  } restore{this, use_points_, solve_index_};
  use_points_ = false;
  solve_index_ = kMaxSolveHints;
  int rc = krylov(r_.get(), z_.get(), rtol, maxit, its, &refs);
  k::pack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), z_.get(), packed_.get(), stream_);
  DCG_HIP(hipMemcpyAsync(x_packed, packed_.get(), L_.ndof * sizeof(double), hipMemcpyDeviceToHost, stream_));
  synchronize();
  return rc;
}

// dcg_linear_solve without device allocations: J, b and x are the caller's
// device buffers (BCSR values, packed state layout), every other buffer the
// FP64 solve needs (Krylov vectors, the unpacked right-hand side and
// solution, preconditioner, partial sums, solver scalars) is carved out of
// the library's fixed work space on this device.  Calls that overlap on one
// device share that space and race.
int Model::linear_solve_noalloc(const double* jac, const double* b_packed, double* x_packed, double rtol, int maxit,
                                int* its) {
  require_device();
  if (cfg_.matrix_free) throw Error("a matrix-free model holds no assembled Jacobian");
  if (cfg_.precision != Precision::FP64)
    throw NotImplemented("the allocation-free linear solve runs FP64 models only (linear_solver.precision = fp64)");
  if (!jac || !b_packed || !x_packed) throw Error("null device buffer");
  const size_t nv = static_cast<size_t>(L_.nslot) * L_.B;
  const size_t align = 32;  // doubles: every region starts on 256 bytes
  auto round = [&](size_t w) { return (w + align - 1) / align * align; };
  const bool ilu = cfg_.precond == Preconditioner::ILU0;
  const size_t sc_words = round((sizeof(k::BicgScalars) + sizeof(double) - 1) / sizeof(double));
  const size_t words = 11 * round(nv) + round(static_cast<size_t>(L_.nslot) * L_.B * L_.B) +
                       (ilu ? round(static_cast<size_t>(matrix_values())) : 0) + round(part_.size()) + sc_words +
                       round(2);
  const size_t limit = k::kNoallocScratchBytes / sizeof(double);
  if (words > limit)
    throw Error("the allocation-free linear solve needs " + std::to_string(words * sizeof(double)) +
                " bytes of work space for this model; the fixed work space holds " +
                std::to_string(k::kNoallocScratchBytes) + " bytes");
  double* base = k::noalloc_scratch(device_);
  size_t off = 0;
  auto take = [&](size_t w) {
    double* q = base + off;
    off += round(w);
    return q;
  };
  KrylovWork w;
  for (auto& v : w.kv) v = take(nv);
  double* rhs = take(nv);
  double* sol = take(nv);
  w.dinv = take(static_cast<size_t>(L_.nslot) * L_.B * L_.B);
  w.lu = ilu ? take(static_cast<size_t>(matrix_values())) : nullptr;
  w.part = take(part_.size());
  w.spart = w.part + spart_off_;
  w.sc = reinterpret_cast<k::BicgScalars*>(base + off);
  off += sc_words;
  w.status = reinterpret_cast<int*>(take(2));
  w.val = jac;
  // the given matrix is the operator; the solve keeps no per-step hint
  struct Restore {
    Model* m;
    bool points;
    int index;
    ~Restore() {
      m->use_points_ = points;
      m->solve_index_ = index;
    }
  } restore{this, use_points_, solve_index_};
  use_points_ = false;
  solve_index_ = kMaxSolveHints;
  k::fill(w.status, 0, 4, stream_);
  k::unpack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), b_packed, rhs, stream_);
  precond_setup(w);
  int rc = bicgstab64(rhs, sol, rtol, 0.0, maxit, its, w);
  k::pack_state(L_.nslot, L_.B, slot_b_.get(), slot_dof_.get(), sol, x_packed, stream_);
  synchronize();
  return rc;
}

// ------------------------------------------------------------------ parameters, footprint, batch view
void Model::set_params(const double* values, size_t n) {
  require_device();
  if (!opt_.runtime_parameters) throw Error("model was created with folded constants: no run-time parameters");
  if (n != spec_.params.size())
    throw Error("expected " + std::to_string(spec_.params.size()) + " parameters, got " + std::to_string(n));
  spec_.param_values.assign(values, values + n);
  if (n) DCG_HIP(hipMemcpyAsync(prm_.get() + kGeoWords, values, n * sizeof(double), hipMemcpyHostToDevice, stream_));
  synchronize();
}

void Model::get_params(double* values, size_t n) const {
  if (n != spec_.param_values.size())
    throw Error("expected room for " + std::to_string(spec_.param_values.size()) + " parameters");
  std::copy(spec_.param_values.begin(), spec_.param_values.end(), values);
}

size_t Model::device_bytes() const {
  size_t b = row_ptr_.bytes() + col_.bytes() + own_start_.bytes() + diag_pos_.bytes() + slot_b_.bytes() +
             chunk_row_.bytes() + own_mask_.bytes() + slot_dof_.bytes() + extra_pos_.bytes() + weights_.bytes() +
             prm_.bytes() + u_.bytes() + hist_.bytes() + r_.bytes() + z_.bytes() + xs_.bytes() + packed_.bytes() +
             val_.bytes() + lu_.bytes() + dinv_.bytes() + val32_.bytes() + part_.bytes() + scalar_out_.bytes() +
             sc_.bytes() + status_.bytes() + lo_rows_.bytes() + up_rows_.bytes() + pdiag_.bytes() + diag_id_.bytes() +
             lin_x_.bytes() + lo_ent_.bytes() + lo_blk_.bytes() + lo_src_.bytes() + up_ent_.bytes() + up_blk_.bytes() +
             up_src_.bytes() + lo_chunk_.bytes() + up_chunk_.bytes() +
             ilu_frow_.bytes() + ilu_fub_.bytes() + ilu_fm_.bytes() + ilu_fpush_.bytes() + ilu_packed_.bytes() +
             ilu_inbox_.bytes();
  for (const auto& v : pc_) b += v.bytes();
  for (const auto& v : pc32_) b += v.bytes();
  for (const auto& v : stage_) b += v.bytes();
  for (const auto& v : kv_) b += v.bytes();
  for (const auto& v : kf_) b += v.bytes();
  for (const auto& c : comp_dev_)
    b += c.cells.bytes() + c.cell_slots.bytes() + c.cell_mask.bytes() + c.g2l.bytes() + c.tiles.bytes() +
         c.cell_mask_grid.bytes();
  for (const auto& f : facet_dev_) b += f.fint.bytes() + f.fdbl.bytes();
  return b;
}

// Device bytes of a model with layout L, before anything is allocated: the
// matrix (FP64, its FP32 copy when mixed, the ILU factor when used), the
// block-diagonal inverse, the per-slot vectors and the mesh tables.
size_t Model::estimate_device_bytes(const Layout& L, int stages, bool mixed, bool ilu, bool matrix_free) {
  const double nval = matrix_free ? 0.0 : static_cast<double>(L.nnzb()) * L.B * L.B;
  const double ns = static_cast<double>(L.nslot), nv = ns * L.B;
  double b = 8.0 * nval * (ilu ? 2.0 : 1.0) + (mixed ? 4.0 * nval : 0.0) + (matrix_free ? 8.0 * nv : 8.0 * ns * L.B * L.B);
  b += 8.0 * nv * (5 + stages + 1 + 9) + (mixed ? 4.0 * nv * 10 : 0.0) + 8.0 * L.ndof;
  b += 4.0 * (ns + 1) + 4.0 * L.nnzb() + 4.0 * ns * 4 + 8.0 * ns * 2;
  double cells = 0;
  for (const auto& C : L.comps) cells += static_cast<double>(C.cells.size());
  b += cells * (4.0 + 1.0 + 4.0 * L.grid.ncorner) + (4.0 + 1.0) * static_cast<double>(L.grid.nvert) * L.comps.size();
  return static_cast<size_t>(b);
}

k::MemberView Model::view() const {
  k::MemberView v;
  v.nslot = static_cast<int>(L_.nslot);
  v.nchunk = plan_.nchunk;
  v.rp = row_ptr_.get();
  v.col = col_.get();
  v.diag_pos = diag_pos_.get();
  v.chunk_row = chunk_row_.get();
  v.val = val_.get();
  v.val32 = val32_.get();
  v.dinv = dinv_.get();
  if (cfg_.points) {
    // the block inverse reads the rows' diagonal blocks through an identity
    // position map; the FP32 copy converts the point coefficients
    v.val = pdiag_.get();
    v.diag_pos = diag_id_.get();
    v.nchunk = 0;
    for (size_t c = 0; c < pc_.size(); ++c)
      if (pc_[c].size()) {
        v.pc = pc_[c].get();
        v.pc32 = c < pc32_.size() ? pc32_[c].get() : nullptr;
        v.npc = static_cast<long>(pc_[c].size());
      }
  }
  v.v[k::kVU] = u_.get();
  v.v[k::kVHist] = hist_.get();
  v.v[k::kVR] = r_.get();
  v.v[k::kVZ] = z_.get();
  v.v[k::kVXs] = xs_.get();
  for (size_t j = 0; j < stage_.size() && j < 3; ++j) v.v[k::kVStage + j] = stage_[j].get();
  for (int i = 0; i < 9; ++i) v.v[k::kVKry + i] = kv_[i].get();
  for (int i = 0; i < k::kBatchVecF; ++i) v.f[i] = kf_[i].get();
  return v;
}

}  // namespace dcopasi
