#include "batch.hh"

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cmath>

namespace dcopasi {

BatchOptions BatchOptions::from_ini(const std::string& text) {
  BatchOptions o;
  auto c = ParameterTree::from_ini(text, "");
  const std::string prec = c.get("precision", std::string("mixed"));
  if (prec == "fp64") o.precision = Precision::FP64;
  else if (prec == "mixed") o.precision = Precision::Mixed;
  else throw IOError("batch precision must be 'fp64' or 'mixed'");
  const std::string pre = c.get("preconditioner", std::string("Jacobi"));
  if (pre == "Jacobi") o.precond = Preconditioner::Jacobi;
  else if (pre == "VertexBlockJacobi") o.precond = Preconditioner::VertexBlockJacobi;
  else if (pre == "None") o.precond = Preconditioner::None;
  else throw IOError("batch preconditioner must be 'Jacobi', 'VertexBlockJacobi' or 'None'");
  o.ir_inner_rtol = c.get("refinement.inner_relative_tolerance", o.ir_inner_rtol);
  o.ir_floor = c.get("refinement.tolerance_floor", o.ir_floor);
  if (o.ir_floor < 0.0 || o.ir_floor >= 1.0) throw IOError("refinement.tolerance_floor must be in [0, 1)");
  o.ir_max = c.get("refinement.max_steps", o.ir_max);
  o.lin_maxit = c.get("iteration_limit", o.lin_maxit);
  o.graph_iterations = c.get("graph_iterations", o.graph_iterations);
  {
    const std::string pt = c.get("point_tile", std::string("auto"));
    if (pt != "auto") {
      char* end = nullptr;
      const long k2 = std::strtol(pt.c_str(), &end, 10);
      if (pt.empty() || *end || k2 < 0 || k2 > 15) throw IOError("point_tile must be 'auto' or a candidate index in [0, 15]");
      o.point_tile = static_cast<int>(k2);
    }
  }
  if (o.graph_iterations < 0 || o.graph_iterations > 64) throw IOError("graph_iterations must be in [0, 64]");
  return o;
}

bool Model::needs_own_assembly() const {
  for (size_t c = 0; c < L_.comps.size(); ++c)
    if ((vol_fn_[c] || tile_fn_[c]) &&
        (c >= static_cast<size_t>(k::kAsmComps) || gen_.tile_batch_kernel[c].empty()))
      return true;
  return !facet_fn_.empty() || extra_pos_.size() > 0;
}

Batch::Batch(int device, const BatchOptions& opt) : device_(device), opt_(opt) {
  DCG_HIP(hipSetDevice(device_));
  DCG_HIP(hipStreamCreate(&stream_));
  status_.resize(4);
}

Batch::~Batch() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
  destroy_graphs();
  members_.clear();
  sgroups_.clear();
  agroups_.clear();
  status_.release();
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::destroy_graphs() {
  for (auto& g : sgroups_)
    for (auto& e : g.graph)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
}

ModelOptions Batch::member_options() const {
  ModelOptions mo;
  mo.runtime_parameters = true;
  mo.batch_member = true;
  return mo;
}

// the batch decides the linear solver: members allocate what it needs
std::string Batch::member_overrides(const std::string& overrides) const {
  std::string ov = overrides;
  ov += "\nmodel.time_step_operator.linear_solver.precision = ";
  ov += opt_.precision == Precision::Mixed ? "mixed" : "fp64";
  ov += "\nmodel.time_step_operator.linear_solver.preconditioner.type = ";
  ov += opt_.precond == Preconditioner::Jacobi            ? "Jacobi"
        : opt_.precond == Preconditioner::VertexBlockJacobi ? "VertexBlockJacobi"
                                                             : "None";
  ov += "\n";
  return ov;
}

int Batch::add(const std::string& ini, const std::string& overrides) {
  DCG_HIP(hipSetDevice(device_));
  ModelOptions mo = member_options();
  mo.stream = stream_;
  auto m = std::make_unique<Model>(ini, member_overrides(overrides), device_, mo);
  if (m->rk().stages > 2) throw NotImplemented("batch members use schemes of at most two stages");
  Entry e;
  e.id = next_id_++;
  e.m = std::move(m);
  e.dt = e.m->time_step();
  members_.push_back(std::move(e));
  dirty_ = true;
  return members_.back().id;
}

int Batch::add(std::unique_ptr<Model> m) {
  if (!m) throw Error("null model");
  if (!m->opt_.runtime_parameters || !m->opt_.batch_member)
    throw Error("batch members are built with Batch::member_options()");
  if (m->rk().stages > 2) throw NotImplemented("batch members use schemes of at most two stages");
  DCG_HIP(hipSetDevice(device_));
  if (!m->on_device()) m->attach(device_, stream_);
  else if (m->stream() != stream_) throw Error("model lives on another stream");
  Entry e;
  e.id = next_id_++;
  e.m = std::move(m);
  e.dt = e.m->time_step();
  members_.push_back(std::move(e));
  dirty_ = true;
  return members_.back().id;
}

void Batch::remove(int id) {
  for (size_t i = 0; i < members_.size(); ++i)
    if (members_[i].id == id) {
      DCG_HIP(hipStreamSynchronize(stream_));
      destroy_graphs();
      members_.erase(members_.begin() + static_cast<long>(i));
      dirty_ = true;
      return;
    }
  throw Error("no batch member with id " + std::to_string(id));
}

Model& Batch::member(int id) {
  for (auto& e : members_)
    if (e.id == id) return *e.m;
  throw Error("no batch member with id " + std::to_string(id));
}

std::vector<int> Batch::ids() const {
  std::vector<int> v;
  for (const auto& e : members_) v.push_back(e.id);
  return v;
}

size_t Batch::device_bytes() const {
  size_t b = 0;
  for (const auto& e : members_) b += e.m->device_bytes();
  return b;
}

void Batch::set_time_step(int id, double dt) {
  for (auto& e : members_)
    if (e.id == id) {
      e.dt = dt;
      return;
    }
  throw Error("no batch member with id " + std::to_string(id));
}

double Batch::time_step(int id) {
  for (auto& e : members_)
    if (e.id == id) return e.dt;
  throw Error("no batch member with id " + std::to_string(id));
}

// ------------------------------------------------------------------ work lists
void Batch::rebuild() {
  DCG_HIP(hipStreamSynchronize(stream_));
  destroy_graphs();
  sgroups_.clear();
  agroups_.clear();
  // solver groups by block size and Jacobian storage
  std::map<std::pair<int, bool>, int> by_b;
  for (size_t i = 0; i < members_.size(); ++i) {
    Entry& e = members_[i];
    const int B = e.m->layout().B;
    const bool pts = e.m->point_storage();
    auto it = by_b.find({B, pts});
    if (it == by_b.end()) {
      it = by_b.emplace(std::make_pair(B, pts), static_cast<int>(sgroups_.size())).first;
      sgroups_.emplace_back();
      sgroups_.back().B = B;
      sgroups_.back().points = pts;
    }
    e.sg = it->second;
    e.si = static_cast<int>(sgroups_[e.sg].mem.size());
    sgroups_[e.sg].mem.push_back(static_cast<int>(i));
  }
  for (auto& g : sgroups_) {
    const int M = static_cast<int>(g.mem.size());
    std::vector<k::MemberView> mv(M);
    std::vector<int2> vwork, vrange(M), vwork4, vrange4(M), crange(M);
    std::vector<int4> cwork;
    for (int j = 0; j < M; ++j) {
      const Model& m = *members_[g.mem[j]].m;
      mv[j] = m.view();
      const int nblk = (mv[j].nslot + 255) / 256;
      vrange[j].x = static_cast<int>(vwork.size());
      for (int b = 0; b < nblk; ++b) vwork.push_back(make_int2(j, b));
      vrange[j].y = static_cast<int>(vwork.size());
      const int nblk4 = (mv[j].nslot + 256 * k::kBVecSlots - 1) / (256 * k::kBVecSlots);
      vrange4[j].x = static_cast<int>(vwork4.size());
      for (int b = 0; b < nblk4; ++b) vwork4.push_back(make_int2(j, b));
      vrange4[j].y = static_cast<int>(vwork4.size());
      crange[j].x = static_cast<int>(cwork.size());
      const auto& ch = m.chunk_host_;
      const auto& rp = m.L_.row_ptr;
      for (int c = 0; c < mv[j].nchunk; ++c) {
        const int r0 = ch[c], r1 = ch[c + 1], k0 = rp[r0], k1 = rp[r1];
        if (r1 - r0 > 0xffff || k1 - k0 > 0xffff) throw Error("internal: SpMV chunk too large for the work list");
        cwork.push_back(make_int4(j, r0, k0, ((r1 - r0) << 16) | (k1 - k0)));
      }
      crange[j].y = static_cast<int>(cwork.size());
    }
    g.h_act.assign(M, 0);
    g.h_rtol.assign(M, 0.0);
    if (g.points) cwork.clear();  // no matrix chunks: the point apply is the operator
    g.mv.upload(mv, stream_);
    g.vwork.upload(vwork, stream_);
    g.vrange.upload(vrange, stream_);
    g.vwork4.upload(vwork4, stream_);
    g.vrange4.upload(vrange4, stream_);
    g.cwork.upload(cwork, stream_);
    g.crange.upload(crange, stream_);
    g.sc.resize(M);
    g.ir.resize(M);
    g.act.resize(M);
    g.rtol.resize(M);
    g.sums.resize(M);
    g.part.resize(2 * std::max(vwork.size(), cwork.size()) + 64);
    g.w.vwork = g.vwork.get();
    g.w.nvwork = static_cast<int>(vwork.size());
    g.w.vrange = g.vrange.get();
    g.w.vwork4 = g.vwork4.get();
    g.w.nvwork4 = static_cast<int>(vwork4.size());
    g.w.vrange4 = g.vrange4.get();
    g.w.cwork = g.cwork.get();
    g.w.ncwork = static_cast<int>(cwork.size());
    g.w.crange = g.crange.get();
    g.w.members = M;
    if (g.points) {
      // per member: its compartment's tables, coefficients and vectors; per
      // generated module: one launch over its members' (member, tile) items
      std::vector<k::PcMember> pcm(M);
      std::vector<int2> prange(M);
      std::map<const RtcModule*, size_t> by_mod;
      // per launch and candidate tile shape: the (member, tile) items and
      // each owner's launch-local [first, last)
      std::vector<std::vector<std::vector<int2>>> items;
      std::vector<std::vector<std::vector<int2>>> ranges;
      std::vector<std::vector<int>> owners;
      for (int j = 0; j < M; ++j) {
        Model& m = *members_[g.mem[j]].m;
        size_t c0 = 0;
        while (c0 < m.pc_.size() && !m.pc_[c0].size()) ++c0;
        if (c0 >= m.pc_.size()) throw Error("internal: point member without coefficients");
        k::PcMember& pm = pcm[j];
        pm.g2l = m.comp_dev_[c0].g2l_identity ? nullptr : m.comp_dev_[c0].g2l.get();
        pm.cmask = m.comp_dev_[c0].cell_mask_grid.get();
        pm.pc = m.pc_[c0].get();
        pm.pc32 = c0 < m.pc32_.size() ? m.pc32_[c0].get() : nullptr;
        pm.tca = m.pc_tca_.get();
        pm.prm = m.prm_.get();
        const Grid& gr = m.L_.grid;
        pm.n[0] = static_cast<int>(gr.n[0]);
        pm.n[1] = gr.dim > 1 ? static_cast<int>(gr.n[1]) : 1;
        pm.n[2] = gr.dim > 2 ? static_cast<int>(gr.n[2]) : 1;
        for (int v2 = 0; v2 < k::kBatchVec; ++v2) pm.v[v2] = mv[j].v[v2];
        for (int v2 = 0; v2 < k::kBatchVecF; ++v2) pm.f[v2] = mv[j].f[v2];
        pm.done = reinterpret_cast<const int*>(reinterpret_cast<const char*>(g.sc.get() + j) +
                                               offsetof(k::BicgScalars, done));
        pm.irdone = reinterpret_cast<const int*>(reinterpret_cast<const char*>(g.ir.get() + j) +
                                                 offsetof(k::IrState, done));
        const RtcModule* key = m.module_.get();
        const size_t nk = m.gen_.apply_tiles[c0].size();
        auto it = by_mod.find(key);
        if (it == by_mod.end()) {
          it = by_mod.emplace(key, g.plaunch.size()).first;
          g.plaunch.emplace_back();
          auto& pl = g.plaunch.back();
          const std::string name = m.gen_.tile_kernel[c0];
          for (size_t k2 = 0; k2 < nk; ++k2) {
            const std::string id = std::to_string(k2);
            pl.f64s.push_back(m.module_->function(name + "_pbatch" + id));
            pl.f32s.push_back(opt_.precision == Precision::Mixed ? m.module_->function(name + "_pfbatch" + id)
                                                                 : nullptr);
            pl.tiles.push_back(m.gen_.apply_tiles[c0][k2]);
            pl.nts.push_back(apply_tile_threads(m.L_.grid.dim, m.gen_.apply_tiles[c0][k2], m.L_.comps[c0].b));
          }
          pl.module = m.module_;
          items.emplace_back(nk);
          ranges.emplace_back(nk);
          owners.emplace_back();
        }
        const size_t l = it->second;
        for (size_t k2 = 0; k2 < nk; ++k2) {
          const auto& cd = m.comp_dev_[c0];
          std::vector<int> tiles(cd.natiles[k2]);
          if (!tiles.empty())
            DCG_HIP(hipMemcpy(tiles.data(), cd.atiles[k2].get(), tiles.size() * sizeof(int), hipMemcpyDeviceToHost));
          auto& it2 = items[l][k2];
          int2 r = make_int2(static_cast<int>(it2.size()), 0);
          for (int t : tiles) it2.push_back(make_int2(j, t));
          r.y = static_cast<int>(it2.size());
          ranges[l][k2].push_back(r);
        }
        owners[l].push_back(j);
      }
      g.pcm.upload(pcm, stream_);
      // the tile shape of each launch: the candidate with the shortest
      // batched apply over the launch's members (no skips, no dot products;
      // the candidates run in turn over several rounds)
      hipEvent_t e0, e1;
      DCG_HIP(hipEventCreate(&e0));
      DCG_HIP(hipEventCreate(&e1));
      int off = 0;
      for (size_t l = 0; l < g.plaunch.size(); ++l) {
        auto& pl = g.plaunch[l];
        const size_t nk = pl.f64s.size();
        size_t best = 0;
        // a module timed at an earlier rebuild keeps its choice (members join
        // and leave between steps; timing again each time would cost more
        // than it finds)
        auto known = tuned_.find(pl.module.get());
        if (known != tuned_.end() && (known->second.first.expired() || known->second.second >= nk)) {
          tuned_.erase(known);
          known = tuned_.end();
        }
        if (opt_.point_tile >= 0) {
          best = static_cast<size_t>(opt_.point_tile) < nk ? static_cast<size_t>(opt_.point_tile) : 0;
        } else if (known != tuned_.end()) {
          best = known->second.second;
        } else if (nk > 1) {
          std::vector<DeviceArray<int2>> dw(nk);
          for (size_t k2 = 0; k2 < nk; ++k2) dw[k2].upload(items[l][k2], stream_);
          std::vector<double> ms(nk, 0.0);
          const k::PcMember* mem = g.pcm.get();
          for (int round = 0; round < 4; ++round)
            for (size_t k2 = 0; k2 < nk; ++k2) {
              const int2* work = dw[k2].get();
              int nwork = static_cast<int>(items[l][k2].size()), o0 = 0, xs = k::kVKry, ys = k::kVKry + 8, a1 = -1,
                  a2 = -1, sk = -1;
              double* part = nullptr;
              void* args[] = {&work, &nwork, &o0, &mem, &xs, &ys, &a1, &a2, &sk, &part};
              const unsigned grid = static_cast<unsigned>((nwork + 7) / 8 * 8);
              if (nwork <= 0) continue;
              DCG_HIP(hipEventRecord(e0, stream_));
              for (int rep = 0; rep < 2; ++rep)
                DCG_HIP(hipModuleLaunchKernel(pl.f64s[k2], grid, 1, 1, static_cast<unsigned>(pl.nts[k2]), 1, 1, 0,
                                              stream_, args, nullptr));
              DCG_HIP(hipEventRecord(e1, stream_));
              DCG_HIP(hipEventSynchronize(e1));
              float t = 0.f;
              DCG_HIP(hipEventElapsedTime(&t, e0, e1));
              if (round) ms[k2] += t;
            }
          for (size_t k2 = 1; k2 < nk; ++k2)
            if (ms[k2] < ms[best]) best = k2;
          tuned_[pl.module.get()] = {pl.module, best};
        }
        pl.f64 = pl.f64s[best];
        pl.f32 = pl.f32s[best];
        pl.tile = pl.tiles[best];
        pl.nt = pl.nts[best];
        pl.work.upload(items[l][best], stream_);
        pl.nwork = static_cast<int>(items[l][best].size());
        pl.off = off;
        for (size_t o = 0; o < owners[l].size(); ++o) {
          const int j = owners[l][o];
          prange[j] = make_int2(ranges[l][best][o].x + off, ranges[l][best][o].y + off);
        }
        off += pl.nwork;
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      g.nitems = off;
      g.prange.upload(prange, stream_);
      g.part.resize(2 * std::max<size_t>(std::max(vwork.size(), static_cast<size_t>(g.nitems)), 1) + 64);
      const int gi = static_cast<int>(&g - sgroups_.data());
      g.op64.range = g.prange.get();
      g.op64.apply = [this, gi](int x, int y, int w1, int w2, double* part, int skip) {
        point_apply(gi, false, x, y, w1, w2, part, skip);
      };
      g.op32.range = g.prange.get();
      g.op32.apply = [this, gi](int x, int y, int w1, int w2, double* part, int skip) {
        point_apply(gi, true, x, y, w1, w2, part, skip);
      };
    }
    tune(g);
  }
  // assembly groups by shared module
  std::map<const RtcModule*, int> by_mod;
  for (size_t i = 0; i < members_.size(); ++i) {
    Entry& e = members_[i];
    e.per_member_asm = e.m->needs_own_assembly();
    const RtcModule* key = e.m->module_.get();
    auto it = by_mod.find(key);
    if (it == by_mod.end()) {
      it = by_mod.emplace(key, static_cast<int>(agroups_.size())).first;
      agroups_.emplace_back();
      AsmGroup& a = agroups_.back();
      a.module = e.m->module_;
      const auto& names = e.m->gen_.tile_batch_kernel;
      a.fn.assign(names.size(), nullptr);
      a.qfn.assign(names.size(), nullptr);
      a.points = e.m->point_storage();
      for (size_t c = 0; c < names.size(); ++c)
        if (!names[c].empty()) {
          a.fn[c] = a.module->function(names[c]);
          if (a.points) a.qfn[c] = a.module->function(e.m->gen_.tile_kernel[c] + "_qbatch");
        }
      a.work.resize(names.size());
      a.nwork.assign(names.size(), 0);
    }
    e.ag = it->second;
    e.ai = static_cast<int>(agroups_[e.ag].mem.size());
    agroups_[e.ag].mem.push_back(static_cast<int>(i));
  }
  for (auto& a : agroups_) {
    const int M = static_cast<int>(a.mem.size());
    std::vector<k::AsmMember> am(M);
    std::vector<std::vector<int2>> work(a.fn.size());
    for (int j = 0; j < M; ++j) {
      Model& m = *members_[a.mem[j]].m;
      const auto v = m.view();
      for (int s = 0; s < 8; ++s) am[j].vec[s] = v.v[s];
      am[j].val = m.val_.get();
      am[j].own_start = m.own_start_.get();
      am[j].own_mask = m.own_mask_.get();
      am[j].prm = m.prm_.get();
      const Grid& gr = m.L_.grid;
      am[j].n[0] = static_cast<int>(gr.n[0]);
      am[j].n[1] = gr.dim > 1 ? static_cast<int>(gr.n[1]) : 1;
      am[j].n[2] = gr.dim > 2 ? static_cast<int>(gr.n[2]) : 1;
      for (int k2 = 0; k2 < 3; ++k2) {
        am[j].origin[k2] = gr.origin[k2];
        am[j].h[k2] = k2 < gr.dim ? gr.h[k2] : 0.0;
      }
      if (a.points) {
        for (size_t c = 0; c < m.pc_.size() && c < static_cast<size_t>(k::kAsmComps); ++c)
          am[j].pc[c] = m.pc_[c].size() ? m.pc_[c].get() : nullptr;
        am[j].pdiag = m.pdiag_.get();
        am[j].tca = m.pc_tca_.get();
      }
      for (size_t c = 0; c < a.fn.size(); ++c) {
        if (!a.fn[c] || c >= static_cast<size_t>(k::kAsmComps)) continue;
        am[j].g2l[c] = m.comp_dev_[c].g2l.get();
        am[j].cmask[c] = m.comp_dev_[c].cell_mask_grid.get();
        // the member's tile list, read back once from its set-up tables
        std::vector<int> tiles(m.comp_dev_[c].ntiles);
        if (!tiles.empty())
          DCG_HIP(hipMemcpy(tiles.data(), m.comp_dev_[c].tiles.get(), tiles.size() * sizeof(int),
                            hipMemcpyDeviceToHost));
        for (int t : tiles) work[c].push_back(make_int2(j, t));
      }
    }
    a.am.upload(am, stream_);
    for (size_t c = 0; c < a.fn.size(); ++c) {
      a.work[c].upload(work[c], stream_);
      a.nwork[c] = static_cast<int>(work[c].size());
    }
    a.coef.resize(static_cast<size_t>(kCalls) * M * 4);
    a.act.resize(2 * static_cast<size_t>(M));
    a.h_act.assign(2 * static_cast<size_t>(M), 0);
    a.h_coef.assign(static_cast<size_t>(kCalls) * M * 4, 0.0);
  }
  DCG_HIP(hipStreamSynchronize(stream_));
  dirty_ = false;
}

// Times the LDS-staged SpMV variants (entry and block form) on the group's
// batched work list and keeps the faster one per precision.
void Batch::tune(SolverGroup& g) {
  if (g.points) return;  // the point apply has no SpMV variants
  const k::Skip none{};
  // zeroed operands: x = ph, y = v of the Krylov set
  k::b_copy(g.B, g.w, g.mv.get(), none, k::kVKry + 4, -1, stream_);
  hipEvent_t e0, e1;
  DCG_HIP(hipEventCreate(&e0));
  DCG_HIP(hipEventCreate(&e1));
  auto time_variant = [&](int variant, bool fp32) {
    auto run = [&]() {
      if (fp32) k::b_spmv32(g.B, variant, g.w, g.mv.get(), none, k::kVKry + 4, k::kVKry + 3, stream_);
      else k::b_spmv64(g.B, variant, g.w, g.mv.get(), none, k::kVKry + 4, k::kVKry + 3, stream_);
    };
    for (int w = 0; w < 2; ++w) run();
    DCG_HIP(hipEventRecord(e0, stream_));
    for (int w = 0; w < 4; ++w) run();
    DCG_HIP(hipEventRecord(e1, stream_));
    DCG_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DCG_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms;
  };
  g.variant64 = time_variant(5, false) < time_variant(0, false) ? 5 : 0;
  if (opt_.precision == Precision::Mixed) {
    k::b_copy_f(g.B, g.w, g.mv.get(), none, k::kVKry + 4, stream_);
    g.variant32 = time_variant(5, true) < time_variant(0, true) ? 5 : 0;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void Batch::capture(SolverGroup& g, bool fp32) {
  hipGraph_t graph;
  DCG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < opt_.graph_iterations; ++it) {
    if (fp32)
      k::b_bicg_iteration<float>(g.B, g.variant32, g.w, g.mv.get(), k::kVKry + 9, jacobi_kind(), g.part.get(),
                                 g.sc.get(), stream_, nullptr, g.points ? &g.op32 : nullptr);
    else
      k::b_bicg_iteration<double>(g.B, g.variant64, g.w, g.mv.get(), k::kVZ, jacobi_kind(), g.part.get(), g.sc.get(),
                                  stream_, nullptr, g.points ? &g.op64 : nullptr);
  }
  DCG_HIP(hipStreamEndCapture(stream_, &graph));
  hipGraphExec_t exec = nullptr;
  DCG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  (void)hipGraphDestroy(graph);
  g.graph[fp32 ? 1 : 0] = exec;
}

void Batch::set_profiling(bool on) {
  DCG_HIP(hipStreamSynchronize(stream_));
  if (on && !ev_[0])
    for (auto& e : ev_) DCG_HIP(hipEventCreate(&e));
  prof_ = on;
  for (auto& p : prof_data_) p = Prof{};
}

void Batch::profile(int kind, long* launches, double* ms, double* bytes) const {
  if (kind < 0 || kind > 1) throw Error("batch profile kind must be 0 (FP64 SpMV) or 1 (FP32 SpMV)");
  *launches = prof_data_[kind].launches;
  *ms = prof_data_[kind].ms;
  *bytes = prof_data_[kind].bytes;
}

size_t Batch::n_dof() const {
  size_t n = 0;
  for (const auto& e : members_) n += e.m->n_dof();
  return n;
}

std::vector<std::array<int, 3>> Batch::spmv_choice() {
  if (dirty_) rebuild();
  std::vector<std::array<int, 3>> v;
  for (const auto& g : sgroups_) v.push_back({g.B, g.variant64, g.variant32});
  return v;
}

// the preconditioner as the batched updates take it: 0 none, 1 vertex
// blocks, 2 the scalar diagonal (B inverse values per slot)
int Batch::jacobi_kind() const {
  return opt_.precond == Preconditioner::None ? 0 : opt_.precond == Preconditioner::Jacobi ? 2 : 1;
}

std::vector<std::array<int, 3>> Batch::point_tiles() {
  if (dirty_) rebuild();
  std::vector<std::array<int, 3>> v;
  for (const auto& g : sgroups_)
    for (const auto& pl : g.plaunch) v.push_back(pl.tile);
  return v;
}

// y = A x over the point members of group gi: one launch per generated module
// (grid a multiple of 8; the kernel deals items to the XCDs in contiguous
// runs).  FP32 solves index the members' FP32 vectors (selector - kVKry).
void Batch::point_apply(int gi, bool fp32, int x, int y, int w1, int w2, double* part, int skip) {
  SolverGroup& g = sgroups_[gi];
  if (fp32) {
    x -= k::kVKry;
    y -= k::kVKry;
    if (w1 >= 0) w1 -= k::kVKry;
    if (w2 >= 0) w2 -= k::kVKry;
  }
  const k::PcMember* mem = g.pcm.get();
  for (auto& pl : g.plaunch) {
    if (pl.nwork <= 0) continue;
    const int2* work = pl.work.get();
    int nwork = pl.nwork, off = pl.off, xs = x, ys = y, a1 = w1, a2 = w2, sk = skip;
    void* args[] = {&work, &nwork, &off, &mem, &xs, &ys, &a1, &a2, &sk, &part};
    const unsigned grid = static_cast<unsigned>((pl.nwork + 7) / 8 * 8);
    DCG_HIP(hipModuleLaunchKernel(fp32 ? pl.f32 : pl.f64, grid, 1, 1, static_cast<unsigned>(pl.nt), 1, 1, 0, stream_,
                                  args, nullptr));
  }
}

// ------------------------------------------------------------------ batched operations
void Batch::push_flags(const std::vector<char>& res, const std::vector<char>& jac) {
  // flags go to the device only when they change (no kernel writes them)
  for (auto& g : sgroups_) {
    for (size_t j = 0; j < g.mem.size(); ++j) g.h_act[j] = res[g.mem[j]] ? 1 : 0;
    if (g.h_act == g.dev_act) continue;
    DCG_HIP(hipMemcpyAsync(g.act.get(), g.h_act.data(), g.h_act.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    g.dev_act = g.h_act;
  }
  for (auto& a : agroups_) {
    const size_t M = a.mem.size();
    for (size_t j = 0; j < M; ++j) {
      a.h_act[j] = res[a.mem[j]] ? 1 : 0;
      a.h_act[M + j] = jac[a.mem[j]] ? 1 : 0;
    }
    if (a.h_act == a.dev_act) continue;
    DCG_HIP(hipMemcpyAsync(a.act.get(), a.h_act.data(), a.h_act.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    a.dev_act = a.h_act;
  }
}

// One assembly pass for the members with on[i] (their flags must already be
// pushed: `on` is the residual flag set for mode 1, the Jacobian set for mode 2).
void Batch::assemble(int call, const std::vector<char>& on, int xsel, int bsel, int rsel, int mode) {
  // per-member preparation (interface blocks, coloured-kernel starts)
  for (size_t i = 0; i < members_.size(); ++i) {
    if (!on[i] || !members_[i].per_member_asm) continue;
    Entry& e = members_[i];
    const auto v = e.m->view();
    const auto& cf = coef_[i][call];
    e.m->launch_assembly(v.v[xsel], cf[0], cf[1], cf[2], bsel >= 0 ? v.v[bsel] : nullptr,
                         rsel >= 0 ? v.v[rsel] : nullptr, (mode & 2) ? e.m->val_.get() : nullptr, mode, true, 1);
  }
  // batched tiles
  for (auto& a : agroups_) {
    const int* act = a.act.get() + ((mode & 2) ? a.mem.size() : 0);
    const double* coef = a.coef.get() + static_cast<size_t>(call) * a.mem.size() * 4;
    for (size_t c = 0; c < a.fn.size(); ++c) {
      if (!a.fn[c] || a.nwork[c] == 0) continue;
      const int2* work = a.work[c].get();
      const k::AsmMember* mem = a.am.get();
      int xs = xsel, bs = bsel, rs = rsel, md = mode;
      if (a.points && mode == 2) {
        // the point coefficients and diagonal blocks replace the matrix
        void* qargs[] = {&work, &mem, &coef, &act, &xs};
        DCG_HIP(hipModuleLaunchKernel(a.qfn[c], static_cast<unsigned>(a.nwork[c]), 1, 1, 256, 1, 1, 0, stream_,
                                      qargs, nullptr));
        continue;
      }
      void* args[] = {&work, &mem, &coef, &act, &xs, &bs, &rs, &md};
      DCG_HIP(hipModuleLaunchKernel(a.fn[c], static_cast<unsigned>(a.nwork[c]), 1, 1, 256, 1, 1, 0, stream_, args,
                                    nullptr));
    }
  }
  // per-member coloured volume kernels and facets
  for (size_t i = 0; i < members_.size(); ++i) {
    if (!on[i] || !members_[i].per_member_asm) continue;
    Entry& e = members_[i];
    const auto v = e.m->view();
    const auto& cf = coef_[i][call];
    e.m->launch_assembly(v.v[xsel], cf[0], cf[1], cf[2], bsel >= 0 ? v.v[bsel] : nullptr,
                         rsel >= 0 ? v.v[rsel] : nullptr, (mode & 2) ? e.m->val_.get() : nullptr, mode, true, 6);
  }
}

void Batch::copy(const std::vector<char>& on, int dst, int src) {
  std::vector<char> none(on.size(), 0);
  push_flags(on, none);
  for (auto& g : sgroups_) {
    const k::Skip sk{g.act.get(), 1, 0};
    k::b_copy(g.B, g.w, g.mv.get(), sk, dst, src, stream_);
  }
}

void Batch::norms(const std::vector<char>& on, int x, std::vector<double>& out) {
  out.assign(members_.size(), 0.0);
  // members choose their norm: l_2 (two_norm2), l_inf or l_1
  for (int op = 0; op < 3; ++op) {
    const int want = op == 0 ? 0 : op == 2 ? 1 : 2;  // op 2 = max |x| is l_inf (norm 1)
    std::vector<char> sel(members_.size(), 0);
    bool any = false;
    for (size_t i = 0; i < members_.size(); ++i)
      if (on[i] && members_[i].m->solver().norm == want) sel[i] = 1, any = true;
    if (!any) continue;
    std::vector<char> none(members_.size(), 0);
    push_flags(sel, none);
    std::vector<std::vector<double>> host(sgroups_.size());
    for (size_t gi = 0; gi < sgroups_.size(); ++gi) {
      auto& g = sgroups_[gi];
      const k::Skip sk{g.act.get(), 1, 0};
      k::b_norm(g.B, g.w, g.mv.get(), sk, x, op, g.part.get(), g.sums.get(), stream_);
      host[gi].resize(g.mem.size());
      DCG_HIP(hipMemcpyAsync(host[gi].data(), g.sums.get(), g.mem.size() * sizeof(double), hipMemcpyDeviceToHost,
                             stream_));
    }
    DCG_HIP(hipStreamSynchronize(stream_));
    for (size_t gi = 0; gi < sgroups_.size(); ++gi)
      for (size_t j = 0; j < sgroups_[gi].mem.size(); ++j)
        if (sel[sgroups_[gi].mem[j]]) out[sgroups_[gi].mem[j]] = host[gi][j];
  }
}

// BiCGSTAB iterations until every member of the group is done: blocks of
// iterations (captured graphs or plain launches), then one read-back of the
// per-member scalars.
void Batch::iterate(SolverGroup& g, bool fp32) {
  const int K = std::max(1, opt_.graph_iterations);
  const size_t M = g.mem.size();
  std::vector<k::BicgScalars> sc(M);
  int done_its = 0;
  const int slot = g.call++;
  const int est = slot < static_cast<int>(g.hints.size()) ? g.hints[slot] : g.est_its;
  int blocks = std::max(1, (est + K - 1) / K);
  int most = 0;
  auto remember = [&]() {
    g.est_its = std::max(2, most);
    if (slot >= static_cast<int>(g.hints.size())) g.hints.resize(slot + 1, 0);
    g.hints[slot] = g.est_its;
  };
  if (prof_) {
    // one iteration at a time, the SpMVs bracketed by events; the bytes are
    // those of the members still iterating
    for (;;) {
      DCG_HIP(hipMemcpyAsync(sc.data(), g.sc.get(), M * sizeof(k::BicgScalars), hipMemcpyDeviceToHost, stream_));
      DCG_HIP(hipStreamSynchronize(stream_));
      double bytes = 0.0;
      bool all = true;
      for (size_t j = 0; j < M; ++j) {
        if (sc[j].done) continue;
        all = false;
        const Model& mj = *members_[g.mem[j]].m;
        bytes += mj.point_storage() ? mj.point_apply_bytes(fp32 ? 4 : 8, 1) : mj.spmv_bytes(fp32 ? 4 : 8, fp32 ? 4 : 8, 1);
      }
      if (all || done_its > opt_.lin_maxit + 4) break;
      if (fp32)
        k::b_bicg_iteration<float>(g.B, g.variant32, g.w, g.mv.get(), k::kVKry + 9, jacobi_kind(), g.part.get(),
                                   g.sc.get(), stream_, ev_, g.points ? &g.op32 : nullptr);
      else
        k::b_bicg_iteration<double>(g.B, g.variant64, g.w, g.mv.get(), k::kVZ, jacobi_kind(), g.part.get(),
                                    g.sc.get(), stream_, ev_, g.points ? &g.op64 : nullptr);
      DCG_HIP(hipEventSynchronize(ev_[3]));
      float a = 0.f, b = 0.f;
      DCG_HIP(hipEventElapsedTime(&a, ev_[0], ev_[1]));
      DCG_HIP(hipEventElapsedTime(&b, ev_[2], ev_[3]));
      Prof& p = prof_data_[fp32 ? 1 : 0];
      p.launches += 2;
      p.ms += a + b;
      p.bytes += 2.0 * bytes;
      ++done_its;
    }
    for (const auto& s2 : sc) most = std::max(most, s2.iter);
    remember();
    return;
  }
  for (;;) {
    for (int b = 0; b < blocks; ++b) {
      if (opt_.graph_iterations > 0) {
        if (!g.graph[fp32 ? 1 : 0]) capture(g, fp32);
        DCG_HIP(hipGraphLaunch(g.graph[fp32 ? 1 : 0], stream_));
      } else if (fp32) {
        k::b_bicg_iteration<float>(g.B, g.variant32, g.w, g.mv.get(), k::kVKry + 9, jacobi_kind(), g.part.get(),
                                   g.sc.get(), stream_, nullptr, g.points ? &g.op32 : nullptr);
      } else {
        k::b_bicg_iteration<double>(g.B, g.variant64, g.w, g.mv.get(), k::kVZ, jacobi_kind(), g.part.get(), g.sc.get(),
                                    stream_, nullptr, g.points ? &g.op64 : nullptr);
      }
    }
    done_its += blocks * K;
    DCG_HIP(hipMemcpyAsync(sc.data(), g.sc.get(), M * sizeof(k::BicgScalars), hipMemcpyDeviceToHost, stream_));
    DCG_HIP(hipStreamSynchronize(stream_));
    bool all = true;
    most = 0;
    for (const auto& s : sc) {
      all = all && s.done != 0;
      most = std::max(most, s.iter);
    }
    if (all) break;
    if (done_its > opt_.lin_maxit + 4) break;  // unfinished members report "not converged"
    blocks = 1;
  }
  remember();
}

void Batch::solve(const std::vector<char>& on, const std::vector<double>& rtol, std::vector<int>& status,
                  std::vector<int>& its, std::vector<int>& refs) {
  status.assign(members_.size(), 0);
  its.assign(members_.size(), 0);
  refs.assign(members_.size(), 0);
  std::vector<char> none(members_.size(), 0);
  push_flags(on, none);
  const bool mixed = opt_.precision == Precision::Mixed;
  for (auto& g : sgroups_) {
    const size_t M = g.mem.size();
    bool any = false;
    for (size_t j = 0; j < M; ++j) {
      g.h_rtol[j] = mixed ? std::max(rtol[g.mem[j]], opt_.ir_floor) : rtol[g.mem[j]];
      any = any || on[g.mem[j]];
    }
    if (!any) continue;
    if (g.h_rtol != g.dev_rtol) {
      DCG_HIP(hipMemcpyAsync(g.rtol.get(), g.h_rtol.data(), M * sizeof(double), hipMemcpyHostToDevice, stream_));
      g.dev_rtol = g.h_rtol;
    }
    const k::Skip off{g.act.get(), 1, 0};
    if (opt_.precond != Preconditioner::None)
      k::b_block_inverse(g.B, g.w, g.mv.get(), off, status_.get(), opt_.precond == Preconditioner::Jacobi, stream_);
    if (!mixed) {
      k::b_bicg_setup(g.w, off, g.rtol.get(), nullptr, 0.0, opt_.lin_maxit, g.sc.get(), stream_);
      k::b_bicg_init<double>(g.B, g.w, g.mv.get(), k::kVR, k::kVZ, g.part.get(), g.sc.get(), stream_);
      iterate(g, false);
      std::vector<k::BicgScalars> sc(M);
      DCG_HIP(hipMemcpyAsync(sc.data(), g.sc.get(), M * sizeof(k::BicgScalars), hipMemcpyDeviceToHost, stream_));
      DCG_HIP(hipStreamSynchronize(stream_));
      for (size_t j = 0; j < M; ++j) {
        const int i = g.mem[j];
        if (!on[i]) continue;
        its[i] = sc[j].iter;
        status[i] = sc[j].done == 1 ? 0 : sc[j].done == 0 ? 2 : sc[j].done;
      }
      continue;
    }
    // FP32 BiCGSTAB inside FP64 iterative refinement, per member
    k::b_to_float(g.B, g.w, g.mv.get(), off, stream_);
    k::b_copy(g.B, g.w, g.mv.get(), off, k::kVZ, -1, stream_);
    k::b_copy(g.B, g.w, g.mv.get(), off, k::kVKry + 8, k::kVR, stream_);
    k::b_norm(g.B, g.w, g.mv.get(), off, k::kVR, 0, g.part.get(), g.sums.get(), stream_);
    k::b_ir_begin(g.w, off, g.rtol.get(), g.sums.get(), g.ir.get(), stream_);
    std::vector<k::IrState> ir(M);
    const k::Skip ir_done{&g.ir.get()->done, static_cast<int>(sizeof(k::IrState) / sizeof(int)), 1};
    bool finished = false;
    for (int r = 0; r < opt_.ir_max; ++r) {
      k::b_bicg_setup(g.w, off, nullptr, g.ir.get(), opt_.ir_inner_rtol, opt_.lin_maxit, g.sc.get(), stream_);
      k::b_bicg_init<float>(g.B, g.w, g.mv.get(), k::kVKry + 8, k::kVKry + 9, g.part.get(), g.sc.get(), stream_);
      iterate(g, true);
      k::b_ir_update(g.B, g.w, g.mv.get(), g.ir.get(), k::kVZ, stream_);
      if (g.points) g.op64.apply(k::kVZ, k::kVKry + 7, -1, -1, nullptr, 1);
      else k::b_spmv64(g.B, g.variant64, g.w, g.mv.get(), ir_done, k::kVZ, k::kVKry + 7, stream_);
      k::b_ir_residual(g.B, g.w, g.mv.get(), g.ir.get(), k::kVR, k::kVKry + 7, k::kVKry + 8, g.part.get(),
                       g.sums.get(), stream_);
      k::b_ir_check(g.w, g.sums.get(), g.sc.get(), g.ir.get(), stream_);
      DCG_HIP(hipMemcpyAsync(ir.data(), g.ir.get(), M * sizeof(k::IrState), hipMemcpyDeviceToHost, stream_));
      DCG_HIP(hipStreamSynchronize(stream_));
      finished = true;
      for (const auto& q : ir) finished = finished && q.done;
      if (finished) break;
    }
    for (size_t j = 0; j < M; ++j) {
      const int i = g.mem[j];
      if (!on[i]) continue;
      its[i] = ir[j].its;
      refs[i] = ir[j].refs;
      status[i] = ir[j].status ? ir[j].status : (ir[j].done ? 0 : 2);
    }
  }
}

// ------------------------------------------------------------------ time step
void Batch::step(std::vector<StepStats>* stats) {
  DCG_HIP(hipSetDevice(device_));
  if (dirty_) rebuild();
  const size_t N = members_.size();
  std::vector<StepStats> S(N);
  if (N == 0) {
    if (stats) stats->clear();
    return;
  }
  for (auto& g : sgroups_) g.call = 0;
  // per-member coefficients of every assembly call of this step
  coef_.assign(N, {});
  int max_stages = 0;
  for (size_t i = 0; i < N; ++i) {
    const Model& m = *members_[i].m;
    const RkTable& rk = m.rk();
    const double t = m.time(), dt = members_[i].dt;
    max_stages = std::max(max_stages, rk.stages);
    for (int s = 1; s <= rk.stages; ++s) {
      for (int j = 0; j < s; ++j) {
        const double am = rk.a[s - 1][j], ab = dt * rk.b[s - 1][j];
        const bool on = am != 0.0 || ab != 0.0;
        coef_[i][hist_call(s, j)] = {t + rk.d[j] * dt, am, ab, on ? 1.0 : 0.0};
      }
      coef_[i][stage_call(s)] = {t + rk.d[s] * dt, rk.a[s - 1][s], dt * rk.b[s - 1][s], 1.0};
    }
  }
  for (auto& a : agroups_) {
    const size_t M = a.mem.size();
    for (int c = 0; c < kCalls; ++c)
      for (size_t j = 0; j < M; ++j)
        for (int q = 0; q < 4; ++q) a.h_coef[(c * M + j) * 4 + q] = coef_[a.mem[j]][c][q];
    DCG_HIP(hipMemcpyAsync(a.coef.get(), a.h_coef.data(), a.h_coef.size() * sizeof(double), hipMemcpyHostToDevice,
                           stream_));
  }
  std::vector<char> alive(N, 1), none(N, 0);
  copy(alive, k::kVStage + 0, k::kVU);
  for (int s = 1; s <= max_stages; ++s) {
    std::vector<char> part(N, 0);
    for (size_t i = 0; i < N; ++i) part[i] = alive[i] && members_[i].m->rk().stages >= s;
    // history of earlier stages
    copy(part, k::kVHist, -1);
    push_flags(part, none);
    for (int j = 0; j < s; ++j) assemble(hist_call(s, j), part, k::kVStage + j, k::kVHist, k::kVHist, 1);
    copy(part, k::kVXs, k::kVStage + s - 1);
    push_flags(part, none);
    assemble(stage_call(s), part, k::kVXs, k::kVHist, k::kVR, 1);
    std::vector<double> nrm;
    norms(part, k::kVR, nrm);
    // Newton per member with the control of the reference's configuration
    struct NS {
      double nrm0 = 0, nrm = 0, prev = 0, stop = 0;
      int it = 0;
      bool running = false;
    };
    std::vector<NS> ns(N);
    for (size_t i = 0; i < N; ++i) {
      if (!part[i]) continue;
      const SolverConfig& c = members_[i].m->solver();
      ns[i].nrm0 = ns[i].nrm = ns[i].prev = nrm[i];
      ns[i].stop = std::max(c.nl_atol, c.nl_rtol * nrm[i]);
      ns[i].running = true;
      S[i].stages = s;
    }
    for (;;) {
      std::vector<char> act(N, 0), jac(N, 0);
      std::vector<double> rtol(N, 0.0);
      bool any = false;
      for (size_t i = 0; i < N; ++i) {
        if (!ns[i].running) continue;
        const Model& m = *members_[i].m;
        const SolverConfig& c = m.solver();
        NS& q = ns[i];
        S[i].residual_norm = q.nrm;
        if (m.spec().is_linear) {
          // one defect correction (make_step_operator.hh:210-245)
          if (q.it >= 1) {
            q.running = false;
            continue;
          }
          act[i] = jac[i] = 1;
          rtol[i] = c.lin_rtol;
          any = true;
          continue;
        }
        if (!std::isfinite(q.nrm)) {
          S[i].status = 3;
          q.running = false;
          alive[i] = 0;
          continue;
        }
        const bool conv = q.nrm <= q.stop || q.nrm == 0.0;
        if (conv && q.it >= c.nl_min) {
          q.running = false;
          continue;
        }
        if (q.it >= c.nl_max) {
          S[i].status = 1;
          q.running = false;
          alive[i] = 0;
          continue;
        }
        act[i] = 1;
        jac[i] = (q.it == 0 || c.linearization_threshold <= 0.0 || q.nrm / q.prev > c.linearization_threshold) ? 1 : 0;
        double rt = c.lin_rtol;
        if (!c.fixed_linear_tolerance) {
          const double ratio = q.it == 0 ? 1.0 : (q.nrm * q.nrm) / (q.prev * q.prev);
          rt = std::max(q.stop / (10.0 * q.nrm), std::min(c.min_linear_reduction, ratio));
        }
        rtol[i] = rt;
        any = true;
      }
      if (!any) break;
      push_flags(act, jac);
      assemble(stage_call(s), jac, k::kVXs, -1, -1, 2);
      std::vector<int> st, its, refs;
      solve(act, rtol, st, its, refs);
      std::vector<char> upd(N, 0);
      for (size_t i = 0; i < N; ++i) {
        if (!act[i]) continue;
        S[i].linear_iterations += its[i];
        S[i].refinement_steps += refs[i];
        if (st[i]) {
          S[i].status = st[i] == 3 ? 3 : st[i] == 4 ? 4 : 2;
          ns[i].running = false;
          alive[i] = 0;
          continue;
        }
        upd[i] = 1;
      }
      push_flags(upd, none);
      for (auto& g : sgroups_) {
        const k::Skip sk{g.act.get(), 1, 0};
        k::b_axpy(g.B, g.w, g.mv.get(), sk, -1.0, k::kVZ, k::kVXs, stream_);
      }
      assemble(stage_call(s), upd, k::kVXs, k::kVHist, k::kVR, 1);
      std::vector<double> nn;
      norms(upd, k::kVR, nn);
      for (size_t i = 0; i < N; ++i) {
        if (!upd[i]) continue;
        ns[i].it += 1;
        S[i].newton_iterations += 1;
        ns[i].prev = ns[i].nrm;
        ns[i].nrm = nn[i];
      }
    }
    std::vector<char> fin(N, 0), last(N, 0);
    for (size_t i = 0; i < N; ++i) {
      fin[i] = part[i] && alive[i];
      last[i] = fin[i] && members_[i].m->rk().stages == s;
    }
    copy(fin, k::kVStage + s, k::kVXs);
    copy(last, k::kVU, k::kVXs);
  }
  DCG_HIP(hipStreamSynchronize(stream_));
  for (size_t i = 0; i < N; ++i)
    if (alive[i]) members_[i].m->set_time(members_[i].m->time() + members_[i].dt);
  if (stats) *stats = S;
}

}  // namespace dcopasi
