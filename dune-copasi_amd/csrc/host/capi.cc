// extern "C" boundary (include/dcopasi.h) over dcopasi::Model.
#include <cstring>
#include <memory>
#include <string>

#include "../../../include/dcopasi.h"
#include "batch.hh"
#include "model.hh"
#include "serve.hh"
#include "stepper.hh"

// owning handle, or a view of a batch member (own empty)
struct dcg_model {
  std::unique_ptr<dcopasi::Model> own;
  dcopasi::Model* m = nullptr;
  std::vector<dcopasi::ReduceResult> reduced;  // last dcg_reduce
};

struct dcg_batch {
  std::unique_ptr<dcopasi::Batch> b;
};

struct dcg_server {
  std::unique_ptr<dcopasi::Server> s;
};

namespace {
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  } catch (...) {
    g_error = "unknown error";
    return -1;
  }
}

int check(dcg_model_t h) {
  if (!h || !h->m) {
    g_error = "null model handle";
    return -1;
  }
  return 0;
}

void copy_name(const std::string& s, char* buf, size_t len) {
  if (!buf || !len) return;
  std::strncpy(buf, s.c_str(), len - 1);
  buf[len - 1] = '\0';
}
}  // namespace

extern "C" {

const char* dcg_last_error(void) { return g_error.c_str(); }

int dcg_model_create(const char* ini_text, const char* overrides, int device, dcg_model_t* out) {
  return guarded([&] {
    if (!out) throw dcopasi::Error("null output handle");
    *out = nullptr;
    auto h = std::make_unique<dcg_model>();
    h->own = std::make_unique<dcopasi::Model>(ini_text ? ini_text : "", overrides ? overrides : "", device);
    h->m = h->own.get();
    *out = h.release();
    return 0;
  });
}

int dcg_model_create_ex(const char* ini_text, const char* overrides, int device, unsigned flags, dcg_model_t* out) {
  return guarded([&] {
    if (!out) throw dcopasi::Error("null output handle");
    *out = nullptr;
    if (flags & ~DCG_RUNTIME_PARAMETERS) throw dcopasi::Error("unknown creation flags");
    dcopasi::ModelOptions opt;
    opt.runtime_parameters = (flags & DCG_RUNTIME_PARAMETERS) != 0;
    auto h = std::make_unique<dcg_model>();
    h->own = std::make_unique<dcopasi::Model>(ini_text ? ini_text : "", overrides ? overrides : "", device, opt);
    h->m = h->own.get();
    *out = h.release();
    return 0;
  });
}

void dcg_model_destroy(dcg_model_t model) { delete model; }

int dcg_state_size(dcg_model_t h, size_t* n_dof, size_t* n_species, size_t* n_compartments) {
  if (check(h)) return -1;
  if (n_dof) *n_dof = h->m->n_dof();
  if (n_species) *n_species = h->m->n_species();
  if (n_compartments) *n_compartments = h->m->n_compartments();
  return 0;
}

int dcg_set_state(dcg_model_t h, const double* state, size_t n, double time) {
  if (check(h)) return -1;
  return guarded([&] {
    if (n != h->m->n_dof()) throw dcopasi::Error("state size mismatch");
    h->m->set_state(state, time);
    return 0;
  });
}

int dcg_get_state(dcg_model_t h, double* state, size_t n, double* time) {
  if (check(h)) return -1;
  return guarded([&] {
    if (n != h->m->n_dof()) throw dcopasi::Error("state size mismatch");
    h->m->get_state(state, time);
    return 0;
  });
}

int dcg_initial_state(dcg_model_t h, double* state, size_t n) {
  if (check(h)) return -1;
  return guarded([&] {
    if (n != h->m->n_dof()) throw dcopasi::Error("state size mismatch");
    h->m->initial_state(state);
    return 0;
  });
}

int dcg_step(dcg_model_t h, double t, double dt, dcg_step_stats* stats) {
  if (check(h)) return -1;
  return guarded([&] {
    dcopasi::StepStats st;
    int rc = h->m->step(t, dt, &st);
    if (stats) {
      stats->status = st.status;
      stats->stages = st.stages;
      stats->newton_iterations = st.newton_iterations;
      stats->linear_iterations = st.linear_iterations;
      stats->refinement_steps = st.refinement_steps;
      stats->residual_norm = st.residual_norm;
    }
    return rc;
  });
}

int dcg_mass(dcg_model_t h, double* per_species) {
  if (check(h)) return -1;
  return guarded([&] {
    h->m->mass(per_species);
    return 0;
  });
}

int dcg_reduce(dcg_model_t h, const char* config_ini, size_t* count) {
  if (check(h)) return -1;
  return guarded([&] {
    h->reduced = h->m->reduce(config_ini ? config_ini : "");
    if (count) *count = h->reduced.size();
    for (const auto& r : h->reduced)
      if (r.flag == 2) return 1;
    return 0;
  });
}

int dcg_write_vtk(dcg_model_t h, const char* path, int append) {
  if (check(h)) return -1;
  return guarded([&] {
    if (!path || !*path) throw dcopasi::IOError("empty output path");
    h->m->write_vtk(path, append != 0);
    return 0;
  });
}

int dcg_reduce_source(dcg_model_t h, const char* config_ini, char* buf, size_t len, size_t* needed) {
  if (check(h)) return -1;
  return guarded([&] {
    const std::string s = h->m->reduce_source(config_ini ? config_ini : "");
    if (needed) *needed = s.size() + 1;
    if (buf && len) copy_name(s, buf, len);
    return 0;
  });
}

int dcg_reduce_result(dcg_model_t h, size_t i, char* name, size_t name_len, double* value, int* flag) {
  if (check(h)) return -1;
  return guarded([&] {
    if (i >= h->reduced.size()) throw dcopasi::Error("no reduce result " + std::to_string(i));
    const auto& r = h->reduced[i];
    if (name) copy_name(r.name, name, name_len);
    if (value) *value = r.value;
    if (flag) *flag = r.flag;
    return 0;
  });
}

int dcg_species_info(dcg_model_t h, size_t s, char* name, size_t name_len, int* compartment, int* local_index) {
  if (check(h)) return -1;
  return guarded([&] {
    const auto& sp = h->m->spec().species;
    if (s >= sp.size()) throw dcopasi::Error("species index out of range");
    copy_name(sp[s].name, name, name_len);
    if (compartment) *compartment = sp[s].comp;
    if (local_index) *local_index = sp[s].local;
    return 0;
  });
}

int dcg_compartment_info(dcg_model_t h, size_t c, char* name, size_t name_len, size_t* n_nodes, size_t* dof_offset,
                         int* n_species) {
  if (check(h)) return -1;
  return guarded([&] {
    const auto& comps = h->m->layout().comps;
    if (c >= comps.size()) throw dcopasi::Error("compartment index out of range");
    copy_name(comps[c].name, name, name_len);
    if (n_nodes) *n_nodes = static_cast<size_t>(comps[c].nslot);
    if (dof_offset) *dof_offset = static_cast<size_t>(comps[c].dof_off);
    if (n_species) *n_species = comps[c].b;
    return 0;
  });
}

int dcg_compartment_nodes(dcg_model_t h, size_t c, long* vertex_ids) {
  if (check(h)) return -1;
  return guarded([&] {
    const auto& comps = h->m->layout().comps;
    if (c >= comps.size()) throw dcopasi::Error("compartment index out of range");
    std::copy(comps[c].nodes.begin(), comps[c].nodes.end(), vertex_ids);
    return 0;
  });
}

int dcg_pattern_info(dcg_model_t h, size_t* n_slots, size_t* n_blocks, int* block_size) {
  if (check(h)) return -1;
  const auto& L = h->m->layout();
  if (n_slots) *n_slots = static_cast<size_t>(L.nslot);
  if (n_blocks) *n_blocks = static_cast<size_t>(L.nnzb());
  if (block_size) *block_size = L.B;
  return 0;
}

int dcg_pattern(dcg_model_t h, int* row_ptr, int* col) {
  if (check(h)) return -1;
  const auto& L = h->m->layout();
  if (row_ptr) std::copy(L.row_ptr.begin(), L.row_ptr.end(), row_ptr);
  if (col) std::copy(L.col.begin(), L.col.end(), col);
  return 0;
}

int dcg_generated_source(dcg_model_t h, char* buf, size_t len, size_t* needed) {
  if (check(h)) return -1;
  const std::string& s = h->m->source();
  if (needed) *needed = s.size() + 1;
  if (buf && len) copy_name(s, buf, len);
  return 0;
}

int dcg_assemble(dcg_model_t h, const double* x, double t, double mass_weight, double stiffness_weight,
                 double* residual, double* jacobian) {
  if (check(h)) return -1;
  return guarded([&] {
    h->m->assemble(x, t, mass_weight, stiffness_weight, residual, jacobian);
    return 0;
  });
}

int dcg_jacobian_apply(dcg_model_t h, const double* x, double t, double mass_weight, double stiffness_weight,
                       const double* v, double* y) {
  if (check(h)) return -1;
  return guarded([&] {
    if (!x || !v || !y) throw dcopasi::Error("null vector");
    h->m->jacobian_apply(x, t, mass_weight, stiffness_weight, v, y);
    return 0;
  });
}

int dcg_spmv(dcg_model_t h, const double* jacobian, const double* x, double* y) {
  if (check(h)) return -1;
  return guarded([&] {
    h->m->spmv(jacobian, x, y);
    return 0;
  });
}

int dcg_linear_solve(dcg_model_t h, const double* jacobian, const double* b, double* x, double rtol, int maxit,
                     int* iterations) {
  if (check(h)) return -1;
  return guarded([&] {
    int its = 0;
    int rc = h->m->linear_solve(jacobian, b, x, rtol, maxit, &its);
    if (iterations) *iterations = its;
    return rc;
  });
}

static_assert(DCG_NOALLOC_WORKSPACE_BYTES == dcopasi::k::kNoallocScratchBytes, "work space size of the header");

int dcg_linear_solve_noalloc(dcg_model_t h, const double* jacobian, const double* b, double* x, double rtol, int maxit,
                             int* iterations) {
  if (check(h)) return -1;
  return guarded([&] {
    int its = 0;
    const int rc = h->m->linear_solve_noalloc(jacobian, b, x, rtol, maxit, &its);
    if (iterations) *iterations = its;
    return rc;
  });
}

int dcg_profile_enable(dcg_model_t h, int on) {
  if (check(h)) return -1;
  return guarded([&] {
    h->m->set_profiling(on != 0);
    return 0;
  });
}

int dcg_profile_read(dcg_model_t h, int kind, long* launches, double* milliseconds, double* bytes) {
  if (check(h)) return -1;
  return guarded([&] {
    long l = 0;
    double ms = 0, b = 0;
    h->m->profile(kind, &l, &ms, &b);
    if (launches) *launches = l;
    if (milliseconds) *milliseconds = ms;
    if (bytes) *bytes = b;
    return 0;
  });
}

int dcg_probe_copy(dcg_model_t h, size_t bytes, double* milliseconds) {
  if (check(h)) return -1;
  return guarded([&] {
    double ms = h->m->probe_copy(static_cast<long>(bytes));
    if (milliseconds) *milliseconds = ms;
    return 0;
  });
}

int dcg_spmv_tuning(dcg_model_t h, int* count, int* variant, double* microseconds, int* chosen) {
  if (check(h)) return -1;
  const auto& t = h->m->spmv_tuning();
  const int n = static_cast<int>(std::min<size_t>(t.size(), 16));
  if (count) *count = n;
  for (int k = 0; k < n; ++k) {
    if (variant) variant[k] = t[k].first;
    if (microseconds) microseconds[k] = t[k].second;
  }
  if (chosen) *chosen = h->m->spmv_variant();
  return 0;
}

int dcg_point_tuning(dcg_model_t h, int* count, int* comp, int* tile, double* microseconds, int* chosen) {
  if (check(h)) return -1;
  const auto& t = h->m->point_tuning();
  const int n = static_cast<int>(std::min<size_t>(t.size(), 64));
  if (count) *count = n;
  for (int k = 0; k < n; ++k) {
    if (comp) comp[k] = t[k].comp;
    if (tile) {
      for (int d = 0; d < 3; ++d) tile[4 * k + d] = t[k].tile[d];
      tile[4 * k + 3] = t[k].waves;
    }
    if (microseconds) microseconds[k] = t[k].us;
    if (chosen) chosen[k] = t[k].chosen ? 1 : 0;
  }
  return 0;
}

int dcg_evolve_defaults(dcg_model_t h, dcg_evolve_options* o) {
  if (check(h)) return -1;
  return guarded([&] {
    if (!o) throw dcopasi::Error("null options");
    const auto e = dcopasi::EvolveOptions::from_model(*h->m);
    o->t_end = e.t_end;
    o->dt_initial = e.dt;
    o->decrease_factor = e.decrease_factor;
    o->increase_factor = e.increase_factor;
    o->min_step = e.min_step;
    o->max_step = e.max_step;
    o->snap_to_end_time = e.snap_to_end_time ? 1 : 0;
    return 0;
  });
}

int dcg_evolve(dcg_model_t h, const dcg_evolve_options* o, dcg_evolve_stats* stats,
               void (*on_step)(void* user, double time), void* user, double* accepted_dt, size_t capacity) {
  if (check(h)) return -1;
  return guarded([&] {
    dcopasi::EvolveOptions e = dcopasi::EvolveOptions::from_model(*h->m);
    if (o) {
      e.t_end = o->t_end;
      e.dt = o->dt_initial;
      e.decrease_factor = o->decrease_factor;
      e.increase_factor = o->increase_factor;
      e.min_step = o->min_step;
      e.max_step = o->max_step;
      e.snap_to_end_time = o->snap_to_end_time != 0;
    }
    std::function<void(double)> cb;
    if (on_step) cb = [on_step, user](double t) { on_step(user, t); };
    const auto st = dcopasi::evolve(*h->m, e, cb);
    if (stats) {
      stats->status = st.status;
      stats->steps = st.steps;
      stats->failed_steps = st.failed_steps;
      stats->time = st.time;
      stats->dt = st.dt;
      stats->newton_iterations = st.newton_iterations;
      stats->linear_iterations = st.linear_iterations;
    }
    if (accepted_dt)
      for (size_t k = 0; k < st.accepted_dt.size() && k < capacity; ++k) accepted_dt[k] = st.accepted_dt[k];
    return st.status;
  });
}

int dcg_param_info(dcg_model_t h, size_t i, char* name, size_t name_len, size_t* count) {
  if (check(h)) return -1;
  return guarded([&] {
    const auto& p = h->m->param_names();
    if (count) *count = p.size();
    if (name && name_len) {
      if (i >= p.size()) throw dcopasi::Error("parameter index out of range");
      copy_name(p[i], name, name_len);
    }
    return 0;
  });
}

int dcg_set_params(dcg_model_t h, const double* values, size_t n) {
  if (check(h)) return -1;
  return guarded([&] {
    h->m->set_params(values, n);
    return 0;
  });
}

int dcg_get_params(dcg_model_t h, double* values, size_t n) {
  if (check(h)) return -1;
  return guarded([&] {
    h->m->get_params(values, n);
    return 0;
  });
}

int dcg_device_bytes(dcg_model_t h, size_t* bytes) {
  if (check(h)) return -1;
  return guarded([&] {
    if (!bytes) throw dcopasi::Error("null output");
    *bytes = h->m->device_bytes();
    return 0;
  });
}

int dcg_estimate_device_bytes(const char* ini_text, const char* overrides, size_t* bytes) {
  return guarded([&] {
    if (!bytes) throw dcopasi::Error("null output");
    dcopasi::Model m(ini_text ? ini_text : "", overrides ? overrides : "", -1);
    const auto& c = m.solver();
    *bytes = dcopasi::Model::estimate_device_bytes(m.layout(), m.rk().stages, c.precision == dcopasi::Precision::Mixed,
                                                   c.precond == dcopasi::Preconditioner::ILU0, c.matrix_free);
    return 0;
  });
}

int dcg_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
  return guarded([&] {
    DCG_HIP(hipSetDevice(device));
    size_t f = 0, t = 0;
    DCG_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return 0;
  });
}

static int check_batch(dcg_batch_t b) {
  if (!b || !b->b) {
    g_error = "null batch handle";
    return -1;
  }
  return 0;
}

int dcg_batch_create(const char* options_ini, int device, dcg_batch_t* out) {
  return guarded([&] {
    if (!out) throw dcopasi::Error("null output handle");
    *out = nullptr;
    if (device < 0) throw dcopasi::Error("a batch needs a device");
    auto h = std::make_unique<dcg_batch>();
    h->b = std::make_unique<dcopasi::Batch>(device, dcopasi::BatchOptions::from_ini(options_ini ? options_ini : ""));
    *out = h.release();
    return 0;
  });
}

void dcg_batch_destroy(dcg_batch_t batch) { delete batch; }

int dcg_batch_add(dcg_batch_t b, const char* ini_text, const char* overrides, int* id) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    const int i = b->b->add(ini_text ? ini_text : "", overrides ? overrides : "");
    if (id) *id = i;
    return 0;
  });
}

int dcg_batch_remove(dcg_batch_t b, int id) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    b->b->remove(id);
    return 0;
  });
}

int dcg_batch_ids(dcg_batch_t b, int* ids, size_t capacity, size_t* count) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    const auto v = b->b->ids();
    if (count) *count = v.size();
    if (ids) {
      if (capacity < v.size()) throw dcopasi::Error("id buffer too small");
      for (size_t i = 0; i < v.size(); ++i) ids[i] = v[i];
    }
    return 0;
  });
}

int dcg_batch_member(dcg_batch_t b, int id, dcg_model_t* out) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    if (!out) throw dcopasi::Error("null output handle");
    auto h = std::make_unique<dcg_model>();
    h->m = &b->b->member(id);
    *out = h.release();
    return 0;
  });
}

int dcg_batch_set_time_step(dcg_batch_t b, int id, double dt) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    if (!(dt > 0.0)) throw dcopasi::Error("time step must be positive");
    b->b->set_time_step(id, dt);
    return 0;
  });
}

int dcg_batch_step(dcg_batch_t b, dcg_step_stats* stats, size_t capacity) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    if (stats && capacity < b->b->size()) throw dcopasi::Error("stats buffer too small");
    std::vector<dcopasi::StepStats> st;
    b->b->step(&st);
    int worst = 0;
    for (size_t i = 0; i < st.size(); ++i) {
      if (stats) {
        stats[i].status = st[i].status;
        stats[i].stages = st[i].stages;
        stats[i].newton_iterations = st[i].newton_iterations;
        stats[i].linear_iterations = st[i].linear_iterations;
        stats[i].refinement_steps = st[i].refinement_steps;
        stats[i].residual_norm = st[i].residual_norm;
      }
      worst = std::max(worst, st[i].status);
    }
    return worst;
  });
}

int dcg_batch_device_bytes(dcg_batch_t b, size_t* bytes) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    if (!bytes) throw dcopasi::Error("null output");
    *bytes = b->b->device_bytes();
    return 0;
  });
}

int dcg_batch_profile_enable(dcg_batch_t b, int on) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    b->b->set_profiling(on != 0);
    return 0;
  });
}

int dcg_batch_profile_read(dcg_batch_t b, int kind, long* launches, double* milliseconds, double* bytes) {
  if (check_batch(b)) return -1;
  return guarded([&] {
    long l = 0;
    double ms = 0, by = 0;
    b->b->profile(kind, &l, &ms, &by);
    if (launches) *launches = l;
    if (milliseconds) *milliseconds = ms;
    if (bytes) *bytes = by;
    return 0;
  });
}

static int check_server(dcg_server_t s) {
  if (!s || !s->s) {
    g_error = "null server handle";
    return -1;
  }
  return 0;
}

int dcg_server_create(const char* options_ini, int device, dcg_server_t* out) {
  return guarded([&] {
    if (!out) throw dcopasi::Error("null output handle");
    *out = nullptr;
    if (device < 0) throw dcopasi::Error("a server needs a device");
    auto h = std::make_unique<dcg_server>();
    h->s = std::make_unique<dcopasi::Server>(device, dcopasi::ServerOptions::from_ini(options_ini ? options_ini : ""));
    *out = h.release();
    return 0;
  });
}

void dcg_server_destroy(dcg_server_t server) { delete server; }

int dcg_server_submit(dcg_server_t h, const char* ini_text, const char* overrides, int steps, int* id) {
  if (check_server(h)) return -1;
  return guarded([&] {
    const int i = h->s->submit(ini_text ? ini_text : "", overrides ? overrides : "", steps);
    if (id) *id = i;
    return 0;
  });
}

int dcg_server_tick(dcg_server_t h, int* open) {
  if (check_server(h)) return -1;
  return guarded([&] {
    const int o = h->s->tick();
    if (open) *open = o;
    return 0;
  });
}

int dcg_server_result(dcg_server_t h, int id, double* state, size_t n, dcg_request_info* info) {
  if (check_server(h)) return -1;
  return guarded([&] {
    std::vector<double> u;
    dcopasi::RequestInfo ri;
    const bool done = h->s->result(id, state ? &u : nullptr, &ri);
    if (info) {
      info->status = ri.status;
      info->finished = ri.finished ? 1 : 0;
      info->steps_done = ri.steps_done;
      info->steps = ri.steps;
      info->newton_iterations = ri.newton_iterations;
      info->linear_iterations = ri.linear_iterations;
      info->n_dof = ri.n_dof;
      info->device_bytes = ri.device_bytes;
      info->t_arrival = ri.t_arrival;
      info->t_ready = ri.t_ready;
      info->t_admitted = ri.t_admitted;
      info->t_finished = ri.t_finished;
    }
    if (ri.status < 0) g_error = ri.error;
    if (!done) return 1;
    if (state) {
      if (n != u.size()) throw dcopasi::Error("state size mismatch");
      std::copy(u.begin(), u.end(), state);
    }
    return 0;
  });
}

int dcg_server_forget(dcg_server_t h, int id) {
  if (check_server(h)) return -1;
  return guarded([&] {
    h->s->forget(id);
    return 0;
  });
}

int dcg_server_stats(dcg_server_t h, int* queued, int* ready, int* resident, int* finished, size_t* used_bytes,
                     size_t* budget_bytes, double* seconds) {
  if (check_server(h)) return -1;
  return guarded([&] {
    const auto st = h->s->stats();
    if (queued) *queued = st.queued;
    if (ready) *ready = st.ready;
    if (resident) *resident = st.resident;
    if (finished) *finished = st.finished;
    if (used_bytes) *used_bytes = st.used_bytes;
    if (budget_bytes) *budget_bytes = st.budget_bytes;
    if (seconds) *seconds = h->s->now();
    return 0;
  });
}

}  // extern "C"
