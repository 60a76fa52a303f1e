#include "serve.hh"

#include <algorithm>

namespace dcopasi {

ServerOptions ServerOptions::from_ini(const std::string& text) {
  ServerOptions o;
  o.batch = BatchOptions::from_ini(text);
  auto c = ParameterTree::from_ini(text, "");
  o.hbm_budget = static_cast<size_t>(c.get("hbm_budget_bytes", 0.0));
  o.max_members = c.get("max_members", o.max_members);
  o.workers = c.get("workers", o.workers);
  if (o.max_members < 1) throw IOError("max_members must be positive");
  if (o.workers < 1 || o.workers > 64) throw IOError("workers must be in [1, 64]");
  return o;
}

Server::Server(int device, const ServerOptions& opt) : device_(device), opt_(opt) {
  t0_ = std::chrono::steady_clock::now();
  batch_ = std::make_unique<Batch>(device, opt.batch);
  if (!opt_.hbm_budget) {
    size_t f = 0, t = 0;
    DCG_HIP(hipMemGetInfo(&f, &t));
    opt_.hbm_budget = static_cast<size_t>(0.9 * static_cast<double>(f));
  }
  for (int i = 0; i < opt_.workers; ++i) threads_.emplace_back([this] { worker(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto& t : threads_) t.join();
  batch_.reset();
}

double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

int Server::submit(const std::string& ini, const std::string& overrides, int steps) {
  if (steps < 1) throw Error("a request needs at least one step");
  auto r = std::make_unique<Request>();
  std::lock_guard<std::mutex> lock(mu_);
  r->id = next_++;
  r->ini = ini;
  r->overrides = overrides;
  r->steps = steps;
  r->info.steps = steps;
  r->info.t_arrival = now();
  const int id = r->id;
  req_[id] = std::move(r);
  todo_.push_back(id);
  cv_.notify_one();
  return id;
}

// Host side of admission, off the GPU's critical path: parse, grid and
// pattern, code generation, hipRTC compilation into the cache, footprint.
void Server::worker() {
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lock(mu_);
      cv_.wait(lock, [this] { return stop_ || !todo_.empty(); });
      if (stop_) return;
      r = req_[todo_.front()].get();
      todo_.pop_front();
      ++building_;
    }
    std::unique_ptr<Model> m;
    size_t est = 0;
    std::string err;
    try {
      m = std::make_unique<Model>(r->ini, batch_->member_overrides(r->overrides), -1, batch_->member_options());
      if (m->rk().stages > 2) throw NotImplemented("served models use schemes of at most two stages");
      m->precompile();
      const auto& c = m->solver();
      est = Model::estimate_device_bytes(m->layout(), m->rk().stages, c.precision == Precision::Mixed, false);
    } catch (const std::exception& e) {
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lock(mu_);
      --building_;
      r->info.t_ready = now();
      if (!err.empty()) {
        r->info.status = -1;
        r->info.error = err;
        r->info.finished = true;
        r->info.t_finished = r->info.t_ready;
      } else {
        r->model = std::move(m);
        r->estimate = est;
        r->info.n_dof = r->model->n_dof();
        ready_.push_back(r->id);
      }
    }
    ready_cv_.notify_all();
  }
}

// Only the thread that calls tick() touches the batch; everything other
// threads read (the request table, infos, the member map, byte and member
// counts) changes under mu_, so submit, result, forget and stats may run
// concurrently with a tick.  Device work (adding, stepping, reading back,
// removing members) happens outside the lock.
int Server::tick() {
  if (batch_->size() == 0) {
    std::unique_lock<std::mutex> lock(mu_);
    ready_cv_.wait_for(lock, std::chrono::milliseconds(100),
                       [this] { return !ready_.empty() || (todo_.empty() && building_ == 0); });
  }
  // admission in arrival order while the footprint fits and a slot is free
  for (;;) {
    Request* r = nullptr;
    std::unique_ptr<Model> m;
    {
      std::lock_guard<std::mutex> lock(mu_);
      if (ready_.empty()) break;
      Request* front = req_.at(ready_.front()).get();
      if (resident_ >= opt_.max_members) break;
      if (used_ + front->estimate > opt_.hbm_budget) {
        if (resident_ > 0) break;  // wait for memory to free up
        if (front->estimate > opt_.hbm_budget) {
          front->info.status = -1;
          front->info.error = "request does not fit the HBM budget";
          front->info.finished = true;
          front->info.t_finished = now();
          ready_.pop_front();
          continue;
        }
      }
      r = front;
      m = std::move(r->model);
      ready_.pop_front();
    }
    int member = -1;
    try {
      Model* mp = m.get();
      member = batch_->add(std::move(m));
      std::vector<double> u0(mp->n_dof());
      mp->initial_state(u0.data());
      mp->set_state(u0.data(), mp->time());
      const size_t bytes = mp->device_bytes();
      std::lock_guard<std::mutex> lock(mu_);
      r->member = member;
      r->info.device_bytes = bytes;
      used_ += bytes;
      ++resident_;
      by_member_[member] = r->id;
      r->info.t_admitted = now();
    } catch (const std::exception& e) {
      if (member >= 0) batch_->remove(member);
      std::lock_guard<std::mutex> lock(mu_);
      r->member = -1;
      r->info.status = -1;
      r->info.error = e.what();
      r->info.finished = true;
      r->info.t_finished = now();
    }
  }
  if (batch_->size() > 0) {
    std::vector<StepStats> st;
    batch_->step(&st);
    const auto ids = batch_->ids();
    std::vector<int> retire;
    {
      std::lock_guard<std::mutex> lock(mu_);
      ++ticks_;
      for (size_t k = 0; k < ids.size(); ++k) {
        Request* r = req_.at(by_member_.at(ids[k])).get();
        if (st[k].status) {
          r->info.status = st[k].status;
          retire.push_back(ids[k]);
          continue;
        }
        r->info.steps_done += 1;
        r->info.newton_iterations += st[k].newton_iterations;
        r->info.linear_iterations += st[k].linear_iterations;
        if (r->info.steps_done >= r->steps) retire.push_back(ids[k]);
      }
    }
    for (int mid : retire) {
      Model& m = batch_->member(mid);
      std::vector<double> state(m.n_dof());
      double t = 0;
      m.get_state(state.data(), &t);
      batch_->remove(mid);
      std::lock_guard<std::mutex> lock(mu_);
      Request* r = req_.at(by_member_.at(mid)).get();
      r->state = std::move(state);
      used_ -= std::min(used_, r->info.device_bytes);
      --resident_;
      by_member_.erase(mid);
      r->member = -1;
      r->info.finished = true;
      r->info.t_finished = now();
    }
  }
  std::lock_guard<std::mutex> lock(mu_);
  int open = 0;
  for (const auto& [id, r] : req_)
    if (!r->info.finished) ++open;
  return open;
}

bool Server::result(int id, std::vector<double>* state, RequestInfo* info) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = req_.find(id);
  if (it == req_.end()) throw Error("no request with id " + std::to_string(id));
  if (info) *info = it->second->info;
  if (!it->second->info.finished) return false;
  if (state) *state = it->second->state;
  return true;
}

void Server::forget(int id) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = req_.find(id);
  if (it == req_.end()) throw Error("no request with id " + std::to_string(id));
  if (!it->second->info.finished) throw Error("request " + std::to_string(id) + " is not finished");
  req_.erase(it);
}

Server::Stats Server::stats() {
  std::lock_guard<std::mutex> lock(mu_);
  Stats s;
  s.queued = static_cast<int>(todo_.size());
  s.ready = static_cast<int>(ready_.size());
  s.resident = resident_;
  for (const auto& [id, r] : req_)
    if (r->info.finished) ++s.finished;
  s.used_bytes = used_;
  s.budget_bytes = opt_.hbm_budget;
  s.ticks = ticks_;
  return s;
}

}  // namespace dcopasi
