// A request loop over one GPU: parameter-sweep and fitting solves arrive as
// INI models with a step count; they are built and their kernels compiled on
// host worker threads, admitted to a resident Batch while their estimated
// HBM footprint fits the budget and a member slot is free, advanced one time
// step per tick together with everything else on the device (continuous
// batching: arrivals join between ticks, finished requests leave and free
// their memory), and their final state is kept for the caller.  Every
// request's mesh, matrix and state stay in HBM from admission to completion.
// The reference solves one model per process (src/dune_copasi.cc:229-290);
// this has no counterpart there.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "batch.hh"

namespace dcopasi {

struct ServerOptions {
  BatchOptions batch;
  size_t hbm_budget = 0;  // bytes for resident requests; 0: 90% of the device's free memory at start
  int max_members = 64;
  int workers = 4;  // host threads that build models and compile their kernels
  // keys: the batch keys plus hbm_budget_bytes, max_members, workers
  static ServerOptions from_ini(const std::string& text);
};

struct RequestInfo {
  int status = 0;  // 0 ok, > 0 numerical failure of a step (StepStats codes), < 0 set-up error
  int steps_done = 0, steps = 0;
  long newton_iterations = 0, linear_iterations = 0;
  size_t n_dof = 0, device_bytes = 0;
  // seconds since the server started
  double t_arrival = 0, t_ready = 0, t_admitted = 0, t_finished = 0;
  bool finished = false;
  std::string error;
};

class Server {
 public:
  Server(int device, const ServerOptions& opt);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // queue a request: `steps` time steps of the model from its initial state
  int submit(const std::string& ini, const std::string& overrides, int steps);
  // one scheduling round: admit what fits, one step of every resident
  // request, retire finished ones; returns the requests not yet finished.
  // One thread drives tick(); submit, result, forget and stats may be called
  // from other threads meanwhile.
  // With nothing resident and nothing built yet it first waits (at most
  // 100 ms) for a worker to finish a request.
  int tick();
  // false while the request is unfinished; state may be null
  bool result(int id, std::vector<double>* state, RequestInfo* info);
  void forget(int id);  // drop a finished request's stored result
  double now() const;   // seconds since start

  struct Stats {
    int queued = 0, ready = 0, resident = 0, finished = 0;
    size_t used_bytes = 0, budget_bytes = 0;
    long ticks = 0;
  };
  Stats stats();

 private:
  struct Request {
    int id;
    std::string ini, overrides;
    int steps;
    std::unique_ptr<Model> model;  // host-built, kernels compiled
    size_t estimate = 0;
    RequestInfo info;
    std::vector<double> state;
    int member = -1;  // batch member id while resident
  };
  void worker();

  int device_;
  ServerOptions opt_;
  std::chrono::steady_clock::time_point t0_;
  std::unique_ptr<Batch> batch_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::condition_variable ready_cv_;  // a worker finished building a request
  bool stop_ = false;
  std::deque<int> todo_;   // submitted, waiting for a worker
  std::deque<int> ready_;  // built, waiting for admission (FIFO)
  std::map<int, std::unique_ptr<Request>> req_;
  std::map<int, int> by_member_;  // batch member id -> request id
  std::vector<std::thread> threads_;
  int next_ = 0;
  int building_ = 0;  // taken by a worker, not yet ready
  size_t used_ = 0;
  int resident_ = 0;  // requests in the batch (read by stats() on any thread)
  long ticks_ = 0;
};

}  // namespace dcopasi
