// Serving loop for Poisson solve requests: continuous batching over the batch
// engine. Requests queue on submit (any thread); a worker thread admits them in
// arrival order whenever the device memory they need fits the HBM budget, runs
// one lockstep iteration of every admitted solve at a time, and retires solves
// as they converge, so requests that arrive while others are on the GPU join
// the running batch between iterations.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"

namespace insp {

struct ServerConfig {
  size_t hbm_budget = size_t(64) << 30;  // device bytes the resident solves may use
  BatchConfig cfg2d, cfg3d;
  int max_admit_per_round = 64;
};

struct RequestResult {
  ProbResult result;
  double latency_s = 0;     // arrival -> result on host
  double queue_s = 0;       // arrival -> admission
  int dim = 2;
  bool done = false;
  bool rejected = false;    // larger than the whole budget
};

class Server {
 public:
  explicit Server(const ServerConfig& cfg);
  ~Server();
  int64_t submit(ProblemSpec&& spec);
  bool poll(int64_t id);
  // Blocks until the request is finished and hands its result over; a result
  // is collected once (the server keeps no copy afterwards).
  RequestResult wait(int64_t id);
  void wait_all();
  int64_t completed() const { return completed_.load(); }
  size_t peak_bytes() const { return peak_bytes_; }

 private:
  struct Pending {
    int64_t id;
    int dim;                 // kept apart from spec: the spec is moved into the engine
    ProblemSpec spec;
    std::chrono::steady_clock::time_point arrival;
    size_t bytes;
  };
  struct Active {
    int64_t id;
    int dim;
    std::chrono::steady_clock::time_point arrival, admitted;
    size_t bytes;            // charged against the budget at admission
  };
  void loop();

  ServerConfig cfg_;
  std::unique_ptr<BatchEngine> eng_[2];
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<Pending> queue_;
  std::map<int64_t, RequestResult> results_;
  std::map<int, Active> active_[2];  // by engine problem id
  std::atomic<int64_t> next_id_{0}, completed_{0};
  size_t in_use_ = 0, peak_bytes_ = 0;
  bool stop_ = false;
  std::string error_;
  std::thread worker_;
};

}  // namespace insp
