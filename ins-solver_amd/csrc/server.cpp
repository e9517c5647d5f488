#include "server.h"

#include <stdexcept>

#include "solver.h"

namespace insp {

Server::Server(const ServerConfig& cfg) : cfg_(cfg) {
  cfg_.cfg2d.dim = 2;
  cfg_.cfg3d.dim = 3;
  eng_[0] = std::make_unique<BatchEngine>(cfg_.cfg2d);
  eng_[1] = std::make_unique<BatchEngine>(cfg_.cfg3d);
  worker_ = std::thread([this] { loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (worker_.joinable()) worker_.join();
}

int64_t Server::submit(ProblemSpec&& spec) {
  if (!spec.mesh) throw InvalidArgument("request without a mesh");
  const int dim = spec.mesh->dim;
  if (dim != 2 && dim != 3) throw InvalidArgument("mesh dimension must be 2 or 3");
  Pending p;
  p.id = next_id_++;
  p.dim = dim;
  p.arrival = std::chrono::steady_clock::now();
  const BatchConfig& bc = dim == 2 ? cfg_.cfg2d : cfg_.cfg3d;
  p.bytes = BatchEngine::arena_bytes(dim, spec.mesh->n_cells, spec.order, !spec.factor.empty(), bc.fp32);
  if (bc.amg) p.bytes += BatchEngine::amg_bytes_estimate(dim, spec.mesh->n_nodes, spec.mesh->n_cells, bc.fp32);
  p.spec = std::move(spec);
  {
    std::lock_guard<std::mutex> g(mu_);
    if (!error_.empty()) throw std::runtime_error("server stopped: " + error_);
    results_[p.id] = RequestResult{};
    queue_.push_back(std::move(p));
  }
  cv_.notify_all();
  return p.id;
}

bool Server::poll(int64_t id) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = results_.find(id);
  if (it == results_.end()) throw InvalidArgument("unknown request id");
  return it->second.done;
}

RequestResult Server::wait(int64_t id) {
  std::unique_lock<std::mutex> g(mu_);
  done_cv_.wait(g, [&] {
    auto it = results_.find(id);
    return it == results_.end() || it->second.done || !error_.empty();
  });
  if (!error_.empty()) throw std::runtime_error("server stopped: " + error_);
  auto it = results_.find(id);
  if (it == results_.end()) throw InvalidArgument("unknown request id");
  RequestResult r = std::move(it->second);
  results_.erase(it);
  return r;
}

void Server::wait_all() {
  std::unique_lock<std::mutex> g(mu_);
  done_cv_.wait(g, [&] {
    if (!error_.empty()) return true;
    for (auto& kv : results_)
      if (!kv.second.done) return false;
    return true;
  });
  if (!error_.empty()) throw std::runtime_error("server stopped: " + error_);
}

void Server::loop() {
  try {
    for (;;) {
      std::vector<Pending> admit;
      bool rejected = false;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] {
          return stop_ || !queue_.empty() || eng_[0]->running() > 0 || eng_[1]->running() > 0;
        });
        if (stop_) return;
        // admission by HBM budget, in arrival order
        int room[2] = {eng_[0]->free_slots(), eng_[1]->free_slots()};
        while (!queue_.empty() && (int)admit.size() < cfg_.max_admit_per_round) {
          Pending& p = queue_.front();
          const int di = p.dim - 2;
          if (p.bytes > cfg_.hbm_budget) {
            RequestResult& r = results_[p.id];
            r.rejected = true;
            r.done = true;
            rejected = true;
            queue_.pop_front();
            continue;
          }
          if (in_use_ + p.bytes > cfg_.hbm_budget || room[di] == 0) break;
          --room[di];
          in_use_ += p.bytes;
          peak_bytes_ = std::max(peak_bytes_, in_use_);
          admit.push_back(std::move(p));
          queue_.pop_front();
        }
      }
      if (rejected) done_cv_.notify_all();
      for (int d = 0; d < 2; ++d) {
        std::vector<ProblemSpec> specs;
        std::vector<Pending*> which;
        for (auto& p : admit)
          if (p.dim == d + 2) {
            which.push_back(&p);
          }
        for (auto* p : which) specs.push_back(std::move(p->spec));
        if (specs.empty()) continue;
        const auto now = std::chrono::steady_clock::now();
        const std::vector<int> ids = eng_[d]->add_many(std::move(specs));
        for (size_t k = 0; k < ids.size(); ++k) {
          Active a{which[k]->id, d + 2, which[k]->arrival, now, which[k]->bytes};
          active_[d][ids[k]] = a;
        }
      }
      // both engines' iterations in flight together (one stream each), then collected
      for (int d = 0; d < 2; ++d) eng_[d]->iterate_launch(std::max(1, eng_[d]->config().chunk));
      for (int d = 0; d < 2; ++d) eng_[d]->iterate_collect();
      for (int d = 0; d < 2; ++d) {
        for (int pid : eng_[d]->take_finished()) {
          auto it = active_[d].find(pid);
          if (it == active_[d].end()) continue;
          const Active a = it->second;
          active_[d].erase(it);
          const auto now = std::chrono::steady_clock::now();
          RequestResult rr;
          rr.result = eng_[d]->result(pid);
          rr.latency_s = std::chrono::duration<double>(now - a.arrival).count();
          rr.queue_s = std::chrono::duration<double>(a.admitted - a.arrival).count();
          rr.dim = d + 2;
          rr.done = true;
          {
            std::lock_guard<std::mutex> g(mu_);
            eng_[d]->release(pid);
            in_use_ -= a.bytes;  // what admission charged
            results_[a.id] = std::move(rr);
            ++completed_;
          }
          done_cv_.notify_all();
        }
      }
    }
  } catch (const std::exception& e) {
    std::lock_guard<std::mutex> g(mu_);
    error_ = e.what();
    done_cv_.notify_all();
  }
}

}  // namespace insp
