// cc_serve.hip — a serving loop with continuous batching for PBILS solves.
//
// Requests (a host graph plus PBILS parameters) are queued; each lane (a stream and a host thread) keeps a set of
// active runs whose state stays on the device and advances all of them one iteration per round with one launch set.
// Between rounds a lane retires the runs that have stopped (results copied out, device memory released) and admits
// queued requests while their HBM footprint — the n^2/8 bitset plus the per-vertex state of the whole population —
// fits the budget left. Large requests go to lane 0 so that a long wide-path run never gates the small ones; the
// lanes' kernels share the device. Every served run follows its own trajectory, identical to a solo solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cc_hip.h"
#include "cc_internal.hpp"
#include "cc_runtime.hpp"

namespace cc::hip {
namespace {

struct Request {
  uint64_t id = 0;
  std::vector<uint64_t> bits;
  uint32_t n = 0;
  uint64_t edges = 0;
  cc_pbils_params params{};
  size_t footprint = 0;
  Clock::time_point arrival;
};

struct Result {
  int code = CC_HIP_OK;
  std::string error;
  std::vector<int32_t> labels;
  std::vector<int64_t> trace;
  cc_pbils_stats stats{};
  double latency = 0.0;  // arrival -> result ready
  double queued = 0.0;   // arrival -> admitted
};

// A request admitted to a lane: its uploaded graph and run state.
struct Active {
  Request request;
  Clock::time_point admitted;
  std::unique_ptr<DeviceGraph> graph;
  std::unique_ptr<RunState> state;
  bool done = false;  // result delivered; the run stays in the launch set (skipped by every kernel) until a rebuild
};

// Iterations a lane advances between host synchronisations while no request waits for admission.
constexpr int kIdleChunk = 4;
// Slabs of the server's device pool; a larger request gets a slab of its own size.
constexpr size_t kSlabBytes = size_t(1) << 30;

size_t GraphBytes(uint32_t n) {
  return static_cast<size_t>(n) * Words64(n) * 8 + kRowTailBytes + static_cast<size_t>(n) * 4;
}

}  // namespace

class Server {
 public:
  Server(const cc_serve_options& options) : options_(options) {
    CC_HIP_CHECK(hipGetDevice(&device_));
    if (options_.budget_bytes <= 0) options_.budget_bytes = static_cast<int64_t>(FreeDeviceBytes() * 0.85);
    if (options_.lanes < 1) options_.lanes = 2;
    if (options_.max_active < 1) options_.max_active = 4096;
    if (options_.large_bytes <= 0) options_.large_bytes = int64_t(1) << 30;
    budget_left_ = options_.budget_bytes;
    lanes_.resize(options_.lanes);
    for (int i = 0; i < options_.lanes; ++i) lanes_[i].thread = std::thread([this, i] { LaneLoop(i); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> lock(mutex_);
      stop_ = true;
    }
    wake_.notify_all();
    for (Lane& l : lanes_) l.thread.join();
  }

  uint64_t Submit(const uint64_t* bits, uint32_t n, uint64_t edges, const cc_pbils_params& params) {
    CheckParams(params);
    if (n == 0 || n > static_cast<uint32_t>(kMaxVerticesWide)) {
      Fail(CC_HIP_EINVAL, "n must be in [1, %d]", kMaxVerticesWide);
    }
    Request r;
    r.n = n;
    r.edges = edges;
    r.params = params;
    r.params.verbose = 0;
    r.params.profile = 0;
    r.bits.assign(bits, bits + static_cast<size_t>(n) * Words64(n));
    r.footprint = Footprint(n, params);
    if (static_cast<int64_t>(r.footprint) > options_.budget_bytes) {
      Fail(CC_HIP_ENOMEM, "request needs %s, the server budget is %s", Mebibytes(r.footprint).c_str(),
           Mebibytes(static_cast<size_t>(options_.budget_bytes)).c_str());
    }
    r.arrival = Clock::now();
    const int lane = (options_.lanes > 1 && static_cast<int64_t>(r.footprint) >= options_.large_bytes)
                         ? 0
                         : (options_.lanes > 1 ? 1 + static_cast<int>(next_small_++ % (options_.lanes - 1)) : 0);
    uint64_t id;
    {
      std::lock_guard<std::mutex> lock(mutex_);
      id = r.id = ++last_id_;
      lanes_[lane].pending.push_back(std::move(r));
      pending_ids_.insert({id, lane});
    }
    wake_.notify_all();
    return id;
  }

  // Blocks until request id is done (timeout_s <= 0: no limit). Returns false on timeout.
  bool Wait(uint64_t id, double timeout_s, Result* out) {
    std::unique_lock<std::mutex> lock(mutex_);
    auto ready = [&] { return results_.count(id) > 0; };
    if (timeout_s > 0) {
      if (!done_.wait_for(lock, std::chrono::duration<double>(timeout_s), ready)) return false;
    } else {
      done_.wait(lock, ready);
    }
    *out = std::move(results_[id]);
    results_.erase(id);
    return true;
  }

  static size_t Footprint(uint32_t n, const cc_pbils_params& params) {
    DeviceGraph shape;
    shape.n = n;
    shape.words = Words64(n);
    const RunSpec spec{&shape, params, PlanDescent(n, params.k, params, 0), params.population, true};
    return GraphBytes(n) + RunState::Footprint(spec);
  }

  void Counters(int64_t* admitted, int64_t* completed, int64_t* budget_left, int64_t* peak_active,
                int64_t* peak_bytes) {
    std::lock_guard<std::mutex> lock(mutex_);
    *admitted = admitted_;
    *completed = completed_;
    *budget_left = budget_left_;
    *peak_active = peak_active_;
    *peak_bytes = peak_bytes_;
  }

 private:
  struct Lane {
    std::thread thread;
    std::deque<Request> pending;
    int64_t active = 0;
  };

  void Complete(uint64_t id, Result&& r) {
    {
      std::lock_guard<std::mutex> lock(mutex_);
      results_[id] = std::move(r);
      ++completed_;
    }
    done_.notify_all();
  }

  // Takes the queued requests of a lane that fit the budget, in arrival order (no overtaking within a lane).
  std::vector<Request> Admit(int lane, int active) {
    std::vector<Request> out;
    std::unique_lock<std::mutex> lock(mutex_);
    auto& q = lanes_[lane].pending;
    while (!q.empty() && active + static_cast<int>(out.size()) < options_.max_active &&
           static_cast<int64_t>(q.front().footprint) <= budget_left_) {
      budget_left_ -= static_cast<int64_t>(q.front().footprint);
      peak_bytes_ = std::max(peak_bytes_, options_.budget_bytes - budget_left_);
      pending_ids_.erase(q.front().id);
      out.push_back(std::move(q.front()));
      q.pop_front();
    }
    admitted_ += static_cast<int64_t>(out.size());
    return out;
  }

  void Release(size_t bytes) {
    std::lock_guard<std::mutex> lock(mutex_);
    budget_left_ += static_cast<int64_t>(bytes);
  }

  void LaneLoop(int lane) {
    (void)hipSetDevice(device_);
    Stream stream;
    const hipStream_t s = stream.get();
    PinnedBuffer pinned;  // the launch sets' control blocks
    std::vector<Active> active;
    std::unique_ptr<LaunchSet> set;
    bool changed = false;
    size_t finished = 0;  // runs of `active` whose result is delivered
    while (true) {
      // admit
      std::vector<Request> admitted = Admit(lane, static_cast<int>(active.size()));
      std::vector<RunState*> fresh;
      for (Request& req : admitted) {
        Active a;
        a.admitted = Clock::now();
        try {
          a.graph = std::make_unique<DeviceGraph>();
          Upload(*a.graph, req.bits.data(), req.n, req.edges, s, &pool_);
          req.bits.clear();
          req.bits.shrink_to_fit();
          RunSpec spec{a.graph.get(), req.params, PlanDescent(req.n, req.params.k, req.params, req.edges),
                       req.params.population, true};
          UseMixedDescent(spec);  // the lane's register runs share the mixed launches
          a.state = std::make_unique<RunState>(spec, s, &pool_);
        } catch (const std::exception& e) {
          Result r;
          r.code = CC_HIP_EDEVICE;
          r.error = e.what();
          Release(req.footprint);
          Complete(req.id, std::move(r));
          continue;
        }
        a.request = std::move(req);
        fresh.push_back(a.state.get());
        active.push_back(std::move(a));
      }
      if (!fresh.empty()) {
        try {
          LaunchSet start(fresh, s, &pool_, &pinned);
          start.Start();
          // its tables go back to the pool when `start` ends: the launches that read them must be done
          CC_HIP_CHECK(hipStreamSynchronize(s));
        } catch (const std::exception& e) {
          (void)hipStreamSynchronize(s);
          const size_t first = active.size() - fresh.size();
          for (size_t i = first; i < active.size(); ++i) {
            Result r;
            r.code = CC_HIP_EDEVICE;
            r.error = e.what();
            Release(active[i].request.footprint);
            Complete(active[i].request.id, std::move(r));
          }
          active.resize(first);
        }
        changed = true;
      }
      // Finished runs stay in the launch set (every kernel skips a stopped run) until new runs rebuild it anyway,
      // they make up half of it, or queued requests wait for their memory; then it goes back to the budget.
      bool waiting = false;
      {
        std::lock_guard<std::mutex> lock(mutex_);
        waiting = !lanes_[lane].pending.empty();
      }
      if (changed || 2 * finished > active.size() || (finished > 0 && waiting)) {
        std::vector<Active> keep;
        for (Active& a : active) {
          if (!a.done) {
            keep.push_back(std::move(a));
            continue;
          }
          const size_t bytes = a.request.footprint;
          a.state.reset();
          a.graph.reset();
          Release(bytes);
        }
        active = std::move(keep);
        finished = 0;
        set.reset();
        changed = !active.empty();
      }
      {
        std::lock_guard<std::mutex> lock(mutex_);
        lanes_[lane].active = static_cast<int64_t>(active.size());
        int64_t total = 0;
        for (const Lane& l : lanes_) total += l.active;
        peak_active_ = std::max(peak_active_, total);
      }
      if (active.empty()) {
        std::unique_lock<std::mutex> lock(mutex_);
        if (stop_ && lanes_[lane].pending.empty()) return;
        wake_.wait_for(lock, std::chrono::milliseconds(5));
        continue;
      }
      if (changed) {
        std::vector<RunState*> states;
        for (Active& a : active) states.push_back(a.state.get());
        set = std::make_unique<LaunchSet>(states, s, &pool_, &pinned);
        changed = false;
      }
      // one round: one iteration of every active run, or a few while nobody waits for admission
      int chunk = kIdleChunk;
      {
        std::lock_guard<std::mutex> lock(mutex_);
        if (!lanes_[lane].pending.empty()) chunk = 1;
      }
      try {
        for (int i = 0; i < chunk; ++i) set->EnqueueIteration(nullptr, nullptr);
        CC_HIP_CHECK(hipGetLastError());
        const Ctl* ctl = set->PullCtl();
        for (size_t i = 0; i < active.size(); ++i) {
          if (active[i].done || !ctl[i].stopped) continue;
          Result r;
          Active& a = active[i];
          const uint32_t n = a.request.n;
          if (ctl[i].error) {
            r.code = CC_HIP_EDEVICE;
            r.error = std::string(DeviceErrorText(ctl[i].error)) + "; the run was stopped";
          } else {
            r.labels.resize(n);
            set->PullLabels(static_cast<int>(i), r.labels.data());
            r.trace.resize(ctl[i].iterations_done);
            set->PullTrace(static_cast<int>(i), ctl[i].iterations_done, r.trace.data());
            r.stats.objective = ctl[i].rec_f;
            r.stats.iterations_done = ctl[i].iterations_done;
            r.stats.local_searches = ctl[i].local_searches;
            r.stats.accepted_moves = ctl[i].accepted_moves;
            std::vector<char> seen(static_cast<size_t>(a.request.params.k), 0);
            for (int32_t label : r.labels) seen[label] = 1;
            for (char c : seen) r.stats.clusters_used += c;
            const DescentPlan plan = set->Plan(static_cast<int>(i));
            r.stats.descent = static_cast<int32_t>(plan.descent);
            r.stats.threads = plan.threads;
          }
          r.latency = Since(a.request.arrival);
          r.queued = std::chrono::duration<double>(a.admitted - a.request.arrival).count();
          r.stats.seconds = std::chrono::duration<double>(Clock::now() - a.admitted).count();
          a.done = true;
          ++finished;
          Complete(a.request.id, std::move(r));
        }
      } catch (const std::exception& e) {
        (void)hipStreamSynchronize(s);  // nothing in flight may still use the memory given back below
        for (Active& a : active) {
          if (!a.done) {
            Result r;
            r.code = CC_HIP_EDEVICE;
            r.error = e.what();
            Complete(a.request.id, std::move(r));
          }
          Release(a.request.footprint);
        }
        active.clear();
        finished = 0;
        set.reset();
      }
    }
  }

  cc_serve_options options_;
  int device_ = 0;
  // Every graph, run and launch table of the lanes. Blocks go back to it when a lane retires a run, after the lane's
  // stream has finished the work that used them; the slabs are freed when the server ends.
  DevicePool pool_{kSlabBytes};
  std::vector<Lane> lanes_;
  std::mutex mutex_;
  std::condition_variable wake_;
  std::condition_variable done_;
  bool stop_ = false;
  uint64_t last_id_ = 0;
  std::atomic<uint64_t> next_small_{0};
  int64_t budget_left_ = 0;
  int64_t admitted_ = 0;
  int64_t completed_ = 0;
  int64_t peak_active_ = 0;
  int64_t peak_bytes_ = 0;
  std::map<uint64_t, int> pending_ids_;
  std::map<uint64_t, Result> results_;
};

}  // namespace cc::hip

struct cc_hip_server {
  std::unique_ptr<cc::hip::Server> server;
};

extern "C" {

void cc_serve_default_options(cc_serve_options* options) {
  std::memset(options, 0, sizeof(*options));
  options->lanes = 2;
  options->max_active = 4096;
  options->large_bytes = int64_t(1) << 30;
}

int cc_hip_server_create(const cc_serve_options* options, cc_hip_server** out, char* err, size_t err_len) {
  *out = nullptr;
  return cc::hip::Guard(err, err_len, [&] {
    cc_serve_options o;
    cc_serve_default_options(&o);
    if (options) o = *options;
    auto h = std::make_unique<cc_hip_server>();
    h->server = std::make_unique<cc::hip::Server>(o);
    *out = h.release();
  });
}

void cc_hip_server_destroy(cc_hip_server* server) { delete server; }

int cc_hip_server_submit(cc_hip_server* server, const uint64_t* bits, uint32_t n, uint64_t edges,
                         const cc_pbils_params* params, uint64_t* id_out, char* err, size_t err_len) {
  return cc::hip::Guard(err, err_len, [&] { *id_out = server->server->Submit(bits, n, edges, *params); });
}

int cc_hip_server_wait(cc_hip_server* server, uint64_t id, double timeout_s, int32_t* labels_out, int64_t* trace_out,
                       cc_pbils_stats* stats, double* latency_s, double* queued_s, char* err, size_t err_len) {
  if (stats) std::memset(stats, 0, sizeof(*stats));
  cc::hip::Result r;
  int code = cc::hip::Guard(err, err_len, [&] {
    if (!server->server->Wait(id, timeout_s, &r)) cc::hip::Fail(CC_HIP_ETIMEOUT, "request %llu not done", (unsigned long long)id);
  });
  if (code != CC_HIP_OK) return code;
  if (r.code != CC_HIP_OK) {
    if (err && err_len) {
      std::strncpy(err, r.error.c_str(), err_len - 1);
      err[err_len - 1] = 0;
    }
    return r.code;
  }
  if (labels_out) std::copy(r.labels.begin(), r.labels.end(), labels_out);
  if (trace_out) std::copy(r.trace.begin(), r.trace.end(), trace_out);
  if (stats) *stats = r.stats;
  if (latency_s) *latency_s = r.latency;
  if (queued_s) *queued_s = r.queued;
  return CC_HIP_OK;
}

int64_t cc_hip_footprint(uint32_t n, const cc_pbils_params* params) {
  return static_cast<int64_t>(cc::hip::Server::Footprint(n, *params));
}

void cc_hip_server_counters(cc_hip_server* server, int64_t* admitted, int64_t* completed, int64_t* budget_left,
                            int64_t* peak_active, int64_t* peak_bytes) {
  server->server->Counters(admitted, completed, budget_left, peak_active, peak_bytes);
}

}  // extern "C"
