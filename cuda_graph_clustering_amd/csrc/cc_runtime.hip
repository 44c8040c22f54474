// cc_runtime.hip — host side of the HIP path: device graph, run memory, descent plan, the PBILS iteration loop,
// Neighborhood batching, and the C ABI of include/cc_hip.h.
//
// Orchestration follows the reference's phase order (src/cc_pbils_cpu.cpp:183-252; cuda/cc_pbils_gpu.cu:347-387):
// init -> rebuild -> record; per iteration: select + descent -> record -> GWW -> perturb -> rebuild. The record,
// early-stop test and GWW ranking run on the device (KernelRecord), so iterations are queued in chunks and the host
// synchronises once per chunk instead of twice per iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "cc_hip.h"
#include "cc_internal.hpp"

namespace cc::hip {
namespace {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

#define CC_HIP_CHECK(call)                                                                                     \
  do {                                                                                                         \
    const hipError_t status_ = (call);                                                                         \
    if (status_ != hipSuccess) {                                                                               \
      throw ::cc::hip::Error(CC_HIP_EDEVICE, std::string("hip error at ") + __FILE__ + ":" +                \
                             std::to_string(__LINE__) +                                                       \
                                      ": " + hipGetErrorString(status_));                                      \
    }                                                                                                          \
  } while (0)

void Fail(int code, const char* fmt, ...) {
  char text[512];
  va_list args;
  va_start(args, fmt);
  std::vsnprintf(text, sizeof(text), fmt, args);
  va_end(args);
  throw Error(code, text);
}

std::string Mebibytes(size_t bytes) { return std::to_string((bytes + (1 << 20) - 1) >> 20) + " MiB"; }

// Device allocations of one call, released together.
class DeviceMemory {
 public:
  DeviceMemory() = default;
  DeviceMemory(const DeviceMemory&) = delete;
  DeviceMemory& operator=(const DeviceMemory&) = delete;
  ~DeviceMemory() {
    for (void* p : blocks_) (void)hipFree(p);
  }
  template <typename T>
  T* Allocate(size_t count) {
    void* p = nullptr;
    if (count == 0) return nullptr;
    CC_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
    blocks_.push_back(p);
    return static_cast<T*>(p);
  }

 private:
  std::vector<void*> blocks_;
};

}  // namespace

// The graph on the device: the bitset and the degrees, uploaded once.
struct DeviceGraph {
  uint32_t n = 0;
  uint32_t words = 0;
  uint64_t edges = 0;
  int device = 0;
  uint64_t* bits = nullptr;
  int32_t* deg = nullptr;
  DeviceMemory memory;
};

namespace {

uint32_t Words64(uint32_t n) { return std::max<uint32_t>(1, (n + 63) / 64); }

void Upload(DeviceGraph& g, const uint64_t* bits, uint32_t n, uint64_t edges) {
  if (n == 0) Fail(CC_HIP_EINVAL, "graph has no vertices");
  if (n > static_cast<uint32_t>(kMaxVerticesWide)) {
    Fail(CC_HIP_EINVAL, "n exceeds %d, the range of the packed move", kMaxVerticesWide);
  }
  CC_HIP_CHECK(hipGetDevice(&g.device));
  g.n = n;
  g.words = Words64(n);
  g.edges = edges;
  const size_t count = static_cast<size_t>(n) * g.words;
  size_t free_bytes = 0, total_bytes = 0;
  CC_HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
  if (count * 8 + n * 4 > free_bytes) {
    Fail(CC_HIP_ENOMEM, "graph needs %s of device memory, %s free", Mebibytes(count * 8).c_str(),
         Mebibytes(free_bytes).c_str());
  }
  g.bits = g.memory.Allocate<uint64_t>(count);
  g.deg = g.memory.Allocate<int32_t>(n);
  CC_HIP_CHECK(hipMemcpy(g.bits, bits, count * sizeof(uint64_t), hipMemcpyHostToDevice));
  RunDev r{};
  r.bits = g.bits;
  r.deg = g.deg;
  r.n = static_cast<int32_t>(n);
  r.words = static_cast<int32_t>(g.words);
  LaunchDegrees(r, nullptr);
  CC_HIP_CHECK(hipGetLastError());
  CC_HIP_CHECK(hipDeviceSynchronize());
}

// reference CheckParams (src/cc_pbils_cpu.cpp:177-181) plus the device limits of SolveGpu (cuda/cc_pbils_gpu.cu:441-454)
void CheckParams(const cc_pbils_params& p) {
  if (p.k < 1) Fail(CC_HIP_EINVAL, "k must be >= 1");
  if (p.population < 1) Fail(CC_HIP_EINVAL, "population must be >= 1");
  if (p.gww < 0.0 || p.gww > 0.5) Fail(CC_HIP_EINVAL, "gww must be in [0, 0.5]");
  if (p.k > kMaxClusters) Fail(CC_HIP_EINVAL, "k exceeds kMaxClusters (%d)", kMaxClusters);
  if (p.ls_kernel < CC_LS_AUTO || p.ls_kernel > CC_LS_WIDE) Fail(CC_HIP_EINVAL, "unknown ls_kernel %d", p.ls_kernel);
}

bool WidePath(uint32_t n, const cc_pbils_params& p) {
  return n > static_cast<uint32_t>(kMaxVerticesNarrow) || p.ls_kernel == CC_LS_WIDE;
}

int GenericThreads(int n) {
  for (int t : {256, 512}) {
    if (t * 8 >= n) return t;
  }
  return 1024;
}

// The descent kernel and its launch shape (reference PlanDescent, cuda/cc_pbils_gpu.cu:69-102).
DescentPlan PlanDescent(uint32_t n_, int k, const cc_pbils_params& p, int wide) {
  const int n = static_cast<int>(n_);
  DescentPlan plan;
  plan.k = k;
  plan.blocks = p.population;
  plan.threads = GenericThreads(n);
  const size_t lds = GenericLdsBytes(n, k, wide);
  const bool fits = lds <= DescentLdsLimit();
  if (p.ls_kernel == CC_LS_SHARED && !fits) {
    Fail(CC_HIP_EINVAL, "state needs %zu bytes of LDS, the device allows %zu", lds, DescentLdsLimit());
  }
  // Wide path: the packed k = 2 / k = 3 kernels unless the generic kernel is asked for
  // (reference cuda/cc_pbils_gpu.cu:84-96).
  if (wide && p.ls_kernel != CC_LS_GLOBAL && (k == 2 || k == 3)) {
    plan.descent = k == 2 ? Descent::kK2Wide : Descent::kK3Wide;
    plan.threads = n >= 8192 ? 1024 : 256;
    return plan;
  }
  if (p.ls_kernel == CC_LS_AUTO && !wide && k >= 3) {
    int threads = 0;
    const int vpt = RegistersVpt(n, k, &threads);
    if (vpt > 0) {
      plan.descent = Descent::kRegisters;
      plan.threads = threads;
      plan.vpt = vpt;
      return plan;
    }
  }
  if (p.ls_kernel == CC_LS_AUTO && k == 2 && !wide) {
    plan.descent = Descent::kK2Registers;
    plan.threads = 1024;
    for (int t : {64, 128, 256, 512}) {
      if (t * 4 >= n) {
        plan.threads = t;
        break;
      }
    }
    plan.vpt = 32;
    for (int v : {1, 2, 4, 8, 16, 32}) {
      if (v * plan.threads >= n) {
        plan.vpt = v;
        break;
      }
    }
    const size_t rows = static_cast<size_t>(n) * Words64(n_) * sizeof(uint64_t);
    plan.rows_in_lds = rows <= DescentLdsLimit();
    plan.lds_bytes = plan.rows_in_lds ? rows : 0;
    return plan;
  }
  if (p.ls_kernel != CC_LS_GLOBAL && fits) {
    plan.descent = Descent::kGenericLds;
    plan.lds_bytes = lds;
  }
  return plan;
}

using Clock = std::chrono::steady_clock;

double Since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// HIP event pairs around launches, summed after the chunk's synchronisation.
class Timer {
 public:
  explicit Timer(bool on) : on_(on) {}
  ~Timer() {
    for (auto& e : events_) (void)hipEventDestroy(e);
  }
  void Begin(hipStream_t s) {
    if (on_) Record(s);
  }
  void End(hipStream_t s, double* total_ms, int64_t* launches) {
    if (!on_) return;
    Record(s);
    pending_.push_back({used_ - 2, total_ms, launches});
  }
  void Collect() {
    for (const auto& p : pending_) {
      float ms = 0.f;
      CC_HIP_CHECK(hipEventElapsedTime(&ms, events_[p.first], events_[p.first + 1]));
      *p.total += ms;
      *p.launches += 1;
    }
    pending_.clear();
    used_ = 0;
  }

 private:
  void Record(hipStream_t s) {
    if (used_ == events_.size()) {
      hipEvent_t e;
      CC_HIP_CHECK(hipEventCreate(&e));
      events_.push_back(e);
    }
    CC_HIP_CHECK(hipEventRecord(events_[used_++], s));
  }
  struct Pending {
    size_t first;
    double* total;
    int64_t* launches;
  };
  bool on_;
  std::vector<hipEvent_t> events_;
  std::vector<Pending> pending_;
  size_t used_ = 0;
};

class Stream {
 public:
  Stream() { CC_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
  ~Stream() { (void)hipStreamDestroy(s_); }
  hipStream_t get() const { return s_; }

 private:
  hipStream_t s_ = nullptr;
};

// Device memory of one run (reference Run, cuda/cc_pbils_gpu.cu:146-246).
struct RunMemory {
  DeviceMemory memory;
  RunDev r{};
  int32_t* order = nullptr;
  Ctl* host_ctl = nullptr;  // pinned

  RunMemory(const DeviceGraph& g, const cc_pbils_params& p, int population, bool wide, bool has_next,
            const DescentPlan& plan) {
    const size_t n = g.n;
    const size_t k = static_cast<size_t>(p.k);
    const size_t P = static_cast<size_t>(population);
    const size_t gsize = wide ? 4 : 2;
    const size_t iterations = static_cast<size_t>(std::max(p.iterations, 1));
    const size_t pop_bytes = P * n + P * k * n * gsize + P * k * 4 + P * 8;
    const size_t work_words = plan.descent == Descent::kK3Wide ? P * n : 0;
    const size_t bytes = pop_bytes * (has_next ? 2 : 1) + P * 8 * 2 + P * 4 * 2 + n + iterations * 8 +
                         P * k * g.words * 8 + work_words * 8 + sizeof(Ctl) + 4096;
    size_t free_bytes = 0, total_bytes = 0;
    CC_HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
    if (bytes > free_bytes) {
      Fail(CC_HIP_ENOMEM, "run needs %s of device memory, %s free", Mebibytes(bytes).c_str(),
           Mebibytes(free_bytes).c_str());
    }
    auto alloc_pop = [&](PopDev& d) {
      d.labels = memory.Allocate<uint8_t>(P * n);
      d.g = wide ? static_cast<void*>(memory.Allocate<int32_t>(P * k * n))
                 : static_cast<void*>(memory.Allocate<int16_t>(P * k * n));
      d.sizes = memory.Allocate<int32_t>(P * k);
      d.f = memory.Allocate<int64_t>(P);
    };
    alloc_pop(r.pop);
    if (has_next) {
      alloc_pop(r.next);
    } else {
      r.next = r.pop;
    }
    r.bits = g.bits;
    r.deg = g.deg;
    r.n = static_cast<int32_t>(g.n);
    r.words = static_cast<int32_t>(g.words);
    r.k = p.k;
    r.population = population;
    r.edges = static_cast<int64_t>(g.edges);
    r.moves = memory.Allocate<int64_t>(P);
    r.ctl = memory.Allocate<Ctl>(1);
    r.rec_labels = memory.Allocate<uint8_t>(n);
    r.trace = memory.Allocate<int64_t>(iterations);
    r.src = memory.Allocate<int32_t>(P);
    order = memory.Allocate<int32_t>(P);
    r.masks = memory.Allocate<uint64_t>(P * k * g.words);
    r.intra = memory.Allocate<int64_t>(P);
    r.work = memory.Allocate<uint64_t>(work_words);
    CC_HIP_CHECK(hipMemset(r.moves, 0, P * sizeof(int64_t)));
    CC_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_ctl), sizeof(Ctl), hipHostMallocDefault));
  }
  ~RunMemory() {
    if (host_ctl) (void)hipHostFree(host_ctl);
  }

  void PullCtl(hipStream_t s) {
    CC_HIP_CHECK(hipMemcpyAsync(host_ctl, r.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
    CC_HIP_CHECK(hipStreamSynchronize(s));
  }

  void PullLabels(int32_t* labels_out, hipStream_t s) {
    std::vector<uint8_t> labels(r.n);
    CC_HIP_CHECK(hipMemcpyAsync(labels.data(), r.rec_labels, labels.size(), hipMemcpyDeviceToHost, s));
    CC_HIP_CHECK(hipStreamSynchronize(s));
    for (int v = 0; v < r.n; ++v) labels_out[v] = labels[v];
  }
};

int ClustersUsed(const int32_t* labels, uint32_t n, int k) {
  std::vector<char> seen(static_cast<size_t>(std::max(k, 1)), 0);
  for (uint32_t v = 0; v < n; ++v) {
    if (labels[v] >= 0 && labels[v] < k) seen[labels[v]] = 1;
  }
  int used = 0;
  for (char c : seen) used += c;
  return used;
}

void FillPlan(const DescentPlan& plan, cc_pbils_stats* stats) {
  stats->descent = static_cast<int32_t>(plan.descent);
  stats->threads = plan.threads;
  stats->blocks = plan.blocks;
  stats->vertices_per_thread = plan.vpt;
  stats->lds_bytes = static_cast<int64_t>(plan.lds_bytes);
}

const char* DescentName(Descent d) {
  switch (d) {
    case Descent::kGenericGlobal: return "generic, global memory";
    case Descent::kGenericLds: return "generic, LDS";
    case Descent::kK2Registers: return "k = 2, registers";
    case Descent::kK2Wide: return "k = 2, packed words";
    case Descent::kK3Wide: return "k = 3, packed words";
    case Descent::kRegisters: return "k = 3..8, registers";
  }
  return "?";
}

void PrintPlan(const DescentPlan& plan, size_t run_bytes) {
  std::printf("  local search: %s, %d threads per block, %d blocks", DescentName(plan.descent), plan.threads,
              plan.blocks);
  if (plan.descent == Descent::kK2Registers) {
    std::printf(", %d vertices per thread%s", plan.vpt, plan.rows_in_lds ? ", adjacency in LDS" : "");
  }
  if (plan.descent == Descent::kGenericLds) std::printf(", %zu bytes of LDS", plan.lds_bytes);
  std::printf("\n");
  (void)run_bytes;
}

// PBILS on a device graph (reference gpu::Solve, cuda/cc_pbils_gpu.cu:347-387).
void SolvePbils(const DeviceGraph& g, const cc_pbils_params& p, int32_t* labels_out, cc_pbils_stats* stats,
                int64_t* trace_out) {
  const auto started = Clock::now();
  CheckParams(p);
  const int wide = WidePath(g.n, p) ? 1 : 0;
  const DescentPlan plan = PlanDescent(g.n, p.k, p, wide);
  if (!DescentSupported(plan, wide)) Fail(CC_HIP_EINVAL, "no descent kernel for this plan");
  Stream stream;
  const hipStream_t s = stream.get();
  RunMemory run(g, p, p.population, wide != 0, true, plan);
  RunDev& r = run.r;
  r.init_stream = StreamSeed(p.seed, kInitStream, 0);
  r.select_stream = StreamSeed(p.seed, kSelectStream, 0);
  r.perturb_stream = StreamSeed(p.seed, kPerturbStream, 0);
  r.tournament = p.tournament;
  r.copies = static_cast<int32_t>(p.gww * p.population);  // WinnerCopies, src/cc_pbils_cpu.cpp:168
  r.early_stop = p.early_stop;
  r.iterations = p.iterations;
  r.perturbation = static_cast<float>(p.perturbation);
  r.do_perturb = (p.k >= 2 && p.perturbation > 0.0) ? 1 : 0;
  r.select = 1;
  if (p.verbose) PrintPlan(plan, 0);

  Timer timer(p.profile != 0);
  Timer iteration_timer(p.profile != 0);
  int64_t iteration_spans = 0;
  LaunchResetCtl(r, s);
  LaunchInitLabels(r, s);
  LaunchRebuild(r, wide, s);
  LaunchRecord(r, kRecordInitial, 0, run.order, s);
  CC_HIP_CHECK(hipGetLastError());

  int launched = 0;
  int chunk = p.time_limit_sec > 0.0 ? 1 : 4;
  int printed = 0;
  int64_t last_record = -1;
  int stall = 0;
  if (p.verbose) {
    run.PullCtl(s);
    last_record = run.host_ctl->rec_f;
  }
  iteration_timer.Begin(s);
  while (launched < p.iterations) {
    const int count = std::min(chunk, p.iterations - launched);
    for (int i = 0; i < count; ++i) {
      timer.Begin(s);
      LaunchDescent(r, plan, wide, s);
      timer.End(s, &stats->search_ms, &stats->search_launches);
      LaunchRecord(r, kRecordIteration, 1, run.order, s);
      LaunchPerturb(r, s);
      timer.Begin(s);
      LaunchRebuild(r, wide, s);
      timer.End(s, &stats->rebuild_ms, &stats->rebuild_launches);
    }
    CC_HIP_CHECK(hipGetLastError());
    launched += count;
    run.PullCtl(s);
    timer.Collect();
    const Ctl& c = *run.host_ctl;
    if (p.verbose && c.iterations_done > printed) {
      std::vector<int64_t> trace(c.iterations_done);
      CC_HIP_CHECK(hipMemcpy(trace.data(), r.trace, trace.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
      for (int it = printed; it < c.iterations_done; ++it) {
        stall = trace[it] < last_record ? 0 : stall + 1;
        last_record = trace[it];
        std::printf("  iter %3d  record %lld  stall %d  %.3fs\n", it + 1, static_cast<long long>(trace[it]), stall,
                    Since(started));
      }
      printed = c.iterations_done;
    }
    if (c.stopped) break;
    if (p.time_limit_sec > 0.0 && Since(started) >= p.time_limit_sec) break;
    chunk = std::min(chunk * 2, 32);
  }
  iteration_timer.End(s, &stats->iteration_ms, &iteration_spans);
  run.PullCtl(s);
  iteration_timer.Collect();
  const Ctl c = *run.host_ctl;
  if (c.error) Fail(CC_HIP_EDEVICE, "a descent exceeded its move cap; the run was stopped");
  run.PullLabels(labels_out, s);
  if (trace_out && c.iterations_done > 0) {
    CC_HIP_CHECK(hipMemcpy(trace_out, r.trace, static_cast<size_t>(c.iterations_done) * sizeof(int64_t),
                           hipMemcpyDeviceToHost));
  }
  stats->objective = c.rec_f;
  stats->iterations_done = c.iterations_done;
  stats->local_searches = c.local_searches;
  stats->accepted_moves = c.accepted_moves;
  stats->clusters_used = ClustersUsed(labels_out, g.n, p.k);
  FillPlan(plan, stats);
  stats->seconds = Since(started);
}

// Neighborhood for k = 2 in batches of at most `population` start vertices (reference SolveNeighborhood,
// cuda/cc_pbils_gpu.cu:389-436). Starts are processed in vertex order and the record only moves on a strictly
// smaller f, so ties keep the smaller start vertex across batches.
void SolveNeighborhood(const DeviceGraph& g, const cc_pbils_params& params, bool local_search, int32_t* labels_out,
                       cc_pbils_stats* stats) {
  const auto started = Clock::now();
  if (params.k != 2) Fail(CC_HIP_EINVAL, "Neighborhood algorithms require k = 2");
  if (params.population < 1) Fail(CC_HIP_EINVAL, "population must be >= 1");
  if (params.ls_kernel < CC_LS_AUTO || params.ls_kernel > CC_LS_WIDE) {
    Fail(CC_HIP_EINVAL, "unknown ls_kernel %d", params.ls_kernel);
  }
  cc_pbils_params p = params;
  const int n = static_cast<int>(g.n);
  p.population = std::min(params.population, n);
  p.gww = 0.0;
  p.iterations = 1;
  const int wide = WidePath(g.n, p) ? 1 : 0;
  DescentPlan plan = PlanDescent(g.n, 2, p, wide);
  if (!DescentSupported(plan, wide)) Fail(CC_HIP_EINVAL, "no descent kernel for this plan");
  Stream stream;
  const hipStream_t s = stream.get();
  RunMemory run(g, p, p.population, wide != 0, false, plan);
  RunDev& r = run.r;
  r.select = 0;
  r.iterations = 1;
  r.early_stop = 1;
  if (p.verbose) PrintPlan(plan, 0);
  Timer timer(p.profile != 0);
  LaunchResetCtl(r, s);
  const int capacity = p.population;
  for (int first = 0; first < n; first += capacity) {
    RunDev b = r;
    b.population = std::min(capacity, n - first);
    LaunchNeighborhoodLabels(b, first, b.population, s);
    timer.Begin(s);
    LaunchRebuild(b, wide, s);
    timer.End(s, &stats->rebuild_ms, &stats->rebuild_launches);
    if (local_search) {
      timer.Begin(s);
      LaunchDescent(b, plan, wide, s);
      timer.End(s, &stats->search_ms, &stats->search_launches);
    }
    LaunchRecord(b, kRecordBatch, 1, run.order, s);
    CC_HIP_CHECK(hipGetLastError());
    if (p.verbose) {
      run.PullCtl(s);
      std::printf("  neighborhoods %d/%d  record %lld  %.3fs\n", first + b.population, n,
                  static_cast<long long>(run.host_ctl->rec_f), Since(started));
    }
  }
  run.PullCtl(s);
  timer.Collect();
  const Ctl c = *run.host_ctl;
  if (c.error) Fail(CC_HIP_EDEVICE, "a descent exceeded its move cap; the run was stopped");
  run.PullLabels(labels_out, s);
  stats->objective = c.rec_f;
  stats->iterations_done = 0;
  stats->local_searches = local_search ? n : 0;
  stats->accepted_moves = c.accepted_moves;
  stats->clusters_used = ClustersUsed(labels_out, g.n, 2);
  FillPlan(plan, stats);
  stats->seconds = Since(started);
}

int Report(const std::exception& e, char* err, size_t err_len, int code) {
  if (err && err_len > 0) {
    std::strncpy(err, e.what(), err_len - 1);
    err[err_len - 1] = 0;
  }
  return code;
}

template <typename Body>
int Guard(char* err, size_t err_len, Body&& body) {
  try {
    body();
    return CC_HIP_OK;
  } catch (const Error& e) {
    return Report(e, err, err_len, e.code);
  } catch (const std::exception& e) {
    return Report(e, err, err_len, CC_HIP_EDEVICE);
  }
}

void ClearStats(cc_pbils_stats* stats) { std::memset(stats, 0, sizeof(*stats)); }

}  // namespace
}  // namespace cc::hip

struct cc_hip_graph {
  cc::hip::DeviceGraph g;
};

extern "C" {

void cc_pbils_default_params(cc_pbils_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->k = 2;  // reference include/cc_pbils.hpp:37-48
  p->population = 128;
  p->tournament = 5;
  p->iterations = 100;
  p->early_stop = 6;
  p->threads = 1;
  p->perturbation = 0.4;
  p->gww = 0.1;
  p->seed = 1;
  p->time_limit_sec = 0.0;
  p->verbose = 0;
  p->ls_kernel = CC_LS_AUTO;
  p->profile = 0;
}

int cc_hip_graph_upload(const uint64_t* bits, uint32_t n, uint64_t edges, cc_hip_graph** out, char* err,
                        size_t err_len) {
  *out = nullptr;
  return cc::hip::Guard(err, err_len, [&] {
    auto* h = new cc_hip_graph();
    try {
      cc::hip::Upload(h->g, bits, n, edges);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

void cc_hip_graph_free(cc_hip_graph* graph) { delete graph; }

int cc_hip_graph_solve_pbils(cc_hip_graph* graph, const cc_pbils_params* params, int32_t* labels_out,
                             cc_pbils_stats* stats, int64_t* trace_out, char* err, size_t err_len) {
  cc::hip::ClearStats(stats);
  return cc::hip::Guard(err, err_len, [&] {
    CC_HIP_CHECK(hipSetDevice(graph->g.device));
    cc::hip::SolvePbils(graph->g, *params, labels_out, stats, trace_out);
  });
}

int cc_hip_graph_solve_neighborhood(cc_hip_graph* graph, const cc_pbils_params* params, int32_t local_search,
                                    int32_t* labels_out, cc_pbils_stats* stats, char* err, size_t err_len) {
  cc::hip::ClearStats(stats);
  return cc::hip::Guard(err, err_len, [&] {
    CC_HIP_CHECK(hipSetDevice(graph->g.device));
    cc::hip::SolveNeighborhood(graph->g, *params, local_search != 0, labels_out, stats);
  });
}

int cc_hip_solve_pbils(const uint64_t* bits, uint32_t n, uint64_t edges, const cc_pbils_params* params,
                       int32_t* labels_out, cc_pbils_stats* stats, int64_t* trace_out, char* err, size_t err_len) {
  cc::hip::ClearStats(stats);
  return cc::hip::Guard(err, err_len, [&] {
    const auto started = cc::hip::Clock::now();  // the solver clock includes allocation and upload
    cc::hip::CheckParams(*params);
    cc::hip::DeviceGraph g;
    cc::hip::Upload(g, bits, n, edges);
    cc::hip::SolvePbils(g, *params, labels_out, stats, trace_out);
    stats->seconds = cc::hip::Since(started);
  });
}

int cc_hip_solve_neighborhood(const uint64_t* bits, uint32_t n, uint64_t edges, const cc_pbils_params* params,
                              int32_t local_search, int32_t* labels_out, cc_pbils_stats* stats, char* err,
                              size_t err_len) {
  cc::hip::ClearStats(stats);
  return cc::hip::Guard(err, err_len, [&] {
    const auto started = cc::hip::Clock::now();
    if (params->k != 2) cc::hip::Fail(CC_HIP_EINVAL, "Neighborhood algorithms require k = 2");
    if (params->population < 1) cc::hip::Fail(CC_HIP_EINVAL, "population must be >= 1");
    if (n == 0 || n > static_cast<uint32_t>(cc::hip::kMaxVerticesWide)) {
      cc::hip::Fail(CC_HIP_EINVAL, "Neighborhood requires 1 <= n <= %d", cc::hip::kMaxVerticesWide);
    }
    cc::hip::DeviceGraph g;
    cc::hip::Upload(g, bits, n, edges);
    cc::hip::SolveNeighborhood(g, *params, local_search != 0, labels_out, stats);
    stats->seconds = cc::hip::Since(started);
  });
}

int cc_hip_device_info(cc_device_info* info, char* err, size_t err_len) {
  std::memset(info, 0, sizeof(*info));
  return cc::hip::Guard(err, err_len, [&] {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) cc::hip::Fail(CC_HIP_ENODEV, "no HIP device");
    int id = 0;
    CC_HIP_CHECK(hipGetDevice(&id));
    hipDeviceProp_t prop;
    CC_HIP_CHECK(hipGetDeviceProperties(&prop, id));
    std::strncpy(info->name, prop.name, sizeof(info->name) - 1);
    std::strncpy(info->arch, prop.gcnArchName, sizeof(info->arch) - 1);
    info->compute_units = prop.multiProcessorCount;
    info->lds_per_block = static_cast<int32_t>(prop.sharedMemPerBlock);
    info->l2_bytes = prop.l2CacheSize;
    info->clock_khz = prop.clockRate;
    size_t free_bytes = 0, total_bytes = 0;
    CC_HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
    info->total_bytes = static_cast<int64_t>(total_bytes);
    info->free_bytes = static_cast<int64_t>(free_bytes);
  });
}

}  // extern "C"
