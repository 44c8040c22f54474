// cc_runtime.hpp — host runtime internals shared by cc_runtime.hip (solves, C ABI) and cc_serve.hip (serving).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "cc_hip.h"
#include "cc_internal.hpp"

namespace cc::hip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what);
};

#define CC_HIP_CHECK(call)                                                                                      \
  do {                                                                                                          \
    const hipError_t status_ = (call);                                                                          \
    if (status_ != hipSuccess) {                                                                                \
      throw ::cc::hip::Error(CC_HIP_EDEVICE, std::string("hip error at ") + __FILE__ + ":" +                    \
                                                 std::to_string(__LINE__) + ": " + hipGetErrorString(status_)); \
    }                                                                                                           \
  } while (0)

[[noreturn]] void Fail(int code, const char* fmt, ...);
std::string Mebibytes(size_t bytes);
size_t FreeDeviceBytes();
uint32_t Words64(uint32_t n);

// Device memory for many short-lived blocks (a server's graphs, runs and launch tables): slabs taken from hipMalloc
// on demand and kept until the pool is destroyed, handed out first-fit in 256-byte units. Returning a block is host
// bookkeeping only. hipFree waits for the work of every stream on the device, so freeing through it would stall one
// serving lane behind another lane's long kernels. The owner returns a block only after the work that used it has
// finished on its stream. Thread-safe.
class DevicePool {
 public:
  explicit DevicePool(size_t slab_bytes) : slab_bytes_(slab_bytes) {}
  ~DevicePool();
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;
  void* Take(size_t bytes);
  void Give(void* p);
  size_t reserved() const;  // bytes of the slabs
  size_t in_use() const;    // bytes of the blocks handed out

 private:
  mutable std::mutex mutex_;
  size_t slab_bytes_;
  std::vector<std::pair<uint8_t*, size_t>> slabs_;
  std::map<uint8_t*, size_t> free_;           // address -> bytes; neighbours within a slab are merged
  std::map<uint8_t*, size_t> used_;
  size_t reserved_ = 0, in_use_ = 0;
};

// A fixed device region handed out front to back in 256-byte units: the static scratch region of the solve that
// allocates nothing. Blocks are not returned; the next call starts over at the region's start.
class ScratchArena {
 public:
  ScratchArena(void* base, size_t bytes) : base_(static_cast<uint8_t*>(base)), bytes_(bytes) {}
  void* Take(size_t bytes);  // CC_HIP_ENOMEM beyond the region
  size_t used() const { return used_; }

 private:
  uint8_t* base_;
  size_t bytes_;
  size_t used_ = 0;
};

// Device allocations released together: by hipFree, or back to a DevicePool when one is set before the first
// allocation; blocks taken from a ScratchArena are not released at all.
class DeviceMemory {
 public:
  DeviceMemory() = default;
  DeviceMemory(const DeviceMemory&) = delete;
  DeviceMemory& operator=(const DeviceMemory&) = delete;
  ~DeviceMemory();
  void UsePool(DevicePool* pool) { pool_ = pool; }
  void UseArena(ScratchArena* arena) { arena_ = arena; }
  void* AllocateBytes(size_t bytes);

 private:
  DevicePool* pool_ = nullptr;
  ScratchArena* arena_ = nullptr;
  std::vector<void*> blocks_;
};

// Pinned host memory that only grows; earlier buffers are kept until destruction (hipHostFree, like hipFree, waits
// for the whole device).
class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  ~PinnedBuffer();
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  void* Reserve(size_t bytes);

 private:
  std::vector<void*> buffers_;
  size_t capacity_ = 0;
};

// A graph on the device: the bitset and the degrees, uploaded once.
struct DeviceGraph {
  uint32_t n = 0;
  uint32_t words = 0;
  uint64_t edges = 0;
  int device = 0;
  uint64_t* bits = nullptr;
  int32_t* deg = nullptr;
  DeviceMemory memory;
};

// With a pool, the graph's memory comes from it (and the free-memory check is left to the pool's slab allocation).
void Upload(DeviceGraph& g, const uint64_t* bits, uint32_t n, uint64_t edges, hipStream_t stream,
            DevicePool* pool = nullptr);
void CheckParams(const cc_pbils_params& p);
bool WidePath(uint32_t n, const cc_pbils_params& p);
// `edges` (of the graph) decides between the two k = 3 wide kernels; 0 for a plan made from the shape alone, whose
// memory footprint does not depend on that choice.
DescentPlan PlanDescent(uint32_t n, int k, const cc_pbils_params& p, uint64_t edges);
// The autotuner's recorded plan for a shape, if any (cc_tune.hip).
bool TunedPlan(uint32_t n, int k, int population, bool wide, DescentPlan* plan);

using Clock = std::chrono::steady_clock;
double Since(Clock::time_point t0);

// HIP event pairs around launches on one stream, summed after the next synchronisation.
class Timer {
 public:
  explicit Timer(bool on) : on_(on) {}
  ~Timer();
  void Begin(hipStream_t s) {
    if (on_) Record(s);
  }
  void End(hipStream_t s, double* total_ms, int64_t* launches);
  void Collect();

 private:
  void Record(hipStream_t s);
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

// A non-blocking stream of default priority.
class Stream {
 public:
  Stream();
  ~Stream();
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  hipStream_t get() const { return s_; }

 private:
  hipStream_t s_ = nullptr;
};

// One run of a batch, as the host sees it.
struct RunSpec {
  const DeviceGraph* graph;
  cc_pbils_params params;
  DescentPlan plan;
  int population;  // slots (PBILS population, or the Neighborhood batch of starts)
  bool pbils;      // PBILS: tournament selection and a second population; Neighborhood: descent in place
  int variant = -1;  // mixed-descent variant when plan.descent is kMixed
};

// A register-kernel run (ls_kernel auto) moved to the mixed descent of its block size (256 threads up to 2048
// vertices, 512 up to 4096, else 1024), so that the runs of a batch or a serving lane share one or two launches; same trajectory.
void UseMixedDescent(RunSpec& spec);

struct DescentGroup {
  DescentPlan plan;
  ItemList list;
};

// Device state of one run (populations, record, trace and scratch) in one zeroed allocation. It outlives the launch
// sets it takes part in, so a serving lane can add and drop runs between iterations.
class RunState {
 public:
  // Memory from hipMalloc, or from `pool`, or from `arena` (then nothing is allocated).
  RunState(const RunSpec& spec, hipStream_t stream, DevicePool* pool = nullptr, ScratchArena* arena = nullptr);
  RunState(const RunState&) = delete;
  RunState& operator=(const RunState&) = delete;
  const RunSpec& spec() const { return spec_; }
  const RunDev& dev() const { return dev_; }
  size_t bytes() const { return bytes_; }
  // Device bytes a run of this spec allocates (the graph excluded).
  static size_t Footprint(const RunSpec& spec);

 private:
  RunSpec spec_;
  RunDev dev_{};
  DeviceMemory memory_;
  size_t bytes_ = 0;
};

// The launches over a set of runs: the device run table, the item tables of every launch (slots, rebuild
// tiles, descent items grouped by kernel instance and ordered largest run first) and the persistent-queue
// counters. An iteration of all runs is one fixed launch sequence. With a pool and a pinned buffer (both kept by the
// caller), the set allocates and frees nothing through the HIP runtime; with an arena it allocates no device or
// pinned memory at all (the control blocks come back into pageable host memory).
class LaunchSet {
 public:
  LaunchSet(const std::vector<RunState*>& states, hipStream_t stream, DevicePool* pool = nullptr,
            PinnedBuffer* pinned = nullptr, ScratchArena* arena = nullptr);
  // Device bytes of the tables a set over runs of these specs takes (what it draws from a pool or an arena).
  static size_t TableBytes(const std::vector<RunSpec>& specs);
  ~LaunchSet();
  LaunchSet(const LaunchSet&) = delete;
  LaunchSet& operator=(const LaunchSet&) = delete;

  void Start();  // reset, initial labels, rebuild, initial record
  void EnqueueIteration(Timer* timer, cc_pbils_stats* stats);
  void Rebuild(Timer* timer, cc_pbils_stats* stats);
  void Descend(Timer* timer, cc_pbils_stats* stats);
  void SetLdsAttributes();

  const Ctl* PullCtl();
  void PullLabels(int run, int32_t* labels_out);
  void PullTrace(int run, int count, int64_t* out);
  void SetPopulation(int run, int population);

  int Size() const { return static_cast<int>(states_.size()); }
  const RunDev* Runs() const { return runs_; }
  const ItemList& AllItems() const { return all_; }
  const std::vector<RebuildGroup>& RebuildGroups() const { return rebuild_; }
  const std::vector<DescentGroup>& DescentGroups() const { return descent_; }
  ItemList Prefix(const ItemList& list, int count) const;
  DescentPlan Plan(int run) const;

 private:
  std::vector<RunState*> states_;
  hipStream_t stream_;
  DeviceMemory memory_;
  std::vector<RunDev> host_runs_;
  RunDev* runs_ = nullptr;
  ItemList all_{};
  std::vector<RebuildGroup> rebuild_;
  std::vector<DescentGroup> descent_;
  int32_t* queues_ = nullptr;
  Ctl* ctl_table_ = nullptr;  // [runs] device copy of every run's Ctl, gathered for one transfer
  int max_n_ = 0;
  Ctl* host_ctl_ = nullptr;
  bool own_host_ctl_ = false;
  std::vector<Ctl> pageable_ctl_;  // the control blocks' host copy when no pinned memory may be allocated
};

// A fixed set of runs with its launch set; iterations are replayed from one captured HIP graph.
class Batch {
 public:
  // With an arena every piece of device state comes from it, and Iterate must be called with use_graph false.
  Batch(const std::vector<RunSpec>& specs, hipStream_t stream, ScratchArena* arena = nullptr);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  void Start() { set_->Start(); }
  void Iterate(int count, bool use_graph, Timer* timer, cc_pbils_stats* stats);
  const Ctl* PullCtl() { return set_->PullCtl(); }
  void PullLabels(int run, int32_t* labels_out) { set_->PullLabels(run, labels_out); }
  void PullTrace(int run, int count, int64_t* out) { set_->PullTrace(run, count, out); }
  void SetPopulation(int run, int population) { set_->SetPopulation(run, population); }
  const RunDev* Runs() const { return set_->Runs(); }
  const ItemList& AllItems() const { return set_->AllItems(); }
  const std::vector<RebuildGroup>& RebuildGroups() const { return set_->RebuildGroups(); }
  const std::vector<DescentGroup>& DescentGroups() const { return set_->DescentGroups(); }
  ItemList Prefix(const ItemList& list, int count) const { return set_->Prefix(list, count); }
  DescentPlan Plan(int run) const { return set_->Plan(run); }
  size_t Bytes() const { return bytes_; }

 private:
  hipStream_t stream_;
  std::vector<std::unique_ptr<RunState>> states_;
  std::unique_ptr<LaunchSet> set_;
  size_t bytes_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
};

void RunToCompletion(Batch& batch, const std::vector<RunSpec>& specs, bool verbose, bool use_graph, double time_limit,
                     Timer* timer, cc_pbils_stats* stats, Clock::time_point started);

void SolvePbils(const DeviceGraph& g, const cc_pbils_params& p, int32_t* labels_out, cc_pbils_stats* stats,
                int64_t* trace_out, Clock::time_point started);
// SolvePbils with all device state in the static scratch region of the current device (no allocation, no HIP graph,
// no event timing), on `stream`.
void SolvePbilsStatic(const DeviceGraph& g, const cc_pbils_params& p, hipStream_t stream, int32_t* labels_out,
                      cc_pbils_stats* stats, int64_t* trace_out, Clock::time_point started);
void SolveNeighborhood(const DeviceGraph& g, const cc_pbils_params& params, bool local_search, int32_t* labels_out,
                       cc_pbils_stats* stats, Clock::time_point started);

struct BatchRun {
  const DeviceGraph* graph;
  cc_pbils_params params;
  int32_t* labels;
  int64_t* trace;
  cc_pbils_stats* stats;
};
void SolveBatch(std::vector<BatchRun>& runs, bool use_graph);

int Report(const std::exception& e, char* err, size_t err_len, int code);

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

}  // namespace cc::hip

struct cc_hip_graph {
  cc::hip::DeviceGraph g;
};
