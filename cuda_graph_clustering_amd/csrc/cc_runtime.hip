// cc_runtime.hip — host side of the HIP path: device graphs, the descent plan, the batch engine (any number of
// independent PBILS runs of different n and k advanced together), the solves built on it, and the C ABI of
// include/cc_hip.h.
//
// Orchestration follows the reference's phase order (src/cc_pbils_cpu.cpp:183-252; cuda/cc_pbils_gpu.cu:347-387):
// init -> rebuild -> record; per iteration: select + descent -> record -> GWW -> perturb -> rebuild. The record,
// early-stop test and GWW ranking run on the device (KernelRecord), so an iteration is a fixed sequence of launches
// that is captured once in a HIP graph and replayed; the host synchronises once per chunk of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "cc_hip.h"
#include "cc_internal.hpp"
#include "cc_runtime.hpp"

namespace cc::hip {

Error::Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}

void Fail(int code, const char* fmt, ...) {
  char text[512];
  va_list args;
  va_start(args, fmt);
  std::vsnprintf(text, sizeof(text), fmt, args);
  va_end(args);
  throw Error(code, text);
}

std::string Mebibytes(size_t bytes) { return std::to_string((bytes + (1 << 20) - 1) >> 20) + " MiB"; }

DevicePool::~DevicePool() {
  for (auto& [p, bytes] : slabs_) (void)hipFree(p);
}

void* DevicePool::Take(size_t bytes) {
  const size_t need = (bytes + 255) & ~static_cast<size_t>(255);
  std::lock_guard<std::mutex> lock(mutex_);
  auto it = free_.begin();
  while (it != free_.end() && it->second < need) ++it;
  if (it == free_.end()) {
    // a new slab: the default size, or a whole one for a larger block (rounded to 2 MiB pages)
    const size_t slab = std::max(slab_bytes_, (need + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1));
    void* p = nullptr;
    const hipError_t status = hipMalloc(&p, slab);
    if (status != hipSuccess) {
      (void)hipGetLastError();
      Fail(CC_HIP_ENOMEM, "device pool: a new slab of %s failed (%s reserved, %s in use): %s", Mebibytes(slab).c_str(),
           Mebibytes(reserved_).c_str(), Mebibytes(in_use_).c_str(), hipGetErrorString(status));
    }
    slabs_.push_back({static_cast<uint8_t*>(p), slab});
    reserved_ += slab;
    it = free_.insert({static_cast<uint8_t*>(p), slab}).first;
  }
  uint8_t* block = it->first;
  const size_t left = it->second - need;
  free_.erase(it);
  if (left > 0) free_.insert({block + need, left});
  used_.insert({block, need});
  in_use_ += need;
  return block;
}

void DevicePool::Give(void* p) {
  std::lock_guard<std::mutex> lock(mutex_);
  auto u = used_.find(static_cast<uint8_t*>(p));
  if (u == used_.end()) return;
  uint8_t* block = u->first;
  size_t bytes = u->second;
  used_.erase(u);
  in_use_ -= bytes;
  auto is_slab_start = [&](const uint8_t* a) {
    for (const auto& s : slabs_) {
      if (s.first == a) return true;
    }
    return false;
  };
  // merge with the free block after it and the one before it, never across the start of a slab
  auto next = free_.find(block + bytes);
  if (next != free_.end() && !is_slab_start(next->first)) {
    bytes += next->second;
    free_.erase(next);
  }
  auto after = free_.lower_bound(block);
  if (after != free_.begin() && !is_slab_start(block)) {
    auto prev = std::prev(after);
    if (prev->first + prev->second == block) {
      prev->second += bytes;
      return;
    }
  }
  free_.insert({block, bytes});
}

size_t DevicePool::reserved() const {
  std::lock_guard<std::mutex> lock(mutex_);
  return reserved_;
}

size_t DevicePool::in_use() const {
  std::lock_guard<std::mutex> lock(mutex_);
  return in_use_;
}

void* ScratchArena::Take(size_t bytes) {
  const size_t need = (bytes + 255) & ~static_cast<size_t>(255);
  if (need > bytes_ - used_) {
    Fail(CC_HIP_ENOMEM, "the static scratch region of %s is exhausted (%s taken, %s more needed)",
         Mebibytes(bytes_).c_str(), Mebibytes(used_).c_str(), Mebibytes(need).c_str());
  }
  void* p = base_ + used_;
  used_ += need;
  return p;
}

DeviceMemory::~DeviceMemory() {
  for (void* p : blocks_) {
    if (pool_) {
      pool_->Give(p);
    } else {
      (void)hipFree(p);
    }
  }
}

void* DeviceMemory::AllocateBytes(size_t bytes) {
  if (bytes == 0) return nullptr;
  if (arena_) return arena_->Take(bytes);  // not recorded: an arena's blocks are never released
  void* p = nullptr;
  if (pool_) {
    p = pool_->Take(bytes);
  } else {
    CC_HIP_CHECK(hipMalloc(&p, bytes));
  }
  blocks_.push_back(p);
  return p;
}

PinnedBuffer::~PinnedBuffer() {
  for (void* p : buffers_) (void)hipHostFree(p);
}

void* PinnedBuffer::Reserve(size_t bytes) {
  if (bytes > capacity_ || buffers_.empty()) {
    const size_t grown = std::max(bytes, 2 * capacity_);
    void* p = nullptr;
    CC_HIP_CHECK(hipHostMalloc(&p, std::max<size_t>(grown, 64), hipHostMallocDefault));
    buffers_.push_back(p);
    capacity_ = std::max<size_t>(grown, 64);
  }
  return buffers_.back();
}

uint32_t Words64(uint32_t n) { return std::max<uint32_t>(1, (n + 63) / 64); }

size_t FreeDeviceBytes() {
  size_t free_bytes = 0, total_bytes = 0;
  CC_HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
  return free_bytes;
}

void Upload(DeviceGraph& g, const uint64_t* bits, uint32_t n, uint64_t edges, hipStream_t stream, DevicePool* pool) {
  if (n == 0) Fail(CC_HIP_EINVAL, "graph has no vertices");
  if (n > static_cast<uint32_t>(kMaxVerticesWide)) {
    Fail(CC_HIP_EINVAL, "n exceeds %d, the range of the packed move", kMaxVerticesWide);
  }
  CC_HIP_CHECK(hipGetDevice(&g.device));
  g.n = n;
  g.words = Words64(n);
  g.edges = edges;
  const size_t count = static_cast<size_t>(n) * g.words;
  if (pool) {
    g.memory.UsePool(pool);
  } else {
    const size_t bytes = count * 8 + kRowTailBytes + static_cast<size_t>(n) * 4 + sizeof(RunDev);
    const size_t free_bytes = FreeDeviceBytes();
    if (bytes > free_bytes) {
      Fail(CC_HIP_ENOMEM, "graph needs %s of device memory, %s free", Mebibytes(bytes).c_str(),
           Mebibytes(free_bytes).c_str());
    }
  }
  g.bits = static_cast<uint64_t*>(g.memory.AllocateBytes(count * sizeof(uint64_t) + kRowTailBytes));
  g.deg = static_cast<int32_t*>(g.memory.AllocateBytes(static_cast<size_t>(n) * sizeof(int32_t)));
  RunDev* probe = static_cast<RunDev*>(g.memory.AllocateBytes(sizeof(RunDev)));
  CC_HIP_CHECK(hipMemcpyAsync(g.bits, bits, count * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
  CC_HIP_CHECK(hipMemsetAsync(g.bits + count, 0, kRowTailBytes, stream));
  RunDev r{};
  r.bits = g.bits;
  r.deg = g.deg;
  r.n = static_cast<int32_t>(n);
  r.words = static_cast<int32_t>(g.words);
  CC_HIP_CHECK(hipMemcpyAsync(probe, &r, sizeof(RunDev), hipMemcpyHostToDevice, stream));
  LaunchDegrees(probe, static_cast<int>(n), stream);
  CC_HIP_CHECK(hipGetLastError());
  CC_HIP_CHECK(hipStreamSynchronize(stream));
}

// reference CheckParams (src/cc_pbils_cpu.cpp:177-181) plus the device limits of SolveGpu (cuda/cc_pbils_gpu.cu:441-454)
void CheckParams(const cc_pbils_params& p) {
  if (p.k < 1) Fail(CC_HIP_EINVAL, "k must be >= 1");
  if (p.population < 1) Fail(CC_HIP_EINVAL, "population must be >= 1");
  if (p.gww < 0.0 || p.gww > 0.5) Fail(CC_HIP_EINVAL, "gww must be in [0, 0.5]");
  if (p.k > kMaxClusters) Fail(CC_HIP_EINVAL, "k exceeds kMaxClusters (%d)", kMaxClusters);
  if (p.ls_kernel < CC_LS_AUTO || p.ls_kernel > CC_LS_WIDE) Fail(CC_HIP_EINVAL, "unknown ls_kernel %d", p.ls_kernel);
  if (p.rebuild != CC_REBUILD_POPCOUNT && p.rebuild != CC_REBUILD_MFMA) {
    Fail(CC_HIP_EINVAL, "unknown rebuild %d", p.rebuild);
  }
}

bool WidePath(uint32_t n, const cc_pbils_params& p) {
  return n > static_cast<uint32_t>(kMaxVerticesNarrow) || p.ls_kernel == CC_LS_WIDE;
}

namespace {

// Threads of the generic kernels: a whole workgroup of 16 waves once n passes 512. With one individual per CU the
// waves are what hides the latency of the per-vertex count loads: at n = 2000 the generic kernels run about twice as
// fast with 1024 threads (two vertices each) as with 256 (eight each).
int GenericThreads(int n) {
  for (int t : {256, 512}) {
    if (t >= n) return t;
  }
  return 1024;
}

// Whether the k = 3 chunk-summary descent suits the graph: a move recomputes the 64-vertex chunks that hold the
// moved vertex's neighbours, about min(mean degree, chunks) of them, while the streaming kernel passes over every
// vertex word. Measured at n = 40000 (625 chunks, ER, pop 64, 2 iterations): chunk / streaming descent time 0.26 at
// mean degree 16, 0.38 at 64, 0.91 at 256, and with the chunk kernel 2.2 and 3.5 times the streaming kernel's at
// 1024 and 4096. The chunk kernel is taken while twice the mean degree is at most the chunk count; with no edge count
// (a plan from the shape alone) the graph is taken as sparse.
bool SparseForChunks(uint32_t n, uint64_t edges) {
  const double mean_degree = 2.0 * static_cast<double>(edges) / std::max<double>(n, 1);
  return 2.0 * mean_degree <= static_cast<double>(Words64(n));
}

}  // namespace

// The descent kernel and its launch shape (reference PlanDescent, cuda/cc_pbils_gpu.cu:69-102).
DescentPlan PlanDescent(uint32_t n_, int k, const cc_pbils_params& p, uint64_t edges) {
  const int n = static_cast<int>(n_);
  const int wide = WidePath(n_, p) ? 1 : 0;
  DescentPlan plan;
  // A tuned plan serves `auto`, and `wide` when it is a wide-path plan (the tuner's choice among the wide kernels).
  if ((p.ls_kernel == CC_LS_AUTO || (p.ls_kernel == CC_LS_WIDE && wide)) &&
      TunedPlan(n_, k, p.population, wide != 0, &plan)) {
    return plan;
  }
  plan.k = k;
  plan.wide = wide;
  plan.threads = GenericThreads(n);
  const size_t lds = GenericLdsBytes(n, k, wide);
  const bool fits = lds <= DescentLdsLimit();
  if (p.ls_kernel == CC_LS_SHARED && !fits) {
    Fail(CC_HIP_EINVAL, "state needs %zu bytes of LDS, the device allows %zu", lds, DescentLdsLimit());
  }
  // Wide path: the packed k = 2 / k = 3 kernels unless the generic kernel is asked for
  // (reference cuda/cc_pbils_gpu.cu:84-96); for k = 3 the chunk-summary variant while its table fits in LDS and the
  // graph is sparse enough that a move leaves most 64-vertex chunks untouched (SparseForChunks).
  if (wide && p.ls_kernel != CC_LS_GLOBAL && (k == 2 || k == 3)) {
    plan.descent = k == 2 ? Descent::kK2Wide : Descent::kK3Wide;
    plan.threads = n >= 8192 ? 1024 : 256;
    plan.lds_bytes = WideResidentLdsBytes(n, k, plan.threads);
    if (k == 3 && K3ChunkLdsBytes(n) <= DescentLdsLimit() && SparseForChunks(n_, edges)) {
      plan.descent = Descent::kK3Chunk;
      plan.lds_bytes = K3ChunkLdsBytes(n);
    }
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
    if (k == 3) {  // beyond the register kernel's reach: two registers per vertex
      plan.descent = Descent::kK3Packed;
      plan.threads = 1024;
      plan.vpt = K3PackedVpt(n);
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

double Since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

// ---- timing ------------------------------------------------------------------------------------------------------

Timer::~Timer() {
  for (auto& e : events_) (void)hipEventDestroy(e);
}

void Timer::Record(hipStream_t s) {
  if (used_ == events_.size()) {
    hipEvent_t e;
    CC_HIP_CHECK(hipEventCreate(&e));
    events_.push_back(e);
  }
  CC_HIP_CHECK(hipEventRecord(events_[used_++], s));
}

void Timer::End(hipStream_t s, double* total_ms, int64_t* launches) {
  if (!on_) return;
  Record(s);
  pending_.push_back({used_ - 2, total_ms, launches});
}

void Timer::Collect() {
  for (const auto& p : pending_) {
    float ms = 0.f;
    CC_HIP_CHECK(hipEventElapsedTime(&ms, events_[p.first], events_[p.first + 1]));
    *p.total += ms;
    *p.launches += 1;
  }
  pending_.clear();
  used_ = 0;
}

Stream::Stream() { CC_HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
Stream::~Stream() { (void)hipStreamDestroy(s_); }

// ---- run state and launch sets --------------------------------------------------------------------------------

namespace {

// Two passes over the same layout: the first sums the bytes, the second hands out 256-byte-aligned pointers of one
// allocation.
class Arena {
 public:
  template <typename T>
  void Take(T*& ptr, size_t count) {
    const size_t bytes = (count * sizeof(T) + 255) & ~static_cast<size_t>(255);
    ptr = (base_ && count) ? reinterpret_cast<T*>(base_ + used_) : nullptr;
    used_ += bytes;
  }
  size_t used() const { return used_; }
  void Start(uint8_t* base) {
    base_ = base;
    used_ = 0;
  }

 private:
  uint8_t* base_ = nullptr;
  size_t used_ = 0;
};

// Expected relative cost of one descent of a run, to order the persistent queues largest first.
double DescentCost(const RunSpec& s) {
  const double n = s.graph->n;
  return n * n * static_cast<double>(s.params.k);
}

// The buffers of one run (reference Run::ForEachBuffer, cuda/cc_pbils_gpu.cu:165-184).
void LayoutRun(const RunSpec& s, RunDev& r, Arena& a) {
  const size_t n = s.graph->n, k = static_cast<size_t>(s.params.k), P = static_cast<size_t>(s.population);
  const size_t iterations = static_cast<size_t>(std::max(s.params.iterations, 1));
  auto pop = [&](PopDev& d) {
    a.Take(d.labels, P * n);
    if (s.plan.wide) {
      int32_t* g = nullptr;
      a.Take(g, P * k * n);
      d.g = g;
    } else {
      int16_t* g = nullptr;
      a.Take(g, P * k * n);
      d.g = g;
    }
    a.Take(d.sizes, P * k);
    a.Take(d.f, P);
  };
  pop(r.pop);
  if (s.pbils) {
    pop(r.next);
  } else {
    r.next = r.pop;
  }
  a.Take(r.moves, P);
  a.Take(r.ctl, 1);
  a.Take(r.rec_labels, n);
  a.Take(r.trace, iterations);
  a.Take(r.src, P);
  a.Take(r.order, P);
  a.Take(r.masks, P * k * s.graph->words);
  a.Take(r.intra, P);
  size_t work = 0;  // in 64-bit words; the same for both k = 3 kernels, so a footprint needs no edge count
  if (s.plan.descent == Descent::kK3Wide || s.plan.descent == Descent::kK3Chunk) {
    work = P * static_cast<size_t>(WideStride(static_cast<int>(n), 3));  // the chunk kernel's u64 per vertex
  }
  if (s.plan.descent == Descent::kK2Wide) work = P * static_cast<size_t>(WideSlotBytes(static_cast<int>(n), 2)) / 8;
  a.Take(r.work, work);
}

}  // namespace

size_t RunState::Footprint(const RunSpec& spec) {
  Arena a;
  RunDev r{};
  LayoutRun(spec, r, a);
  return a.used();
}

RunState::RunState(const RunSpec& spec, hipStream_t stream, DevicePool* pool, ScratchArena* arena) : spec_(spec) {
  bytes_ = Footprint(spec_);
  if (arena) {
    memory_.UseArena(arena);
  } else if (pool) {
    memory_.UsePool(pool);
  } else {
    const size_t free_bytes = FreeDeviceBytes();
    if (bytes_ > free_bytes) {
      Fail(CC_HIP_ENOMEM, "run needs %s of device memory, %s free", Mebibytes(bytes_).c_str(),
           Mebibytes(free_bytes).c_str());
    }
  }
  uint8_t* base = static_cast<uint8_t*>(memory_.AllocateBytes(bytes_));
  // Start from zeros: e.g. a Neighborhood batch without local search sums a moves[] that no descent writes.
  CC_HIP_CHECK(hipMemsetAsync(base, 0, bytes_, stream));
  Arena layout;
  layout.Start(base);
  RunDev& r = dev_;
  LayoutRun(spec_, r, layout);
  const cc_pbils_params& p = spec_.params;
  r.bits = spec_.graph->bits;
  r.deg = spec_.graph->deg;
  r.n = static_cast<int32_t>(spec_.graph->n);
  r.words = static_cast<int32_t>(spec_.graph->words);
  r.k = p.k;
  r.population = spec_.population;
  r.edges = static_cast<int64_t>(spec_.graph->edges);
  r.init_stream = StreamSeed(p.seed, kInitStream, 0);
  r.select_stream = StreamSeed(p.seed, kSelectStream, 0);
  r.perturb_stream = StreamSeed(p.seed, kPerturbStream, 0);
  r.tournament = p.tournament;
  r.copies = spec_.pbils ? static_cast<int32_t>(p.gww * p.population) : 0;  // WinnerCopies, src/cc_pbils_cpu.cpp:168
  r.early_stop = p.early_stop;
  r.iterations = p.iterations;
  r.perturbation = static_cast<float>(p.perturbation);
  r.do_perturb = (p.k >= 2 && p.perturbation > 0.0) ? 1 : 0;
  r.select = spec_.pbils ? 1 : 0;
  r.wide = spec_.plan.wide;
  r.variant = spec_.variant;
}

LaunchSet::LaunchSet(const std::vector<RunState*>& states, hipStream_t stream, DevicePool* pool, PinnedBuffer* pinned,
                     ScratchArena* arena)
    : states_(states), stream_(stream) {
  const int nruns = static_cast<int>(states_.size());
  if (nruns == 0) Fail(CC_HIP_EINVAL, "empty batch");
  if (arena) {
    memory_.UseArena(arena);
  } else if (pool) {
    memory_.UsePool(pool);
  }
  host_runs_.resize(nruns);
  for (int i = 0; i < nruns; ++i) {
    host_runs_[i] = states_[i]->dev();
    max_n_ = std::max(max_n_, static_cast<int>(states_[i]->spec().graph->n));
  }
  auto spec = [&](int i) -> const RunSpec& { return states_[i]->spec(); };

  // item tables (host)
  std::vector<Item> all;
  std::map<int, std::vector<Item>> tiles;  // wide -> (run, first column) of the rebuild tiles
  std::vector<std::pair<DescentPlan, std::vector<Item>>> groups;
  for (int i = 0; i < nruns; ++i) {
    const RunSpec& s = spec(i);
    for (int slot = 0; slot < s.population; ++slot) all.push_back({i, slot});
    auto& tl = tiles[s.plan.wide];
    for (int q0 = 0; q0 < s.population * s.params.k; q0 += kRebuildCols) tl.push_back({i, q0});
    size_t g = 0;
    while (g < groups.size() && !SameLaunch(groups[g].first, s.plan)) ++g;
    if (g == groups.size()) groups.push_back({s.plan, {}});
    groups[g].first.lds_bytes = std::max(groups[g].first.lds_bytes, s.plan.lds_bytes);
    for (int slot = 0; slot < s.population; ++slot) groups[g].second.push_back({i, slot});
  }
  for (auto& g : groups) {
    std::stable_sort(g.second.begin(), g.second.end(),
                     [&](const Item& a, const Item& b) { return DescentCost(spec(a.run)) > DescentCost(spec(b.run)); });
  }

  // one allocation for the tables
  std::vector<Item*> tile_tables(tiles.size()), group_tables(groups.size());
  Item* all_table = nullptr;
  auto layout = [&](Arena& a) {
    a.Take(runs_, nruns);
    a.Take(all_table, all.size());
    size_t i = 0;
    for (auto& kv : tiles) a.Take(tile_tables[i++], kv.second.size());
    for (size_t g = 0; g < groups.size(); ++g) a.Take(group_tables[g], groups[g].second.size());
    a.Take(queues_, std::max<size_t>(groups.size(), 1));
    a.Take(ctl_table_, nruns);
  };
  Arena sizing;
  layout(sizing);
  Arena tables;
  tables.Start(static_cast<uint8_t*>(memory_.AllocateBytes(sizing.used())));
  layout(tables);

  CC_HIP_CHECK(hipMemcpyAsync(runs_, host_runs_.data(), sizeof(RunDev) * nruns, hipMemcpyHostToDevice, stream_));
  all_ = ItemList{all_table, static_cast<int32_t>(all.size()), nullptr};
  CC_HIP_CHECK(hipMemcpyAsync(all_table, all.data(), sizeof(Item) * all.size(), hipMemcpyHostToDevice, stream_));
  size_t ti = 0;
  for (auto& [wide, list] : tiles) {
    RebuildGroup g{};
    g.wide = wide;
    g.tiles = ItemList{tile_tables[ti], static_cast<int32_t>(list.size()), nullptr};
    CC_HIP_CHECK(hipMemcpyAsync(tile_tables[ti], list.data(), sizeof(Item) * list.size(), hipMemcpyHostToDevice,
                                stream_));
    rebuild_.push_back(g);
    ++ti;
  }
  int cus = 0;
  CC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, spec(0).graph->device));
  for (size_t g = 0; g < groups.size(); ++g) {
    DescentGroup d{groups[g].first, ItemList{group_tables[g], static_cast<int32_t>(groups[g].second.size()), queues_ + g}};
    // Persistent workgroups: enough to occupy every CU (a surplus finds the queue empty and exits at once), unless
    // the plan names a resident count (the autotuner may keep fewer individuals' state in cache at a time).
    const int per_cu = d.plan.threads >= 1024 ? 1 : (d.plan.threads >= 512 ? 2 : 4);
    d.plan.blocks = std::min(d.list.count, d.plan.blocks > 0 ? d.plan.blocks : cus * per_cu);
    if (!DescentSupported(d.plan)) Fail(CC_HIP_EINVAL, "no descent kernel for this plan");
    CC_HIP_CHECK(hipMemcpyAsync(group_tables[g], groups[g].second.data(), sizeof(Item) * groups[g].second.size(),
                                hipMemcpyHostToDevice, stream_));
    descent_.push_back(d);
  }
  if (arena) {
    pageable_ctl_.resize(nruns);
    host_ctl_ = pageable_ctl_.data();
  } else if (pinned) {
    host_ctl_ = static_cast<Ctl*>(pinned->Reserve(sizeof(Ctl) * nruns));
  } else {
    CC_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_ctl_), sizeof(Ctl) * nruns, hipHostMallocDefault));
    own_host_ctl_ = true;
  }
  CC_HIP_CHECK(hipStreamSynchronize(stream_));
}

LaunchSet::~LaunchSet() {
  if (host_ctl_ && own_host_ctl_) (void)hipHostFree(host_ctl_);
}

// The table layout of the constructor, counted from the specs alone: the run table, all items, the rebuild tiles of
// each G width, the descent items of each kernel instance, the queue counters and the control-block table.
size_t LaunchSet::TableBytes(const std::vector<RunSpec>& specs) {
  size_t all = 0;
  std::map<int, size_t> tiles;
  std::vector<std::pair<DescentPlan, size_t>> groups;
  for (const RunSpec& s : specs) {
    all += static_cast<size_t>(s.population);
    tiles[s.plan.wide] += (static_cast<size_t>(s.population) * s.params.k + kRebuildCols - 1) / kRebuildCols;
    size_t g = 0;
    while (g < groups.size() && !SameLaunch(groups[g].first, s.plan)) ++g;
    if (g == groups.size()) groups.push_back({s.plan, 0});
    groups[g].second += static_cast<size_t>(s.population);
  }
  Arena a;
  RunDev* runs = nullptr;
  Item* items = nullptr;
  int32_t* queues = nullptr;
  Ctl* ctl = nullptr;
  a.Take(runs, specs.size());
  a.Take(items, all);
  for (const auto& kv : tiles) a.Take(items, kv.second);
  for (const auto& g : groups) a.Take(items, g.second);
  a.Take(queues, std::max<size_t>(groups.size(), 1));
  a.Take(ctl, specs.size());
  return a.used();
}

void LaunchSet::Rebuild(Timer* timer, cc_pbils_stats* stats) {
  if (timer) timer->Begin(stream_);
  LaunchMasks(runs_, all_, stream_);
  for (const RebuildGroup& g : rebuild_) LaunchRebuildG(runs_, g, max_n_, stream_);
  LaunchFinalize(runs_, all_, stream_);
  if (timer) timer->End(stream_, &stats->rebuild_ms, &stats->rebuild_launches);
}

void LaunchSet::Descend(Timer* timer, cc_pbils_stats* stats) {
  LaunchResetQueues(queues_, static_cast<int>(descent_.size()), stream_);
  if (timer) timer->Begin(stream_);
  for (const DescentGroup& d : descent_) LaunchDescent(runs_, d.list, d.plan, stream_);
  if (timer) timer->End(stream_, &stats->search_ms, &stats->search_launches);
}

void LaunchSet::Start() {
  LaunchResetCtl(runs_, Size(), stream_);
  LaunchInitLabels(runs_, all_, max_n_, stream_);
  Rebuild(nullptr, nullptr);
  LaunchRecord(runs_, Size(), kRecordInitial, 0, stream_);
  CC_HIP_CHECK(hipGetLastError());
}

void LaunchSet::EnqueueIteration(Timer* timer, cc_pbils_stats* stats) {
  Descend(timer, stats);
  LaunchRecord(runs_, Size(), kRecordIteration, 1, stream_);
  LaunchPerturb(runs_, all_, max_n_, stream_);
  Rebuild(timer, stats);
}

void LaunchSet::SetLdsAttributes() {
  for (const DescentGroup& d : descent_) SetDescentAttributes(d.plan);
}

const Ctl* LaunchSet::PullCtl() {
  if (host_runs_.size() == 1) {
    CC_HIP_CHECK(hipMemcpyAsync(host_ctl_, host_runs_[0].ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
  } else {
    LaunchGatherCtl(runs_, Size(), ctl_table_, stream_);
    CC_HIP_CHECK(hipGetLastError());
    CC_HIP_CHECK(hipMemcpyAsync(host_ctl_, ctl_table_, sizeof(Ctl) * host_runs_.size(), hipMemcpyDeviceToHost,
                                stream_));
  }
  CC_HIP_CHECK(hipStreamSynchronize(stream_));
  return host_ctl_;
}

void LaunchSet::PullLabels(int run, int32_t* labels_out) {
  const RunDev& r = host_runs_[run];
  std::vector<uint8_t> labels(r.n);
  CC_HIP_CHECK(hipMemcpyAsync(labels.data(), r.rec_labels, labels.size(), hipMemcpyDeviceToHost, stream_));
  CC_HIP_CHECK(hipStreamSynchronize(stream_));
  for (int v = 0; v < r.n; ++v) labels_out[v] = labels[v];
}

void LaunchSet::PullTrace(int run, int count, int64_t* out) {
  if (count <= 0 || !out) return;
  CC_HIP_CHECK(hipMemcpyAsync(out, host_runs_[run].trace, sizeof(int64_t) * count, hipMemcpyDeviceToHost, stream_));
  CC_HIP_CHECK(hipStreamSynchronize(stream_));
}

// Neighborhood uses one run whose population is the batch of starts; the last batch may be smaller.
void LaunchSet::SetPopulation(int run, int population) {
  host_runs_[run].population = population;
  CC_HIP_CHECK(hipMemcpyAsync(runs_ + run, &host_runs_[run], sizeof(RunDev), hipMemcpyHostToDevice, stream_));
}

ItemList LaunchSet::Prefix(const ItemList& list, int count) const {
  ItemList l = list;
  l.count = std::min(l.count, count);
  return l;
}

DescentPlan LaunchSet::Plan(int run) const {
  for (const DescentGroup& d : descent_) {
    if (SameLaunch(d.plan, states_[run]->spec().plan)) return d.plan;
  }
  return states_[run]->spec().plan;
}

Batch::Batch(const std::vector<RunSpec>& specs, hipStream_t stream, ScratchArena* arena) : stream_(stream) {
  std::vector<RunState*> raw;
  for (const RunSpec& s : specs) {
    states_.push_back(std::make_unique<RunState>(s, stream, nullptr, arena));
    raw.push_back(states_.back().get());
    bytes_ += states_.back()->bytes();
  }
  set_ = std::make_unique<LaunchSet>(raw, stream, nullptr, nullptr, arena);
}

Batch::~Batch() {
  // The executable graph's kernel nodes point into the run states and launch tables that the members release after
  // this body; every replay on the stream ends before any of them goes, also on an error path that left work queued.
  (void)hipStreamSynchronize(stream_);
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
}

void Batch::Iterate(int count, bool use_graph, Timer* timer, cc_pbils_stats* stats) {
  if (!use_graph || timer) {
    for (int i = 0; i < count; ++i) set_->EnqueueIteration(timer, stats);
    CC_HIP_CHECK(hipGetLastError());
    return;
  }
  if (!graph_exec_) {
    set_->SetLdsAttributes();  // outside the capture
    CC_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    set_->EnqueueIteration(nullptr, nullptr);
    CC_HIP_CHECK(hipStreamEndCapture(stream_, &graph_));
    CC_HIP_CHECK(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  }
  for (int i = 0; i < count; ++i) CC_HIP_CHECK(hipGraphLaunch(graph_exec_, stream_));
}

// ---- solves ------------------------------------------------------------------------------------------------------

namespace {

int ClustersUsed(const int32_t* labels, uint32_t n, int k) {
  std::vector<char> seen(static_cast<size_t>(std::max(k, 1)), 0);
  for (uint32_t v = 0; v < n; ++v) {
    if (labels[v] >= 0 && labels[v] < k) seen[labels[v]] = 1;
  }
  int used = 0;
  for (char c : seen) used += c;
  return used;
}

const char* DescentName(Descent d) {
  switch (d) {
    case Descent::kGenericGlobal: return "generic, global memory";
    case Descent::kGenericLds: return "generic, LDS";
    case Descent::kK2Registers: return "k = 2, registers";
    case Descent::kK2Wide: return "k = 2, packed words";
    case Descent::kK3Wide: return "k = 3, packed words";
    case Descent::kRegisters: return "k = 3..32, registers";
    case Descent::kMixed: return "mixed register kernels";
    case Descent::kK3Chunk: return "k = 3, packed words with chunk summaries";
    case Descent::kK3Packed: return "k = 3, packed counts in registers";
  }
  return "?";
}

void PrintPlan(const DescentPlan& plan, size_t bytes) {
  std::printf("  local search: %s, %d threads per block, %d blocks", DescentName(plan.descent), plan.threads,
              plan.blocks);
  if (plan.vpt) std::printf(", %d vertices per thread", plan.vpt);
  if (plan.rows_in_lds) std::printf(", adjacency in LDS");
  if (plan.lds_bytes) std::printf(", %zu bytes of LDS", plan.lds_bytes);
  std::printf("\n  device memory: %s\n", Mebibytes(bytes).c_str());
}

void FillPlan(const DescentPlan& plan, cc_pbils_stats* stats) {
  stats->descent = static_cast<int32_t>(plan.descent);
  stats->threads = plan.threads;
  stats->blocks = plan.blocks;
  stats->vertices_per_thread = plan.vpt;
  stats->lds_bytes = static_cast<int64_t>(plan.lds_bytes);
}

void FillStats(const Ctl& c, const RunSpec& spec, const int32_t* labels, cc_pbils_stats* stats) {
  stats->objective = c.rec_f;
  stats->iterations_done = c.iterations_done;
  stats->local_searches = c.local_searches;
  stats->accepted_moves = c.accepted_moves;
  stats->clusters_used = ClustersUsed(labels, spec.graph->n, spec.params.k);
}

}  // namespace

// Advances every run of the batch until all have stopped; chunks of iterations between host synchronisations grow
// from 4 to 32 (1 with a time limit, which is checked between chunks).
void RunToCompletion(Batch& batch, const std::vector<RunSpec>& specs, bool verbose, bool use_graph, double time_limit,
                     Timer* timer, cc_pbils_stats* stats, Clock::time_point started) {
  int max_iterations = 0;
  for (const RunSpec& s : specs) max_iterations = std::max(max_iterations, s.params.iterations);
  int launched = 0;
  int chunk = time_limit > 0.0 ? 1 : 4;
  int printed = 0;
  int64_t last_record = -1;
  int stall = 0;
  if (verbose) last_record = batch.PullCtl()[0].rec_f;
  while (launched < max_iterations) {
    const int count = std::min(chunk, max_iterations - launched);
    batch.Iterate(count, use_graph, timer, stats);
    launched += count;
    const Ctl* ctl = batch.PullCtl();
    if (timer) timer->Collect();
    if (verbose && ctl[0].iterations_done > printed) {
      std::vector<int64_t> trace(ctl[0].iterations_done);
      batch.PullTrace(0, ctl[0].iterations_done, trace.data());
      for (int it = printed; it < ctl[0].iterations_done; ++it) {
        stall = trace[it] < last_record ? 0 : stall + 1;
        last_record = trace[it];
        std::printf("  iter %3d  record %lld  stall %d  %.3fs\n", it + 1, static_cast<long long>(trace[it]), stall,
                    Since(started));
      }
      printed = ctl[0].iterations_done;
    }
    bool all_stopped = true;
    for (size_t i = 0; i < specs.size(); ++i) all_stopped &= ctl[i].stopped != 0;
    if (all_stopped) break;
    if (time_limit > 0.0 && Since(started) >= time_limit) break;
    chunk = std::min(chunk * 2, 32);
  }
}

// PBILS on a device graph (reference gpu::Solve, cuda/cc_pbils_gpu.cu:347-387).
void SolvePbils(const DeviceGraph& g, const cc_pbils_params& p, int32_t* labels_out, cc_pbils_stats* stats,
                int64_t* trace_out, Clock::time_point started) {
  CheckParams(p);
  RunSpec spec{&g, p, PlanDescent(g.n, p.k, p, g.edges), p.population, true};
  Stream stream;
  Batch batch({spec}, stream.get());
  if (p.verbose) PrintPlan(batch.Plan(0), batch.Bytes());
  Timer timer(p.profile != 0);
  Timer span(p.profile != 0);
  int64_t spans = 0;
  batch.Start();
  span.Begin(stream.get());
  RunToCompletion(batch, {spec}, p.verbose != 0, true, p.time_limit_sec, p.profile ? &timer : nullptr, stats, started);
  span.End(stream.get(), &stats->iteration_ms, &spans);
  const Ctl c = batch.PullCtl()[0];
  span.Collect();
  if (c.error) Fail(CC_HIP_EDEVICE, "%s; the run was stopped", DeviceErrorText(c.error));
  batch.PullLabels(0, labels_out);
  batch.PullTrace(0, c.iterations_done, trace_out);
  FillStats(c, spec, labels_out, stats);
  FillPlan(batch.Plan(0), stats);
}

// ---- the solve that allocates nothing ------------------------------------------------------------------------

namespace {

// The static scratch region: every piece of device state of SolvePbilsStatic (populations, G, record, trace, GWW and
// rebuild scratch, launch tables, queue counters). Each GPU holds its own copy at its own address, from the load of
// this library's code object until the process ends. In host code `scratch` names a host-side stand-in, never the
// device memory: the device address of the current GPU's copy comes from hipGetSymbolAddress, once per device.
constexpr size_t kStaticScratchBytes = CC_HIP_STATIC_SCRATCH_BYTES;
__device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

std::mutex scratch_mutex;
std::map<int, void*> scratch_address;  // device -> address of its copy of `scratch`

// The current device's copy of the region (the caller has made `device` current).
void* ScratchAddress(int device) {
  std::lock_guard<std::mutex> lock(scratch_mutex);
  const auto found = scratch_address.find(device);
  if (found != scratch_address.end()) return found->second;
  void* p = nullptr;
  CC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) Fail(CC_HIP_EDEVICE, "no device address for the static scratch region on device %d", device);
  scratch_address[device] = p;
  return p;
}

}  // namespace

// SolvePbils without an allocation: the run's state and launch tables are laid out in the static scratch region
// (zeroed by the run state's own hipMemsetAsync), iterations are launched one by one on `stream` (instantiating a HIP
// graph would allocate), the control blocks come back into pageable host memory, and no events are created.
void SolvePbilsStatic(const DeviceGraph& g, const cc_pbils_params& p, hipStream_t stream, int32_t* labels_out,
                      cc_pbils_stats* stats, int64_t* trace_out, Clock::time_point started) {
  CheckParams(p);
  if (p.profile) Fail(CC_HIP_EINVAL, "the solve without allocation takes no event timings (params.profile)");
  const RunSpec spec{&g, p, PlanDescent(g.n, p.k, p, g.edges), p.population, true};
  const size_t need = RunState::Footprint(spec) + LaunchSet::TableBytes({spec});
  if (need > kStaticScratchBytes) {
    Fail(CC_HIP_ENOMEM, "the run needs %s of device state, more than the %s static scratch region", Mebibytes(need).c_str(),
         Mebibytes(kStaticScratchBytes).c_str());
  }
  ScratchArena arena(ScratchAddress(g.device), kStaticScratchBytes);
  Batch batch({spec}, stream, &arena);
  if (p.verbose) PrintPlan(batch.Plan(0), batch.Bytes());
  batch.Start();
  RunToCompletion(batch, {spec}, p.verbose != 0, false, p.time_limit_sec, nullptr, stats, started);
  const Ctl c = batch.PullCtl()[0];
  if (c.error) Fail(CC_HIP_EDEVICE, "%s; the run was stopped", DeviceErrorText(c.error));
  batch.PullLabels(0, labels_out);
  batch.PullTrace(0, c.iterations_done, trace_out);
  FillStats(c, spec, labels_out, stats);
  FillPlan(batch.Plan(0), stats);
}

// Neighborhood for k = 2 in batches of at most `population` start vertices (reference SolveNeighborhood,
// cuda/cc_pbils_gpu.cu:389-436). Starts are processed in vertex order and the record only moves on a strictly
// smaller f, so ties keep the smaller start vertex across batches.
void SolveNeighborhood(const DeviceGraph& g, const cc_pbils_params& params, bool local_search, int32_t* labels_out,
                       cc_pbils_stats* stats, Clock::time_point started) {
  if (params.k != 2) Fail(CC_HIP_EINVAL, "Neighborhood algorithms require k = 2");
  if (params.population < 1) Fail(CC_HIP_EINVAL, "population must be >= 1");
  if (params.ls_kernel < CC_LS_AUTO || params.ls_kernel > CC_LS_WIDE) {
    Fail(CC_HIP_EINVAL, "unknown ls_kernel %d", params.ls_kernel);
  }
  if (params.rebuild != CC_REBUILD_POPCOUNT && params.rebuild != CC_REBUILD_MFMA) {
    Fail(CC_HIP_EINVAL, "unknown rebuild %d", params.rebuild);
  }
  cc_pbils_params p = params;
  const int n = static_cast<int>(g.n);
  p.population = std::min(params.population, n);
  p.gww = 0.0;
  p.iterations = 1;
  p.early_stop = 1;
  RunSpec spec{&g, p, PlanDescent(g.n, 2, p, g.edges), p.population, false};
  Stream stream;
  const hipStream_t s = stream.get();
  Batch batch({spec}, s);
  if (p.verbose) PrintPlan(batch.Plan(0), batch.Bytes());
  Timer timer(p.profile != 0);
  Timer* t = p.profile ? &timer : nullptr;
  LaunchResetCtl(batch.Runs(), 1, s);
  const int capacity = p.population;
  // The matrix-core rebuild: the adjacency expanded once to bytes, a start matrix per batch.
  const bool mfma = p.rebuild == CC_REBUILD_MFMA;
  DeviceMemory mfma_memory;
  int8_t* adj8 = nullptr;
  int8_t* z = nullptr;
  const int n_pad = static_cast<int>(MfmaPad(n));
  if (mfma) {
    if (spec.plan.wide) Fail(CC_HIP_EINVAL, "the matrix-core rebuild needs n <= %d", kMaxVerticesNarrow);
    const size_t bytes = static_cast<size_t>(n_pad) * n_pad + MfmaPad(capacity) * n_pad;
    if (bytes > FreeDeviceBytes()) Fail(CC_HIP_ENOMEM, "the matrix-core rebuild needs %s", Mebibytes(bytes).c_str());
    adj8 = static_cast<int8_t*>(mfma_memory.AllocateBytes(static_cast<size_t>(n_pad) * n_pad));
    z = static_cast<int8_t*>(mfma_memory.AllocateBytes(MfmaPad(capacity) * n_pad));
    if (t) t->Begin(s);
    LaunchExpandAdjacency(g.bits, n, static_cast<int>(g.words), n_pad, adj8, s);
    if (t) t->End(s, &stats->rebuild_ms, &stats->rebuild_launches);
  }
  for (int first = 0; first < n; first += capacity) {
    const int count = std::min(capacity, n - first);
    if (count != capacity) batch.SetPopulation(0, count);
    LaunchNeighborhoodLabels(batch.Runs(), n, first, count, s);
    if (t) t->Begin(s);
    LaunchMasks(batch.Runs(), batch.Prefix(batch.AllItems(), count), s);
    if (mfma) {
      LaunchRebuildMfmaK2(batch.Runs(), count, n_pad, adj8, z, s);
    } else {
      for (const RebuildGroup& rg : batch.RebuildGroups()) {
        RebuildGroup prefix = rg;
        prefix.tiles = batch.Prefix(rg.tiles, (count * params.k + kRebuildCols - 1) / kRebuildCols);
        LaunchRebuildG(batch.Runs(), prefix, n, s);
      }
    }
    LaunchFinalize(batch.Runs(), batch.Prefix(batch.AllItems(), count), s);
    if (t) t->End(s, &stats->rebuild_ms, &stats->rebuild_launches);
    if (local_search) {
      const DescentGroup& d = batch.DescentGroups()[0];
      DescentPlan plan = d.plan;
      plan.blocks = std::min(plan.blocks, count);
      LaunchResetQueues(d.list.queue, 1, s);
      if (t) t->Begin(s);
      LaunchDescent(batch.Runs(), batch.Prefix(d.list, count), plan, s);
      if (t) t->End(s, &stats->search_ms, &stats->search_launches);
    }
    LaunchRecord(batch.Runs(), 1, kRecordBatch, 1, s);
    CC_HIP_CHECK(hipGetLastError());
    if (p.verbose) {
      std::printf("  neighborhoods %d/%d  record %lld  %.3fs\n", first + count, n,
                  static_cast<long long>(batch.PullCtl()[0].rec_f), Since(started));
    }
  }
  const Ctl c = batch.PullCtl()[0];
  if (t) t->Collect();
  if (c.error) Fail(CC_HIP_EDEVICE, "%s; the run was stopped", DeviceErrorText(c.error));
  batch.PullLabels(0, labels_out);
  FillStats(c, spec, labels_out, stats);
  stats->iterations_done = 0;
  stats->local_searches = local_search ? n : 0;
  FillPlan(batch.Plan(0), stats);
}

void UseMixedDescent(RunSpec& spec) {
  const bool registers = spec.plan.descent == Descent::kK2Registers || spec.plan.descent == Descent::kRegisters;
  if (!registers || spec.params.ls_kernel != CC_LS_AUTO) return;
  DescentPlan variant_plan;
  const int variant = MixedVariant(static_cast<int>(spec.graph->n), spec.params.k, &variant_plan);
  if (variant < 0) return;
  spec.variant = variant;
  spec.plan = DescentPlan{};
  spec.plan.descent = Descent::kMixed;
  spec.plan.threads = variant_plan.threads;
}

// Many independent PBILS runs advanced together: one launch per phase for all runs, descents grouped by kernel
// instance and claimed from persistent work queues, iterations replayed from one HIP graph.
void SolveBatch(std::vector<BatchRun>& runs, bool use_graph) {
  std::vector<RunSpec> specs;
  for (BatchRun& b : runs) {
    CheckParams(b.params);
    RunSpec spec{b.graph, b.params, PlanDescent(b.graph->n, b.params.k, b.params, b.graph->edges),
                 b.params.population, true};
    if (runs.size() > 1) UseMixedDescent(spec);
    specs.push_back(spec);
  }
  Stream stream;
  Batch batch(specs, stream.get());
  batch.Start();
  cc_pbils_stats unused{};
  RunToCompletion(batch, specs, false, use_graph, 0.0, nullptr, &unused, Clock::now());
  const Ctl* ctl = batch.PullCtl();
  std::vector<Ctl> copy(ctl, ctl + specs.size());
  for (size_t i = 0; i < runs.size(); ++i) {
    if (copy[i].error) Fail(CC_HIP_EDEVICE, "run %zu: %s", i, DeviceErrorText(copy[i].error));
    batch.PullLabels(static_cast<int>(i), runs[i].labels);
    batch.PullTrace(static_cast<int>(i), copy[i].iterations_done, runs[i].trace);
    FillStats(copy[i], specs[i], runs[i].labels, runs[i].stats);
    FillPlan(batch.Plan(static_cast<int>(i)), runs[i].stats);
  }
}

int Report(const std::exception& e, char* err, size_t err_len, int code) {
  if (err && err_len > 0) {
    std::strncpy(err, e.what(), err_len - 1);
    err[err_len - 1] = 0;
  }
  return code;
}

}  // namespace cc::hip

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
    auto h = std::make_unique<cc_hip_graph>();
    cc::hip::Stream stream;
    cc::hip::Upload(h->g, bits, n, edges, stream.get());
    *out = h.release();
  });
}

void cc_hip_graph_free(cc_hip_graph* graph) { delete graph; }

int cc_hip_graph_solve_pbils(cc_hip_graph* graph, const cc_pbils_params* params, int32_t* labels_out,
                             cc_pbils_stats* stats, int64_t* trace_out, char* err, size_t err_len) {
  std::memset(stats, 0, sizeof(*stats));
  return cc::hip::Guard(err, err_len, [&] {
    const auto started = cc::hip::Clock::now();
    CC_HIP_CHECK(hipSetDevice(graph->g.device));
    cc::hip::SolvePbils(graph->g, *params, labels_out, stats, trace_out, started);
    stats->seconds = cc::hip::Since(started);
  });
}

int cc_hip_graph_solve_pbils_static(cc_hip_graph* graph, const cc_pbils_params* params, void* stream,
                                    int32_t* labels_out, cc_pbils_stats* stats, int64_t* trace_out, char* err,
                                    size_t err_len) {
  std::memset(stats, 0, sizeof(*stats));
  return cc::hip::Guard(err, err_len, [&] {
    const auto started = cc::hip::Clock::now();
    CC_HIP_CHECK(hipSetDevice(graph->g.device));
    cc::hip::SolvePbilsStatic(graph->g, *params, static_cast<hipStream_t>(stream), labels_out, stats, trace_out,
                              started);
    stats->seconds = cc::hip::Since(started);
  });
}

int cc_hip_graph_solve_neighborhood(cc_hip_graph* graph, const cc_pbils_params* params, int32_t local_search,
                                    int32_t* labels_out, cc_pbils_stats* stats, char* err, size_t err_len) {
  std::memset(stats, 0, sizeof(*stats));
  return cc::hip::Guard(err, err_len, [&] {
    const auto started = cc::hip::Clock::now();
    CC_HIP_CHECK(hipSetDevice(graph->g.device));
    cc::hip::SolveNeighborhood(graph->g, *params, local_search != 0, labels_out, stats, started);
    stats->seconds = cc::hip::Since(started);
  });
}

int cc_hip_solve_pbils(const uint64_t* bits, uint32_t n, uint64_t edges, const cc_pbils_params* params,
                       int32_t* labels_out, cc_pbils_stats* stats, int64_t* trace_out, char* err, size_t err_len) {
  std::memset(stats, 0, sizeof(*stats));
  return cc::hip::Guard(err, err_len, [&] {
    const auto started = cc::hip::Clock::now();  // the solver clock includes allocation and upload
    cc::hip::CheckParams(*params);
    cc::hip::DeviceGraph g;
    {
      cc::hip::Stream stream;
      cc::hip::Upload(g, bits, n, edges, stream.get());
    }
    cc::hip::SolvePbils(g, *params, labels_out, stats, trace_out, started);
    stats->seconds = cc::hip::Since(started);
  });
}

int cc_hip_solve_neighborhood(const uint64_t* bits, uint32_t n, uint64_t edges, const cc_pbils_params* params,
                              int32_t local_search, int32_t* labels_out, cc_pbils_stats* stats, char* err,
                              size_t err_len) {
  std::memset(stats, 0, sizeof(*stats));
  return cc::hip::Guard(err, err_len, [&] {
    const auto started = cc::hip::Clock::now();
    if (params->k != 2) cc::hip::Fail(CC_HIP_EINVAL, "Neighborhood algorithms require k = 2");
    if (params->population < 1) cc::hip::Fail(CC_HIP_EINVAL, "population must be >= 1");
    if (n == 0 || n > static_cast<uint32_t>(cc::hip::kMaxVerticesWide)) {
      cc::hip::Fail(CC_HIP_EINVAL, "Neighborhood requires 1 <= n <= %d", cc::hip::kMaxVerticesWide);
    }
    cc::hip::DeviceGraph g;
    {
      cc::hip::Stream stream;
      cc::hip::Upload(g, bits, n, edges, stream.get());
    }
    cc::hip::SolveNeighborhood(g, *params, local_search != 0, labels_out, stats, started);
    stats->seconds = cc::hip::Since(started);
  });
}

int cc_hip_solve_batch(cc_batch_run* runs, int32_t count, int32_t use_graph, double* seconds, char* err,
                       size_t err_len) {
  return cc::hip::Guard(err, err_len, [&] {
    const auto started = cc::hip::Clock::now();
    if (count < 1) cc::hip::Fail(CC_HIP_EINVAL, "empty batch");
    std::vector<cc::hip::BatchRun> batch(count);
    for (int i = 0; i < count; ++i) {
      if (!runs[i].graph || !runs[i].labels_out) cc::hip::Fail(CC_HIP_EINVAL, "run %d: graph and labels required", i);
      std::memset(&runs[i].stats, 0, sizeof(runs[i].stats));
      batch[i] = {&runs[i].graph->g, runs[i].params, runs[i].labels_out, runs[i].trace_out, &runs[i].stats};
    }
    CC_HIP_CHECK(hipSetDevice(runs[0].graph->g.device));
    cc::hip::SolveBatch(batch, use_graph != 0);
    const double s = cc::hip::Since(started);
    for (int i = 0; i < count; ++i) runs[i].stats.seconds = s;
    if (seconds) *seconds = s;
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
