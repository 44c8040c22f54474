// cc_tune.hip — the descent autotuner: for a graph shape (n, k, population, G width) it times every descent plan
// the kernels support (kernel kind, threads per workgroup, vertices per thread, adjacency or state in LDS or global
// memory, persistent workgroup count) on a short PBILS run of the caller's own graph, checks that all of them follow
// the same trajectory, and records the fastest in a process-wide table that PlanDescent consults for
// ls_kernel = auto. The table can be saved to and loaded from a text file. The reference fixes these choices at
// compile time for one GPU (cuda/cc_gpu_tuning.cuh:13-77, cuda/cc_pbils_gpu.cu:61-102).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cc_hip.h"
#include "cc_internal.hpp"
#include "cc_runtime.hpp"

namespace cc::hip {
namespace {

using TuneKey = std::tuple<uint32_t, int, int, int>;  // n, k, population, wide

std::mutex& TableMutex() {
  static std::mutex m;
  return m;
}

std::map<TuneKey, DescentPlan>& Table() {
  static std::map<TuneKey, DescentPlan> t;
  return t;
}

int DeviceCus(int device) {
  int cus = 0;
  CC_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  return cus;
}

void AddIfSupported(std::vector<DescentPlan>& out, const DescentPlan& plan) {
  if (DescentSupported(plan)) out.push_back(plan);
}

// Every plan the kernels can run for this shape; `blocks` 0 means one resident workgroup set per CU.
std::vector<DescentPlan> Candidates(uint32_t n_, int k, int population, bool wide, int cus) {
  const int n = static_cast<int>(n_);
  std::vector<DescentPlan> out;
  DescentPlan base;
  base.k = k;
  base.wide = wide ? 1 : 0;
  if (!wide && k == 2) {
    const size_t rows = static_cast<size_t>(n) * Words64(n_) * sizeof(uint64_t);
    for (int t : {64, 128, 256, 512, 1024}) {
      for (int v : {1, 2, 4, 8, 16, 32}) {
        if (v * t < n) continue;
        for (bool lds : {false, true}) {
          if (lds && rows > DescentLdsLimit()) continue;
          DescentPlan p = base;
          p.descent = Descent::kK2Registers;
          p.threads = t;
          p.vpt = v;
          p.rows_in_lds = lds;
          p.lds_bytes = lds ? rows : 0;
          AddIfSupported(out, p);
        }
        break;  // the smallest vertices-per-thread that covers n
      }
    }
  }
  if (!wide && k >= 3 && k <= 32) {
    for (int t : {256, 512, 1024}) {
      for (int v : {1, 2, 4, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32}) {
        if (v * t < n) continue;
        DescentPlan p = base;
        p.descent = Descent::kRegisters;
        p.threads = t;
        p.vpt = v;
        if (DescentSupported(p)) {
          out.push_back(p);
          break;
        }
      }
    }
  }
  if (!wide && k == 3 && n > 8192) {
    DescentPlan p = base;
    p.descent = Descent::kK3Packed;
    p.threads = 1024;
    p.vpt = K3PackedVpt(n);
    if (p.vpt > 0 && DescentSupported(p)) out.push_back(p);
  }
  if (wide && (k == 2 || k == 3)) {
    std::vector<int> block_counts = {0};  // 0: one workgroup per item
    for (int b : {population / 2, population / 4}) {
      if (b >= 8 && b < std::min(population, cus)) block_counts.push_back(b);
    }
    for (int t : {256, 1024}) {
      for (int b : block_counts) {
        DescentPlan p = base;
        p.descent = k == 2 ? Descent::kK2Wide : Descent::kK3Wide;
        p.threads = t;
        p.blocks = b;
        AddIfSupported(out, p);  // every tile streamed
        p.lds_bytes = WideResidentLdsBytes(n, k, t);
        AddIfSupported(out, p);  // the first tiles in LDS
      }
      if (k == 3 && K3ChunkLdsBytes(n) <= DescentLdsLimit()) {
        DescentPlan p = base;
        p.descent = Descent::kK3Chunk;
        p.threads = t;
        p.lds_bytes = K3ChunkLdsBytes(n);
        AddIfSupported(out, p);
      }
    }
  }
  const size_t lds = GenericLdsBytes(n, k, wide ? 1 : 0);
  for (int t : {256, 512, 1024}) {
    DescentPlan p = base;
    p.threads = t;
    if (lds <= DescentLdsLimit()) {
      p.descent = Descent::kGenericLds;
      p.lds_bytes = lds;
      AddIfSupported(out, p);
    }
    p.descent = Descent::kGenericGlobal;
    p.lds_bytes = 0;
    AddIfSupported(out, p);
  }
  return out;
}

}  // namespace

bool TunedPlan(uint32_t n, int k, int population, bool wide, DescentPlan* plan) {
  std::lock_guard<std::mutex> lock(TableMutex());
  const auto it = Table().find(TuneKey{n, k, population, wide ? 1 : 0});
  if (it == Table().end()) return false;
  *plan = it->second;
  return true;
}

// Times each candidate on `iterations` PBILS iterations of the caller's graph and parameters (device time of the
// descent launches, HIP events); every candidate must reproduce the first one's record, accepted moves and labels.
void Autotune(const DeviceGraph& g, const cc_pbils_params& params, int iterations, std::vector<cc_tune_candidate>& report,
              int* chosen) {
  CheckParams(params);
  if (iterations < 1) Fail(CC_HIP_EINVAL, "autotune needs at least one iteration");
  cc_pbils_params p = params;
  p.iterations = iterations;
  p.early_stop = iterations;
  p.time_limit_sec = 0.0;
  p.verbose = 0;
  const bool wide = WidePath(g.n, p);
  const std::vector<DescentPlan> plans = Candidates(g.n, p.k, p.population, wide, DeviceCus(g.device));
  if (plans.empty()) Fail(CC_HIP_EINVAL, "no descent plan supports n = %u, k = %d", g.n, p.k);
  report.clear();
  std::vector<int32_t> labels(g.n), first_labels;
  int64_t first_f = 0, first_moves = 0;
  int best = -1;
  for (size_t i = 0; i < plans.size(); ++i) {
    RunSpec spec{&g, p, plans[i], p.population, true};
    Stream stream;
    Batch batch({spec}, stream.get());
    Timer timer(true);
    cc_pbils_stats stats{};
    batch.Start();
    RunToCompletion(batch, {spec}, false, false, 0.0, &timer, &stats, Clock::now());
    const Ctl c = batch.PullCtl()[0];
    timer.Collect();
    if (c.error) Fail(CC_HIP_EDEVICE, "autotune: a descent exceeded its move cap");
    batch.PullLabels(0, labels.data());
    if (i == 0) {
      first_f = c.rec_f;
      first_moves = c.accepted_moves;
      first_labels = labels;
    } else if (c.rec_f != first_f || c.accepted_moves != first_moves || labels != first_labels) {
      Fail(CC_HIP_EDEVICE, "autotune: descent plan %zu left the trajectory of plan 0", i);
    }
    cc_tune_candidate t{};
    const DescentPlan applied = batch.Plan(0);
    t.descent = static_cast<int32_t>(applied.descent);
    t.threads = applied.threads;
    t.vertices_per_thread = applied.vpt;
    t.rows_in_lds = applied.rows_in_lds ? 1 : 0;
    t.blocks = applied.blocks;
    t.lds_bytes = static_cast<int64_t>(applied.lds_bytes);
    t.search_ms = stats.search_ms;
    t.accepted_moves = c.accepted_moves;
    report.push_back(t);
    if (best < 0 || t.search_ms < report[best].search_ms) best = static_cast<int>(i);
  }
  {
    std::lock_guard<std::mutex> lock(TableMutex());
    Table()[TuneKey{g.n, p.k, p.population, wide ? 1 : 0}] = plans[best];
  }
  *chosen = best;
}

void ClearTuned() {
  std::lock_guard<std::mutex> lock(TableMutex());
  Table().clear();
}

// One line per entry: n k population wide descent threads vpt rows_in_lds lds_bytes blocks
void SaveTuned(const char* path) {
  std::lock_guard<std::mutex> lock(TableMutex());
  FILE* f = std::fopen(path, "w");
  if (!f) Fail(CC_HIP_EINVAL, "cannot write %s", path);
  for (const auto& [key, plan] : Table()) {
    std::fprintf(f, "%u %d %d %d %d %d %d %d %zu %d\n", std::get<0>(key), std::get<1>(key), std::get<2>(key),
                 std::get<3>(key), static_cast<int>(plan.descent), plan.threads, plan.vpt, plan.rows_in_lds ? 1 : 0,
                 plan.lds_bytes, plan.blocks);
  }
  std::fclose(f);
}

int LoadTuned(const char* path) {
  FILE* f = std::fopen(path, "r");
  if (!f) Fail(CC_HIP_EINVAL, "cannot read %s", path);
  std::vector<std::pair<TuneKey, DescentPlan>> entries;
  unsigned n = 0;
  int k = 0, population = 0, wide = 0, descent = 0, threads = 0, vpt = 0, rows = 0, blocks = 0;
  size_t lds = 0;
  int line = 0;
  while (std::fscanf(f, "%u %d %d %d %d %d %d %d %zu %d", &n, &k, &population, &wide, &descent, &threads, &vpt,
                     &rows, &lds, &blocks) == 10) {
    ++line;
    DescentPlan p;
    p.descent = static_cast<Descent>(descent);
    p.k = k;
    p.threads = threads;
    p.vpt = vpt;
    p.rows_in_lds = rows != 0;
    p.lds_bytes = lds;
    p.wide = wide;
    p.blocks = blocks;
    // any single-run kernel (the mixed batch kernel is chosen per batch, not per shape)
    const bool valid = descent >= 0 && descent <= static_cast<int>(Descent::kK3Packed) &&
                       p.descent != Descent::kMixed && lds <= DescentLdsLimit() && DescentSupported(p) &&
                       (p.descent != Descent::kK3Chunk || lds >= K3ChunkLdsBytes(static_cast<int>(n)));
    if (!valid) {
      std::fclose(f);
      Fail(CC_HIP_EINVAL, "%s: entry %d names no supported descent plan", path, line);
    }
    entries.push_back({TuneKey{n, k, population, wide}, p});
  }
  const bool clean_end = std::feof(f) != 0;
  std::fclose(f);
  if (!clean_end) Fail(CC_HIP_EINVAL, "%s: malformed entry after %d entries", path, line);
  std::lock_guard<std::mutex> lock(TableMutex());
  for (const auto& [key, plan] : entries) Table()[key] = plan;
  return static_cast<int>(entries.size());
}

}  // namespace cc::hip

extern "C" {

int cc_hip_autotune(cc_hip_graph* graph, const cc_pbils_params* params, int32_t iterations, cc_tune_candidate* out,
                    int32_t capacity, int32_t* count, int32_t* chosen, char* err, size_t err_len) {
  if (count) *count = 0;
  if (chosen) *chosen = -1;
  return cc::hip::Guard(err, err_len, [&] {
    if (!graph || !params) cc::hip::Fail(CC_HIP_EINVAL, "graph and params required");
    CC_HIP_CHECK(hipSetDevice(graph->g.device));
    std::vector<cc_tune_candidate> report;
    int best = -1;
    cc::hip::Autotune(graph->g, *params, iterations, report, &best);
    const int n = static_cast<int>(report.size());
    for (int i = 0; i < std::min(n, capacity); ++i) out[i] = report[i];
    if (count) *count = n;
    if (chosen) *chosen = best;
  });
}

void cc_hip_tune_clear(void) { cc::hip::ClearTuned(); }

int cc_hip_tune_save(const char* path, char* err, size_t err_len) {
  return cc::hip::Guard(err, err_len, [&] { cc::hip::SaveTuned(path); });
}

int cc_hip_tune_load(const char* path, int32_t* entries, char* err, size_t err_len) {
  return cc::hip::Guard(err, err_len, [&] {
    const int loaded = cc::hip::LoadTuned(path);
    if (entries) *entries = loaded;
  });
}

}  // extern "C"
