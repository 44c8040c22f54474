// cc_pbils.hpp — PBILS parameters and result, and the host helpers shared by the GPU front end: the record
// bookkeeping rules, GWW ranking and the independent objective recounts that verify a result.
//
// Same names, fields and defaults as the reference (include/cc_pbils.hpp:26-137 of the reference), so that the
// reference's own callers (its command line, include/cc_cli.hpp, and cuda/main_gpu.cu) build against this header
// unchanged. The CPU solver of the reference (State, LocalSearch, SolveCpu) has no counterpart in this library: the
// CPU path lives in the test oracle only.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "cc_graph.hpp"

namespace cc {

struct PbilsParams {
  // Which descent kernels run the local search; all follow one trajectory (include/cc_pbils.hpp:30-35).
  enum LocalSearchKernel {
    kLocalSearchAuto = 0,
    kLocalSearchGlobal = 1,
    kLocalSearchShared = 2,
    kLocalSearchWide = 3,
  };

  int k = 2;
  int population = 128;
  int tournament = 5;
  int iterations = 100;
  int early_stop = 6;
  double perturbation = 0.4;
  double gww = 0.1;
  int threads = 1;  // CPU only in the reference; unused here
  uint64_t seed = 1;
  double time_limit_sec = 0.0;
  bool verbose = false;
  int ls_kernel = kLocalSearchAuto;
  bool profile = false;  // time the descent and rebuild kernels with HIP events
};

struct PbilsResult {
  long long objective = -1;
  std::vector<int> labels;
  int iterations_done = 0;
  int clusters_used = 0;
  double seconds = 0.0;
  long long local_searches = 0;
  long long accepted_moves = 0;
  // device-side detail (not in the reference)
  int descent = 0;
  int descent_threads = 0;
  double search_ms = 0.0;
  double rebuild_ms = 0.0;
};

// The solver clock, the stall count and the stop rule of a PBILS run, as both of the reference's backends keep them
// (include/cc_pbils.hpp:62-85 of the reference). The GPU runtime applies the same rule on the device (the record
// kernel); this class is the host-side form for callers that drive iterations themselves, and the clock starts at
// construction.
class Progress {
 public:
  explicit Progress(const PbilsParams& params) : params_(params), start_(std::chrono::steady_clock::now()) {}

  double Seconds() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - start_).count();
  }

  // Books one finished iteration (improved: it produced a new record). Returns true when the run must stop: early_stop
  // iterations in a row without a record, or the time limit (when positive) reached.
  bool Next(PbilsResult& result, bool improved) {
    result.iterations_done += 1;
    result.local_searches += params_.population;
    stall_ = improved ? 0 : stall_ + 1;
    if (params_.verbose) {
      std::printf("  iter %3d  record %lld  stall %d  %.3fs\n", result.iterations_done, result.objective, stall_,
                  Seconds());
    }
    const bool stalled = stall_ >= params_.early_stop;
    const bool out_of_time = params_.time_limit_sec > 0.0 && Seconds() >= params_.time_limit_sec;
    return stalled || out_of_time;
  }

 private:
  const PbilsParams& params_;
  std::chrono::steady_clock::time_point start_;
  int stall_ = 0;
};

// f by the definition, pairwise, O(n^2) (reference src/cc_pbils_cpu.cpp:12-23).
long long ObjectiveDirect(const Graph& graph, const std::vector<int>& labels);
// f over bit rows: popcount(row XOR own-cluster mask) - 1 per vertex, halved; O(n^2/64) on all cores
// (reference src/cc_pbils_cpu.cpp:25-53).
long long ObjectiveBitwise(const Graph& graph, const std::vector<int>& labels);
int CountClustersUsed(const std::vector<int>& labels, int k);

// Worker threads of a CPU solve: params.threads, or every hardware thread when it is <= 0, and never more than one per
// individual (reference include/cc_pbils.hpp:119, src/cc_pbils_cpu.cpp:163-166). The GPU path does not use it; the
// reference's command line prints it for its CPU backend.
int ResolveThreads(const PbilsParams& params);

constexpr double kMaxGwwShare = 0.5;
// The worst floor(gww * population) slots receive copies of the best (reference src/cc_pbils_cpu.cpp:168).
int WinnerCopies(double gww, int population);
// Slots by f, equal f by slot (reference src/cc_pbils_cpu.cpp:170-175).
std::vector<int> RankSlots(const long long* f, int population);
// Throws std::runtime_error like the reference CheckParams (src/cc_pbils_cpu.cpp:177-181).
void CheckParams(const PbilsParams& params);

}  // namespace cc
