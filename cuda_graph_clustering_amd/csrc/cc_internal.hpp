// cc_internal.hpp — definitions shared by the kernels (cc_kernels.hip) and the host runtime (cc_runtime.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cc_math.hpp"

namespace cc::hip {

constexpr int kWave = 64;
constexpr int kMaxClusters = 64;
constexpr int kMaxVerticesNarrow = 32767;
constexpr int kMaxVerticesWide = 1 << 20;

// Run state that the iteration kernels read and advance on the device, so that no host round trip is needed per
// iteration: the record, the early-stop counter and the iteration number (reference: PullRecord, Progress::Next
// and RankSlots run on the host, cuda/cc_pbils_gpu.cu:324-345 and include/cc_pbils.hpp:69-78).
struct Ctl {
  int32_t iteration;          // iteration index used by the next descent (tournament step)
  int32_t perturb_iteration;  // iteration index of the pending perturbation
  int32_t stall;              // iterations without a new record
  int32_t iterations_done;
  int32_t stopped;            // set once the run is over: every later kernel of the run returns at once
  int32_t have_record;
  int32_t rec_slot;
  int32_t error;              // a descent exceeded MoveCap(): the run is stopped and reported as failed
  int64_t rec_f;
  int64_t local_searches;
  int64_t accepted_moves;
};

// One population on the device. G is [slot][cluster][vertex] of GT (int16 narrow, int32 wide).
struct PopDev {
  uint8_t* labels;  // [P][n]
  void* g;          // [P][k][n]
  int32_t* sizes;   // [P][k]
  int64_t* f;       // [P]
};

// Everything the kernels of one run see; passed by value.
struct RunDev {
  const uint64_t* bits;  // [n][words]
  int32_t* deg;          // [n], written once by KernelDegrees
  int32_t n, words, k, population;
  int64_t edges;
  PopDev pop;   // the population at iteration boundaries (and the target of perturb + rebuild)
  PopDev next;  // written by the descent, read by the record, GWW and perturb
  int64_t* moves;        // [P] accepted moves of the last descent per slot
  Ctl* ctl;
  uint8_t* rec_labels;   // [n]
  int64_t* trace;        // [iterations] record after each iteration
  int32_t* src;          // [P] GWW: which slot of `next` each slot of the new population starts from
  uint64_t* masks;       // [P][k][words] rebuild scratch
  int64_t* intra;        // [P] rebuild scratch: sum_v G[label v][v]
  uint64_t* work;        // [P][n] packed vertex words of the k = 3 wide descent (null for other plans)
  uint64_t init_stream, select_stream, perturb_stream;
  int32_t tournament, copies, early_stop, iterations;
  float perturbation;
  int32_t do_perturb;  // k >= 2 and perturbation > 0 (src/cc_pbils_cpu.cpp:241)
  int32_t select;      // 1: the descent of slot s starts from Tournament(pop); 0: from pop[s] itself
};

// Every accepted move lowers f by at least 1, so a descent ends after at most f moves; in practice it takes well
// under n moves. A descent that reaches this cap means a broken invariant: it stops, sets ctl->error and the run
// fails instead of spinning on the device.
__host__ __device__ inline int64_t MoveCap(int64_t f_start, int n) {
  const int64_t practical = 64 * static_cast<int64_t>(n) + 4096;
  return f_start < practical ? f_start + 1 : practical;
}

// Kernel launchers (cc_kernels.hip). All take the stream of the run.
void LaunchInitLabels(const RunDev& r, hipStream_t stream);
void LaunchDegrees(const RunDev& r, hipStream_t stream);
void LaunchRebuild(const RunDev& r, int wide, hipStream_t stream);
void LaunchRecord(const RunDev& r, int mode, int read_next, int32_t* order, hipStream_t stream);
void LaunchPerturb(const RunDev& r, hipStream_t stream);
void LaunchNeighborhoodLabels(const RunDev& r, int first, int count, hipStream_t stream);
void LaunchResetCtl(const RunDev& r, hipStream_t stream);

// Record kernel modes.
constexpr int kRecordInitial = 0;    // offer the population after init (no iteration bookkeeping)
constexpr int kRecordIteration = 1;  // offer `next` after a descent, then stall / stop / GWW ranking
constexpr int kRecordBatch = 2;      // Neighborhood batch: offer `next`, add moves, strict < keeps earlier starts

// Descent kinds (values match CC_DESCENT_* of cc_hip.h).
enum class Descent : int {
  kGenericGlobal = 0,
  kGenericLds = 1,
  kK2Registers = 2,
  kK2Wide = 3,
  kK3Wide = 4,
  kRegisters = 5,  // k in [3, 8], narrow, state in registers
};

struct DescentPlan {
  Descent descent = Descent::kGenericGlobal;
  int k = 0;
  int threads = 256;
  int vpt = 0;             // vertices per thread (k = 2 registers)
  bool rows_in_lds = false;  // k = 2 registers: the whole adjacency staged in LDS
  size_t lds_bytes = 0;    // dynamic LDS
  int blocks = 0;
};

// Descent launcher; `wide` selects 32-bit G.
void LaunchDescent(const RunDev& r, const DescentPlan& plan, int wide, hipStream_t stream);
// Whether the descent kernels exist for this plan (template instance present).
bool DescentSupported(const DescentPlan& plan, int wide);
// Largest dynamic LDS the descent kernels may request.
size_t DescentLdsLimit();
// Dynamic LDS of the generic descent with G and labels in LDS.
size_t GenericLdsBytes(int n, int k, int wide);
// Vertices per thread of the register descent for (n, k), and its threads per block; 0 if no instance fits.
int RegistersVpt(int n, int k, int* threads);

}  // namespace cc::hip
