// cc_internal.hpp — definitions shared by the kernels (cc_kernels.hip, cc_descent.hip) and the host runtime
// (cc_runtime.hip, cc_serve.hip).
//
// Every kernel works on a batch: it reads the run descriptors (RunDev) of many independent PBILS runs from a device
// array and finds its run and slot through an item table. A single solve is a batch of one run.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cc_math.hpp"

namespace cc::hip {

constexpr int kWave = 64;
constexpr int kMaxClusters = 64;
constexpr int kMaxVerticesNarrow = 32767;
constexpr int kMaxVerticesWide = 1 << 20;
// Zero bytes after the last adjacency row of a device graph. The register descents read a 64-bit window of the row
// at each thread's first vertex and ignore the bits of slots beyond n; those reads stay within vpt T / 8 + 4 bytes of
// the row's start (at most 32768 / 8 + 4 = 4100, of which a row holds at least 8).
constexpr size_t kRowTailBytes = 4096;

// Run state that the iteration kernels read and advance on the device, so that no host round trip is needed per
// iteration: the record, the early-stop counter and the iteration number (reference: PullRecord, Progress::Next
// and RankSlots run on the host, cuda/cc_pbils_gpu.cu:324-345 and include/cc_pbils.hpp:69-78).
struct Ctl {
  int32_t iteration;          // iteration index used by the next descent (tournament step)
  int32_t perturb_iteration;  // iteration index of the pending perturbation
  int32_t stall;              // iterations without a new record
  int32_t iterations_done;
  int32_t stopped;            // set once the run is over: every later kernel skips the run
  int32_t have_record;
  int32_t rec_slot;
  int32_t error;              // kErrorMoveCap / kErrorLds: the run is stopped and reported as failed
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

// One run (one graph, one parameter set). Lives in a device array; kernels read it through a pointer.
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
  int32_t* order;        // [P] GWW ranking scratch
  uint64_t* masks;       // [P][k][words] rebuild scratch
  int64_t* intra;        // [P] rebuild scratch: sum_v G[label v][v]
  uint64_t* work;        // packed vertex words of the wide descents (null for other plans): k = 3 chunk kernel
                         // [P][WideStride(n, 3)] u64; streaming kernels [P][WideSlotBytes(n, k)] bytes (k = 3:
                         // within the chunk kernel's size, so a footprint needs no edge count)
  uint64_t init_stream, select_stream, perturb_stream;
  int32_t tournament, copies, early_stop, iterations;
  float perturbation;
  int32_t do_perturb;  // k >= 2 and perturbation > 0 (src/cc_pbils_cpu.cpp:241)
  int32_t select;      // 1: the descent of slot s starts from Tournament(pop); 0: from pop[s] itself
  int32_t wide;        // 32-bit G
  int32_t variant;     // index of the run's kernel in the table of its mixed descent (register kernels), or -1
};

// A slot of a run: the unit of work of the slot-level kernels and of the descent.
struct Item {
  int32_t run;
  int32_t slot;
};

// The items of one launch, with the persistent-grid work counter of the descent.
struct ItemList {
  const Item* items;
  int32_t count;
  int32_t* queue;  // descent: next item to claim (reset in-stream by KernelResetQueues)
};

// Every accepted move lowers f by at least 1 and f never goes below 0, so a descent from f_start ends after at most
// f_start moves, exactly as the uncapped descent of src/cc_pbils_cpu.cpp:141-149 does. The cap f_start + 1 can
// therefore only be reached through a broken invariant (a wrong delta or G): the descent then stops, sets
// ctl->error, and the run fails instead of spinning on the device. No valid descent is cut short.
__host__ __device__ inline int64_t MoveCap(int64_t f_start) { return f_start + 1; }

// Ctl::error: why a descent stopped its run on the device.
constexpr int32_t kErrorMoveCap = 1;  // it reached MoveCap()
constexpr int32_t kErrorLds = 2;      // its launch allocated less LDS than its kernel needs
inline const char* DeviceErrorText(int32_t error) {
  return error == kErrorLds ? "a descent was launched with less LDS than it needs" : "a descent exceeded its move cap";
}

// The streaming wide-path descents load and store a group of consecutive vertices per lane: k = 2 four 24-bit fields
// in 12 bytes, k = 3 two 48-bit words in 12 bytes. A wave's tile is kStreamLoads such groups of its 64 lanes (1024
// vertices at k = 2, 512 at k = 3, 3 KiB either way); each slot is padded to whole tiles (WideSlotBytes per slot in
// RunDev::work, 16-byte aligned).
constexpr int kStreamLoads = 4;
__host__ __device__ constexpr int WideGroupVertices(int k) { return k == 2 ? 4 : 2; }
__host__ __device__ constexpr int WideGroupBytes(int) { return 12; }
__host__ __device__ constexpr int WideTileVertices(int k) { return 64 * WideGroupVertices(k) * kStreamLoads; }
__host__ __device__ constexpr int WideTileBytes(int k) { return 64 * WideGroupBytes(k) * kStreamLoads; }
// n rounded up to whole tiles
__host__ __device__ constexpr int64_t WideStride(int n, int k) {
  return (static_cast<int64_t>(n) + WideTileVertices(k) - 1) / WideTileVertices(k) * WideTileVertices(k);
}
__host__ __device__ constexpr int64_t WideSlotBytes(int n, int k) {
  return WideStride(n, k) / WideGroupVertices(k) * WideGroupBytes(k);
}
// The first tiles of a slot may live in the workgroup's LDS instead of HBM (at most kResidentPerWave per wave, as
// many as the LDS holds): they are never streamed. A resident tile holds its row dwords in registers through the
// move (four VGPRs per tile).
constexpr int kResidentPerWave = 3;

// ---- population kernels (cc_kernels.hip); max_n is the largest n among the runs of the items -------------------

void LaunchResetCtl(const RunDev* runs, int nruns, hipStream_t stream);
void LaunchResetQueues(int32_t* queues, int count, hipStream_t stream);
void LaunchGatherCtl(const RunDev* runs, int nruns, Ctl* out, hipStream_t stream);
void LaunchDegrees(const RunDev* run, int n, hipStream_t stream);
void LaunchInitLabels(const RunDev* runs, const ItemList& items, int max_n, hipStream_t stream);
// G, sizes and f of `pop` for the items; `tiles` lists (run, first column) of the kRebuildCols-column tiles of every
// run with this G width, column q = slot * k + cluster, in order of run and column.
constexpr int kRebuildCols = 64;
struct RebuildGroup {
  ItemList tiles;
  int wide;
};
void LaunchMasks(const RunDev* runs, const ItemList& items, hipStream_t stream);
void LaunchRebuildG(const RunDev* runs, const RebuildGroup& group, int max_n, hipStream_t stream);
void LaunchFinalize(const RunDev* runs, const ItemList& items, hipStream_t stream);
void LaunchRecord(const RunDev* runs, int nruns, int mode, int read_next, hipStream_t stream);
void LaunchPerturb(const RunDev* runs, const ItemList& items, int max_n, hipStream_t stream);
void LaunchNeighborhoodLabels(const RunDev* run, int n, int first, int count, hipStream_t stream);

// Record kernel modes.
constexpr int kRecordInitial = 0;    // offer the population after init (no iteration bookkeeping)
constexpr int kRecordIteration = 1;  // offer `next` after a descent, then stall / stop / GWW ranking
constexpr int kRecordBatch = 2;      // Neighborhood batch: offer `next`, add moves, strict < keeps earlier starts

// ---- k = 2 rebuild on the matrix cores (cc_mfma.hip) -----------------------------------------------------------

// n rounded up to the 128-byte K slabs (and 64-row tiles) of the int8 matrices.
size_t MfmaPad(int n);
// adj8: n_pad x n_pad bytes, adj8[v][u] = A[v][u] (zero padding).
void LaunchExpandAdjacency(const uint64_t* bits, int n, int words, int n_pad, int8_t* adj8, hipStream_t stream);
// G (int16), sizes untouched, intra[s] accumulated (zeroed before by LaunchMasks) for slots 0..count-1 of `run`;
// z: MfmaPad(count) x n_pad bytes of scratch.
void LaunchRebuildMfmaK2(const RunDev* run, int count, int n_pad, const int8_t* adj8, int8_t* z, hipStream_t stream);

// ---- descent (cc_descent.hip) --------------------------------------------------------------------------------

// Descent kinds (values match CC_DESCENT_* of cc_hip.h).
enum class Descent : int {
  kGenericGlobal = 0,
  kGenericLds = 1,
  kK2Registers = 2,
  kK2Wide = 3,
  kK3Wide = 4,
  kRegisters = 5,  // k in [3, 32], narrow, state in registers
  kMixed = 6,      // a batch launch over runs of different register kernels (256/512/1024 threads), by RunDev::variant
  kK3Chunk = 7,    // k = 3 wide, packed words plus per-chunk pair summaries in LDS (K3ChunkLdsBytes)
  kK3Packed = 8,   // k = 3 narrow, packed counts in registers, consecutive vertices per thread (n up to 32767)
};

struct DescentPlan {
  Descent descent = Descent::kGenericGlobal;
  int k = 0;
  int threads = 256;
  int vpt = 0;               // vertices per thread (register kernels)
  bool rows_in_lds = false;  // k = 2 registers: the whole adjacency staged in LDS
  size_t lds_bytes = 0;      // dynamic LDS
  int wide = 0;
  int blocks = 0;            // persistent workgroups
};

// Same kernel instance and launch shape: the items of such plans can share one launch.
inline bool SameLaunch(const DescentPlan& a, const DescentPlan& b) {
  return a.descent == b.descent && a.threads == b.threads && a.vpt == b.vpt && a.rows_in_lds == b.rows_in_lds &&
         a.wide == b.wide && (a.descent != Descent::kRegisters || a.k == b.k);
}

// A persistent launch: plan.blocks workgroups claim the items of `items` through items.queue.
void LaunchDescent(const RunDev* runs, const ItemList& items, const DescentPlan& plan, hipStream_t stream);
// Sets the dynamic-LDS attribute of the plan's kernel instance without launching it (LaunchDescent sets it too): a
// caller that records launches in a stream capture calls this first, so no attribute call happens inside the capture.
void SetDescentAttributes(const DescentPlan& plan);
bool DescentSupported(const DescentPlan& plan);
size_t DescentLdsLimit();
size_t GenericLdsBytes(int n, int k, int wide);
// LDS of the k = 3 chunk-summary descent for n vertices (64-vertex chunks): the pair table (6 x u32 per chunk) and
// the dirty-chunk list of each move parity (2 x u32 per chunk).
__host__ __device__ constexpr size_t K3ChunkLdsBytes(int n) { return (static_cast<size_t>(n) + 63) / 64 * (6 + 2) * 4; }
// LDS of the streaming wide-path descents (k = 2, k = 3) with `threads` per workgroup: the slot's first tiles, as
// many as fit (at most kResidentPerWave per wave and no more than the slot's whole tiles; a partial last tile is
// always streamed).
size_t WideResidentLdsBytes(int n, int k, int threads);
int RegistersVpt(int n, int k, int* threads);
// Vertices per thread of the packed k = 3 descent (1024 threads) for n <= 32767: the smallest of 16, 20, ..., 32
// that covers n.
int K3PackedVpt(int n);
// The mixed-descent variant that runs a register plan of (k, n), or -1 if none: an index into the table of the launch
// with the fewest threads (256, 512, 1024) whose variant holds n, plan->threads receiving that count; `plan` receives
// the equivalent register plan.
int MixedVariant(int n, int k, DescentPlan* plan);

}  // namespace cc::hip
