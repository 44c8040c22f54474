// cc_kernels.hip — population kernels of the PBILS path for gfx950: initial labels, G rebuild, record and GWW
// ranking, perturbation, Neighborhood start partitions. The descent kernels live in cc_descent.hip.
//
// All kernels are batched: a block finds its run and slot through an item table (or, for the record, one block per
// run), so one launch serves any number of independent runs of different n and k. Every kernel of an iteration
// skips runs whose ctl->stopped is set, so the host can queue (or replay from a HIP graph) whole chunks of iterations
// without a round trip per iteration and still end each run on exactly the iteration where the reference's early
// stop fires.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "cc_internal.hpp"
#include "cc_wave.hpp"

namespace cc::hip {
namespace {

constexpr int kBlock = 256;

__global__ void KernelResetCtl(const RunDev* runs, int nruns) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nruns) return;
  Ctl c{};
  c.rec_f = -1;
  *runs[i].ctl = c;
}

// out[i] = *runs[i].ctl, so that the host reads every run's control block with one transfer.
__global__ void KernelGatherCtl(const RunDev* runs, int nruns, Ctl* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nruns) out[i] = *runs[i].ctl;
}

__global__ void KernelResetQueues(int32_t* queues, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) queues[i] = 0;
}

// labels[slot][v] = InitialLabel(init, slot * n + v, k)  (reference: KernelInitLabels, src/cc_pbils_cpu.cpp:198-205)
__global__ void __launch_bounds__(kBlock) KernelInitLabels(const RunDev* runs, const Item* items) {
  const Item it = items[blockIdx.y];
  const RunDev& r = runs[it.run];
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= r.n) return;
  const uint64_t index = static_cast<uint64_t>(it.slot) * static_cast<uint64_t>(r.n) + static_cast<uint64_t>(v);
  r.pop.labels[index] = static_cast<uint8_t>(InitialLabel(r.init_stream, index, r.k));
}

__global__ void __launch_bounds__(kBlock) KernelDegrees(const RunDev* run) {
  const RunDev& r = *run;
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= r.n) return;
  const uint64_t* row = r.bits + static_cast<size_t>(v) * r.words;
  int acc = 0;
  for (int w = 0; w < r.words; ++w) acc += __popcll(row[w]);
  r.deg[v] = acc;
}

// Rebuild, part 1: per slot, the cluster masks (one ballot per cluster per 64 labels) and the cluster sizes.
__global__ void __launch_bounds__(kBlock) KernelMasks(const RunDev* runs, const Item* items) {
  const Item it = items[blockIdx.x];
  const RunDev& r = runs[it.run];
  if (r.ctl->stopped) return;
  const int s = it.slot;
  const int n = r.n, k = r.k, words = r.words;
  const int lane = threadIdx.x & 63;
  const uint8_t* labels = r.pop.labels + static_cast<size_t>(s) * n;
  uint64_t* masks = r.masks + static_cast<size_t>(s) * k * words;
  __shared__ int32_t sizes[kMaxClusters];
  for (int c = threadIdx.x; c < k; c += kBlock) sizes[c] = 0;
  __syncthreads();
  for (int w = WaveIndex(); w < words; w += kBlock / kWave) {
    const int v = w * kWave + lane;
    const int label = v < n ? labels[v] : -1;
    for (int c = 0; c < k; ++c) {
      const uint64_t m = __ballot(label == c);
      if (lane == c % kWave) {
        masks[static_cast<size_t>(c) * words + w] = m;
        if (m) atomicAdd(&sizes[c], __popcll(m));
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < k; c += kBlock) r.pop.sizes[static_cast<size_t>(s) * k + c] = sizes[c];
  if (threadIdx.x == 0) r.intra[s] = 0;
}

// Rebuild, part 2: G[s][c][v] = sum_w popc(row_v[w] & mask_{s,c}[w]), an AND-popcount product of the adjacency
// (n rows of W words) with the cluster masks of the columns q = s k + c, which KernelMasks lays out as [q][w], so
// that G[s][c] is row q of G. A block computes kRebuildRows vertices x kRebuildCols columns of one run (any k),
// staging kRebuildWords-word slices of both operands through LDS; thread (tx, ty) owns vertices v0 + tx + 16 i and
// columns q0 + ty + 16 j, i, j < 4, so the adjacency is read once per 64 columns instead of once per slot pair.
// The diagonal G[s][label v][v] is summed per slot into intra[s] for the objective.
constexpr int kRebuildRows = 64;
constexpr int kRebuildWords = 32;
constexpr int kRebuildThreads = 256;
static_assert(kRebuildCols == 64 && kRebuildThreads == 16 * 16, "4 x 4 outputs per thread");

template <typename GT>
__global__ void __launch_bounds__(kRebuildThreads) KernelRebuildG(const RunDev* runs, const Item* tiles) {
  const Item tile = tiles[blockIdx.y];
  const RunDev& r = runs[tile.run];
  if (r.ctl->stopped) return;
  const int n = r.n, k = r.k, words = r.words;
  const int v0 = blockIdx.x * kRebuildRows;
  const int q0 = tile.slot;                       // first column of the tile
  const int columns = r.population * k - q0;      // columns of the run from q0 on
  if (v0 >= n || columns <= 0) return;            // uniform per block
  __shared__ uint64_t a_lds[kRebuildWords][kRebuildRows + 1];
  __shared__ uint64_t m_lds[kRebuildWords][kRebuildCols + 1];
  __shared__ unsigned long long tile_intra[kRebuildCols];  // per slot of the tile (k >= 1: at most 64 slots)
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  if (tid < kRebuildCols) tile_intra[tid] = 0;

  int acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0;

  for (int w0 = 0; w0 < words; w0 += kRebuildWords) {
    const int span = words - w0 < kRebuildWords ? words - w0 : kRebuildWords;
    __syncthreads();
    // consecutive threads read consecutive words of one row (or column); zero outside the run and the row
    for (int e = tid; e < kRebuildRows * kRebuildWords; e += kRebuildThreads) {
      const int row = e / kRebuildWords, w = e % kRebuildWords;
      const int v = v0 + row;
      a_lds[w][row] = (v < n && w < span) ? r.bits[static_cast<size_t>(v) * words + w0 + w] : 0ull;
      m_lds[w][row] = (row < columns && w < span) ? r.masks[static_cast<size_t>(q0 + row) * words + w0 + w] : 0ull;
    }
    __syncthreads();
    for (int w = 0; w < span; ++w) {
      uint64_t a[4], m[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = a_lds[w][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) m[j] = m_lds[w][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += __popcll(a[i] & m[j]);
    }
  }

  const int s_first = q0 / k;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int qt = ty + 16 * j;
    if (qt >= columns) continue;
    const int q = q0 + qt;
    const int s = q / k, c = q - s * k;
    GT* g = static_cast<GT*>(r.pop.g) + static_cast<size_t>(q) * n;
    const uint8_t* labels = r.pop.labels + static_cast<size_t>(s) * n;
    int diag = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int v = v0 + tx + 16 * i;
      if (v < n) {
        g[v] = static_cast<GT>(acc[i][j]);
        if (labels[v] == c) diag += acc[i][j];
      }
    }
    if (diag) atomicAdd(&tile_intra[s - s_first], static_cast<unsigned long long>(diag));
  }
  __syncthreads();
  if (tid < kRebuildCols && tile_intra[tid]) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&r.intra[s_first + tid]), tile_intra[tid]);
  }
}

// Rebuild, part 3: f = m + sum_c C(s_c, 2) - 2W with 2W = intra (reference include/cc_math.hpp:66-68).
__global__ void __launch_bounds__(kBlock) KernelFinalize(const RunDev* runs, const Item* items, int count) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= count) return;
  const Item it = items[i];
  const RunDev& r = runs[it.run];
  if (r.ctl->stopped) return;
  const int32_t* sizes = r.pop.sizes + static_cast<size_t>(it.slot) * r.k;
  r.pop.f[it.slot] = r.edges + SamePairs(sizes, r.k) - r.intra[it.slot];
}

// Best (f, slot) of a block: the first minimum.
struct FSlot {
  int64_t f;
  int32_t slot;
};

__device__ __forceinline__ bool Better(const FSlot& a, const FSlot& b) {
  return a.f < b.f || (a.f == b.f && a.slot < b.slot);
}

__device__ FSlot BlockArgMin(FSlot x, FSlot* scratch /*[16]*/) {
  for (int offset = 32; offset > 0; offset >>= 1) {
    FSlot o;
    o.f = __shfl_xor(x.f, offset, 64);
    o.slot = __shfl_xor(x.slot, offset, 64);
    if (Better(o, x)) x = o;
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) scratch[wave] = x;
  __syncthreads();
  FSlot best = scratch[0];
  const int waves = (blockDim.x + 63) >> 6;
  for (int w = 1; w < waves; ++w) {
    if (Better(scratch[w], best)) best = scratch[w];
  }
  return best;
}

__device__ int64_t BlockSum(int64_t x, int64_t* scratch /*[16]*/) {
  for (int offset = 32; offset > 0; offset >>= 1) x += __shfl_xor(x, offset, 64);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) scratch[wave] = x;
  __syncthreads();
  int64_t total = 0;
  const int waves = (blockDim.x + 63) >> 6;
  for (int w = 0; w < waves; ++w) total += scratch[w];
  return total;
}

constexpr int kRecordThreads = 1024;
constexpr int kRankLdsSlots = 4096;

// The record after a phase, one block per run: the first slot of minimal f, taken only if strictly better
// (reference src/cc_pbils_cpu.cpp:208-214, cuda/cc_pbils_gpu.cu:324-334). In iteration mode it also does
// Progress::Next (include/cc_pbils.hpp:69-78) and the GWW ranking by (f, slot) (src/cc_pbils_cpu.cpp:168-175,
// 234-239), leaving in src[] the slot of `next` that each new slot copies.
__global__ void __launch_bounds__(kRecordThreads) KernelRecord(const RunDev* runs, int mode, int read_next) {
  const RunDev& r = runs[blockIdx.x];
  Ctl* ctl = r.ctl;
  if (ctl->stopped) return;
  const PopDev& p = read_next ? r.next : r.pop;
  const int population = r.population;
  const int tid = threadIdx.x;
  __shared__ FSlot scratch[16];
  __shared__ int64_t sums[16];
  __shared__ int32_t improved_sh, stop_sh;
  __shared__ int64_t f_lds[kRankLdsSlots];

  FSlot mine{INT64_MAX, INT32_MAX};
  for (int s = tid; s < population; s += kRecordThreads) {
    const FSlot c{p.f[s], s};
    if (Better(c, mine)) mine = c;
  }
  const FSlot best = BlockArgMin(mine, scratch);
  if (tid == 0) {
    const bool improved = !(ctl->have_record && r.n > 0 && best.f >= ctl->rec_f);
    if (improved) {
      ctl->rec_f = best.f;
      ctl->rec_slot = best.slot;
      ctl->have_record = 1;
    }
    improved_sh = improved;
  }
  __syncthreads();
  const bool improved = improved_sh;
  if (improved) {
    const uint8_t* from = p.labels + static_cast<size_t>(best.slot) * r.n;
    for (int v = tid; v < r.n; v += kRecordThreads) r.rec_labels[v] = from[v];
  }
  if (mode == kRecordInitial) return;

  int64_t moved = 0;
  for (int s = tid; s < population; s += kRecordThreads) moved += r.moves[s];
  moved = BlockSum(moved, sums);
  if (tid == 0) ctl->accepted_moves += moved;
  if (mode == kRecordBatch) return;

  if (tid == 0) {
    ctl->iterations_done += 1;
    ctl->local_searches += population;
    ctl->stall = improved ? 0 : ctl->stall + 1;
    r.trace[ctl->iterations_done - 1] = ctl->rec_f;
    const bool stop = ctl->stall >= r.early_stop || ctl->iterations_done >= r.iterations;
    stop_sh = stop;
    if (stop) ctl->stopped = 1;
  }
  __syncthreads();
  if (stop_sh) return;

  for (int s = tid; s < population; s += kRecordThreads) r.src[s] = s;
  if (r.copies > 0) {
    const bool staged = population <= kRankLdsSlots;
    if (staged) {
      for (int s = tid; s < population; s += kRecordThreads) f_lds[s] = p.f[s];
    }
    __syncthreads();
    const int64_t* f = staged ? f_lds : p.f;
    // rank(s) = #{t : (f_t, t) < (f_s, s)}: a stable sort by f, ties by slot.
    for (int s = tid; s < population; s += kRecordThreads) {
      const int64_t fs = f[s];
      int rank = 0;
      for (int t = 0; t < population; ++t) {
        const int64_t ft = f[t];
        rank += (ft < fs) | ((ft == fs) & (t < s));
      }
      r.order[rank] = s;
    }
    __syncthreads();
    for (int i = tid; i < r.copies; i += kRecordThreads) r.src[r.order[population - 1 - i]] = r.order[i];
  }
  if (tid == 0) {
    ctl->perturb_iteration = ctl->iteration;
    ctl->iteration += 1;
  }
}

// The next population: slot s starts from next[src[s]] (GWW) and is perturbed with the per-label stream
// (reference KernelPerturb and Perturb, src/cc_pbils_cpu.cpp:151-161: index = slot * n + v).
__global__ void __launch_bounds__(kBlock) KernelPerturb(const RunDev* runs, const Item* items) {
  const Item it = items[blockIdx.y];
  const RunDev& r = runs[it.run];
  if (r.ctl->stopped) return;
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= r.n) return;
  const int s = it.slot;
  const int from = r.src[s];
  int label = r.next.labels[static_cast<size_t>(from) * r.n + v];
  if (r.do_perturb) {
    const uint64_t index = static_cast<uint64_t>(s) * static_cast<uint64_t>(r.n) + static_cast<uint64_t>(v);
    label = PerturbedLabel(r.perturb_stream, index, r.ctl->perturb_iteration, label, r.k, r.perturbation);
  }
  r.pop.labels[static_cast<size_t>(s) * r.n + v] = static_cast<uint8_t>(label);
}

// Start partition of start vertex first + b: v and its neighbours -> 0, the rest -> 1 (reference
// KernelNeighborhoodLabels, cuda/cc_population_kernels.cuh:15-22).
__global__ void __launch_bounds__(kBlock) KernelNeighborhoodLabels(const RunDev* run, int first) {
  const RunDev& r = *run;
  const int u = blockIdx.x * kBlock + threadIdx.x;
  const int b = blockIdx.y;
  if (u >= r.n) return;
  const int v = first + b;
  const uint64_t* row = r.bits + static_cast<size_t>(v) * r.words;
  const bool joined = (row[u >> 6] >> (u & 63)) & 1ull;
  r.pop.labels[static_cast<size_t>(b) * r.n + u] = (u == v || joined) ? 0 : 1;
}

int Blocks(int count, int per) { return (count + per - 1) / per; }

// Kernels that take their item from blockIdx.y are launched in chunks of at most kMaxGridY items, each with the
// item table advanced to its first item.
constexpr int kMaxGridY = 65535;

template <typename Launch>
void ForChunksY(const Item* items, int count, Launch&& launch) {
  for (int first = 0; first < count; first += kMaxGridY) launch(items + first, std::min(kMaxGridY, count - first));
}

}  // namespace

void LaunchResetCtl(const RunDev* runs, int nruns, hipStream_t stream) {
  KernelResetCtl<<<Blocks(nruns, kBlock), kBlock, 0, stream>>>(runs, nruns);
}

void LaunchGatherCtl(const RunDev* runs, int nruns, Ctl* out, hipStream_t stream) {
  KernelGatherCtl<<<Blocks(nruns, kBlock), kBlock, 0, stream>>>(runs, nruns, out);
}

void LaunchResetQueues(int32_t* queues, int count, hipStream_t stream) {
  if (count > 0) KernelResetQueues<<<Blocks(count, kBlock), kBlock, 0, stream>>>(queues, count);
}

void LaunchDegrees(const RunDev* run, int n, hipStream_t stream) {
  KernelDegrees<<<Blocks(n, kBlock), kBlock, 0, stream>>>(run);
}

void LaunchInitLabels(const RunDev* runs, const ItemList& items, int max_n, hipStream_t stream) {
  ForChunksY(items.items, items.count, [&](const Item* chunk, int count) {
    KernelInitLabels<<<dim3(Blocks(max_n, kBlock), count), kBlock, 0, stream>>>(runs, chunk);
  });
}

void LaunchMasks(const RunDev* runs, const ItemList& items, hipStream_t stream) {
  if (items.count == 0) return;
  KernelMasks<<<items.count, kBlock, 0, stream>>>(runs, items.items);
}

template <typename GT>
static void LaunchRebuildGT(const RunDev* runs, const RebuildGroup& g, int max_n, hipStream_t stream) {
  ForChunksY(g.tiles.items, g.tiles.count, [&](const Item* chunk, int count) {
    KernelRebuildG<GT><<<dim3(Blocks(max_n, kRebuildRows), count), kRebuildThreads, 0, stream>>>(runs, chunk);
  });
}

void LaunchRebuildG(const RunDev* runs, const RebuildGroup& group, int max_n, hipStream_t stream) {
  if (group.tiles.count == 0) return;
  if (group.wide) {
    LaunchRebuildGT<int32_t>(runs, group, max_n, stream);
  } else {
    LaunchRebuildGT<int16_t>(runs, group, max_n, stream);
  }
}

void LaunchFinalize(const RunDev* runs, const ItemList& items, hipStream_t stream) {
  if (items.count == 0) return;
  KernelFinalize<<<Blocks(items.count, kBlock), kBlock, 0, stream>>>(runs, items.items, items.count);
}

void LaunchRecord(const RunDev* runs, int nruns, int mode, int read_next, hipStream_t stream) {
  KernelRecord<<<nruns, kRecordThreads, 0, stream>>>(runs, mode, read_next);
}

void LaunchPerturb(const RunDev* runs, const ItemList& items, int max_n, hipStream_t stream) {
  ForChunksY(items.items, items.count, [&](const Item* chunk, int count) {
    KernelPerturb<<<dim3(Blocks(max_n, kBlock), count), kBlock, 0, stream>>>(runs, chunk);
  });
}

void LaunchNeighborhoodLabels(const RunDev* run, int n, int first, int count, hipStream_t stream) {
  KernelNeighborhoodLabels<<<dim3(Blocks(n, kBlock), count), kBlock, 0, stream>>>(run, first);
}

}  // namespace cc::hip
