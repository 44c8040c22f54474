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

// Rebuild, part 2: G[s][c][v] = sum_w popc(row_v[w] & mask_{s,c}[w]) for a tile of kRebuildRows vertices x two
// slots of one run x kClusters clusters. The masks of the tile are staged through LDS in chunks of kWordTile words;
// every thread owns one row. The diagonal G[s][label v][v] is summed into intra[s] for the objective.
constexpr int kRebuildRows = 256;
constexpr int kRebuildSlots = 2;
constexpr int kWordTile = 256;

template <typename GT, int kClusters>
__global__ void __launch_bounds__(kRebuildRows) KernelRebuildG(const RunDev* runs, const Item* pairs) {
  const Item pair = pairs[blockIdx.y];
  const RunDev& r = runs[pair.run];
  if (r.ctl->stopped) return;
  const int n = r.n, k = r.k, words = r.words, population = r.population;
  const int c0 = blockIdx.z * kClusters;
  if (c0 >= k || static_cast<int>(blockIdx.x) * kRebuildRows >= n) return;  // uniform per block
  const int v = blockIdx.x * kRebuildRows + threadIdx.x;
  const int s0 = pair.slot;
  __shared__ uint64_t tile[kRebuildSlots][kClusters][kWordTile];
  __shared__ int64_t block_intra[kRebuildSlots];
  if (threadIdx.x < kRebuildSlots) block_intra[threadIdx.x] = 0;

  int acc[kRebuildSlots][kClusters];
#pragma unroll
  for (int i = 0; i < kRebuildSlots; ++i)
#pragma unroll
    for (int c = 0; c < kClusters; ++c) acc[i][c] = 0;

  const uint64_t* row = r.bits + static_cast<size_t>(v < n ? v : 0) * words;
  for (int w0 = 0; w0 < words; w0 += kWordTile) {
    const int span = words - w0 < kWordTile ? words - w0 : kWordTile;
    __syncthreads();
    for (int i = threadIdx.x; i < kRebuildSlots * kClusters * kWordTile; i += kRebuildRows) {
      const int w = i % kWordTile;
      const int c = (i / kWordTile) % kClusters;
      const int si = i / (kWordTile * kClusters);
      const int s = s0 + si;
      uint64_t m = 0;
      if (w < span && s < population && c0 + c < k) {
        m = r.masks[(static_cast<size_t>(s) * k + c0 + c) * words + w0 + w];
      }
      tile[si][c][w] = m;
    }
    __syncthreads();
    if (v < n) {
      for (int w = 0; w < span; ++w) {
        const uint64_t x = row[w0 + w];
#pragma unroll
        for (int i = 0; i < kRebuildSlots; ++i)
#pragma unroll
          for (int c = 0; c < kClusters; ++c) acc[i][c] += __popcll(x & tile[i][c][w]);
      }
    }
  }
  if (v < n) {
#pragma unroll
    for (int i = 0; i < kRebuildSlots; ++i) {
      const int s = s0 + i;
      if (s >= population) continue;
      GT* g = static_cast<GT*>(r.pop.g) + static_cast<size_t>(s) * k * n;
      const int label = r.pop.labels[static_cast<size_t>(s) * n + v];
      int diag = 0;
#pragma unroll
      for (int c = 0; c < kClusters; ++c) {
        if (c0 + c < k) g[static_cast<size_t>(c0 + c) * n + v] = static_cast<GT>(acc[i][c]);
        if (label == c0 + c) diag = acc[i][c];
      }
      if (diag) atomicAdd(reinterpret_cast<unsigned long long*>(&block_intra[i]), static_cast<unsigned long long>(diag));
    }
  }
  __syncthreads();
  if (threadIdx.x < kRebuildSlots && s0 + static_cast<int>(threadIdx.x) < population && block_intra[threadIdx.x]) {
    atomicAdd(reinterpret_cast<unsigned long long*>(&r.intra[s0 + threadIdx.x]),
              static_cast<unsigned long long>(block_intra[threadIdx.x]));
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
  ForChunksY(g.pairs.items, g.pairs.count, [&](const Item* chunk, int count) {
    const dim3 grid(Blocks(max_n, kRebuildRows), count, 1);
    switch (g.kc) {  // k <= 8: exactly k clusters per block; k > 8: chunks of 8 along grid z
      case 2: KernelRebuildG<GT, 2><<<grid, kRebuildRows, 0, stream>>>(runs, chunk); break;
      case 3: KernelRebuildG<GT, 3><<<grid, kRebuildRows, 0, stream>>>(runs, chunk); break;
      case 4: KernelRebuildG<GT, 4><<<grid, kRebuildRows, 0, stream>>>(runs, chunk); break;
      case 5: KernelRebuildG<GT, 5><<<grid, kRebuildRows, 0, stream>>>(runs, chunk); break;
      case 6: KernelRebuildG<GT, 6><<<grid, kRebuildRows, 0, stream>>>(runs, chunk); break;
      case 7: KernelRebuildG<GT, 7><<<grid, kRebuildRows, 0, stream>>>(runs, chunk); break;
      default: {
        const dim3 grid8(grid.x, grid.y, Blocks(g.max_k, 8));
        KernelRebuildG<GT, 8><<<grid8, kRebuildRows, 0, stream>>>(runs, chunk);
      }
    }
  });
}

void LaunchRebuildG(const RunDev* runs, const RebuildGroup& group, int max_n, hipStream_t stream) {
  if (group.pairs.count == 0) return;
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
