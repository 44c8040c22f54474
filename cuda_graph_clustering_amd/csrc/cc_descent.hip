// cc_descent.hip — steepest-descent local search on gfx950, one workgroup per individual.
//
// Both kernels pick, move after move, the improving move of smallest delta, ties to the smaller vertex and then the
// smaller target cluster (reference src/cc_pbils_cpu.cpp:99-118). The choice is a minimum over a packed key whose
// order is exactly that tie-break, so any reduction order gives the same move.
//
//  * KernelDescentGeneric: any k <= 64. G (int16 narrow / int32 wide) and 8-bit labels live in LDS when they fit
//    (k n |G| + n bytes) and otherwise in the slot's global buffers. Thread t owns vertices v = t (mod T): the
//    update of G for the previous move is fused into the scan of the next one, and only the owner touches a
//    vertex, so a move costs one barrier. Rows of the adjacency are read with wave-uniform (scalar) loads.
//  * KernelDescentK2: k = 2, narrow. Per vertex e = G1 - G0 and the label bit stay in registers; a move changes the
//    e of each neighbour by +-2 and nothing else (reference cuda/cc_local_search_k2.cuh:1-8 for the algebra). With
//    n words of adjacency small enough, the whole bitset is staged in LDS once per workgroup.
//
// Cluster sizes use the reference's one-barrier double buffer (cuda/cc_local_search.cuh:46-60): during the scan of
// move i the buffer i & 1 holds the sizes before move i - 1 and the correction for move i - 1 comes from registers.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cc_internal.hpp"
#include "cc_wave.hpp"

namespace cc::hip {
namespace {

constexpr uint64_t kNoMove = ~0ull;
constexpr int kDeltaOffset = 1 << 30;

// (delta + offset) in the high word, then v, from, to: unsigned order = (delta, v, to) order, since v fixes from.
__device__ __forceinline__ uint64_t PackMove(int delta, int v, int from, int to) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(delta + kDeltaOffset)) << 32) |
         (static_cast<uint32_t>(v) << 12) | (static_cast<uint32_t>(from) << 6) | static_cast<uint32_t>(to);
}

__host__ __device__ constexpr size_t Align16(size_t x) { return (x + 15) & ~static_cast<size_t>(15); }

template <int T, typename GT, bool kLds>
__global__ void __launch_bounds__(T) KernelDescentGeneric(RunDev r) {
  Ctl* ctl = r.ctl;
  if (ctl->stopped) return;
  constexpr int kWaves = T / kWave;
  const int s = blockIdx.x;
  const int n = r.n, k = r.k;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = WaveIndex();
  __shared__ uint64_t partial[2][kWaves];
  __shared__ int32_t sz[2][kMaxClusters];
  __shared__ int32_t src_sh;
  extern __shared__ __align__(16) uint8_t dyn[];

  if (tid == 0) {
    src_sh = r.select ? Tournament(r.select_stream, s, ctl->iteration, r.pop.f, r.population, r.tournament) : s;
  }
  __syncthreads();
  const int src = src_sh;

  const uint8_t* lab_in = r.pop.labels + static_cast<size_t>(src) * n;
  const GT* g_in = static_cast<const GT*>(r.pop.g) + static_cast<size_t>(src) * k * n;
  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  GT* g_out = static_cast<GT*>(r.next.g) + static_cast<size_t>(s) * k * n;
  const int32_t* sizes_in = r.pop.sizes + static_cast<size_t>(src) * k;

  GT* G;
  uint8_t* L;
  if constexpr (kLds) {
    G = reinterpret_cast<GT*>(dyn);
    L = dyn + Align16(sizeof(GT) * static_cast<size_t>(k) * n);
  } else {
    G = g_out;
    L = lab_out;
  }
  // Each thread copies exactly the vertices it owns, so no barrier is needed between the copy and the scan.
  if (kLds || g_in != static_cast<const GT*>(g_out)) {
    for (int base = wave * kWave; base < n; base += T) {
      const int v = base + lane;
      if (v < n) {
        for (int c = 0; c < k; ++c) G[static_cast<size_t>(c) * n + v] = g_in[static_cast<size_t>(c) * n + v];
        L[v] = lab_in[v];
      }
    }
  }
  for (int c = tid; c < k; c += T) sz[0][c] = sz[1][c] = sizes_in[c];
  const int64_t f_start = r.pop.f[src];
  __syncthreads();

  int64_t objective = f_start;
  int64_t applied = 0;
  const int64_t cap = MoveCap(f_start, n);
  int parity = 0;
  int prev_from = -1, prev_to = -1;  // the last move, not yet folded into sz[parity]
  int mv_v = -1, mv_from = 0, mv_to = 0;
  const uint64_t* row = r.bits;
  while (true) {
    const int32_t* szb = sz[parity];
    int best_d = 0, best_v = -1, best_from = 0, best_to = 0;
    for (int base = wave * kWave; base < n; base += T) {
      const uint64_t word = mv_v >= 0 ? row[base >> 6] : 0ull;  // wave-uniform
      const int v = base + lane;
      if (v < n) {
        if ((word >> lane) & 1ull) {
          G[static_cast<size_t>(mv_from) * n + v] -= 1;
          G[static_cast<size_t>(mv_to) * n + v] += 1;
        }
        if (v == mv_v) L[v] = static_cast<uint8_t>(mv_to);
        const int a = L[v];
        const int sa = szb[a] + (a == prev_to) - (a == prev_from);
        const int ga = G[static_cast<size_t>(a) * n + v];
        for (int c = 0; c < k; ++c) {
          if (c == a) continue;
          const int sc = szb[c] + (c == prev_to) - (c == prev_from);
          const int d = (sc - sa + 1) - 2 * (static_cast<int>(G[static_cast<size_t>(c) * n + v]) - ga);
          if (d < best_d) {
            best_d = d;
            best_v = v;
            best_from = a;
            best_to = c;
          }
        }
      }
    }
    const uint64_t key = best_v < 0 ? kNoMove : PackMove(best_d, best_v, best_from, best_to);
    const uint64_t best = BlockMin<kWaves>(key, partial[parity]);
    if (best == kNoMove) break;
    const int delta = static_cast<int>(static_cast<uint32_t>(best >> 32)) - kDeltaOffset;
    const uint32_t mv = static_cast<uint32_t>(best);
    const int v = static_cast<int>(mv >> 12);
    const int from = static_cast<int>((mv >> 6) & 63u);
    const int to = static_cast<int>(mv & 63u);
    // Every thread has finished the scan that read sz[parity]: bring it to the sizes after this move.
    for (int c = tid; c < k; c += T) {
      sz[parity][c] += (c == prev_to) - (c == prev_from) + (c == to) - (c == from);
    }
    prev_from = from;
    prev_to = to;
    parity ^= 1;
    row = r.bits + static_cast<size_t>(v) * r.words;
    mv_v = v;
    mv_from = from;
    mv_to = to;
    objective += delta;
    if (++applied >= cap) {
      if (tid == 0) {
        ctl->error = 1;
        ctl->stopped = 1;
      }
      break;
    }
  }

  if constexpr (kLds) {
    for (int base = wave * kWave; base < n; base += T) {
      const int v = base + lane;
      if (v < n) {
        for (int c = 0; c < k; ++c) g_out[static_cast<size_t>(c) * n + v] = G[static_cast<size_t>(c) * n + v];
        lab_out[v] = L[v];
      }
    }
  }
  for (int c = tid; c < k; c += T) {
    r.next.sizes[static_cast<size_t>(s) * k + c] = sz[parity][c] + (c == prev_to) - (c == prev_from);
  }
  if (tid == 0) {
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

// k = 2, narrow path (n <= 32767): key = (delta + 2^15) << 16 | v << 1 | label, a 32-bit minimum.
constexpr uint32_t kNoMove32 = ~0u;

template <int T, int kVpt, bool kRowsLds>
__global__ void __launch_bounds__(T) KernelDescentK2(RunDev r) {
  Ctl* ctl = r.ctl;
  if (ctl->stopped) return;
  constexpr int kWaves = T / kWave;
  static_assert(kVpt <= 32, "labels of a thread are packed into one 32-bit word");
  const int s = blockIdx.x;
  const int n = r.n;
  const int words = r.words;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = WaveIndex();
  __shared__ uint32_t partial[2][kWaves];
  __shared__ int32_t src_sh;
  extern __shared__ __align__(16) uint8_t dyn[];
  uint64_t* rows_lds = reinterpret_cast<uint64_t*>(dyn);

  if (tid == 0) {
    src_sh = r.select ? Tournament(r.select_stream, s, ctl->iteration, r.pop.f, r.population, r.tournament) : s;
  }
  if constexpr (kRowsLds) {
    const size_t total = static_cast<size_t>(n) * words;
    for (size_t i = tid; i < total; i += T) rows_lds[i] = r.bits[i];
  }
  __syncthreads();
  const int src = src_sh;
  const int16_t* g_in = static_cast<const int16_t*>(r.pop.g) + static_cast<size_t>(src) * 2 * n;
  const uint8_t* lab_in = r.pop.labels + static_cast<size_t>(src) * n;

  int e[kVpt];
  uint32_t label = 0;
  uint32_t active = 0;
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = wave * kWave + lane + j * T;
    e[j] = 0;
    if (v < n) {
      e[j] = static_cast<int>(g_in[n + v]) - static_cast<int>(g_in[v]);
      label |= static_cast<uint32_t>(lab_in[v] & 1) << j;
      active |= 1u << j;
    }
  }
  int n0 = r.pop.sizes[static_cast<size_t>(src) * 2];
  int64_t objective = r.pop.f[src];

  // delta(0 -> 1) = sv - 2e, delta(1 -> 0) = 2 - sv + 2e with sv = n1 - n0 + 1.
  auto scan_key = [&](int sv) {
    int best_d = 0;
    int best_j = -1;
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      const int d = ((label >> j) & 1u) ? 2 - sv + 2 * e[j] : sv - 2 * e[j];
      if (((active >> j) & 1u) && d < best_d) {
        best_d = d;
        best_j = j;
      }
    }
    if (best_j < 0) return kNoMove32;
    const uint32_t v = static_cast<uint32_t>(tid + best_j * T);
    return (static_cast<uint32_t>(best_d + 32768) << 16) | (v << 1) | ((label >> best_j) & 1u);
  };

  uint32_t key = scan_key(n - 2 * n0 + 1);
  int64_t applied = 0;
  const int64_t cap = MoveCap(objective, n);
  int parity = 0;
  while (true) {
    const uint32_t best = BlockMin<kWaves>(key, partial[parity]);
    parity ^= 1;
    if (best == kNoMove32) break;
    const int delta = static_cast<int>(best >> 16) - 32768;
    const int v = static_cast<int>((best & 0xFFFFu) >> 1);
    const int from = static_cast<int>(best & 1u);
    n0 += from ? 1 : -1;
    const int step = from ? -2 : 2;
    if (v % T == tid) label ^= 1u << (v / T);
    const uint64_t* row = (kRowsLds ? rows_lds : r.bits) + static_cast<size_t>(v) * words;
    const int sv = n - 2 * n0 + 1;
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      const int base = wave * kWave + j * T;  // wave-uniform
      const uint64_t word = base < n ? row[base >> 6] : 0ull;
      if ((word >> lane) & 1ull) e[j] += step;
    }
    key = scan_key(sv);
    objective += delta;
    if (++applied >= cap) {
      if (tid == 0) {
        ctl->error = 1;
        ctl->stopped = 1;
      }
      break;
    }
  }

  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  int16_t* g_out = static_cast<int16_t*>(r.next.g) + static_cast<size_t>(s) * 2 * n;
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = wave * kWave + lane + j * T;
    if (v < n) {
      const int degree = r.deg[v];
      lab_out[v] = static_cast<uint8_t>((label >> j) & 1u);
      g_out[v] = static_cast<int16_t>((degree - e[j]) / 2);
      g_out[n + v] = static_cast<int16_t>((degree + e[j]) / 2);
    }
  }
  if (tid == 0) {
    r.next.sizes[static_cast<size_t>(s) * 2] = n0;
    r.next.sizes[static_cast<size_t>(s) * 2 + 1] = n - n0;
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

// Adjacency rows are read-only for the whole run: reading them through the constant address space lets uniform
// row words come in with scalar loads.
using ConstWords = const __attribute__((address_space(4))) uint64_t*;

// k in [3, 8], narrow path, per-vertex state in registers. Thread t owns v = t + j T, j < kVpt. Per vertex:
//   gs[c] = 128 G[v][c]   and   keya = key of its own cluster,
// where the key of target c is  key_c = ((2 G_c - s_c + kOff) << 6) | (63 - c) = gs[c] + K_c  with the wave-uniform
// K_c = ((kOff - s_c) << 6) + 63 - c. The largest key gives the largest 2 G_c - s_c, ties to the smaller c, and
//   delta(a -> c) = (s_c - s_a + 1) - 2 (G_c - G_a) = (key_a >> 6) - (key_c >> 6) + 1.
// Including the own cluster a in the maximum is harmless: a move is improving only if its key is strictly above
// key_a's q by more than one, and then a is not the maximum. A move of v: from -> to changes, for each neighbour u,
// gs[from] -= 128 and gs[to] += 128; for every vertex with label a it shifts keya by t_a (128 bit - 64) with
// t_a = [a == to] - [a == from] (G change of a neighbour plus the change of K_a through s_a); t_a is fetched per
// lane with ds_bpermute from a lane table.
template <int T, int K, int kVpt>
__global__ void __launch_bounds__(T) KernelDescentRegs(RunDev r) {
  Ctl* ctl = r.ctl;
  if (ctl->stopped) return;
  constexpr int kWaves = T / kWave;
  constexpr int kOff = 1 << 16;
  constexpr int kIdleKey = 1 << 30;  // own key of a vertex beyond n: never improvable
  static_assert(K >= 3 && K <= 8, "register descent handles k in [3, 8]");
  const int s = blockIdx.x;
  const int n = r.n;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = WaveIndex();
  __shared__ uint64_t partial[2][kWaves];
  __shared__ int32_t src_sh;
  if (tid == 0) {
    src_sh = r.select ? Tournament(r.select_stream, s, ctl->iteration, r.pop.f, r.population, r.tournament) : s;
  }
  __syncthreads();
  const int src = __builtin_amdgcn_readfirstlane(src_sh);
  const int16_t* g_in = static_cast<const int16_t*>(r.pop.g) + static_cast<size_t>(src) * K * n;
  const uint8_t* lab_in = r.pop.labels + static_cast<size_t>(src) * n;
  const int32_t* sizes_in = r.pop.sizes + static_cast<size_t>(src) * K;

  int sz[K];
#pragma unroll
  for (int c = 0; c < K; ++c) sz[c] = __builtin_amdgcn_readfirstlane(sizes_in[c]);
  auto kconst = [&](int c) { return ((kOff - sz[c]) << 6) + (63 - c); };

  int gs[kVpt][K];
  int lab[kVpt];
  int keya[kVpt];
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = tid + j * T;
    lab[j] = 0;
    keya[j] = kIdleKey;
#pragma unroll
    for (int c = 0; c < K; ++c) gs[j][c] = 0;
    if (v < n) {
      lab[j] = lab_in[v];
#pragma unroll
      for (int c = 0; c < K; ++c) {
        gs[j][c] = static_cast<int>(g_in[static_cast<size_t>(c) * n + v]) << 7;
        if (lab[j] == c) keya[j] = gs[j][c] + kconst(c);
      }
    }
  }
  const int64_t f_start = r.pop.f[src];
  int64_t objective = f_start;
  int64_t applied = 0;
  const int64_t cap = MoveCap(f_start, n);

  int kc[K];
  auto scan = [&]() {
#pragma unroll
    for (int c = 0; c < K; ++c) kc[c] = kconst(c);
    int best_d = 0, best_j = -1, best_key = 0, best_a = 0;
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      int kb = gs[j][0] + kc[0];
#pragma unroll
      for (int c = 1; c < K; ++c) kb = max(kb, gs[j][c] + kc[c]);
      const int d = (keya[j] >> 6) - (kb >> 6) + 1;
      if (d < best_d) {
        best_d = d;
        best_j = j;
        best_key = kb;
        best_a = lab[j];
      }
    }
    if (best_j < 0) return kNoMove;
    return PackMove(best_d, tid + best_j * T, best_a, 63 - (best_key & 63));
  };

  uint64_t key = scan();
  int parity = 0;
  const ConstWords bits = (ConstWords)r.bits;  // address-space cast (global -> constant)
  while (true) {
    const uint64_t best = BlockMin<kWaves>(key, partial[parity]);
    parity ^= 1;
    if (best == kNoMove) break;
    const int delta = static_cast<int>(static_cast<uint32_t>(best >> 32)) - kDeltaOffset;
    const uint32_t mv = static_cast<uint32_t>(best);
    const int v = __builtin_amdgcn_readfirstlane(static_cast<int>(mv >> 12));
    const int from = __builtin_amdgcn_readfirstlane(static_cast<int>((mv >> 6) & 63u));
    const int to = __builtin_amdgcn_readfirstlane(static_cast<int>(mv & 63u));
    const ConstWords row = bits + static_cast<size_t>(v) * r.words;
    uint64_t words[kVpt];
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      const int base = wave * kWave + j * T;  // wave-uniform
      words[j] = base < n ? row[base >> 6] : 0ull;
    }
#pragma unroll
    for (int c = 0; c < K; ++c) sz[c] += (c == to) - (c == from);
    // lane c of the table holds t_c = [c == to] - [c == from]
    const int table = (lane == to) - (lane == from);
    int b7[kVpt];
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      const int bit = static_cast<int>((words[j] >> lane) & 1ull);
      b7[j] = bit << 7;
      const int ta = __builtin_amdgcn_ds_bpermute(lab[j] << 2, table);
      keya[j] += ta * (b7[j] - 64);
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
      if (c == from) {
#pragma unroll
        for (int j = 0; j < kVpt; ++j) gs[j][c] -= b7[j];
      }
      if (c == to) {
#pragma unroll
        for (int j = 0; j < kVpt; ++j) gs[j][c] += b7[j];
      }
    }
    // the moved vertex: new label, own key from its (unchanged) G and the new sizes
    const int mv_j = v / T;
    if (tid == v - mv_j * T) {
#pragma unroll
      for (int j = 0; j < kVpt; ++j) {
        if (j == mv_j) {
          lab[j] = to;
#pragma unroll
          for (int c = 0; c < K; ++c) {
            if (c == to) keya[j] = gs[j][c] + kconst(c);
          }
        }
      }
    }
    key = scan();
    objective += delta;
    if (++applied >= cap) {
      if (tid == 0) {
        ctl->error = 1;
        ctl->stopped = 1;
      }
      break;
    }
  }

  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  int16_t* g_out = static_cast<int16_t*>(r.next.g) + static_cast<size_t>(s) * K * n;
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = tid + j * T;
    if (v < n) {
      lab_out[v] = static_cast<uint8_t>(lab[j]);
#pragma unroll
      for (int c = 0; c < K; ++c) g_out[static_cast<size_t>(c) * n + v] = static_cast<int16_t>(gs[j][c] >> 7);
    }
  }
  if (tid < K) {
    int size = 0;
#pragma unroll
    for (int c = 0; c < K; ++c) size = tid == c ? sz[c] : size;
    r.next.sizes[static_cast<size_t>(s) * K + tid] = size;
  }
  if (tid == 0) {
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

constexpr size_t kLdsBytes = 160 * 1024;
// Static LDS of the descent kernels (partials, sizes, source slot), rounded up.
constexpr size_t kStaticLdsReserve = 2048;

template <typename Kernel>
void AllowLds(Kernel kernel, size_t bytes) {
  if (bytes > 0) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(bytes));
  }
}

template <int T, typename GT>
void LaunchGeneric(const RunDev& r, const DescentPlan& plan, hipStream_t stream) {
  const int blocks = r.population;
  if (plan.descent == Descent::kGenericLds) {
    auto kernel = KernelDescentGeneric<T, GT, true>;
    AllowLds(kernel, plan.lds_bytes);
    kernel<<<blocks, T, plan.lds_bytes, stream>>>(r);
  } else {
    KernelDescentGeneric<T, GT, false><<<blocks, T, 0, stream>>>(r);
  }
}

template <typename GT>
void LaunchGenericAny(const RunDev& r, const DescentPlan& plan, hipStream_t stream) {
  switch (plan.threads) {
    case 256: LaunchGeneric<256, GT>(r, plan, stream); break;
    case 512: LaunchGeneric<512, GT>(r, plan, stream); break;
    default: LaunchGeneric<1024, GT>(r, plan, stream); break;
  }
}

template <int T, int kVpt>
void LaunchK2(const RunDev& r, const DescentPlan& plan, hipStream_t stream) {
  if (plan.rows_in_lds) {
    auto kernel = KernelDescentK2<T, kVpt, true>;
    AllowLds(kernel, plan.lds_bytes);
    kernel<<<r.population, T, plan.lds_bytes, stream>>>(r);
  } else {
    KernelDescentK2<T, kVpt, false><<<r.population, T, 0, stream>>>(r);
  }
}

template <int T>
bool LaunchK2Vpt(const RunDev& r, const DescentPlan& plan, hipStream_t stream, bool launch) {
  switch (plan.vpt) {
    case 1: if (launch) LaunchK2<T, 1>(r, plan, stream); return true;
    case 2: if (launch) LaunchK2<T, 2>(r, plan, stream); return true;
    case 4: if (launch) LaunchK2<T, 4>(r, plan, stream); return true;
    case 8: if (launch) LaunchK2<T, 8>(r, plan, stream); return true;
    case 16: if (launch) LaunchK2<T, 16>(r, plan, stream); return true;
    case 32: if (launch) LaunchK2<T, 32>(r, plan, stream); return true;
    default: return false;
  }
}

// Register descent instances: T = 256 with kVpt in {1, 2, 4}; T = 1024 with kVpt in {2, 4, 6, 8, 10, 12},
// bounded by kVpt (K + 5) <= 112 so that the state fits the 128 VGPRs of a 1024-thread workgroup.
constexpr bool RegsFit(int k, int vpt) { return vpt * (k + 5) <= 112; }

template <int T, int K, int kVpt>
bool LaunchRegsOne(const RunDev& r, hipStream_t stream, bool launch) {
  if constexpr (RegsFit(K, kVpt)) {
    if (launch) KernelDescentRegs<T, K, kVpt><<<r.population, T, 0, stream>>>(r);
    return true;
  } else {
    return false;
  }
}

template <int K>
bool LaunchRegsK(const RunDev& r, const DescentPlan& plan, hipStream_t stream, bool launch) {
  if (plan.threads == 256) {
    switch (plan.vpt) {
      case 1: return LaunchRegsOne<256, K, 1>(r, stream, launch);
      case 2: return LaunchRegsOne<256, K, 2>(r, stream, launch);
      case 4: return LaunchRegsOne<256, K, 4>(r, stream, launch);
      default: return false;
    }
  }
  if (plan.threads == 1024) {
    switch (plan.vpt) {
      case 2: return LaunchRegsOne<1024, K, 2>(r, stream, launch);
      case 4: return LaunchRegsOne<1024, K, 4>(r, stream, launch);
      case 6: return LaunchRegsOne<1024, K, 6>(r, stream, launch);
      case 8: return LaunchRegsOne<1024, K, 8>(r, stream, launch);
      case 10: return LaunchRegsOne<1024, K, 10>(r, stream, launch);
      case 12: return LaunchRegsOne<1024, K, 12>(r, stream, launch);
      default: return false;
    }
  }
  return false;
}

bool DispatchRegs(const RunDev& r, const DescentPlan& plan, hipStream_t stream, bool launch) {
  switch (r.k) {
    case 3: return LaunchRegsK<3>(r, plan, stream, launch);
    case 4: return LaunchRegsK<4>(r, plan, stream, launch);
    case 5: return LaunchRegsK<5>(r, plan, stream, launch);
    case 6: return LaunchRegsK<6>(r, plan, stream, launch);
    case 7: return LaunchRegsK<7>(r, plan, stream, launch);
    case 8: return LaunchRegsK<8>(r, plan, stream, launch);
    default: return false;
  }
}

bool DispatchK2(const RunDev& r, const DescentPlan& plan, hipStream_t stream, bool launch) {
  switch (plan.threads) {
    case 64: return LaunchK2Vpt<64>(r, plan, stream, launch);
    case 128: return LaunchK2Vpt<128>(r, plan, stream, launch);
    case 256: return LaunchK2Vpt<256>(r, plan, stream, launch);
    case 512: return LaunchK2Vpt<512>(r, plan, stream, launch);
    case 1024: return LaunchK2Vpt<1024>(r, plan, stream, launch);
    default: return false;
  }
}

}  // namespace

size_t DescentLdsLimit() { return kLdsBytes - kStaticLdsReserve; }

bool DescentSupported(const DescentPlan& plan, int wide) {
  switch (plan.descent) {
    case Descent::kGenericGlobal:
    case Descent::kGenericLds:
      return plan.threads == 256 || plan.threads == 512 || plan.threads == 1024;
    case Descent::kK2Registers: {
      RunDev dummy{};
      return !wide && DispatchK2(dummy, plan, nullptr, false);
    }
    case Descent::kRegisters: {
      RunDev dummy{};
      dummy.k = plan.k;
      return !wide && DispatchRegs(dummy, plan, nullptr, false);
    }
    default:
      return false;
  }
}

int RegistersVpt(int n, int k, int* threads) {
  if (k < 3 || k > 8) return 0;
  if (n <= 1024) {
    *threads = 256;
    for (int v : {1, 2, 4}) {
      if (v * 256 >= n) return v;
    }
  }
  *threads = 1024;
  for (int v : {2, 4, 6, 8, 10, 12}) {
    if (v * 1024 >= n) return RegsFit(k, v) ? v : 0;
  }
  return 0;
}

void LaunchDescent(const RunDev& r, const DescentPlan& plan, int wide, hipStream_t stream) {
  switch (plan.descent) {
    case Descent::kGenericGlobal:
    case Descent::kGenericLds:
      if (wide) {
        LaunchGenericAny<int32_t>(r, plan, stream);
      } else {
        LaunchGenericAny<int16_t>(r, plan, stream);
      }
      break;
    case Descent::kK2Registers:
      DispatchK2(r, plan, stream, true);
      break;
    case Descent::kRegisters:
      DispatchRegs(r, plan, stream, true);
      break;
    default:
      break;
  }
}

size_t GenericLdsBytes(int n, int k, int wide) {
  return Align16((wide ? 4 : 2) * static_cast<size_t>(k) * n) + Align16(static_cast<size_t>(n));
}

}  // namespace cc::hip
