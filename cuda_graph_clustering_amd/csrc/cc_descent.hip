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

#include <algorithm>
#include <cstdint>
#include <type_traits>

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
__device__ __forceinline__ void DescentGenericBody(const RunDev& r, int s) {
  Ctl* ctl = r.ctl;
  constexpr int kWaves = T / kWave;
  // Wide path (int32 counts in global memory, mostly missing the caches): a scan whose count loads do not wait for
  // the label, kChunk of them in flight per vertex (twice as fast at n = 40000, k = 8). With int16 counts, which
  // stay in L2 or LDS, the plain scan over c != a is faster (about 12 % at n = 6000, k = 3).
  constexpr bool kChunkedScan = !kLds && sizeof(GT) == 4;
  constexpr int kChunk = 8;
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
  const int64_t cap = MoveCap(f_start);
  int parity = 0;
  int prev_from = -1, prev_to = -1;  // the last move, not yet folded into sz[parity]
  int mv_v = -1, mv_from = 0, mv_to = 0;
  const uint64_t* const bits = r.bits;
  const int row_words = __builtin_amdgcn_readfirstlane(r.words);
  const uint64_t* row = bits;
  while (true) {
    const int32_t* szb = sz[parity];
    int best_d = 0, best_v = -1, best_from = 0, best_to = 0;
    if constexpr (kChunkedScan) {
      for (int base = wave * kWave; base < n; base += T) {
        const uint64_t word = mv_v >= 0 ? row[base >> 6] : 0ull;  // wave-uniform
        const int v = base + lane;
        if (v < n) {
          // The last move is applied to the counts as they are loaded. With q_c = s_c - 2 G[c][v] the move of v to c
          // has delta q_c - q_a + 1, so the best target is the first c of smallest q_c over all clusters (c = a gives
          // delta 1, never an improving move) and q_a is taken on the way: no load waits for the label, and the
          // counts go out kChunk at a time.
          const bool neighbour = (word >> lane) & 1ull;
          if (v == mv_v) L[v] = static_cast<uint8_t>(mv_to);
          const int a = L[v];
          int qmin = 0x7FFFFFFF, cmin = 0, qa = 0;
          for (int c0 = 0; c0 < k; c0 += kChunk) {
            int g[kChunk];
#pragma unroll
            for (int i = 0; i < kChunk; ++i) {
              g[i] = c0 + i < k ? static_cast<int>(G[static_cast<size_t>(c0 + i) * n + v]) : 0;
            }
#pragma unroll
            for (int i = 0; i < kChunk; ++i) {
              const int c = c0 + i;
              if (c < k) {
                int gc = g[i];
                if (neighbour && (c == mv_from || c == mv_to)) {
                  gc += c == mv_to ? 1 : -1;
                  G[static_cast<size_t>(c) * n + v] = static_cast<GT>(gc);
                }
                const int q = szb[c] + (c == prev_to) - (c == prev_from) - 2 * gc;
                if (q < qmin) {
                  qmin = q;
                  cmin = c;
                }
                qa = c == a ? q : qa;
              }
            }
          }
          const int d = qmin - qa + 1;
          if (d < best_d) {
            best_d = d;
            best_v = v;
            best_from = a;
            best_to = cmin;
          }
        }
      }
    } else {
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
    row = bits + static_cast<size_t>(v) * row_words;
    mv_v = v;
    mv_from = from;
    mv_to = to;
    objective += delta;
    if (++applied >= cap) {
      if (tid == 0) {
        ctl->error = kErrorMoveCap;
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

// Adjacency rows as 32-bit halves in the global address space, for per-lane vector loads.
using GlobalWords32 = const __attribute__((address_space(1))) uint32_t*;

// k = 2, narrow path (n <= 32767). The key of a move is (delta + 2^15) << 16 | v << 1 | label, a 32-bit minimum
// (delta of an improving move lies in [-(n - 1), -1], so it fits).
//
// Thread t owns the kVpt consecutive vertices v_j = t kVpt + j (so the row bits of a move that it needs are one 64-bit
// window). Per slot two VGPRs hold
//   key_j  ((d_j + 2^15) << 16) | (v_j << 1) | label, where d_j is the delta of moving v_j to the other cluster:
//          sv - 2 e (label 0) or 2 - sv + 2 e (label 1), with e = G1 - G0 and sv = n1 - n0 + 1. |d_j| <= n - 1,
//          so key_j is the packed key of v_j's move, and the move improves exactly when key_j < 2^31;
//   sg_j   +2^16 (label 0), -2^16 (label 1), or 0 for a slot beyond n (whose key stays at d = 32767).
// When v moves out of cluster `from`, sv changes by step = (from ? -2 : 2) for everyone and e by step for each
// neighbour, so d_j += +-step (1 - 2 A[v][v_j]), i.e. key_j += sg_j step (1 - 2 A[v][v_j]); v's own key becomes
// that of the reverse move, -delta, with the flipped label.
constexpr uint32_t kImproving = 1u << 31;

template <int T, int kVpt, bool kRowsLds>
__device__ __forceinline__ void DescentK2Body(const RunDev& r, int s) {
  Ctl* ctl = r.ctl;
  constexpr int kWaves = T / kWave;
  static_assert(kVpt * T <= 32768, "v_j << 1 | label must fit in 16 bits");
  static_assert(kVpt <= 32, "a thread's row bits lie in one 64-bit window");
  const int n = r.n;
  const int words = r.words;
  const int tid = threadIdx.x;
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
  int n0 = r.pop.sizes[static_cast<size_t>(src) * 2];
  int64_t objective = r.pop.f[src];

  uint32_t key[kVpt];
  int sg[kVpt];
  {
    // Branch-free: a slot beyond n loads vertex n - 1 and discards it (branches around writes to the slot arrays
    // make the compiler copy whole arrays between blocks).
    const int sv = n - 2 * n0 + 1;
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      const int v = tid * kVpt + j;
      const bool real = v < n;
      const int vc = real ? v : n - 1;
      const int e = static_cast<int>(g_in[n + vc]) - static_cast<int>(g_in[vc]);
      const int label = real ? (lab_in[vc] & 1) : 0;
      const int d = real ? (label ? 2 - sv + 2 * e : sv - 2 * e) : 32767;
      sg[j] = real ? (label ? -65536 : 65536) : 0;
      key[j] = (static_cast<uint32_t>(d + 32768) << 16) | (static_cast<uint32_t>(v) << 1) |
               static_cast<uint32_t>(label);
    }
  }
  auto scan = [&]() {
    uint32_t m = ~0u;
#pragma unroll
    for (int j = 0; j < kVpt; ++j) m = min(m, key[j]);
    return m;
  };

  const int row_words = __builtin_amdgcn_readfirstlane(words);
  const uint64_t* const row_base =
      reinterpret_cast<const uint64_t*>(Uniform(static_cast<int64_t>(reinterpret_cast<uintptr_t>(r.bits))));
  const int window = (tid * kVpt) >> 5;  // the dword holding the thread's first bit; the second dword follows
  const uint32_t shift = static_cast<uint32_t>((tid * kVpt) & 31);
  uint32_t mine = scan();
  int64_t applied = 0;
  const int64_t cap = MoveCap(objective);
  int parity = 0;
  while (true) {
    const uint32_t best = BlockMin<kWaves>(mine, partial[parity]);
    parity ^= 1;
    if (best >= kImproving) break;
    const int delta = static_cast<int>(best >> 16) - 32768;
    const int v = __builtin_amdgcn_readfirstlane(static_cast<int>((best & 0xFFFFu) >> 1));
    const int from = static_cast<int>(best & 1u);
    // Bits [t kVpt, t kVpt + kVpt) of the row lie in its dwords `window` and `window` + 1. In LDS the second index
    // is clamped to the row (a dword past the row holds only bits of slots beyond n); in global memory it may lie in
    // the next row or the zero tail after the last row (kRowTailBytes; at most kVpt T / 8 + 4 bytes past the row's
    // start). Bits of slots beyond n are ignored (sg_j = 0).
    uint32_t w0, w1;
    if constexpr (kRowsLds) {
      const uint32_t* row = reinterpret_cast<const uint32_t*>(rows_lds + static_cast<size_t>(v) * row_words);
      w0 = row[min(window, 2 * row_words - 1)];
      w1 = row[min(window + 1, 2 * row_words - 1)];
    } else {
      const GlobalWords32 row = (GlobalWords32)(row_base + static_cast<size_t>(v) * row_words) + window;
      w0 = row[0];
      w1 = row[1];
    }
    const uint32_t bits = __builtin_amdgcn_alignbit(w1, w0, shift);
    n0 += from ? 1 : -1;
    const int step = from ? -2 : 2;
    // The builtin, not SignedBit: here the compiler folds its shift form into one v_mad_i32_i24 per slot, and an
    // opaque v_bfe_i32 made it fall back to a quarter-rate 32-bit multiply (11 % slower at C3).
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      const int bit = __builtin_amdgcn_sbfe(static_cast<int>(bits), j, 1);  // 0 or -1
      const int t = __mul24(bit, 2 * step) + step;                         // step (1 - 2 A[v][v_j])
      key[j] += static_cast<uint32_t>(__mul24(sg[j], t));
    }
    // The mover: its key becomes the reverse move's (delta -> -delta, label 1 - from). Only the wave that owns it
    // takes the (uniform) branch.
    const int owner = v / kVpt;  // uniform
    if (wave == (owner >> 6)) {
      const int jv = v - owner * kVpt;
      const uint32_t flipped = (static_cast<uint32_t>(32768 - delta) << 16) | (static_cast<uint32_t>(v) << 1) |
                               static_cast<uint32_t>(1 - from);
#pragma unroll
      for (int j = 0; j < kVpt; ++j) {
        if (j == jv && tid == owner) {
          key[j] = flipped;
          sg[j] = from ? 65536 : -65536;
        }
      }
    }
    mine = scan();
    objective += delta;
    if (++applied >= cap) {
      if (tid == 0) {
        ctl->error = kErrorMoveCap;
        ctl->stopped = 1;
      }
      break;
    }
  }

  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  int16_t* g_out = static_cast<int16_t*>(r.next.g) + static_cast<size_t>(s) * 2 * n;
  const int sv = n - 2 * n0 + 1;
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = tid * kVpt + j;
    if (v < n) {
      const int label = static_cast<int>(key[j] & 1u);
      const int d = static_cast<int>(key[j] >> 16) - 32768;
      const int e = label ? (d - 2 + sv) / 2 : (sv - d) / 2;
      const int degree = r.deg[v];
      lab_out[v] = static_cast<uint8_t>(label);
      g_out[v] = static_cast<int16_t>((degree - e) / 2);
      g_out[n + v] = static_cast<int16_t>((degree + e) / 2);
    }
  }
  if (tid == 0) {
    r.next.sizes[static_cast<size_t>(s) * 2] = n0;
    r.next.sizes[static_cast<size_t>(s) * 2 + 1] = n - n0;
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

// The register descent's move term of vertex slot J, m = 2^Q - 2^(Q+1) A[v][slot] = (bit J of `bits`, sign-extended,
// & mask) | bias with mask = -2^(Q+1) and bias = 2^Q, as v_bfe_i32 and one v_and_or_b32. gfx950's three-operand
// encoding takes no literal, so the mask comes in a VGPR and the bias in an SGPR (left to itself the compiler spends
// four instructions).
template <int J>
__device__ __forceinline__ int MoveTerm(uint32_t bits, int mask, int bias) {
  int r;
  asm("v_bfe_i32 %0, %1, %2, 1\n\tv_and_or_b32 %0, %0, %3, %4" : "=&v"(r) : "v"(bits), "n"(J), "v"(mask), "s"(bias));
  return r;
}

// Label field width of the register descent: two bits up to k = 4, four up to k = 16, five up to k = 32.
constexpr int RegsLabelBits(int k) { return k <= 4 ? 2 : k <= 16 ? 4 : 5; }
// Clusters held in registers for a run of k: k itself up to 16, above that k rounded up to a multiple of four; the
// extra clusters stay empty and are never offered as targets.
constexpr int RegsClusters(int k) { return k <= 16 ? k : (k + 3) & ~3; }

// Bit position of q in the register descent's keys: 12, or 13 when the slot index needs a fifth bit above the two
// label fields (the 512-thread instances with up to 32 slots per thread). With |q| <= 2n < 2^16 a key stays within
// 2^29 either way.
constexpr int RegsQShift(int k, int vpt) { return (vpt << (2 * RegsLabelBits(k))) <= 4096 ? 12 : 13; }

// k in [3, 32], narrow path, per-vertex state in registers. Thread t owns the kVpt consecutive vertices
// v = t kVpt + j, so the row bits of a move that it needs are one 64-bit window of the row (two dword loads and one
// v_alignbit) and a wave's loads share one or two cache lines. With b the label field width (RegsLabelBits: 2 bits
// up to k = 4, 4 up to 16, 5 above) the slot index j sits at bit 2b and stays below bit Q = RegsQShift (12, or 13
// for the widest slot counts).
//
// With q_c(v) = s_c - 2 G[v][c], the best move of v goes to the cluster of smallest q_c (ties to the smaller c) and
// has delta = q_c - q_a + 1 (a = label of v). Per (j, c) one register holds y = -key with
//   key(j, c) = (q_c << Q) | (j << 2b) | (a << b) | c,
// and per j the own term qa = q_a << Q. Then z_j = max_c y[j][c] + qa = -x_j with
//   x_j = ((q_c - q_a) << Q) | (j << 2b) | (a << b) | c,
// whose signed order is (delta, j, c): the largest z over the thread's vertices is its smallest-delta move, ties to
// its smaller vertex, and the target comes with it. The scan is thus ceil((K - 1) / 2) v_max3 and one add per
// vertex. Including c = a in the maximum is harmless: it gives delta 1, never an improving move. The workgroup
// minimum of ((delta - 1 + 65535) << 15) | v (32 bits, n <= 32767) is the reference's move order (delta, v); the
// winner's label and target are read from the owner's entry of an LDS table (cand[]).
//
// A move v: from -> to changes, with bit = [u is a neighbour of v] and m = 2^Q - 2^(Q+1) bit,
//   y[from] += m,  y[to] -= m        (s_from - 1, s_to + 1, G[u][from] - bit, G[u][to] + bit)
//   qa += t_a m,   t_a = [a == to] - [a == from]   (fetched per lane with ds_bpermute from a lane table)
// for every vertex u; the mover then rewrites its own label bits and qa. Lanes beyond n hold y = -2^30, qa = 0 and
// a label index (63) whose t is 0, so they never offer a move. For k > 16 the K - k empty clusters also hold
// y = -2^30: no move touches their column, and the mover's label rewrites shift it by (from - to) << b, which sums
// to the vertex's first label minus its last.
template <int T, int K, int kVpt>
__device__ __forceinline__ void DescentRegsBody(const RunDev& r, int s) {
  Ctl* ctl = r.ctl;
  constexpr int kPadY = -(1 << 30);
  constexpr int kPadLabel = 63;
  // Cache warming of the candidates' rows (below): before the barrier for k in [5, 11], after the winner's row loads
  // from k = 12 on, none up to k = 4. Measured at n = 2000 and, for k = 5, at n = 10000: before the barrier is 2.1 %
  // faster than after the row at k = 5, after the row 9 % faster than before the barrier at k >= 12, and either
  // placement is 2-6 % slower than none at k <= 4, whose moves are too short for the extra loads to pay.
  constexpr bool kWarmEarly = K >= 5 && K < 12;
  constexpr bool kWarmLate = K >= 12;
  constexpr int kLb = RegsLabelBits(K);
  constexpr int kLabelMask = (1 << kLb) - 1;
  constexpr int kSlotShift = 2 * kLb;
  constexpr int kQ = RegsQShift(K, kVpt);
  constexpr int kUnit = 1 << kQ;  // one unit of q in a key
  using Cand = std::conditional_t<(kSlotShift <= 8), uint8_t, uint16_t>;
  static_assert(K >= 3 && K <= 32 && RegsClusters(K) == K, "register descent holds 3 to 16 clusters, or 20 to 32 by 4");
  static_assert((kVpt << kSlotShift) <= kUnit && kVpt * T <= 32768, "slot index below bit Q, vertex in 15 bits");
  static_assert(kVpt <= 32, "a thread's row bits lie in one 64-bit window");
  const int n = r.n;
  const int k = r.k;  // == K up to 16; above, K - k empty clusters pad the registers
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  __shared__ uint32_t partial[2][T / kWave];
  __shared__ Cand cand[2][T];
  __shared__ int32_t src_sh;
  if (tid == 0) {
    src_sh = r.select ? Tournament(r.select_stream, s, ctl->iteration, r.pop.f, r.population, r.tournament) : s;
  }
  __syncthreads();
  const int src = __builtin_amdgcn_readfirstlane(src_sh);
  const int16_t* g_in = static_cast<const int16_t*>(r.pop.g) + static_cast<size_t>(src) * k * n;
  const uint8_t* lab_in = r.pop.labels + static_cast<size_t>(src) * n;
  const int32_t* sizes_in = r.pop.sizes + static_cast<size_t>(src) * k;
  auto live = [&](int c) { return K <= 16 || c < k; };  // c known at compile time; k uniform

  // Cluster sizes (< 2^15 on the narrow path) as 16-bit fields, four per 64-bit word, wave-uniform: a move updates
  // them with scalar shifts and adds.
  constexpr int kSizeWords = (K + 3) / 4;
  uint64_t szp[kSizeWords];
#pragma unroll
  for (int i = 0; i < kSizeWords; ++i) szp[i] = 0;
#pragma unroll
  for (int c = 0; c < K; ++c) {
    const int size = live(c) ? sizes_in[c] : 0;
    szp[c >> 2] |= static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(size)) << (16 * (c & 3));
  }
  auto size_of = [&](int c) {  // c known at compile time
    return static_cast<int>((szp[c >> 2] >> (16 * (c & 3))) & 0xFFFFu);
  };
  auto size_add = [&](int c, int64_t d) {  // c uniform
    const uint64_t inc = static_cast<uint64_t>(d) << (16 * (c & 3));
#pragma unroll
    for (int i = 0; i < kSizeWords; ++i) {
      if ((c >> 2) == i) szp[i] += inc;  // uniform
    }
  };
  int kc[K];  // (s_c << Q) + c at the start
#pragma unroll
  for (int c = 0; c < K; ++c) kc[c] = (size_of(c) << kQ) + c;

  int y[kVpt][K];
  int qa[kVpt];
  int lab4[kVpt];  // label * 4: the ds_bpermute address of the lane table
  // Branch-free: a padding slot (v >= n) loads vertex n - 1 and keeps its padding values (branches around writes
  // to the slot arrays make the compiler copy whole arrays between blocks).
  const int first = tid * kVpt;
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = first + j;
    const bool real = v < n;
    const int vc = real ? v : n - 1;
    const int a = lab_in[vc];
    int ga = 0, sa = 0;
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const int g = live(c) ? g_in[static_cast<size_t>(c) * n + vc] : 0;
      y[j][c] = real && live(c) ? (g << (kQ + 1)) - (j << kSlotShift) - (a << kLb) - kc[c] : kPadY;
      ga = c == a ? g : ga;
      sa = c == a ? size_of(c) : sa;
    }
    qa[j] = real ? (sa << kQ) - (ga << (kQ + 1)) : 0;
    lab4[j] = (real ? a : kPadLabel) * 4;
  }
  const int64_t f_start = Uniform(r.pop.f[src]);
  int64_t objective = f_start;
  // Moves counted in 32 bits and compared in scalar registers, the error written after the loop, so that every exit
  // of the move loop is a scalar branch (n <= 32767 keeps f_start, and with it the cap, below 2^30).
  const int cap = static_cast<int>(MoveCap(f_start));
  int applied = 0;
  bool capped = false;

  // The thread's key: ((delta - 1 + 65535) << 15) | v of its best improving move, or ~0.
  auto scan = [&](int parity) {
    int zmax = 0;
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      int ymax = y[j][0];
#pragma unroll
      for (int c = 1; c < K; ++c) ymax = max(ymax, y[j][c]);
      zmax = j == 0 ? ymax + qa[j] : max(zmax, ymax + qa[j]);
    }
    const int x = -zmax;  // ((delta - 1) << Q) | (j << 2b) | (a << b) | c
    cand[parity][tid] = static_cast<Cand>(x & ((1 << kSlotShift) - 1));
    if (x >= -kUnit) return ~0u;  // delta >= 0: no improving move
    const int j = (x >> kSlotShift) & ((kUnit >> kSlotShift) - 1);
    return (static_cast<uint32_t>((x >> kQ) + 65535) << 15) | static_cast<uint32_t>(first + j);
  };

  int parity = 0;
  uint32_t key = scan(parity);
  const int row_words = __builtin_amdgcn_readfirstlane(r.words);
  const uint64_t* const row_base =
      reinterpret_cast<const uint64_t*>(Uniform(static_cast<int64_t>(reinterpret_cast<uintptr_t>(r.bits))));
  const int window = first >> 5;  // the dword holding bit `first` of a row; the window's second dword follows
  const uint32_t shift = static_cast<uint32_t>(first & 31);
  uint32_t warm = 0;  // the previous move's cache-warming load, consumed one move later
  // The wave minimum through lane swaps measured 1.0 % faster per launch than through four v_readlane (k = 5,
  // n = 10000, 20 iterations: 5.89 -> 5.83 ms). A workgroup minimum through an LDS atomic (ds_min_u32 into one of
  // three rotating words, no reduction after the barrier) gained 0.8 % alone but lost 3 % together with the swaps,
  // so the per-wave partials stay.
  while (true) {
    const uint32_t wave_best_v = WaveMinSwapAll(key);
    const uint32_t wave_best = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(wave_best_v)));
    // Each wave touches the cache lines of its own best candidate's row (one dword per 128-byte line) before the
    // barrier: the winner is one of these candidates, so its row is on its way while the workgroup minimum is
    // formed, and the runners-up are often the next moves' winners. The value is consumed, without effect, one move
    // later. Vector loads return in order, so the wait for the row below also covers this load, which was issued
    // earlier. Measured at k = 5, n = 10000: 5.85 ms per launch without, 5.71 with the load issued after the row's,
    // 5.59 with it here.
    uint32_t warm_next = 0;
    if constexpr (kWarmEarly) {
      const int cand_v = wave_best == ~0u ? 0 : static_cast<int>(wave_best & 0x7FFFu);
      const GlobalWords32 cand_row = (GlobalWords32)(row_base + static_cast<size_t>(cand_v) * row_words);
      warm_next = cand_row[min(lane * 32, 2 * row_words - 1)];
    }
    const uint32_t best = BlockMinOfWaves<T / kWave>(wave_best_v, partial[parity]);
    if (best == ~0u) break;
    const int delta = static_cast<int>(best >> 15) - 65535 + 1;
    const int v = __builtin_amdgcn_readfirstlane(static_cast<int>(best & 0x7FFFu));
    // The row loads go out before the LDS read of the winner's labels, so the two latencies overlap. Bits
    // [first, first + kVpt) of the row lie in its dwords window and window + 1 ((first & 31) + kVpt <= 63); for the
    // last rows the second dword may lie in the zero tail after the last row (kRowTailBytes).
    const GlobalWords32 row = (GlobalWords32)(row_base + static_cast<size_t>(v) * row_words) + window;
    const uint32_t w0 = row[0];
    const uint32_t w1 = row[1];
    if constexpr (kWarmLate) {  // after the row loads, so that the wait for the row does not cover it
      const int cand_v = wave_best == ~0u ? v : static_cast<int>(wave_best & 0x7FFFu);
      const GlobalWords32 cand_row = (GlobalWords32)(row_base + static_cast<size_t>(cand_v) * row_words);
      warm_next = cand_row[min(lane * 32, 2 * row_words - 1)];
    }
    const int owner = v / kVpt;
    const int info = __builtin_amdgcn_readfirstlane(static_cast<int>(cand[parity][owner]));
    const int from = (info >> kLb) & kLabelMask;
    const int to = info & kLabelMask;
    parity ^= 1;
    size_add(to, 1);
    size_add(from, -1);
    // lane c of the table holds t_c = [c == to] - [c == from]. (Issuing all the lookups before the first use of
    // the row, held in place by a scheduling barrier, measured 1 % slower at k = 5, n = 10000: the row has usually
    // arrived by then, and the compiler's own interleaving of lookups and neighbour terms is the better one.)
    const int table = (lane == to) - (lane == from);
    const uint32_t bits = __builtin_amdgcn_alignbit(w1, w0, shift);
    if constexpr (kWarmEarly || kWarmLate) {
      asm volatile("" ::"v"(warm));
      warm = warm_next;
    }
    int m[kVpt];
    Unroll<kVpt>([&](auto slot) {
      constexpr int j = decltype(slot)::value;
      // Two instructions through MoveTerm against the compiler's four; measured faster from six slots per thread
      // (C2, ten slots: 1.6 %; sixteen: 12 %, with less register pressure), 2-5 % slower with up to four, where
      // the compiler's own form schedules better around the short move.
      if constexpr (kVpt >= 6) {
        m[j] = MoveTerm<j>(bits, ~(2 * kUnit - 1), kUnit);
      } else {
        m[j] = (__builtin_amdgcn_sbfe(static_cast<int>(bits), j, 1) << (kQ + 1)) + kUnit;
      }
      const int ta = __builtin_amdgcn_ds_bpermute(lab4[j], table);
      qa[j] += __mul24(ta, m[j]);  // |ta| <= 1, |m| = 2^Q: a full-rate 24-bit multiply
    });
#pragma unroll
    for (int c = 0; c < K; ++c) {
      if (c == from) {
#pragma unroll
        for (int j = 0; j < kVpt; ++j) y[j][c] += m[j];
      }
      if (c == to) {
#pragma unroll
        for (int j = 0; j < kVpt; ++j) y[j][c] -= m[j];
      }
    }
    // The moved vertex: new label bits in its y, and its own term qa = (s_to << Q) - (G_to << (Q + 1)), which is
    // -(y[to] + (j << 2b) + (to << b) + to) once y carries the new sizes and label.
    const int mv_j = v - owner * kVpt;
    if (tid == owner) {
#pragma unroll
      for (int j = 0; j < kVpt; ++j) {
        if (j == mv_j) {
          lab4[j] = to * 4;
          int y_to = 0;
#pragma unroll
          for (int c = 0; c < K; ++c) {
            y[j][c] -= (to - from) << kLb;
            if (c == to) y_to = y[j][c];
          }
          qa[j] = -(y_to + (j << kSlotShift) + (to << kLb) + to);
        }
      }
    }
    // Row words of slots beyond n are not masked: a padding lane's y (kPadY) drifts by at most 2^Q per move, so
    // it is reset every 4096 moves (at most 2^25 of drift), long before it could come near a real entry (|y| < 2^29).
    if ((applied & 4095) == 4095) {
#pragma unroll
      for (int j = 0; j < kVpt; ++j) {
        if (first + j >= n) {
#pragma unroll
          for (int c = 0; c < K; ++c) y[j][c] = kPadY;
        }
      }
    }
    key = scan(parity);
    objective += delta;
    if (++applied >= cap) {
      capped = true;
      break;
    }
  }
  if (capped && tid == 0) {
    ctl->error = kErrorMoveCap;
    ctl->stopped = 1;
  }

  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  int16_t* g_out = static_cast<int16_t*>(r.next.g) + static_cast<size_t>(s) * k * n;
#pragma unroll
  for (int c = 0; c < K; ++c) kc[c] = (size_of(c) << kQ) + c;
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = first + j;
    if (v < n) {
      const int a = lab4[j] >> 2;
      lab_out[v] = static_cast<uint8_t>(a);
#pragma unroll
      for (int c = 0; c < K; ++c) {
        if (live(c)) {
          g_out[static_cast<size_t>(c) * n + v] =
              static_cast<int16_t>((y[j][c] + kc[c] + (j << kSlotShift) + (a << kLb)) >> (kQ + 1));
        }
      }
    }
  }
  if (tid < k) {
    int size = 0;
#pragma unroll
    for (int c = 0; c < K; ++c) size = tid == c ? size_of(c) : size;
    r.next.sizes[static_cast<size_t>(s) * k + tid] = size;
  }
  if (tid == 0) {
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

// k = 3, narrow path up to n = 32767, per-vertex state packed in one pair of registers:
//   lo = G0 | G1 << 16,  hi = G2 | a << 16        (G_c = neighbours of v in cluster c < 2^15, a = label of v)
// Thread t owns the kVpt consecutive vertices v = t kVpt + j, so the row bits of a move that it needs are one
// 64-bit window of the row (two dword loads and one v_alignbit), however many vertices it holds. A move adds
// d = [c == to] - [c == from] to field c of every neighbour: one masked add per register.
//
// The scan recomputes the register kernel's key from the counts, with y'_c = (G_c << 13) - ((s_c << 12) + c):
// the largest y'_c is the best target (smallest q_c = s_c - 2 G_c, ties to the smaller c), and
//   z = max_c y'_c - y'_a - 5 a - (j << 4) = -(((q_c - q_a) << 12) | (j << 4) | (a << 2) | c),
// so the largest z over the thread's vertices is its smallest-delta move, ties to its smaller vertex. Vertices
// beyond n hold G = (24576, 4096, 4096) with a = 0: their y'_0 exceeds the other two by more than any size
// difference, so z = -(j << 4) never offers a move. Their row bits are not masked (they come from the zero tail or
// the next row), so their counts drift by at most one per move; they are reset every 1024 moves.
template <int T, int kVpt>
__device__ __forceinline__ void DescentK3PackedBody(const RunDev& r, int s) {
  constexpr int kWaves = T / kWave;
  constexpr uint32_t kPadLo = 24576u | (4096u << 16);
  constexpr uint32_t kPadHi = 4096u;
  static_assert(kVpt >= 1 && kVpt <= 32 && kVpt * T <= 32768, "slot index in 5 bits, vertex in 15 bits");
  Ctl* ctl = r.ctl;
  const int n = r.n;
  const int tid = threadIdx.x;
  __shared__ uint32_t partial[2][kWaves];
  __shared__ uint8_t cand[2][T];
  __shared__ int32_t src_sh;
  if (tid == 0) {
    src_sh = r.select ? Tournament(r.select_stream, s, ctl->iteration, r.pop.f, r.population, r.tournament) : s;
  }
  __syncthreads();
  const int src = __builtin_amdgcn_readfirstlane(src_sh);
  const int16_t* g_in = static_cast<const int16_t*>(r.pop.g) + static_cast<size_t>(src) * 3 * n;
  const uint8_t* lab_in = r.pop.labels + static_cast<size_t>(src) * n;
  const int32_t* sizes_in = r.pop.sizes + static_cast<size_t>(src) * 3;
  int size0 = __builtin_amdgcn_readfirstlane(sizes_in[0]);
  int size1 = __builtin_amdgcn_readfirstlane(sizes_in[1]);
  int size2 = __builtin_amdgcn_readfirstlane(sizes_in[2]);

  const int first = tid * kVpt;
  uint32_t lo[kVpt], hi[kVpt];
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = first + j;
    const bool real = v < n;
    const int vc = real ? v : n - 1;
    const uint32_t g0 = static_cast<uint16_t>(g_in[vc]);
    const uint32_t g1 = static_cast<uint16_t>(g_in[static_cast<size_t>(n) + vc]);
    const uint32_t g2 = static_cast<uint16_t>(g_in[2 * static_cast<size_t>(n) + vc]);
    const uint32_t a = lab_in[vc];
    lo[j] = real ? g0 | (g1 << 16) : kPadLo;
    hi[j] = real ? g2 | (a << 16) : kPadHi;
  }
  const int64_t f_start = Uniform(r.pop.f[src]);
  int64_t objective = f_start;
  // Moves counted in 32 bits and compared in scalar registers, the error written after the loop, so that every exit
  // of the move loop is a scalar branch (n <= 32767 keeps f_start, and with it the cap, below 2^30).
  const int cap = static_cast<int>(MoveCap(f_start));
  int applied = 0;
  bool capped = false;

  // The thread's key: ((delta - 1 + 65535) << 15) | v of its best improving move, or ~0.
  auto scan = [&](int parity) {
    const int y0c = -(size0 << 12);
    const int y1c = -((size1 << 12) + 1);
    const int y2c = -((size2 << 12) + 2);
    int zmax = -(1 << 30);
#pragma unroll
    for (int j = 0; j < kVpt; ++j) {
      const int y0 = static_cast<int>((lo[j] & 0xFFFFu) << 13) + y0c;
      const int y1 = static_cast<int>((lo[j] >> 16) << 13) + y1c;
      const int y2 = static_cast<int>((hi[j] & 0xFFFFu) << 13) + y2c;
      const int a = static_cast<int>(hi[j] >> 16);
      const int ya = a == 0 ? y0 : (a == 1 ? y1 : y2);
      const int z = max(max(y0, y1), y2) - ya - 5 * a - (j << 4);
      zmax = max(zmax, z);
    }
    const int x = -zmax;  // ((delta - 1) << 12) | (j << 4) | (a << 2) | c
    cand[parity][tid] = static_cast<uint8_t>(x & 15);
    if (x >= -4096) return ~0u;  // delta >= 0: no improving move
    return (static_cast<uint32_t>((x >> 12) + 65535) << 15) | static_cast<uint32_t>(first + ((x >> 4) & 31));
  };

  int parity = 0;
  uint32_t key = scan(parity);
  const int row_words = __builtin_amdgcn_readfirstlane(r.words);
  const uint64_t* const row_base =
      reinterpret_cast<const uint64_t*>(Uniform(static_cast<int64_t>(reinterpret_cast<uintptr_t>(r.bits))));
  const int window = first >> 5;  // the dword holding bit `first` of a row; the window's second dword follows
  const uint32_t shift = static_cast<uint32_t>(first & 31);
  const int lane = tid & 63;
  uint32_t warm = 0;  // the previous move's cache-warming load, consumed one move later
  while (true) {
    const uint32_t wave_best = WaveMin(key);
    // warm the lines of this wave's best candidate's row before the barrier, as in the register kernel (at
    // n = 24000 1.4 % faster than after the row loads)
    const int cand_v = wave_best == ~0u ? 0 : static_cast<int>(wave_best & 0x7FFFu);
    const GlobalWords32 cand_row = (GlobalWords32)(row_base + static_cast<size_t>(cand_v) * row_words);
    const uint32_t warm_next = cand_row[min(lane * 32, 2 * row_words - 1)];
    const uint32_t best = BlockMinOfWaves<kWaves>(wave_best, partial[parity]);
    if (best == ~0u) break;
    const int delta = static_cast<int>(best >> 15) - 65535 + 1;
    const int v = __builtin_amdgcn_readfirstlane(static_cast<int>(best & 0x7FFFu));
    // Bits [first, first + kVpt) of the row lie in its dwords window and window + 1 (first & 31 + kVpt <= 63);
    // for the last rows the second dword may lie in the zero tail after the last row (kRowTailBytes).
    const GlobalWords32 row = (GlobalWords32)(row_base + static_cast<size_t>(v) * row_words) + window;
    const uint32_t w0 = row[0];
    const uint32_t w1 = row[1];
    const int owner = v / kVpt;
    const int info = __builtin_amdgcn_readfirstlane(static_cast<int>(cand[parity][owner]));
    const int from = (info >> 2) & 3;
    const int to = info & 3;
    parity ^= 1;
    size0 += (to == 0) - (from == 0);
    size1 += (to == 1) - (from == 1);
    size2 += (to == 2) - (from == 2);
    // per-field increments of a neighbour, uniform; no field leaves [0, 2^15), so the halves never carry
    const uint32_t dlo =
        static_cast<uint32_t>((to == 0) - (from == 0)) + (static_cast<uint32_t>((to == 1) - (from == 1)) << 16);
    const uint32_t dhi = static_cast<uint32_t>((to == 2) - (from == 2));
    const uint32_t bits = __builtin_amdgcn_alignbit(w1, w0, shift);
    asm volatile("" ::"v"(warm));
    warm = warm_next;
    Unroll<kVpt>([&](auto slot) {
      constexpr int j = decltype(slot)::value;
      const uint32_t mask = static_cast<uint32_t>(SignedBit<j>(bits));
      lo[j] += mask & dlo;
      hi[j] += mask & dhi;
    });
    // the moved vertex: new label (its own counts do not change: it is not its own neighbour)
    const int mv_j = v - owner * kVpt;
    if (tid == owner) {
#pragma unroll
      for (int j = 0; j < kVpt; ++j) {
        if (j == mv_j) hi[j] = (hi[j] & 0xFFFFu) | (static_cast<uint32_t>(to) << 16);
      }
    }
    if ((applied & 1023) == 1023) {
#pragma unroll
      for (int j = 0; j < kVpt; ++j) {
        if (first + j >= n) {
          lo[j] = kPadLo;
          hi[j] = kPadHi;
        }
      }
    }
    key = scan(parity);
    objective += delta;
    if (++applied >= cap) {
      capped = true;
      break;
    }
  }
  if (capped && tid == 0) {
    ctl->error = kErrorMoveCap;
    ctl->stopped = 1;
  }

  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  int16_t* g_out = static_cast<int16_t*>(r.next.g) + static_cast<size_t>(s) * 3 * n;
#pragma unroll
  for (int j = 0; j < kVpt; ++j) {
    const int v = first + j;
    if (v < n) {
      lab_out[v] = static_cast<uint8_t>(hi[j] >> 16);
      g_out[v] = static_cast<int16_t>(lo[j] & 0xFFFFu);
      g_out[static_cast<size_t>(n) + v] = static_cast<int16_t>(lo[j] >> 16);
      g_out[2 * static_cast<size_t>(n) + v] = static_cast<int16_t>(hi[j] & 0xFFFFu);
    }
  }
  if (tid < 3) r.next.sizes[static_cast<size_t>(s) * 3 + tid] = tid == 0 ? size0 : (tid == 1 ? size1 : size2);
  if (tid == 0) {
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

template <typename Word>
using GlobalPtr = __attribute__((address_space(1))) Word*;
template <typename Word>
using LdsPtr = __attribute__((address_space(3))) Word*;
// 16 bytes of vertex words, one global_load_dwordx4 / global_store_dwordx4 (ds_read_b128 / ds_write_b128) per lane.
using Quad = uint32_t __attribute__((ext_vector_type(4)));

// k = 3 64-bit vertex words of the chunk-summary descent: G0 | G1 << 20 | G2 << 40 | label << 62 (G < 2^20). A
// neighbour update is one 64-bit add of (1 << 20 to) - (1 << 20 from): G[from] >= 1 there and G[to] + 1 <= deg, so no
// field borrows or carries. Best target of a vertex with label a: the smaller h_c = s_c - 2 G_c of the two other
// clusters, ties to the smaller c; delta = h_c - h_a + 1 (the reference's cuda/cc_local_search_k3.cuh:20-67 states
// the same algebra).
constexpr int kK3Bits = 20;
constexpr uint64_t kK3Mask = (1ull << kK3Bits) - 1;
constexpr int kK3LabelShift = 62;

// The group of vertex words that a lane of a streaming descent loads and stores at once: kV consecutive vertices in
// kBytes bytes of the slot's buffer (group g at byte kBytes * g).
//
// k = 2: four 24-bit fields f = 2 e + label + 2^23 (e = G1 - G0) in 12 bytes (one global_load_dwordx3), the 96-bit
// little-endian value f0 + f1 2^24 + f2 2^48 + f3 2^72. |2 e + label| <= 2 deg + 1 < 2^21 for n <= 2^20, so every
// field stays inside (0, 2^24) and adding to one field never carries into the next: a group's neighbour update is
// one 96-bit add of the fields' increments.
constexpr uint32_t kK2Bias = 1u << 23;
struct GroupK2 {
  static constexpr int kV = 4;
  static constexpr int kBytes = 12;
  using Word = uint32_t;
  using Inc = int;  // +4 (a neighbour of a vertex moving 0 -> 1: e + 2) or -4
  struct Regs {
    uint32_t d0, d1, d2;
  };
  __device__ static Regs Load(GlobalPtr<uint32_t> p, size_t g) {
    p += 3 * g;
    return {p[0], p[1], p[2]};
  }
  __device__ static Regs Load(LdsPtr<uint32_t> p, size_t g) {
    p += 3 * g;
    return {p[0], p[1], p[2]};
  }
  __device__ static void Store(GlobalPtr<uint32_t> p, size_t g, const Regs& r) {
    p += 3 * g;
    p[0] = r.d0;
    p[1] = r.d1;
    p[2] = r.d2;
  }
  __device__ static void Store(LdsPtr<uint32_t> p, size_t g, const Regs& r) {
    p += 3 * g;
    p[0] = r.d0;
    p[1] = r.d1;
    p[2] = r.d2;
  }
  template <int I>
  __device__ static Word Get(const Regs& r) {
    if constexpr (I == 0) return r.d0 & 0xFFFFFFu;
    if constexpr (I == 1) return __builtin_amdgcn_alignbit(r.d1, r.d0, 24) & 0xFFFFFFu;
    if constexpr (I == 2) return __builtin_amdgcn_alignbit(r.d2, r.d1, 16) & 0xFFFFFFu;
    return r.d2 >> 8;
  }
  __device__ static void Add(Regs& r, uint32_t bits, Inc inc) {
    // 4 in field i: bit 2 + 24 i of the 96-bit value
    const uint32_t a0 = ((bits & 1u) << 2) | ((bits & 2u) << 25);
    const uint32_t a1 = (bits & 4u) << 16;
    const uint32_t a2 = (bits & 8u) << 7;
    unsigned __int128 v = static_cast<unsigned __int128>(r.d0) | static_cast<unsigned __int128>(r.d1) << 32 |
                          static_cast<unsigned __int128>(r.d2) << 64;
    const unsigned __int128 a = static_cast<unsigned __int128>(a0) | static_cast<unsigned __int128>(a1) << 32 |
                                static_cast<unsigned __int128>(a2) << 64;
    v = inc > 0 ? v + a : v - a;
    r.d0 = static_cast<uint32_t>(v);
    r.d1 = static_cast<uint32_t>(v >> 32);
    r.d2 = static_cast<uint32_t>(v >> 64);
  }
  // the label bit of field i
  __device__ static void FlipLabel(Regs& r, int i) {
    if (i == 0) r.d0 ^= 1u;
    if (i == 1) r.d0 ^= 1u << 24;
    if (i == 2) r.d1 ^= 1u << 16;
    if (i == 3) r.d2 ^= 1u << 8;
  }
  __device__ static Regs Pack(const Word* f) {
    return {f[0] | f[1] << 24, f[1] >> 8 | f[2] << 16, f[2] >> 16 | f[3] << 8};
  }
};

// k = 3: two 48-bit words in 12 bytes, the 96-bit little-endian value w0 + w1 2^48. A neighbour update adds a signed
// increment to its word; no field of a word borrows or carries, so this is one 96-bit add per group. Two word
// formats (Words: K3Counts, K3Diffs) share the layout.
template <typename Words>
struct GroupK3 {
  static constexpr int kV = 2;
  static constexpr int kBytes = 12;
  using Word = uint64_t;
  using Inc = typename Words::Inc;
  using Regs = GroupK2::Regs;
  __device__ static Regs Load(GlobalPtr<uint32_t> p, size_t g) { return GroupK2::Load(p, g); }
  __device__ static Regs Load(LdsPtr<uint32_t> p, size_t g) { return GroupK2::Load(p, g); }
  __device__ static void Store(GlobalPtr<uint32_t> p, size_t g, const Regs& r) { GroupK2::Store(p, g, r); }
  __device__ static void Store(LdsPtr<uint32_t> p, size_t g, const Regs& r) { GroupK2::Store(p, g, r); }
  template <int I>
  __device__ static Word Get(const Regs& r) {
    if constexpr (I == 0) return static_cast<uint64_t>(r.d1 & 0xFFFFu) << 32 | r.d0;
    return static_cast<uint64_t>(r.d2) << 16 | r.d1 >> 16;
  }
  template <int I>
  __device__ static void Set(Regs& r, Word w) {
    if constexpr (I == 0) {
      r.d0 = static_cast<uint32_t>(w);
      r.d1 = (r.d1 & 0xFFFF0000u) | (static_cast<uint32_t>(w >> 32) & 0xFFFFu);
    } else {
      r.d1 = (r.d1 & 0xFFFFu) | static_cast<uint32_t>(w) << 16;
      r.d2 = static_cast<uint32_t>(w >> 16);
    }
  }
  // inc to the words whose bit i is set in `bits` (bits < 4)
  // Two carry chains of 32-bit adds: word 0's increment sign-extended over the three dwords, then word 1's
  // (shifted by 48 bits: dwords 1 and 2 only). Increments are below 2^31 in magnitude.
  __device__ static void Add(Regs& r, uint32_t bits, const Inc& inc) {
    const int32_t i0 = static_cast<int32_t>(Words::Increment(inc, (r.d1 >> (Words::kLabelShift - 32)) & 3u)) &
                       -static_cast<int32_t>(bits & 1u);
    const int32_t i1 = static_cast<int32_t>(Words::Increment(inc, (r.d2 >> (Words::kLabelShift - 16)) & 3u)) &
                       -static_cast<int32_t>((bits >> 1) & 1u);
    const uint32_t x0 = static_cast<uint32_t>(i0 >> 31);
    unsigned c;
    r.d0 = __builtin_addc(r.d0, static_cast<uint32_t>(i0), 0u, &c);
    r.d1 = __builtin_addc(r.d1, x0, c, &c);
    r.d2 = __builtin_addc(r.d2, x0, c, &c);
    r.d1 = __builtin_addc(r.d1, static_cast<uint32_t>(i1) << 16, 0u, &c);
    r.d2 = __builtin_addc(r.d2, static_cast<uint32_t>(i1 >> 16), c, &c);
  }
  __device__ static Regs Pack(const Word* w) {
    return {static_cast<uint32_t>(w[0]),
            (static_cast<uint32_t>(w[0] >> 32) & 0xFFFFu) | static_cast<uint32_t>(w[1]) << 16,
            static_cast<uint32_t>(w[1] >> 16)};
  }
};

// The two clusters other than a, c1 < c2.
__device__ __forceinline__ int OtherLow3(int a) { return a == 0 ? 1 : 0; }
__device__ __forceinline__ int OtherHigh3(int a) { return a == 2 ? 1 : 2; }

// Best move of a vertex with label a from its two gains over the other clusters c1 < c2, e_c = s_c - s_a + 1 -
// 2 (G_c - G_a) = delta(a -> c): the smaller, ties to c1.
__device__ __forceinline__ void Best3(int a, int e1, int e2, int* d, int* to) {
  const bool take2 = e2 < e1;
  *d = take2 ? e2 : e1;
  *to = take2 ? OtherHigh3(a) : OtherLow3(a);
}

// k = 3 words while every degree is below 2^15: G0 | G1 << 15 | G2 << 30 | label << 45. A neighbour's increment is
// (1 << 15 to) - (1 << 15 from) whatever its label (G[from] >= 1, G[to] + 1 <= deg < 2^15).
struct K3Counts {
  static constexpr int kBits = 15;
  static constexpr uint64_t kMask = (1ull << kBits) - 1;
  static constexpr int kLabelShift = 45;
  using Inc = int64_t;
  __device__ static int64_t Increment(const Inc& inc, uint32_t) { return inc; }
  __device__ static Inc IncrementFor(int from, int to) {
    return static_cast<int64_t>((1ull << (kBits * to)) - (1ull << (kBits * from)));
  }
  __device__ static uint64_t Encode(int g0, int g1, int g2, int a) {
    return static_cast<uint64_t>(g0) | static_cast<uint64_t>(g1) << kBits | static_cast<uint64_t>(g2) << (2 * kBits) |
           static_cast<uint64_t>(a) << kLabelShift;
  }
  // the counts and the label of a word (the degree is not needed)
  __device__ static int Decode(uint64_t w, int, int* g) {
    g[0] = static_cast<int>(w & kMask);
    g[1] = static_cast<int>((w >> kBits) & kMask);
    g[2] = static_cast<int>((w >> (2 * kBits)) & kMask);
    return static_cast<int>(w >> kLabelShift);
  }
  __device__ static void Best(uint64_t w, int s0, int s1, int s2, int* a, int* d, int* to) {
    *a = static_cast<int>(w >> kLabelShift);
    const int h0 = s0 - 2 * static_cast<int>(w & kMask);
    const int h1 = s1 - 2 * static_cast<int>((w >> kBits) & kMask);
    const int h2 = s2 - 2 * static_cast<int>((w >> (2 * kBits)) & kMask);
    const int ha = *a == 0 ? h0 : (*a == 1 ? h1 : h2);
    Best3(*a, (*a == 0 ? h1 : h0) - ha + 1, (*a == 2 ? h1 : h2) - ha + 1, d, to);
  }
  __device__ static uint64_t Relabel(uint64_t w, int, int to) {
    return (w & ~(3ull << kLabelShift)) | static_cast<uint64_t>(to) << kLabelShift;
  }
};

// k = 3 words for any degree: for a vertex with label a and the other clusters c1 < c2, the differences
// x1 = G_c1 - G_a and x2 = G_c2 - G_a (|x| <= deg < 2^20) as 21-bit fields biased by 2^20, and a:
// (x1 + 2^20) | (x2 + 2^20) << 21 | a << 42. delta(a -> c) = (s_c - s_a + 1) - 2 (G_c - G_a) needs only these. A
// neighbour's increment depends on its own label, one of three values per move; G is restored at the end of a
// descent with G_a = (deg - x1 - x2) / 3.
struct K3Diffs {
  static constexpr int kBits = 21;
  static constexpr uint64_t kMask = (1ull << kBits) - 1;
  static constexpr int kBias = 1 << 20;
  static constexpr int kLabelShift = 42;
  struct Inc {
    int64_t by_label[3];
  };
  __device__ static int64_t Increment(const Inc& inc, uint32_t a) {
    return a == 0 ? inc.by_label[0] : (a == 1 ? inc.by_label[1] : inc.by_label[2]);
  }
  // a move from -> to changes a neighbour's G by +1 at `to` and -1 at `from`, so its differences by
  // dG[c] - dG[a] for the neighbour's label a
  __device__ static Inc IncrementFor(int from, int to) {
    Inc inc;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int da = (a == to) - (a == from);
      const int c1 = OtherLow3(a), c2 = OtherHigh3(a);
      const int dx1 = (c1 == to) - (c1 == from) - da;
      const int dx2 = (c2 == to) - (c2 == from) - da;
      inc.by_label[a] = static_cast<int64_t>(dx1) + static_cast<int64_t>(dx2) * (int64_t(1) << kBits);
    }
    return inc;
  }
  // the word of a vertex with label a and counts (or counts less any common offset) g0, g1, g2
  __device__ static uint64_t Encode(int g0, int g1, int g2, int a) {
    const int ga = a == 0 ? g0 : (a == 1 ? g1 : g2);
    const int x1 = (a == 0 ? g1 : g0) - ga;
    const int x2 = (a == 2 ? g1 : g2) - ga;
    return static_cast<uint64_t>(x1 + kBias) | static_cast<uint64_t>(x2 + kBias) << kBits |
           static_cast<uint64_t>(a) << kLabelShift;
  }
  __device__ static int Decode(uint64_t w, int deg, int* g) {
    const int a = static_cast<int>(w >> kLabelShift);
    const int x1 = static_cast<int>(w & kMask) - kBias;
    const int x2 = static_cast<int>((w >> kBits) & kMask) - kBias;
    const int ga = (deg - x1 - x2) / 3;
    g[a] = ga;
    g[OtherLow3(a)] = ga + x1;
    g[OtherHigh3(a)] = ga + x2;
    return a;
  }
  __device__ static void Best(uint64_t w, int s0, int s1, int s2, int* a, int* d, int* to) {
    *a = static_cast<int>(w >> kLabelShift);
    const int x1 = static_cast<int>(w & kMask) - kBias;
    const int x2 = static_cast<int>((w >> kBits) & kMask) - kBias;
    const int sa = *a == 0 ? s0 : (*a == 1 ? s1 : s2);
    const int sc1 = *a == 0 ? s1 : s0;
    const int sc2 = *a == 2 ? s1 : s2;
    Best3(*a, sc1 - sa + 1 - 2 * x1, sc2 - sa + 1 - 2 * x2, d, to);
  }
  // the moved vertex's word for its new label: differences against G_to (its own G does not change)
  __device__ static uint64_t Relabel(uint64_t w, int from, int to) {
    int rel[3];
    rel[from] = 0;
    rel[OtherLow3(from)] = static_cast<int>(w & kMask) - kBias;
    rel[OtherHigh3(from)] = static_cast<int>((w >> kBits) & kMask) - kBias;
    return Encode(rel[0], rel[1], rel[2], to);
  }
};

template <typename G, typename F, int... I>
__device__ __forceinline__ void ForEachWordImpl(const typename G::Regs& r, F& f, std::integer_sequence<int, I...>) {
  (f(G::template Get<I>(r), I), ...);
}
// f(word, i) for the words i = 0 .. kV - 1 of a group, in order
template <typename G, typename F>
__device__ __forceinline__ void ForEachWord(const typename G::Regs& r, F&& f) {
  ForEachWordImpl<G>(r, f, std::make_integer_sequence<int, G::kV>{});
}

// A wave's tile of the streaming layout: kStreamLoads groups per lane, lane l of load j holding group
// tile * kTile / kV + 64 j + l, i.e. the kV consecutive vertices u = tile * kTile + (64 j + l) kV + i. Tile t
// belongs to wave t % kWaves.
template <typename G>
constexpr int kGroupTile = kWave * G::kV * kStreamLoads;

// One pass of a streaming wide-path descent over a slot's vertex words after a move (G: GroupK2 or GroupK3<>). The row
// bits of a lane's group are bits (l kV) % 32 + i of a single row dword (32 / kV lanes read the same one). Tiles
// t < `resident` live in LDS (`res`, the same layout; resident <= n / kTile, so they hold no vertex >= n, and
// resident <= kWaves * kResidentPerWave), the others in HBM (`st`). A wave first issues the row dwords of its resident
// tiles and the loads of its first streamed tile, then updates its resident tiles while those loads are in flight,
// then streams the rest, loading its next tile before it consumes the current one (two tiles of groups and row dwords
// per lane in flight behind the arithmetic). Every word of a neighbour of the moved vertex gets `inc`,
// relabel(group, i) gives the moved vertex (word i of its group) its new label (it is no neighbour of itself), a
// group with a changed word is written back, and visit(word, u) sees every vertex u < n in increasing order within
// the thread. `st` is the slot's base (padded to whole tiles: the padding words take increments and are written
// back, never visited). Row dwords of the last tile may lie beyond the row (the next row, or the zero tail after the
// last row); their bits reach only padding words.
template <int T, typename G, typename Relabel, typename Visit>
__device__ __forceinline__ void StreamMove(GlobalPtr<uint32_t> st, LdsPtr<uint32_t> res, int resident,
                                           GlobalWords32 row, int n, int wave, int lane, typename G::Inc inc,
                                           int mover, Relabel&& relabel, Visit&& visit) {
  using Regs = typename G::Regs;
  constexpr int kV = G::kV;
  constexpr int kJ = kStreamLoads;
  constexpr int kTile = kGroupTile<G>;
  constexpr int kWaves = T / kWave;
  constexpr int kRes = kResidentPerWave;
  const int tiles = (n + kTile - 1) / kTile;
  const uint32_t lane_bit = static_cast<uint32_t>(lane * kV) & 31u;
  const int lane_dword = (lane * kV) >> 5;
  // the moved vertex: its tile and load (uniform), lane and word
  const int mover_tile = mover / kTile;
  const int mover_load = (mover - mover_tile * kTile) / (kWave * kV);
  const int mover_lane = (mover / kV) % kWave;
  const int mover_word = mover % kV;
  auto group = [&](int t, int j) { return static_cast<size_t>(t) * (kTile / kV) + j * kWave + lane; };
  auto load_rows = [&](int t, uint32_t* rr) {
#pragma unroll
    for (int j = 0; j < kJ; ++j) rr[j] = row[((t * kTile + j * kWave * kV) >> 5) + lane_dword];
  };
  auto load = [&](int t, Regs* w, uint32_t* rr) {
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
      w[j] = G::Load(st, group(t, j));
      rr[j] = row[((t * kTile + j * kWave * kV) >> 5) + lane_dword];
    }
  };
  auto consume = [&](int t, const Regs* w, const uint32_t* rr, auto full, auto in_lds) {
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
      const uint32_t bits = (rr[j] >> lane_bit) & ((1u << kV) - 1u);
      Regs x = w[j];
      G::Add(x, bits, inc);
      bool changed = bits != 0;
      if (t == mover_tile && j == mover_load && lane == mover_lane) {  // once per move, in one lane
        relabel(x, mover_word);
        changed = true;
      }
      if (changed) {
        if constexpr (decltype(in_lds)::value) {
          G::Store(res, group(t, j), x);
        } else {
          G::Store(st, group(t, j), x);
        }
      }
      const int u0 = t * kTile + (j * kWave + lane) * kV;
      ForEachWord<G>(x, [&](typename G::Word word, int i) {
        if (decltype(full)::value || u0 + i < n) visit(word, u0 + i);
      });
    }
  };
  uint32_t rres[kRes][kJ];
#pragma unroll
  for (int i = 0; i < kRes; ++i) {
    if (wave + i * kWaves < resident) load_rows(wave + i * kWaves, rres[i]);
  }
  // this wave's first streamed tile: the smallest t >= resident with t % kWaves == wave
  int t = wave + kWaves * ((max(resident - wave, 0) + kWaves - 1) / kWaves);
  Regs cw[kJ], nw[kJ];
  uint32_t cr[kJ], nr[kJ];
  if (t < tiles) load(t, cw, cr);
#pragma unroll
  for (int i = 0; i < kRes; ++i) {
    const int tr = wave + i * kWaves;
    if (tr < resident) {
      Regs w[kJ];
#pragma unroll
      for (int j = 0; j < kJ; ++j) w[j] = G::Load(res, group(tr, j));
      consume(tr, w, rres[i], std::true_type{}, std::true_type{});
    }
  }
  for (; t < tiles; t += kWaves) {
    if (t + kWaves < tiles) load(t + kWaves, nw, nr);
    if ((t + 1) * kTile <= n) {  // wave-uniform
      consume(t, cw, cr, std::true_type{}, std::false_type{});
    } else {
      consume(t, cw, cr, std::false_type{}, std::false_type{});
    }
#pragma unroll
    for (int j = 0; j < kJ; ++j) {
      cw[j] = nw[j];
      cr[j] = nr[j];
    }
  }
}

// The groups of a slot's streaming state that this thread owns (StreamMove's layout), in increasing order:
// visit(g, u0) with the group's first vertex u0 < n (its other vertices may lie beyond n).
template <int T, typename G, typename Visit>
__device__ __forceinline__ void ForEachOwnedGroup(int n, int wave, int lane, Visit&& visit) {
  constexpr int kTile = kGroupTile<G>;
  for (int t = wave; t * kTile < n; t += T / kWave) {
#pragma unroll
    for (int j = 0; j < kStreamLoads; ++j) {
      const int u0 = t * kTile + (j * kWave + lane) * G::kV;
      if (u0 < n) visit(static_cast<size_t>(u0 / G::kV), u0);
    }
  }
}

// Wide path, k = 2: one 24-bit field per vertex (GroupK2), f = 2 e + label + 2^23 with e = G1 - G0, in the run's
// work buffer (reference algebra cuda/cc_local_search_k2.cuh:251-254, w = f - 2^23): delta = sv - w for label 0 and
// w + 1 - sv for label 1; a neighbour of a vertex moving 0 -> 1 gains 4 (e + 2), 1 -> 0 loses 4. G is restored from
// w and the degrees at the end.
template <int T>
__device__ __forceinline__ void DescentK2WideBody(const RunDev& r, int s, int resident_tiles) {
  using G = GroupK2;
  Ctl* ctl = r.ctl;
  constexpr int kWaves = T / kWave;
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
  const int src = src_sh;
  const int32_t* g_in = static_cast<const int32_t*>(r.pop.g) + static_cast<size_t>(src) * 2 * n;
  const uint8_t* lab_in = r.pop.labels + static_cast<size_t>(src) * n;
  const GlobalPtr<uint32_t> st =
      (GlobalPtr<uint32_t>)(reinterpret_cast<uint8_t*>(r.work) + static_cast<size_t>(s) * WideSlotBytes(n, 2));
  // the slot's first whole tiles in LDS (the launch provides resident_tiles tiles of it)
  const int resident = min(resident_tiles, min(n / kGroupTile<G>, kWaves * kResidentPerWave));
  const size_t res_groups = static_cast<size_t>(resident) * (kGroupTile<G> / G::kV);
  extern __shared__ __align__(16) uint8_t dyn[];
  const LdsPtr<uint32_t> res = (LdsPtr<uint32_t>)dyn;
  int n0 = r.pop.sizes[static_cast<size_t>(src) * 2];
  const int64_t f_start = r.pop.f[src];

  int sv_bias = n - 2 * n0 + 1 + static_cast<int>(kK2Bias);  // sv + 2^23
  int best_d = 0, best_u = -1;
  uint32_t best_f = 0;
  auto consider = [&](uint32_t f, int u) {
    const int x = static_cast<int>(f);
    const int d = (f & 1u) ? x + 1 - sv_bias : sv_bias - x;
    if (d < best_d) {
      best_d = d;
      best_u = u;
      best_f = f;
    }
  };
  auto key_of = [&]() {
    return best_u < 0 ? kNoMove
                      : PackMove(best_d, best_u, static_cast<int>(best_f & 1u), static_cast<int>(best_f & 1u) ^ 1);
  };
  ForEachOwnedGroup<T, G>(n, wave, lane, [&](size_t g, int u0) {
    uint32_t f[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int u = u0 + i;
      f[i] = u < n ? static_cast<uint32_t>(2 * (g_in[n + u] - g_in[u]) + (lab_in[u] & 1) + static_cast<int>(kK2Bias))
                   : kK2Bias;
      if (u < n) consider(f[i], u);
    }
    if (g < res_groups) {
      G::Store(res, g, G::Pack(f));
    } else {
      G::Store(st, g, G::Pack(f));
    }
  });
  uint64_t key = key_of();
  int64_t objective = f_start;
  int64_t applied = 0;
  const int64_t cap = MoveCap(f_start);
  int parity = 0;
  const uint64_t* const row_base =
      reinterpret_cast<const uint64_t*>(Uniform(static_cast<int64_t>(reinterpret_cast<uintptr_t>(r.bits))));
  const int row_words = __builtin_amdgcn_readfirstlane(r.words);
  bool capped = false;
  while (true) {
    const uint64_t best = BlockMin<kWaves>(key, partial[parity]);
    parity ^= 1;
    if (best == kNoMove) break;
    const int delta = static_cast<int>(static_cast<uint32_t>(best >> 32)) - kDeltaOffset;
    const uint32_t mv = static_cast<uint32_t>(best);
    const int v = static_cast<int>(mv >> 12);
    const int from = static_cast<int>((mv >> 6) & 63u);
    n0 += from ? 1 : -1;
    sv_bias = n - 2 * n0 + 1 + static_cast<int>(kK2Bias);
    const GlobalWords32 row = (GlobalWords32)(row_base + static_cast<size_t>(v) * row_words);
    best_d = 0;
    best_u = -1;
    StreamMove<T, G>(st, res, resident, row, n, wave, lane, from ? -4 : 4, v,
                     [](G::Regs& x, int i) { G::FlipLabel(x, i); }, consider);
    key = key_of();
    objective += delta;
    if (++applied >= cap) {
      capped = true;
      break;
    }
  }
  if (capped && tid == 0) {
    ctl->error = kErrorMoveCap;
    ctl->stopped = 1;
  }
  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  int32_t* g_out = static_cast<int32_t*>(r.next.g) + static_cast<size_t>(s) * 2 * n;
  ForEachOwnedGroup<T, G>(n, wave, lane, [&](size_t g, int u0) {
    const G::Regs x = g < res_groups ? G::Load(res, g) : G::Load(st, g);
    ForEachWord<G>(x, [&](uint32_t f, int i) {
      const int u = u0 + i;
      if (u < n) {
        const int w = static_cast<int>(f) - static_cast<int>(kK2Bias);
        const int e = w >> 1;
        const int degree = r.deg[u];
        lab_out[u] = static_cast<uint8_t>(w & 1);
        g_out[u] = (degree - e) / 2;
        g_out[n + u] = (degree + e) / 2;
      }
    });
  });
  if (tid == 0) {
    r.next.sizes[static_cast<size_t>(s) * 2] = n0;
    r.next.sizes[static_cast<size_t>(s) * 2 + 1] = n - n0;
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

// Wide path, k = 3: one 48-bit word per vertex (GroupK3<Words>: K3Counts while every degree is below 2^15, else
// K3Diffs) in the run's work buffer.
template <int T, typename Words>
__device__ __forceinline__ void DescentK3WideBody(const RunDev& r, int s, int resident_tiles) {
  using G = GroupK3<Words>;
  Ctl* ctl = r.ctl;
  constexpr int kWaves = T / kWave;
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
  const int src = src_sh;
  const int32_t* g_in = static_cast<const int32_t*>(r.pop.g) + static_cast<size_t>(src) * 3 * n;
  const uint8_t* lab_in = r.pop.labels + static_cast<size_t>(src) * n;
  const int32_t* sizes_in = r.pop.sizes + static_cast<size_t>(src) * 3;
  const GlobalPtr<uint32_t> st =
      (GlobalPtr<uint32_t>)(reinterpret_cast<uint8_t*>(r.work) + static_cast<size_t>(s) * WideSlotBytes(n, 3));
  // the slot's first whole tiles in LDS (the launch provides resident_tiles tiles of it)
  const int resident = min(resident_tiles, min(n / kGroupTile<G>, kWaves * kResidentPerWave));
  const size_t res_groups = static_cast<size_t>(resident) * (kGroupTile<G> / G::kV);
  extern __shared__ __align__(16) uint8_t dyn[];
  const LdsPtr<uint32_t> res = (LdsPtr<uint32_t>)dyn;
  int s0 = sizes_in[0], s1 = sizes_in[1], s2 = sizes_in[2];
  const int64_t f_start = r.pop.f[src];

  int best_d = 0, best_move = -1;
  auto consider = [&](uint64_t w, int u) {
    int a, d, to;
    Words::Best(w, s0, s1, s2, &a, &d, &to);
    if (d < best_d) {
      best_d = d;
      best_move = (u << 4) | (a << 2) | to;
    }
  };
  auto key_of = [&]() {
    return best_move < 0 ? kNoMove : PackMove(best_d, best_move >> 4, (best_move >> 2) & 3, best_move & 3);
  };
  ForEachOwnedGroup<T, G>(n, wave, lane, [&](size_t g, int u0) {
    uint64_t w[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int u = u0 + i;
      w[i] = u < n ? Words::Encode(g_in[u], g_in[n + u], g_in[2 * n + u], lab_in[u]) : 0;
      if (u < n) consider(w[i], u);
    }
    if (g < res_groups) {
      G::Store(res, g, G::Pack(w));
    } else {
      G::Store(st, g, G::Pack(w));
    }
  });
  uint64_t key = key_of();
  int64_t objective = f_start;
  int64_t applied = 0;
  const int64_t cap = MoveCap(f_start);
  int parity = 0;
  const uint64_t* const row_base =
      reinterpret_cast<const uint64_t*>(Uniform(static_cast<int64_t>(reinterpret_cast<uintptr_t>(r.bits))));
  const int row_words = __builtin_amdgcn_readfirstlane(r.words);
  bool capped = false;
  while (true) {
    const uint64_t best = BlockMin<kWaves>(key, partial[parity]);
    parity ^= 1;
    if (best == kNoMove) break;
    const int delta = static_cast<int>(static_cast<uint32_t>(best >> 32)) - kDeltaOffset;
    const uint32_t mv = static_cast<uint32_t>(best);
    const int v = static_cast<int>(mv >> 12);
    const int from = static_cast<int>((mv >> 6) & 63u);
    const int to = static_cast<int>(mv & 63u);
    s0 += (to == 0) - (from == 0);
    s1 += (to == 1) - (from == 1);
    s2 += (to == 2) - (from == 2);
    const GlobalWords32 row = (GlobalWords32)(row_base + static_cast<size_t>(v) * row_words);
    best_d = 0;
    best_move = -1;
    StreamMove<T, G>(
        st, res, resident, row, n, wave, lane, Words::IncrementFor(from, to), v,
        [from, to](typename G::Regs& x, int i) {
          if (i == 0) {
            G::template Set<0>(x, Words::Relabel(G::template Get<0>(x), from, to));
          } else {
            G::template Set<1>(x, Words::Relabel(G::template Get<1>(x), from, to));
          }
        },
        consider);
    key = key_of();
    objective += delta;
    if (++applied >= cap) {
      capped = true;
      break;
    }
  }
  if (capped && tid == 0) {
    ctl->error = kErrorMoveCap;
    ctl->stopped = 1;
  }
  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  int32_t* g_out = static_cast<int32_t*>(r.next.g) + static_cast<size_t>(s) * 3 * n;
  ForEachOwnedGroup<T, G>(n, wave, lane, [&](size_t g, int u0) {
    const typename G::Regs x = g < res_groups ? G::Load(res, g) : G::Load(st, g);
    ForEachWord<G>(x, [&](uint64_t w, int i) {
      const int u = u0 + i;
      if (u < n) {
        int gc[3];
        lab_out[u] = static_cast<uint8_t>(Words::Decode(w, r.deg[u], gc));
        g_out[u] = gc[0];
        g_out[n + u] = gc[1];
        g_out[2 * n + u] = gc[2];
      }
    });
  });
  if (tid == 0) {
    int32_t* sizes = r.next.sizes + static_cast<size_t>(s) * 3;
    sizes[0] = s0;
    sizes[1] = s1;
    sizes[2] = s2;
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

// Wide path, k = 3, with chunk summaries. The vertices are cut into chunks of 64 (one adjacency word, one wave).
// For every chunk and every pair (label a, target c != a) a table in LDS holds the chunk's best candidate of the
// pair,
//   key = ((G_c - G_a + 2^20) << 6) | (63 - u % 64),   the largest G_c - G_a, ties to the smaller vertex u,
// or 0 if no vertex of the chunk has label a. Since delta(u, a -> c) = (s_c - s_a + 1) - 2 (G_c - G_a) and the
// sizes are the same for the whole pair, the pair's best move is its largest key over the chunks (the first chunk
// on ties), and the move of the descent (smallest delta, then vertex, then target; src/cc_pbils_cpu.cpp:99-118) is
// the smallest PackMove over the six pairs' best keys, the sizes applied there. A move of v changes G only for v's
// neighbours and the label only of v, so only the chunks of the non-zero words of v's row and v's own chunk are
// recomputed: a few chunks per move on a sparse graph instead of all n vertices. The packed vertex words are those
// of KernelDescentK3Wide. Pairs: p = 2 a + i, c = the i-th cluster other than a.
constexpr int kPairs3 = 6;

__device__ __forceinline__ constexpr int PairTarget3(int p) {
  return (p >> 1) == 0 ? 1 + (p & 1) : ((p >> 1) == 1 ? 2 * (p & 1) : (p & 1));
}

__device__ __forceinline__ uint32_t WaveMax32(uint32_t x) { return ~WaveMin(~x); }

template <int T>
__device__ __forceinline__ void DescentK3ChunkBody(const RunDev& r, int s) {
  Ctl* ctl = r.ctl;
  constexpr int kWaves = T / kWave;
  const int n = r.n;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = WaveIndex();
  const int chunks = (n + 63) >> 6;
  constexpr int kRows = kWave / 16;  // dirty chunks a wave recomputes at once
  const int row_in_wave = lane >> 4;
  const int quad = lane & 15;
  __shared__ uint64_t partial[2][kWaves];
  __shared__ int32_t dirty_count[2];
  __shared__ int32_t src_sh;
  extern __shared__ __align__(16) uint8_t dyn[];
  uint32_t* table = reinterpret_cast<uint32_t*>(dyn);               // [chunks][kPairs3]
  int32_t* dirty = reinterpret_cast<int32_t*>(table + kPairs3 * chunks);  // [2][chunks]
  if (tid == 0) {
    src_sh = r.select ? Tournament(r.select_stream, s, ctl->iteration, r.pop.f, r.population, r.tournament) : s;
    dirty_count[0] = 0;
    dirty_count[1] = 0;
  }
  __syncthreads();
  const int src = src_sh;
  const int32_t* g_in = static_cast<const int32_t*>(r.pop.g) + static_cast<size_t>(src) * 3 * n;
  const uint8_t* lab_in = r.pop.labels + static_cast<size_t>(src) * n;
  const int32_t* sizes_in = r.pop.sizes + static_cast<size_t>(src) * 3;
  const GlobalPtr<uint64_t> st = (GlobalPtr<uint64_t>)(r.work + static_cast<size_t>(s) * WideStride(n, 3));
  // the three cluster sizes (< 2^21) as 21-bit fields of one uniform word: a move updates them with scalar adds
  uint64_t szp = static_cast<uint64_t>(sizes_in[0]) | static_cast<uint64_t>(sizes_in[1]) << 21 |
                 static_cast<uint64_t>(sizes_in[2]) << 42;
  szp = static_cast<uint64_t>(Uniform(static_cast<int64_t>(szp)));
  auto size_of = [&](int c) { return static_cast<int>((szp >> (21 * c)) & 0x1FFFFFu); };  // c known at compile time
  const int64_t f_start = Uniform(r.pop.f[src]);

  // The pair keys of a chunk from its lanes' words, into the table: the pairs whose source label is in the bit
  // set `labels` (uniform); the entries of the other pairs stay as they are.
  auto summarize = [&](int chunk, uint64_t w, bool valid, int labels) {
    const int a = valid ? static_cast<int>(w >> kK3LabelShift) : 3;
    const int g[3] = {static_cast<int>(w & kK3Mask), static_cast<int>((w >> kK3Bits) & kK3Mask),
                      static_cast<int>((w >> (2 * kK3Bits)) & kK3Mask)};
    const int ga = a == 0 ? g[0] : (a == 1 ? g[1] : g[2]);
    const uint32_t tie = static_cast<uint32_t>(63 - lane);
#pragma unroll
    for (int p = 0; p < kPairs3; ++p) {
      if (!((labels >> (p >> 1)) & 1)) continue;  // uniform
      const uint32_t key =
          a == (p >> 1) ? (static_cast<uint32_t>(g[PairTarget3(p)] - ga + (1 << kK3Bits)) << 6) | tie : 0u;
      const uint32_t best = WaveMax32(key);
      if (lane == 0) table[chunk * kPairs3 + p] = best;
    }
  };

  for (int chunk = wave; chunk < chunks; chunk += kWaves) {
    const int u = chunk * kWave + lane;
    uint64_t w = 0;
    if (u < n) {
      w = static_cast<uint64_t>(g_in[u]) | static_cast<uint64_t>(g_in[n + u]) << kK3Bits |
          static_cast<uint64_t>(g_in[2 * n + u]) << (2 * kK3Bits) | static_cast<uint64_t>(lab_in[u]) << kK3LabelShift;
      st[u] = w;
    }
    summarize(chunk, w, u < n, 7);
  }

  // The thread's candidate: over its chunks of the table, each pair's best entry, as moves. Across chunks only the
  // gain (key >> 6) is compared and the first chunk wins a tie: the low bits order vertices within one chunk only,
  // and every vertex of an earlier chunk is smaller than every vertex of a later one.
  auto candidate = [&]() {
    uint32_t m[kPairs3];
    int mc[kPairs3];
#pragma unroll
    for (int p = 0; p < kPairs3; ++p) {
      m[p] = 0;
      mc[p] = 0;
    }
    // one pass over the thread's chunks reads the six entries of each (three 8-byte reads)
#pragma unroll 2
    for (int c = tid; c < chunks; c += T) {
      const uint2* entry = reinterpret_cast<const uint2*>(table + c * kPairs3);
      const uint2 e01 = entry[0], e23 = entry[1], e45 = entry[2];
      const uint32_t k[kPairs3] = {e01.x, e01.y, e23.x, e23.y, e45.x, e45.y};
#pragma unroll
      for (int p = 0; p < kPairs3; ++p) {
        if ((k[p] >> 6) > (m[p] >> 6)) {
          m[p] = k[p];
          mc[p] = c;
        }
      }
    }
    uint64_t best = kNoMove;
#pragma unroll
    for (int p = 0; p < kPairs3; ++p) {
      if (m[p]) {
        const int a = p >> 1, to = PairTarget3(p);
        const int x = static_cast<int>(m[p] >> 6) - (1 << kK3Bits);
        const int d = (size_of(to) - size_of(a) + 1) - 2 * x;
        if (d < 0) {
          const int u = mc[p] * kWave + 63 - static_cast<int>(m[p] & 63u);
          best = Min(best, PackMove(d, u, a, to));
        }
      }
    }
    return best;
  };

  __syncthreads();  // the table is complete
  uint64_t key = candidate();
  int64_t objective = f_start;
  int64_t applied = 0;
  const int64_t cap = MoveCap(f_start);
  int parity = 0;
  const uint64_t* const row_base =
      reinterpret_cast<const uint64_t*>(Uniform(static_cast<int64_t>(reinterpret_cast<uintptr_t>(r.bits))));
  const int row_words = __builtin_amdgcn_readfirstlane(r.words);
  while (true) {
    const uint64_t best = BlockMin<kWaves>(key, partial[parity]);
    if (best == kNoMove) break;
    const int delta = static_cast<int>(static_cast<uint32_t>(best >> 32)) - kDeltaOffset;
    const uint32_t mv = static_cast<uint32_t>(best);
    const int v = static_cast<int>(mv >> 12);
    const int from = static_cast<int>((mv >> 6) & 63u);
    const int to = static_cast<int>(mv & 63u);
    szp += (uint64_t(1) << (21 * to)) - (uint64_t(1) << (21 * from));
    const uint64_t inc = (1ull << (kK3Bits * to)) - (1ull << (kK3Bits * from));
    const uint64_t* row = row_base + static_cast<size_t>(v) * row_words;
    // the chunks to recompute: every non-zero word of v's row, and v's own chunk
    int32_t* list = dirty + parity * chunks;
    for (int c = tid; c < chunks; c += T) {
      if (row[c] != 0ull || c == (v >> 6)) list[atomicAdd(&dirty_count[parity], 1)] = c;
    }
    __syncthreads();
    const int count = dirty_count[parity];
    // Each row of 16 lanes takes one dirty chunk, so a wave recomputes four chunks at once: lane q of the row holds
    // the chunk's vertices 4 q .. 4 q + 3 (two 16-byte loads), and a pair's chunk maximum is one DPP row maximum
    // shared by the four chunks. The pairs recomputed are those of the labels changed in any of the four chunks (a
    // pair recomputed without need gets its unchanged value).
    for (int base = wave * kRows; base < count; base += kWaves * kRows) {  // wave-uniform
      const int i = base + row_in_wave;
      const bool active = i < count;
      const int chunk = active ? list[i] : 0;
      const uint32_t bits = active ? static_cast<uint32_t>(row[chunk] >> (4 * quad)) & 0xFu : 0u;
      const int u0 = chunk * kWave + 4 * quad;
      const GlobalPtr<Quad> sq = (GlobalPtr<Quad>)(st + u0);
      uint64_t w[4] = {0, 0, 0, 0};
      if (active) {
        const Quad lo = sq[0], hi = sq[1];
        __builtin_memcpy(&w[0], &lo, 16);
        __builtin_memcpy(&w[2], &hi, 16);
      }
      int changed = 0;  // labels whose pairs change: a neighbour keeps its label, the mover leaves `from` for `to`
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const bool valid = active && u0 + e < n;
        if (valid && ((bits >> e) & 1u)) {
          w[e] += inc;
          changed |= 1 << static_cast<int>(w[e] >> kK3LabelShift);
        }
        if (valid && u0 + e == v) {
          w[e] = (w[e] & ~(3ull << kK3LabelShift)) | (static_cast<uint64_t>(to) << kK3LabelShift);
          changed |= (1 << from) | (1 << to);
        }
      }
      if (changed) {
        Quad lo, hi;
        __builtin_memcpy(&lo, &w[0], 16);
        __builtin_memcpy(&hi, &w[2], 16);
        sq[0] = lo;
        sq[1] = hi;
      }
      int labels = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) labels |= __ballot((changed >> c) & 1) ? 1 << c : 0;
#pragma unroll
      for (int p = 0; p < kPairs3; ++p) {
        if (!((labels >> (p >> 1)) & 1)) continue;  // uniform
        const int src = p >> 1, tgt = PairTarget3(p);
        uint32_t best = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int a = static_cast<int>(w[e] >> kK3LabelShift);
          const int ga = static_cast<int>((w[e] >> (kK3Bits * src)) & kK3Mask);
          const int gc = static_cast<int>((w[e] >> (kK3Bits * tgt)) & kK3Mask);
          const uint32_t key = (static_cast<uint32_t>(gc - ga + (1 << kK3Bits)) << 6) |
                               static_cast<uint32_t>(63 - 4 * quad - e);
          best = (active && u0 + e < n && a == src) ? max(best, key) : best;
        }
        best = RowMax(best);
        if (active && quad == 0) table[chunk * kPairs3 + p] = best;
      }
    }
    // the other list's counter was last read before this move's first barrier
    if (tid == 0) dirty_count[parity ^ 1] = 0;
    parity ^= 1;
    __syncthreads();  // the table is current
    key = candidate();
    objective += delta;
    if (++applied >= cap) {
      if (tid == 0) {
        ctl->error = kErrorMoveCap;
        ctl->stopped = 1;
      }
      break;
    }
  }
  uint8_t* lab_out = r.next.labels + static_cast<size_t>(s) * n;
  int32_t* g_out = static_cast<int32_t*>(r.next.g) + static_cast<size_t>(s) * 3 * n;
  for (int u = tid; u < n; u += T) {
    const uint64_t w = st[u];
    g_out[u] = static_cast<int32_t>(w & kK3Mask);
    g_out[n + u] = static_cast<int32_t>((w >> kK3Bits) & kK3Mask);
    g_out[2 * n + u] = static_cast<int32_t>((w >> (2 * kK3Bits)) & kK3Mask);
    lab_out[u] = static_cast<uint8_t>(w >> kK3LabelShift);
  }
  if (tid == 0) {
    int32_t* sizes = r.next.sizes + static_cast<size_t>(s) * 3;
    sizes[0] = size_of(0);
    sizes[1] = size_of(1);
    sizes[2] = size_of(2);
    r.next.f[s] = objective;
    r.moves[s] = applied;
  }
}

constexpr size_t kLdsBytes = 160 * 1024;
// Static LDS of the descent kernels (partials, sizes, source slot, claim), rounded up.
constexpr size_t kStaticLdsReserve = 2048;

// A persistent workgroup: claims items (run, slot) from the launch's queue until none is left and runs the descent
// of each; items of stopped runs are skipped. The barrier before each claim keeps the shared state of the previous
// item alive until every wave is done with it.
//
// Another workgroup of the same run may set ctl->stopped in the middle of this launch (a descent that reaches its
// move cap), so the flag is read once, by the thread that claims the item, and handed to the whole workgroup through
// LDS: every wave then takes the same branch and meets the same barriers.
template <typename Body>
__device__ __forceinline__ void Persistent(const RunDev* runs, const ItemList& list, Body&& body) {
  __shared__ int32_t claimed;
  __shared__ int32_t skip;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const int index = atomicAdd(list.queue, 1);
      claimed = index;
      skip = index < list.count
                 ? __hip_atomic_load(&runs[list.items[index].run].ctl->stopped, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT)
                 : 1;
    }
    __syncthreads();
    const int index = claimed;
    if (index >= list.count) break;
    if (skip) continue;
    const Item it = list.items[index];
    body(runs[it.run], it.slot);
  }
}

template <int T, typename GT, bool kLds>
__global__ void __launch_bounds__(T) KernelDescentGeneric(const RunDev* runs, ItemList list) {
  Persistent(runs, list, [](const RunDev& r, int s) { DescentGenericBody<T, GT, kLds>(r, s); });
}

template <int T, int kVpt, bool kRowsLds>
__global__ void __launch_bounds__(T) KernelDescentK2(const RunDev* runs, ItemList list) {
  Persistent(runs, list, [](const RunDev& r, int s) { DescentK2Body<T, kVpt, kRowsLds>(r, s); });
}

template <int T, int K, int kVpt>
__global__ void __launch_bounds__(T) KernelDescentRegs(const RunDev* runs, ItemList list) {
  Persistent(runs, list, [](const RunDev& r, int s) { DescentRegsBody<T, K, kVpt>(r, s); });
}

template <int T, int kVpt>
__global__ void __launch_bounds__(T) KernelDescentK3Packed(const RunDev* runs, ItemList list) {
  Persistent(runs, list, [](const RunDev& r, int s) { DescentK3PackedBody<T, kVpt>(r, s); });
}

template <int T>
__global__ void __launch_bounds__(T) KernelDescentK2Wide(const RunDev* runs, ItemList list, int resident_tiles) {
  Persistent(runs, list, [resident_tiles](const RunDev& r, int s) { DescentK2WideBody<T>(r, s, resident_tiles); });
}

// The largest degree of a run's graph, in every thread of the workgroup.
template <int T>
__device__ __forceinline__ int BlockMaxDegree(const RunDev& r) {
  __shared__ int32_t max_sh;
  if (threadIdx.x == 0) max_sh = 0;
  __syncthreads();
  int m = 0;
  for (int u = threadIdx.x; u < r.n; u += T) m = max(m, r.deg[u]);
  m = static_cast<int>(~WaveMin(~static_cast<uint32_t>(m)));
  if ((threadIdx.x & 63) == 0) atomicMax(&max_sh, m);
  __syncthreads();
  return max_sh;
}

template <int T>
__global__ void __launch_bounds__(T) KernelDescentK3Wide(const RunDev* runs, ItemList list, int resident_tiles) {
  Persistent(runs, list, [resident_tiles](const RunDev& r, int s) {
    if (BlockMaxDegree<T>(r) < (1 << K3Counts::kBits)) {
      DescentK3WideBody<T, K3Counts>(r, s, resident_tiles);
    } else {
      DescentK3WideBody<T, K3Diffs>(r, s, resident_tiles);
    }
  });
}

// The launch passes the dynamic LDS it allocated: an item whose table and lists would not fit stops its run with an
// error instead of reaching beyond the allocation.
template <int T>
__global__ void __launch_bounds__(T) KernelDescentK3Chunk(const RunDev* runs, ItemList list, int lds_bytes) {
  Persistent(runs, list, [lds_bytes](const RunDev& r, int s) {
    if (K3ChunkLdsBytes(r.n) <= static_cast<size_t>(lds_bytes)) {
      DescentK3ChunkBody<T>(r, s);
    } else if (threadIdx.x == 0) {
      r.ctl->error = kErrorLds;
      r.ctl->stopped = 1;
    }
  });
}

// The mixed descent: one persistent launch for a batch whose runs use different register kernels. The queue holds
// the items of all of them, largest runs first, and each item runs the variant its run names (RunDev::variant), so
// no CU idles while another kernel instance's launch drains its tail. Three block sizes: 1024 threads (one
// workgroup per CU) for any register run, and 512 and 256 threads held to 128 VGPRs (two and four workgroups per
// CU, so one descent's barrier and row-load latency overlaps the others' work) for runs up to 4096 and 2048
// vertices.
struct MixedDesc {
  int kind;  // 0: k = 2 registers, 1: k in [3, 8] registers
  int k;
  int vpt;
};
constexpr MixedDesc kMixedTable[] = {
    {0, 2, 1},  {0, 2, 2},  {0, 2, 4},  {0, 2, 8},  {0, 2, 16}, {0, 2, 32},             //
    {1, 3, 1},  {1, 3, 2},  {1, 3, 4},  {1, 3, 8},  {1, 3, 12},                          //
    {1, 4, 1},  {1, 4, 2},  {1, 4, 4},  {1, 4, 8},  {1, 4, 12},                          //
    {1, 5, 1},  {1, 5, 2},  {1, 5, 4},  {1, 5, 8},                                       //
    {1, 6, 1},  {1, 6, 2},  {1, 6, 4},  {1, 6, 8},                                       //
    {1, 7, 1},  {1, 7, 2},  {1, 7, 4},  {1, 7, 8},                                       //
    {1, 8, 1},  {1, 8, 2},  {1, 8, 4},  {1, 8, 8},                                       //
};
constexpr int kMixedCount = sizeof(kMixedTable) / sizeof(kMixedTable[0]);
constexpr int kMixedThreads = 1024;
constexpr MixedDesc kMixedTable512[] = {
    {0, 2, 1}, {0, 2, 2}, {0, 2, 4}, {0, 2, 8},  //
    {1, 3, 1}, {1, 3, 2}, {1, 3, 4}, {1, 3, 8},  //
    {1, 4, 1}, {1, 4, 2}, {1, 4, 4}, {1, 4, 8},  //
    {1, 5, 1}, {1, 5, 2}, {1, 5, 4}, {1, 5, 8},  //
    {1, 6, 1}, {1, 6, 2}, {1, 6, 4}, {1, 6, 8},  //
    {1, 7, 1}, {1, 7, 2}, {1, 7, 4}, {1, 7, 8},  //
    {1, 8, 1}, {1, 8, 2}, {1, 8, 4}, {1, 8, 8},  //
};
constexpr int kMixedCount512 = sizeof(kMixedTable512) / sizeof(kMixedTable512[0]);
constexpr int kMixedThreads512 = 512;
// the same variants with 256 threads hold n <= 2048
constexpr const MixedDesc* kMixedTable256 = kMixedTable512;
constexpr int kMixedCount256 = kMixedCount512;
constexpr int kMixedThreads256 = 256;

template <int T, const MixedDesc* kTable, int I>
__device__ __forceinline__ void MixedOne(const RunDev& r, int s) {
  constexpr MixedDesc d = kTable[I];
  if constexpr (d.kind == 0) {
    DescentK2Body<T, d.vpt, false>(r, s);
  } else {
    DescentRegsBody<T, d.k, d.vpt>(r, s);
  }
}

template <int T, const MixedDesc* kTable, int... I>
__device__ __forceinline__ void MixedSwitch(const RunDev& r, int s, std::integer_sequence<int, I...>) {
  const int v = r.variant;
  ((v == I ? (MixedOne<T, kTable, I>(r, s), 0) : 0), ...);
}

__global__ void __launch_bounds__(kMixedThreads) KernelDescentMixed(const RunDev* runs, ItemList list) {
  Persistent(runs, list, [](const RunDev& r, int s) {
    MixedSwitch<kMixedThreads, kMixedTable>(r, s, std::make_integer_sequence<int, kMixedCount>{});
  });
}

// Two (four) workgroups per CU: 16 waves, 4 per SIMD, so at most 128 VGPRs each.
__global__ void __launch_bounds__(kMixedThreads512) __attribute__((amdgpu_waves_per_eu(4)))
KernelDescentMixed512(const RunDev* runs, ItemList list) {
  Persistent(runs, list, [](const RunDev& r, int s) {
    MixedSwitch<kMixedThreads512, kMixedTable512>(r, s, std::make_integer_sequence<int, kMixedCount512>{});
  });
}

__global__ void __launch_bounds__(kMixedThreads256) __attribute__((amdgpu_waves_per_eu(4)))
KernelDescentMixed256(const RunDev* runs, ItemList list) {
  Persistent(runs, list, [](const RunDev& r, int s) {
    MixedSwitch<kMixedThreads256, kMixedTable512>(r, s, std::make_integer_sequence<int, kMixedCount256>{});
  });
}

template <typename Kernel>
void AllowLds(Kernel kernel, size_t bytes) {
  if (bytes > 0) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(bytes));
  }
}

// What a dispatch does with the kernel instance it finds: nothing (does the instance exist?), set its dynamic-LDS
// attribute only (outside a stream capture, before the launches are recorded), or set the attribute and launch.
enum class GoMode { kProbe, kAttributes, kLaunch };

// Launch context shared by the dispatchers below.
struct Go {
  const RunDev* runs;
  const ItemList* list;
  const DescentPlan* plan;
  hipStream_t stream;
  GoMode mode;

  int Blocks() const { return plan->blocks > 0 ? plan->blocks : list->count; }

  template <typename Kernel>
  bool Run(Kernel kernel, int threads, size_t lds) const {
    if (mode == GoMode::kProbe) return true;
    AllowLds(kernel, lds);
    if (mode == GoMode::kLaunch) kernel<<<Blocks(), threads, lds, stream>>>(runs, *list);
    return true;
  }

  // A kernel that takes the plan's dynamic LDS size as its last argument.
  template <typename Kernel>
  bool RunWithLds(Kernel kernel, int threads) const {
    if (mode == GoMode::kProbe) return true;
    const size_t lds = plan->lds_bytes;
    AllowLds(kernel, lds);
    if (mode == GoMode::kLaunch) kernel<<<Blocks(), threads, lds, stream>>>(runs, *list, static_cast<int>(lds));
    return true;
  }
};

template <int T, typename GT>
bool GoGeneric(const Go& go) {
  if (go.plan->descent == Descent::kGenericLds) return go.Run(KernelDescentGeneric<T, GT, true>, T, go.plan->lds_bytes);
  return go.Run(KernelDescentGeneric<T, GT, false>, T, 0);
}

template <typename GT>
bool GoGenericT(const Go& go) {
  switch (go.plan->threads) {
    case 256: return GoGeneric<256, GT>(go);
    case 512: return GoGeneric<512, GT>(go);
    case 1024: return GoGeneric<1024, GT>(go);
    default: return false;
  }
}

template <int T, int kVpt>
bool GoK2(const Go& go) {
  if (go.plan->rows_in_lds) return go.Run(KernelDescentK2<T, kVpt, true>, T, go.plan->lds_bytes);
  return go.Run(KernelDescentK2<T, kVpt, false>, T, 0);
}

template <int T>
bool GoK2Vpt(const Go& go) {
  switch (go.plan->vpt) {
    case 1: return GoK2<T, 1>(go);
    case 2: return GoK2<T, 2>(go);
    case 4: return GoK2<T, 4>(go);
    case 8: return GoK2<T, 8>(go);
    case 16: return GoK2<T, 16>(go);
    case 32: return GoK2<T, 32>(go);
    default: return false;
  }
}

bool GoK2Any(const Go& go) {
  switch (go.plan->threads) {
    case 64: return GoK2Vpt<64>(go);
    case 128: return GoK2Vpt<128>(go);
    case 256: return GoK2Vpt<256>(go);
    case 512: return GoK2Vpt<512>(go);
    case 1024: return GoK2Vpt<1024>(go);
    default: return false;
  }
}

// Register descent instances: T = 256 with kVpt in {1, 2, 4}; T = 1024 with kVpt in {2, 4, ..., 16}; T = 512 with
// kVpt in {8, 12, ..., 32} for k up to 16. A thread keeps vpt (K + 3) registers of slot state: at most 104 of the 128
// VGPRs a wave has with 16 waves per CU (256 and 1024 threads), at most 200 of the 256 with 8 waves (512 threads,
// which trade half the waves and their per-move reductions for twice the slots per thread). The slot index stays
// below bit 12 of the key, or bit 13 in the 512-thread instances (RegsQShift).
constexpr bool RegsFit(int clusters, int vpt, int threads = 1024) {
  if (threads == 512) {
    return clusters <= 16 && vpt * (clusters + 3) <= 200 && (vpt << (2 * RegsLabelBits(clusters))) <= 8192;
  }
  return vpt * (clusters + 3) <= 104 && (vpt << (2 * RegsLabelBits(clusters))) <= 4096;
}

template <int T, int K, int kVpt>
bool GoRegsOne(const Go& go) {
  if constexpr (RegsFit(K, kVpt, T)) {
    return go.Run(KernelDescentRegs<T, K, kVpt>, T, 0);
  } else {
    return false;
  }
}

template <int K>
bool GoRegsK(const Go& go) {
  if (go.plan->threads == 256) {
    switch (go.plan->vpt) {
      case 1: return GoRegsOne<256, K, 1>(go);
      case 2: return GoRegsOne<256, K, 2>(go);
      case 4: return GoRegsOne<256, K, 4>(go);
      default: return false;
    }
  }
  if (go.plan->threads == 512) {
    switch (go.plan->vpt) {
      case 8: return GoRegsOne<512, K, 8>(go);
      case 12: return GoRegsOne<512, K, 12>(go);
      case 16: return GoRegsOne<512, K, 16>(go);
      case 20: return GoRegsOne<512, K, 20>(go);
      case 24: return GoRegsOne<512, K, 24>(go);
      case 28: return GoRegsOne<512, K, 28>(go);
      case 32: return GoRegsOne<512, K, 32>(go);
      default: return false;
    }
  }
  if (go.plan->threads == 1024) {
    switch (go.plan->vpt) {
      case 2: return GoRegsOne<1024, K, 2>(go);
      case 4: return GoRegsOne<1024, K, 4>(go);
      case 6: return GoRegsOne<1024, K, 6>(go);
      case 8: return GoRegsOne<1024, K, 8>(go);
      case 10: return GoRegsOne<1024, K, 10>(go);
      case 12: return GoRegsOne<1024, K, 12>(go);
      case 14: return GoRegsOne<1024, K, 14>(go);
      case 16: return GoRegsOne<1024, K, 16>(go);
      default: return false;
    }
  }
  return false;
}

bool GoRegs(const Go& go) {
  switch (go.plan->k) {
    case 3: return GoRegsK<3>(go);
    case 4: return GoRegsK<4>(go);
    case 5: return GoRegsK<5>(go);
    case 6: return GoRegsK<6>(go);
    case 7: return GoRegsK<7>(go);
    case 8: return GoRegsK<8>(go);
    case 9: return GoRegsK<9>(go);
    case 10: return GoRegsK<10>(go);
    case 11: return GoRegsK<11>(go);
    case 12: return GoRegsK<12>(go);
    case 13: return GoRegsK<13>(go);
    case 14: return GoRegsK<14>(go);
    case 15: return GoRegsK<15>(go);
    case 16: return GoRegsK<16>(go);
    case 17: case 18: case 19: case 20: return GoRegsK<20>(go);
    case 21: case 22: case 23: case 24: return GoRegsK<24>(go);
    case 25: case 26: case 27: case 28: return GoRegsK<28>(go);
    case 29: case 30: case 31: case 32: return GoRegsK<32>(go);
    default: return false;
  }
}

// The streaming wide-path kernels take the number of resident tiles their dynamic LDS holds as a launch argument, so
// the tiles a descent keeps in LDS never exceed what the launch allocated (a batch launch allocates the largest LDS
// of its runs' plans; a tuned plan may hold none).
template <template <int> class Wide>
bool GoWide(const Go& go) {
  const size_t lds = go.plan->lds_bytes;
  const int resident = static_cast<int>(lds / WideTileBytes(Wide<256>::kK));
  auto run = [&](auto kernel, int threads) {
    if (go.mode == GoMode::kProbe) return true;
    AllowLds(kernel, lds);
    if (go.mode == GoMode::kLaunch) kernel<<<go.Blocks(), threads, lds, go.stream>>>(go.runs, *go.list, resident);
    return true;
  };
  if (go.plan->threads == 256) return run(Wide<256>::kernel, 256);
  if (go.plan->threads == 1024) return run(Wide<1024>::kernel, 1024);
  return false;
}

bool GoK3Packed(const Go& go) {
  switch (go.plan->vpt) {
    case 16: return go.Run(KernelDescentK3Packed<1024, 16>, 1024, 0);
    case 20: return go.Run(KernelDescentK3Packed<1024, 20>, 1024, 0);
    case 24: return go.Run(KernelDescentK3Packed<1024, 24>, 1024, 0);
    case 28: return go.Run(KernelDescentK3Packed<1024, 28>, 1024, 0);
    case 32: return go.Run(KernelDescentK3Packed<1024, 32>, 1024, 0);
    default: return false;
  }
}

template <int T>
struct WideK2 {
  static constexpr int kK = 2;
  static constexpr auto kernel = KernelDescentK2Wide<T>;
};
template <int T>
struct WideK3 {
  static constexpr int kK = 3;
  static constexpr auto kernel = KernelDescentK3Wide<T>;
};

bool Dispatch(const Go& go) {
  const DescentPlan& p = *go.plan;
  switch (p.descent) {
    case Descent::kGenericGlobal:
    case Descent::kGenericLds:
      return p.wide ? GoGenericT<int32_t>(go) : GoGenericT<int16_t>(go);
    case Descent::kK2Registers:
      return !p.wide && GoK2Any(go);
    case Descent::kRegisters:
      return !p.wide && GoRegs(go);
    case Descent::kK2Wide:
      return p.wide && p.k == 2 && GoWide<WideK2>(go);
    case Descent::kK3Wide:
      return p.wide && p.k == 3 && GoWide<WideK3>(go);
    case Descent::kK3Packed:
      return !p.wide && p.k == 3 && p.threads == 1024 && GoK3Packed(go);
    case Descent::kK3Chunk:
      if (!p.wide || p.k != 3 || p.lds_bytes == 0) return false;
      if (p.threads == 256) return go.RunWithLds(KernelDescentK3Chunk<256>, 256);
      if (p.threads == 1024) return go.RunWithLds(KernelDescentK3Chunk<1024>, 1024);
      return false;
    case Descent::kMixed:
      if (p.wide) return false;
      if (p.threads == kMixedThreads) return go.Run(KernelDescentMixed, kMixedThreads, 0);
      if (p.threads == kMixedThreads512) return go.Run(KernelDescentMixed512, kMixedThreads512, 0);
      if (p.threads == kMixedThreads256) return go.Run(KernelDescentMixed256, kMixedThreads256, 0);
      return false;
  }
  return false;
}

}  // namespace

size_t DescentLdsLimit() { return kLdsBytes - kStaticLdsReserve; }

bool DescentSupported(const DescentPlan& plan) {
  const ItemList none{nullptr, 0, nullptr};
  return Dispatch(Go{nullptr, &none, &plan, nullptr, GoMode::kProbe});
}

void SetDescentAttributes(const DescentPlan& plan) {
  const ItemList none{nullptr, 0, nullptr};
  Dispatch(Go{nullptr, &none, &plan, nullptr, GoMode::kAttributes});
}

void LaunchDescent(const RunDev* runs, const ItemList& items, const DescentPlan& plan, hipStream_t stream) {
  if (items.count == 0) return;
  Dispatch(Go{runs, &items, &plan, stream, GoMode::kLaunch});
}

size_t GenericLdsBytes(int n, int k, int wide) {
  return Align16((wide ? 4 : 2) * static_cast<size_t>(k) * n) + Align16(static_cast<size_t>(n));
}

size_t WideResidentLdsBytes(int n, int k, int threads) {
  const int64_t tiles = n / WideTileVertices(k);  // whole tiles only
  const int64_t fit = static_cast<int64_t>(DescentLdsLimit() / WideTileBytes(k));
  const int64_t per_waves = static_cast<int64_t>(threads / kWave) * kResidentPerWave;
  return static_cast<size_t>(std::min({tiles, fit, per_waves})) * WideTileBytes(k);
}


int MixedVariant(int n, int k, DescentPlan* plan) {
  auto pick = [&](const MixedDesc* table, int count, int threads) {
    for (int i = 0; i < count; ++i) {
      const MixedDesc& d = table[i];
      if (d.k == k && d.vpt * threads >= n) {
        plan->descent = d.kind == 0 ? Descent::kK2Registers : Descent::kRegisters;
        plan->k = k;
        plan->threads = threads;
        plan->vpt = d.vpt;
        plan->rows_in_lds = false;
        plan->lds_bytes = 0;
        plan->wide = 0;
        return i;
      }
    }
    return -1;
  };
  const int smallest = pick(kMixedTable256, kMixedCount256, kMixedThreads256);
  if (smallest >= 0) return smallest;
  const int small = pick(kMixedTable512, kMixedCount512, kMixedThreads512);
  return small >= 0 ? small : pick(kMixedTable, kMixedCount, kMixedThreads);
}

int K3PackedVpt(int n) {
  for (int v : {16, 20, 24, 28, 32}) {
    if (v * 1024 >= n) return v;
  }
  return 0;
}

int RegistersVpt(int n, int k, int* threads) {
  if (k < 3 || k > 32) return 0;
  const int clusters = RegsClusters(k);
  if (n <= 1024) {
    for (int v : {1, 2, 4}) {
      if (v * 256 >= n) {
        if (!RegsFit(clusters, v)) break;
        *threads = 256;
        return v;
      }
    }
  }
  *threads = 1024;
  for (int v : {2, 4, 6, 8, 10, 12, 14, 16}) {
    if (v * 1024 >= n) return RegsFit(clusters, v) ? v : 0;
  }
  return 0;
}

}  // namespace cc::hip
