// cc_math.hpp — the objective, the O(1) move delta and the counter-based RNG, shared by host and device code.
//
// Notation: n vertices, k clusters (labels in [0, k), empty clusters allowed), m edges, sizes[c] vertices with
// label c, G[v][c] neighbours of v with label c (G = A Z), W = intra-cluster edges = tr(Z^T A Z) / 2.
//   f = m + sum_c C(sizes[c], 2) - 2 W                                           (reference include/cc_math.hpp:1-18)
//   moving v from a to c changes f by (s_c - s_a + 1) - 2 (G[v][c] - G[v][a])
//
// The RNG is splitmix64 driven by seeds that the caller carries; every random decision of PBILS is a pure
// function of (seed, stream, index, step), which is what makes the device trajectory identical to the CPU one.
// The constants below are a data contract with the reference (include/cc_math.hpp:84-145) and must not change.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define CC_HD __host__ __device__
#else
#define CC_HD
#endif

namespace cc {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

CC_HD inline uint64_t Mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// One splitmix64 step: advance by the golden gamma, output the mixed state.
CC_HD inline uint64_t SplitMix64(uint64_t& state) {
  state += kGolden;
  return Mix64(state);
}

CC_HD inline uint32_t NextU32(uint64_t& state) { return static_cast<uint32_t>(SplitMix64(state) >> 32); }

// Uniform in [0, bound): the high half of u32 * bound; 0 for bound 0.
CC_HD inline uint32_t RandBelow(uint64_t& state, uint32_t bound) {
  if (bound == 0) return 0;
  return static_cast<uint32_t>((static_cast<uint64_t>(NextU32(state)) * bound) >> 32);
}

// Uniform float in [0, 1) on the 2^-24 grid (the top 24 bits are exactly a float mantissa).
CC_HD inline float RandFloat(uint64_t& state) {
  return static_cast<float>(NextU32(state) >> 8) * (1.0f / 16777216.0f);
}

// Seed of an independent stream: the splitmix64-advanced mix of (seed, stream, step).
CC_HD inline uint64_t StreamSeed(uint64_t seed, uint64_t stream, uint64_t step) {
  uint64_t s = seed ^ (stream * 0xD1B54A32D192ED03ull) ^ (step * 0xA0761D6478BD642Full);
  SplitMix64(s);
  return s;
}

// The t-th output (t >= 0) of a splitmix64 stream that starts at state0, without walking the stream.
CC_HD inline uint64_t SplitMix64At(uint64_t state0, uint64_t t) { return Mix64(state0 + (t + 1) * kGolden); }

// Stream constants (reference include/cc_math.hpp:118-121, src/cc_graph.cpp:21). Changing any changes every
// trajectory.
constexpr uint64_t kInitStream = 0xA11CE;
constexpr uint64_t kSelectStream = 0x5E1EC7;
constexpr uint64_t kPerturbStream = 0xBEEF;
constexpr uint64_t kPerturbFirstStep = 0x5000000;
constexpr uint64_t kErdosRenyiStream = 0xE4D05F;
// Object sample of the tag-similarity graph (reference src/cc_graph.cpp:22).
constexpr uint64_t kSampleStream = 0x7A65;
// Streams of the generators this project adds.
constexpr uint64_t kSbmStream = 0x5B3D10C;
constexpr uint64_t kChungLuStream = 0xC1A9E5;
// Per-instance seeds of a batched mix: StreamSeed(master, kBatchGraphStream | kBatchSearchStream, instance).
constexpr uint64_t kBatchGraphStream = 0xBA7C4;
constexpr uint64_t kBatchSearchStream = 0xBA7C5;
constexpr uint64_t kBatchShapeStream = 0xBA7C6;

// Label of vertex v of individual slot at start: index = slot * n + v.
CC_HD inline int InitialLabel(uint64_t base, uint64_t index, int k) {
  uint64_t rng = StreamSeed(base, index, 0);
  return static_cast<int>(RandBelow(rng, static_cast<uint32_t>(k)));
}

// Tournament winner for slot of the next population: the best f among `tournament` draws, the first drawn among
// equals (strict <).
CC_HD inline int Tournament(uint64_t base, int slot, int iteration, const int64_t* f, int population,
                            int tournament) {
  uint64_t rng = StreamSeed(base, static_cast<uint64_t>(static_cast<int64_t>(slot)),
                            static_cast<uint64_t>(static_cast<int64_t>(iteration) + 1));
  int chosen = static_cast<int>(RandBelow(rng, static_cast<uint32_t>(population)));
  for (int t = 1; t < tournament; ++t) {
    const int candidate = static_cast<int>(RandBelow(rng, static_cast<uint32_t>(population)));
    if (f[candidate] < f[chosen]) chosen = candidate;
  }
  return chosen;
}

// With probability p the label of a uniformly chosen other cluster, otherwise the same label.
CC_HD inline int PerturbedLabel(uint64_t base, uint64_t index, int iteration, int label, int k, float probability) {
  uint64_t rng = StreamSeed(base, index, static_cast<uint64_t>(static_cast<int64_t>(iteration)) + kPerturbFirstStep);
  if (RandFloat(rng) >= probability) return label;
  return (label + 1 + static_cast<int>(RandBelow(rng, static_cast<uint32_t>(k - 1)))) % k;
}

CC_HD inline int64_t SamePairs(const int32_t* sizes, int k) {
  int64_t acc = 0;
  for (int c = 0; c < k; ++c) {
    const int64_t s = sizes[c];
    acc += s * (s - 1) / 2;
  }
  return acc;
}

CC_HD inline int64_t Objective(int64_t edges, const int32_t* sizes, int k, int64_t intra_edges) {
  return edges + SamePairs(sizes, k) - 2 * intra_edges;
}

CC_HD inline int MoveDelta(int size_from, int size_to, int g_from, int g_to) {
  return (size_to - size_from + 1) - 2 * (g_to - g_from);
}

}  // namespace cc
