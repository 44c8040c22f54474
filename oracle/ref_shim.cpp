// ref_shim.cpp — a C ABI over the reference CPU solver, compiled together with the reference's own sources
// (src/cc_graph.cpp, src/cc_pbils_cpu.cpp under the reference tree) into oracle/_ref/libccref.so by
// oracle/Makefile. TEST INFRASTRUCTURE ONLY: it lets the tests check the restatement in cc_oracle.c against the
// real reference on any graph, and bench.py time the reference CPU path as its cpu_baseline.
//
// The graph crosses the boundary as the 64-bit-word bitset used everywhere in this repository and is rebuilt here
// into the reference's cc::Graph with Graph::AddEdge (include/cc_graph.hpp:32).
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "cc_graph.hpp"
#include "cc_math.hpp"
#include "cc_pbils.hpp"

namespace {

cc::Graph FromBits(const uint64_t* bits, uint32_t n, uint32_t words) {
  cc::Graph graph(n);
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t* row = bits + (size_t)i * words;
    for (uint32_t w = (i + 1) / 64; w < words; ++w) {
      for (uint64_t b = row[w]; b != 0; b &= b - 1) {
        const uint32_t j = w * 64 + (uint32_t)__builtin_ctzll(b);
        if (j > i) graph.AddEdge(i, j);
      }
    }
  }
  return graph;
}

void ToBits(const cc::Graph& graph, uint64_t* bits, uint32_t words) {
  const uint32_t n = graph.Size();
  std::memset(bits, 0, sizeof(uint64_t) * (size_t)n * words);
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t j = 0; j < n; ++j) {
      if (graph.IsJoined(i, j)) bits[(size_t)i * words + j / 64] |= 1ull << (j % 64);
    }
  }
}

struct RefResult {
  int64_t objective;
  int32_t iterations_done, clusters_used;
  int64_t local_searches, accepted_moves;
  double seconds;
};

void Export(const cc::PbilsResult& r, int32_t* labels_out, RefResult* out) {
  out->objective = r.objective;
  out->iterations_done = r.iterations_done;
  out->clusters_used = r.clusters_used;
  out->local_searches = r.local_searches;
  out->accepted_moves = r.accepted_moves;
  out->seconds = r.seconds;
  for (size_t v = 0; v < r.labels.size(); ++v) labels_out[v] = r.labels[v];
}

void SetError(char* err, size_t len, const char* what) {
  if (err && len) {
    std::strncpy(err, what, len - 1);
    err[len - 1] = 0;
  }
}

}  // namespace

extern "C" {

uint64_t ref_erdos_renyi(uint32_t n, double p, uint64_t seed, uint64_t* bits, uint32_t words) {
  const cc::Graph graph = cc::Graph::ErdosRenyi(n, p, seed);
  ToBits(graph, bits, words);
  return graph.EdgeCount();
}

int ref_solve_cpu(const uint64_t* bits, uint32_t n, uint32_t words, int32_t k, int32_t population,
                  int32_t tournament, int32_t iterations, int32_t early_stop, double perturbation, double gww,
                  int32_t threads, uint64_t seed, int32_t* labels_out, RefResult* out, char* err, size_t err_len) {
  try {
    const cc::Graph graph = FromBits(bits, n, words);
    cc::PbilsParams params;
    params.k = k;
    params.population = population;
    params.tournament = tournament;
    params.iterations = iterations;
    params.early_stop = early_stop;
    params.perturbation = perturbation;
    params.gww = gww;
    params.threads = threads;
    params.seed = seed;
    Export(cc::SolveCpu(graph, params), labels_out, out);
    return 0;
  } catch (const std::exception& e) {
    SetError(err, err_len, e.what());
    return -1;
  }
}

// The sequential Neighborhood enumeration of the reference's GPU test (cuda/cc_verify_gpu.cu:21-35), written
// against the reference State API.
int ref_neighborhood(const uint64_t* bits, uint32_t n, uint32_t words, int32_t local_search, int32_t* labels_out,
                     RefResult* out) {
  const cc::Graph graph = FromBits(bits, n, words);
  cc::PbilsResult result;
  cc::State state;
  state.Init((int)n, 2);
  for (uint32_t v = 0; v < n; ++v) {
    for (uint32_t u = 0; u < n; ++u) state.labels[u] = (u == v || graph.IsJoined(v, u)) ? 0 : 1;
    state.Rebuild(graph);
    if (local_search) cc::LocalSearch(graph, state, &result.accepted_moves);
    if (result.objective < 0 || state.f < result.objective) {
      result.objective = state.f;
      result.labels = state.labels;
    }
  }
  result.local_searches = local_search ? n : 0;
  result.clusters_used = cc::CountClustersUsed(result.labels, 2);
  Export(result, labels_out, out);
  return 0;
}

int64_t ref_rebuild(const uint64_t* bits, uint32_t n, uint32_t words, int32_t k, const int32_t* labels,
                    int32_t* g_out, int32_t* sizes_out) {
  const cc::Graph graph = FromBits(bits, n, words);
  cc::State state;
  state.Init((int)n, k);
  for (uint32_t v = 0; v < n; ++v) state.labels[v] = labels[v];
  state.Rebuild(graph);
  for (size_t i = 0; i < state.g.size(); ++i) g_out[i] = state.g[i];
  for (int c = 0; c < k; ++c) sizes_out[c] = state.sizes[c];
  return state.f;
}

int64_t ref_local_search(const uint64_t* bits, uint32_t n, uint32_t words, int32_t k, int32_t* labels,
                         int64_t* moves) {
  const cc::Graph graph = FromBits(bits, n, words);
  cc::State state;
  state.Init((int)n, k);
  for (uint32_t v = 0; v < n; ++v) state.labels[v] = labels[v];
  state.Rebuild(graph);
  long long applied = 0;
  cc::LocalSearch(graph, state, &applied);
  for (uint32_t v = 0; v < n; ++v) labels[v] = state.labels[v];
  *moves = applied;
  return state.f;
}

// The reference's tag-similarity graph (Graph::LoadTags) as a 64-bit-word bitset: returns -1 with a message, or the
// vertex count; bits are written only if capacity_words holds them.
int64_t ref_load_tags(const char* path, const char* similarity, double threshold, uint32_t n, uint64_t seed,
                      uint64_t* bits, uint64_t capacity_words, uint64_t* edges_out, char* err, size_t err_len) {
  try {
    const cc::Graph graph = cc::Graph::LoadTags(path, similarity, threshold, n, seed);
    const uint32_t size = graph.Size();
    const uint32_t words = (size + 63) / 64 > 0 ? (size + 63) / 64 : 1;
    *edges_out = graph.EdgeCount();
    if ((uint64_t)size * words <= capacity_words) ToBits(graph, bits, words);
    return size;
  } catch (const std::exception& e) {
    SetError(err, err_len, e.what());
    return -1;
  }
}

// The reference's instance files: format 0 Graph::LoadMatrix, 1 Graph::LoadBaselineJson, 2 Graph::LoadEdgeList.
// Returns -1 with a message, or the vertex count; bits are written only if capacity_words holds them.
int64_t ref_load_graph(const char* path, int32_t format, uint64_t* bits, uint64_t capacity_words, uint64_t* edges_out,
                       char* err, size_t err_len) {
  try {
    const cc::Graph graph = format == 0   ? cc::Graph::LoadMatrix(path)
                            : format == 1 ? cc::Graph::LoadBaselineJson(path)
                                          : cc::Graph::LoadEdgeList(path);
    const uint32_t size = graph.Size();
    const uint32_t words = (size + 63) / 64 > 0 ? (size + 63) / 64 : 1;
    *edges_out = graph.EdgeCount();
    if ((uint64_t)size * words <= capacity_words) ToBits(graph, bits, words);
    return size;
  } catch (const std::exception& e) {
    SetError(err, err_len, e.what());
    return -1;
  }
}

// Graph::SaveMatrix (format 0) or Graph::SaveBaselineJson (format 1) of the reference.
int ref_save_graph(const uint64_t* bits, uint32_t n, uint32_t words, const char* path, int32_t format) {
  try {
    const cc::Graph graph = FromBits(bits, n, words);
    if (format == 0) {
      graph.SaveMatrix(path);
    } else {
      graph.SaveBaselineJson(path);
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

void ref_sample_without_replacement(uint32_t total, uint32_t n, uint64_t seed, uint32_t* out) {
  const std::vector<unsigned> order = cc::SampleWithoutReplacement(total, n, seed);
  for (size_t i = 0; i < order.size(); ++i) out[i] = order[i];
}

int64_t ref_objective_direct(const uint64_t* bits, uint32_t n, uint32_t words, const int32_t* labels) {
  const cc::Graph graph = FromBits(bits, n, words);
  return cc::ObjectiveDirect(graph, std::vector<int>(labels, labels + n));
}

int64_t ref_objective_bitwise(const uint64_t* bits, uint32_t n, uint32_t words, const int32_t* labels) {
  const cc::Graph graph = FromBits(bits, n, words);
  return cc::ObjectiveBitwise(graph, std::vector<int>(labels, labels + n));
}

void ref_rank_slots(const int64_t* f, int32_t population, int32_t* order) {
  std::vector<long long> values(f, f + population);
  const std::vector<int> ranked = cc::RankSlots(values.data(), population);
  for (int i = 0; i < population; ++i) order[i] = ranked[i];
}

int32_t ref_winner_copies(double gww, int32_t population) { return cc::WinnerCopies(gww, population); }

uint64_t ref_stream_seed(uint64_t seed, uint64_t stream, uint64_t step) { return cc::StreamSeed(seed, stream, step); }

int32_t ref_initial_label(uint64_t base, uint64_t index, int32_t k) { return cc::InitialLabel(base, index, k); }

int32_t ref_perturbed_label(uint64_t base, uint64_t index, int32_t iteration, int32_t label, int32_t k, float p) {
  return cc::PerturbedLabel(base, index, iteration, label, k, p);
}

int32_t ref_tournament(uint64_t base, int32_t slot, int32_t iteration, const int64_t* f, int32_t population,
                       int32_t tournament) {
  std::vector<long long> values(f, f + population);
  return cc::Tournament(base, slot, iteration, values.data(), population, tournament);
}

}  // extern "C"
