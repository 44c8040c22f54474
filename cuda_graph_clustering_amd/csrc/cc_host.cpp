// cc_host.cpp — host side of the library: the Graph type with its generators and file formats, the objective
// recounts, the GWW helpers, the C++ entry points of cc_gpu.hpp over the C ABI, and the generator/recount part of
// the C ABI.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cc_gpu.hpp"
#include "cc_graph.hpp"
#include "cc_graph_io.hpp"
#include "cc_hip.h"
#include "cc_math.hpp"
#include "cc_pbils.hpp"

namespace cc {
namespace {

int HardwareThreads() { return std::max(1, static_cast<int>(std::thread::hardware_concurrency())); }

// Items [0, count) handed out through an atomic counter to `threads` workers (the caller is one of them).
void ParallelFor(int count, int threads, const std::function<void(int)>& body) {
  threads = std::max(1, std::min(threads <= 0 ? HardwareThreads() : threads, count));
  std::atomic<int> next{0};
  auto drain = [&] {
    for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) body(i);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(drain);
  drain();
  for (auto& t : pool) t.join();
}

uint32_t WordsFor(unsigned n) { return std::max(1u, (n + kWordBits - 1) / kWordBits); }

// RandFloat of draw number t of a splitmix64 stream that starts at state0.
inline float DrawAt(uint64_t state0, uint64_t t) {
  const uint32_t u = static_cast<uint32_t>(SplitMix64At(state0, t) >> 32);
  return static_cast<float>(u >> 8) * (1.0f / 16777216.0f);
}

// Fills every row of a symmetric random graph whose pair (i, j), j < i, is an edge iff draw number i(i-1)/2 + j
// of the stream is below probability(i, j). Each row computes its own pairs on both sides of the diagonal, so
// rows are independent.
template <typename Probability>
void FillPairs(unsigned n, uint32_t words, uint64_t state0, Word* bits, int threads, Probability probability) {
  ParallelFor(static_cast<int>(n), threads, [&](int item) {
    const unsigned i = static_cast<unsigned>(item);
    Word* row = bits + static_cast<size_t>(i) * words;
    std::fill(row, row + words, Word(0));
    for (unsigned j = 0; j < n; ++j) {
      if (j == i) continue;
      const unsigned hi = std::max(i, j), lo = std::min(i, j);
      const uint64_t t = static_cast<uint64_t>(hi) * (hi - 1) / 2 + lo;
      if (static_cast<double>(DrawAt(state0, t)) < probability(hi, lo)) row[j / kWordBits] |= Word(1) << (j % kWordBits);
    }
  });
}

uint64_t CountEdges(const Word* bits, size_t count) {
  uint64_t ones = 0;
  for (size_t i = 0; i < count; ++i) ones += static_cast<uint64_t>(__builtin_popcountll(bits[i]));
  return ones / 2;
}

void GeneratorErdosRenyi(unsigned n, double p, uint64_t seed, Word* bits, int threads) {
  FillPairs(n, WordsFor(n), StreamSeed(seed, kErdosRenyiStream, 0), bits, threads,
            [p](unsigned, unsigned) { return p; });
}

void GeneratorSbm(unsigned n, int blocks, double p_in, double p_out, uint64_t seed, Word* bits, int threads) {
  const uint64_t b = static_cast<uint64_t>(std::max(blocks, 1));
  FillPairs(n, WordsFor(n), StreamSeed(seed, kSbmStream, 0), bits, threads, [=](unsigned i, unsigned j) {
    return (static_cast<uint64_t>(i) * b) / n == (static_cast<uint64_t>(j) * b) / n ? p_in : p_out;
  });
}

// Chung-Lu in O(n + m) by geometric skips (J. C. Miller, A. Hagberg, "Efficient generation of networks with given
// expected degrees", WAW 2011). The weights fall with the index, so along row u the pair probability
// p(u, v) = min(1, w_u w_v / S) falls with v: from v with current bound p, floor(log r / log(1 - p)) pairs
// (r uniform in (0, 1]) are passed over at once, the pair reached is kept with probability p(u, v) / p, and the bound
// becomes p(u, v) — each pair an edge with probability p(u, v), independently. Row u draws from its own stream
// StreamSeed(seed, kChungLuStream, u + 1) (53-bit uniforms), so rows are independent work items; each writes the
// upper triangle of its row, and the edges it found are mirrored afterwards.
inline double UniformOpenZero(uint64_t& state) {  // (0, 1]
  return static_cast<double>((SplitMix64(state) >> 11) + 1) * (1.0 / 9007199254740992.0);
}
inline double UniformOpenOne(uint64_t& state) {  // [0, 1)
  return static_cast<double>(SplitMix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

void GeneratorChungLu(unsigned n, double gamma, double mean_degree, uint64_t seed, Word* bits, int threads) {
  std::vector<double> w(n);
  const double exponent = -1.0 / (gamma - 1.0);
  double sum = 0.0;
  for (unsigned i = 0; i < n; ++i) {
    w[i] = std::pow(static_cast<double>(i + 1), exponent);
    sum += w[i];
  }
  const double total = mean_degree * static_cast<double>(n);
  const double scale = n ? total / sum : 0.0;
  for (unsigned i = 0; i < n; ++i) w[i] *= scale;
  const double inv_total = n ? 1.0 / total : 0.0;
  const uint32_t words = WordsFor(n);
  auto probability = [&](unsigned u, unsigned v) {
    const double q = w[u] * w[v] * inv_total;
    return q > 1.0 ? 1.0 : q;
  };
  std::vector<std::vector<unsigned>> above(n);  // row u: its neighbours v > u, for the mirror pass
  ParallelFor(static_cast<int>(n), threads, [&](int item) {
    const unsigned u = static_cast<unsigned>(item);
    Word* row = bits + static_cast<size_t>(u) * words;
    std::fill(row, row + words, Word(0));
    uint64_t state = StreamSeed(seed, kChungLuStream, static_cast<uint64_t>(u) + 1);
    unsigned v = u + 1;
    if (v >= n) return;
    double p = probability(u, v);
    while (v < n && p > 0.0) {
      if (p < 1.0) {
        const double skip = std::floor(std::log(UniformOpenZero(state)) / std::log1p(-p));
        if (skip >= static_cast<double>(n - v)) break;
        v += static_cast<unsigned>(skip);
      }
      const double q = probability(u, v);
      if (UniformOpenOne(state) < q / p) {
        row[v / kWordBits] |= Word(1) << (v % kWordBits);
        above[u].push_back(v);
      }
      p = q;
      ++v;
    }
  });
  for (unsigned u = 0; u < n; ++u) {
    for (const unsigned v : above[u]) bits[static_cast<size_t>(v) * words + u / kWordBits] |= Word(1) << (u % kWordBits);
  }
}

}  // namespace

// ---- Graph ------------------------------------------------------------------------------------------------------

Graph::Graph(unsigned n) : n_(n), words_per_row_(WordsFor(n)) {
  bits_.assign(static_cast<size_t>(n_) * words_per_row_, Word(0));
}

void Graph::AddEdge(unsigned i, unsigned j) {
  if (i == j || IsJoined(i, j)) return;
  bits_[static_cast<size_t>(i) * words_per_row_ + j / kWordBits] |= Word(1) << (j % kWordBits);
  bits_[static_cast<size_t>(j) * words_per_row_ + i / kWordBits] |= Word(1) << (i % kWordBits);
  ++edges_;
}

double Graph::Density() const {
  const double pairs = static_cast<double>(n_) * (n_ - 1) / 2.0;
  return pairs > 0 ? static_cast<double>(edges_) / pairs : 0.0;
}

Graph Graph::FromBits(std::vector<Word> bits, unsigned n) {
  Graph g;
  g.n_ = n;
  g.words_per_row_ = WordsFor(n);
  if (bits.size() != static_cast<size_t>(n) * g.words_per_row_) throw std::runtime_error("bitset size mismatch");
  g.bits_ = std::move(bits);
  g.edges_ = CountEdges(g.bits_.data(), g.bits_.size());
  return g;
}

Graph Graph::ErdosRenyi(unsigned n, double density, uint64_t seed, int threads) {
  Graph g(n);
  GeneratorErdosRenyi(n, density, seed, g.bits_.data(), threads);
  g.edges_ = CountEdges(g.bits_.data(), g.bits_.size());
  return g;
}

Graph Graph::StochasticBlock(unsigned n, int blocks, double p_in, double p_out, uint64_t seed, int threads) {
  Graph g(n);
  GeneratorSbm(n, blocks, p_in, p_out, seed, g.bits_.data(), threads);
  g.edges_ = CountEdges(g.bits_.data(), g.bits_.size());
  return g;
}

Graph Graph::ChungLu(unsigned n, double gamma, double mean_degree, uint64_t seed, int threads) {
  Graph g(n);
  GeneratorChungLu(n, gamma, mean_degree, seed, g.bits_.data(), threads);
  g.edges_ = CountEdges(g.bits_.data(), g.bits_.size());
  return g;
}

// ---- tag-similarity graph ----------------------------------------------------------------------------------------

namespace {

enum class TagMeasure { kJaccard, kCosine, kDice, kOverlap };

TagMeasure ParseTagMeasure(const std::string& name) {
  if (name == "jaccard") return TagMeasure::kJaccard;
  if (name == "cosine") return TagMeasure::kCosine;
  if (name == "dice") return TagMeasure::kDice;
  if (name == "overlap") return TagMeasure::kOverlap;
  throw std::runtime_error("unknown similarity '" + name + "', expected jaccard, cosine, dice or overlap");
}

// The baseline's formulas in double precision, evaluated in its order (reference src/cc_graph.cpp:108-118), so the
// comparison with the threshold agrees with it to the last bit; 0/0 of empty sets is NaN and never reaches it.
bool TagsJoined(TagMeasure m, size_t common, size_t a, size_t b, double threshold) {
  const double inter = static_cast<double>(common);
  double s = 0.0;
  switch (m) {
    case TagMeasure::kJaccard: s = inter / static_cast<double>(a + b - common); break;
    case TagMeasure::kCosine: s = inter / std::sqrt(static_cast<double>(a) * static_cast<double>(b)); break;
    case TagMeasure::kDice: s = 2.0 * inter / (static_cast<double>(a) + static_cast<double>(b)); break;
    case TagMeasure::kOverlap: s = inter / std::min(static_cast<double>(a), static_cast<double>(b)); break;
  }
  return s >= threshold;
}

}  // namespace

std::vector<unsigned> SampleWithoutReplacement(unsigned total, unsigned n, uint64_t seed) {
  std::vector<unsigned> order(total);
  for (unsigned i = 0; i < total; ++i) order[i] = i;
  if (n == 0 || n >= total) return order;
  uint64_t state = StreamSeed(seed, kSampleStream, 0);
  for (unsigned i = 0; i < n; ++i) std::swap(order[i], order[i + RandBelow(state, total - i)]);
  order.resize(n);
  return order;
}

// Vertices with the same tag set have the same row (but for their own diagonal bit), so the rows are built once per
// distinct set (a class). The classes that share at least one tag with a class are found through per-tag lists of
// classes, with their common-tag counts, instead of by comparing every pair of classes; pairs without a common tag
// are joined only if the measure of an empty intersection passes the threshold (threshold <= 0), which is decided
// once per pair of set sizes. A row is then the union of the member lists of its joined classes.
Graph Graph::LoadTags(const std::string& path, const std::string& similarity, double threshold, unsigned n,
                      uint64_t seed) {
  const TagMeasure measure = ParseTagMeasure(similarity);
  const std::vector<std::vector<int>> sets = ReadTagFile(path);
  if (sets.empty()) throw std::runtime_error("no objects in " + path);
  const std::vector<unsigned> objects = SampleWithoutReplacement(static_cast<unsigned>(sets.size()), n, seed);
  const unsigned size = static_cast<unsigned>(objects.size());

  // classes, their members (vertices in increasing order) and the per-tag class lists
  std::map<std::vector<int>, int> class_of_set;
  std::vector<const std::vector<int>*> class_tags;
  std::vector<std::vector<unsigned>> members;
  for (unsigned v = 0; v < size; ++v) {
    const auto found = class_of_set.emplace(sets[objects[v]], static_cast<int>(class_tags.size()));
    if (found.second) {
      class_tags.push_back(&found.first->first);
      members.emplace_back();
    }
    members[found.first->second].push_back(v);
  }
  const int classes = static_cast<int>(class_tags.size());
  int tag_count = 0;
  for (const auto* tags : class_tags) tag_count = std::max(tag_count, tags->empty() ? 0 : tags->back() + 1);
  std::vector<std::vector<int>> classes_of_tag(tag_count);
  for (int c = 0; c < classes; ++c) {
    for (int t : *class_tags[c]) classes_of_tag[t].push_back(c);
  }
  // empty intersections: joined iff the measure of (0, a, b) passes, for the distinct set sizes a, b
  std::vector<size_t> sizes;
  for (const auto* tags : class_tags) sizes.push_back(tags->size());
  std::sort(sizes.begin(), sizes.end());
  sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
  bool any_disjoint = false;
  for (size_t a : sizes) {
    for (size_t b : sizes) any_disjoint = any_disjoint || TagsJoined(measure, 0, a, b, threshold);
  }

  Graph g(size);
  const uint32_t words = g.words_per_row_;
  std::atomic<uint64_t> degree_sum{0};
  std::atomic<int> next{0};
  const int workers = std::max(1, std::min(HardwareThreads(), classes));
  ParallelFor(workers, workers, [&](int) {
    std::vector<uint32_t> common(classes, 0);
    std::vector<int> touched;
    std::vector<Word> row(words);
    for (int c = next.fetch_add(1); c < classes; c = next.fetch_add(1)) {
      const size_t a = class_tags[c]->size();
      touched.clear();
      for (int t : *class_tags[c]) {
        for (int other : classes_of_tag[t]) {
          if (common[other]++ == 0) touched.push_back(other);
        }
      }
      std::fill(row.begin(), row.end(), Word(0));
      auto add_members = [&](int other) {
        for (unsigned u : members[other]) row[u / kWordBits] |= Word(1) << (u % kWordBits);
      };
      if (any_disjoint) {
        for (int other = 0; other < classes; ++other) {
          if (TagsJoined(measure, common[other], a, class_tags[other]->size(), threshold)) add_members(other);
        }
        for (int other : touched) common[other] = 0;
      } else {
        for (int other : touched) {
          if (TagsJoined(measure, common[other], a, class_tags[other]->size(), threshold)) add_members(other);
          common[other] = 0;
        }
      }
      uint64_t ones = 0;
      for (Word w : row) ones += static_cast<uint64_t>(__builtin_popcountll(w));
      // the class is joined with itself iff its own bits are set (a set's similarity to itself is 1, or NaN if
      // empty); a vertex has no loop
      const unsigned first = members[c][0];
      const bool self = (row[first / kWordBits] >> (first % kWordBits)) & 1u;
      for (unsigned v : members[c]) {
        Word* dst = g.bits_.data() + static_cast<size_t>(v) * words;
        std::copy(row.begin(), row.end(), dst);
        dst[v / kWordBits] &= ~(Word(1) << (v % kWordBits));
      }
      degree_sum += (ones - (self ? 1 : 0)) * members[c].size();
    }
  });
  g.edges_ = degree_sum.load() / 2;
  return g;
}

// ---- recounts and GWW helpers -----------------------------------------------------------------------------------

namespace {

void CheckLabels(const Graph& graph, const std::vector<int>& labels) {
  if (labels.size() != graph.Size()) {
    throw std::runtime_error("a clustering of " + std::to_string(labels.size()) + " labels for a graph of " +
                             std::to_string(graph.Size()) + " vertices");
  }
  for (const int label : labels) {
    if (label < 0) throw std::runtime_error("labels must be non-negative");
  }
}

// Rows [0, n) split into one contiguous range per thread; body(first, last, sums) adds the range's counts to sums[2].
template <typename Body>
void SumOverRows(unsigned n, long long* out, Body&& body) {
  const int threads = std::max(1, std::min(HardwareThreads(), static_cast<int>(n / 512) + 1));
  std::vector<long long> part(2 * static_cast<size_t>(threads), 0);
  std::vector<std::thread> pool;
  const unsigned step = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const unsigned first = std::min<unsigned>(n, t * step), last = std::min<unsigned>(n, first + step);
    long long* sums = part.data() + 2 * t;
    if (t + 1 == threads) {
      body(first, last, sums);
    } else {
      pool.emplace_back([&body, first, last, sums] { body(first, last, sums); });
    }
  }
  for (auto& th : pool) th.join();
  out[0] = out[1] = 0;
  for (int t = 0; t < threads; ++t) {
    out[0] += part[2 * t];
    out[1] += part[2 * t + 1];
  }
}

}  // namespace

// The definition, pair by pair: a pair disagrees when "same cluster" and "joined" differ. For the 64 partners j of a
// word, "same cluster" is collected into a mask and compared with the row word at once; only j > i is counted.
long long ObjectiveDirect(const Graph& graph, const std::vector<int>& labels) {
  CheckLabels(graph, labels);
  const unsigned n = graph.Size();
  long long disagreements = 0;
  for (unsigned i = 0; i + 1 < n; ++i) {
    const Word* row = graph.Row(i);
    for (unsigned w = (i + 1) / kWordBits; w * kWordBits < n; ++w) {
      Word same = 0;
      const unsigned base = w * kWordBits;
      for (unsigned b = 0; b < kWordBits && base + b < n; ++b) same |= Word(labels[base + b] == labels[i]) << b;
      Word counted = base + kWordBits <= n ? ~Word(0) : (Word(1) << (n - base)) - 1;
      if (base <= i) counted &= ~Word(0) << (i + 1 - base);  // partners above i only
      disagreements += __builtin_popcountll((same ^ row[w]) & counted);
    }
  }
  return disagreements;
}

// f = (m - W) + (sum_c C(n_c, 2) - W): the edges between clusters plus the missing edges inside clusters, with the
// intra-cluster edges W and the edges m counted afresh from the bit rows (popcount of row & own-cluster mask, and of
// the row) on all cores; nothing of the solver's G or f is used.
long long ObjectiveBitwise(const Graph& graph, const std::vector<int>& labels) {
  CheckLabels(graph, labels);
  const unsigned n = graph.Size();
  const unsigned words = graph.WordsPerRow();
  const int k = labels.empty() ? 1 : *std::max_element(labels.begin(), labels.end()) + 1;
  std::vector<Word> member(static_cast<size_t>(k) * words, Word(0));
  std::vector<long long> size(static_cast<size_t>(k), 0);
  for (unsigned v = 0; v < n; ++v) {
    member[static_cast<size_t>(labels[v]) * words + v / kWordBits] |= Word(1) << (v % kWordBits);
    ++size[labels[v]];
  }
  // counts[0]: sum of degrees (2m), counts[1]: sum over v of v's neighbours in its own cluster (2W)
  long long counts[2];
  SumOverRows(n, counts, [&](unsigned first, unsigned last, long long* sums) {
    long long degree = 0, inside = 0;
    for (unsigned v = first; v < last; ++v) {
      const Word* row = graph.Row(v);
      const Word* own = member.data() + static_cast<size_t>(labels[v]) * words;
      for (unsigned w = 0; w < words; ++w) {
        degree += __builtin_popcountll(row[w]);
        inside += __builtin_popcountll(row[w] & own[w]);
      }
    }
    sums[0] = degree;
    sums[1] = inside;
  });
  long long pairs_inside = 0;
  for (const long long c : size) pairs_inside += c * (c - 1) / 2;
  const long long m = counts[0] / 2, w_inside = counts[1] / 2;
  return (m - w_inside) + (pairs_inside - w_inside);
}

int CountClustersUsed(const std::vector<int>& labels, int k) {
  if (k <= 0) return 0;
  std::vector<bool> seen(static_cast<size_t>(k), false);
  for (const int label : labels) {
    if (label >= 0 && label < k) seen[static_cast<size_t>(label)] = true;
  }
  return static_cast<int>(std::count(seen.begin(), seen.end(), true));
}

int ResolveThreads(const PbilsParams& params) {
  const int threads = params.threads >= 1 ? params.threads : HardwareThreads();
  return params.population >= 1 ? std::min(threads, params.population) : threads;
}

int WinnerCopies(double gww, int population) { return static_cast<int>(gww * population); }

std::vector<int> RankSlots(const long long* f, int population) {
  std::vector<int> order(population);
  for (int slot = 0; slot < population; ++slot) order[slot] = slot;
  std::stable_sort(order.begin(), order.end(), [f](int a, int b) { return f[a] < f[b]; });
  return order;
}

// The parameter rules of both backends (reference src/cc_pbils_cpu.cpp:177-181).
void CheckParams(const PbilsParams& params) {
  if (params.k < 1) throw std::runtime_error("the cluster bound k must be at least 1");
  if (params.population < 1) throw std::runtime_error("the population must hold at least one individual");
  if (!(params.gww >= 0.0 && params.gww <= kMaxGwwShare)) {
    throw std::runtime_error("the GWW share must lie in [0, 0.5]");
  }
}

// ---- GPU entry points over the C ABI ----------------------------------------------------------------------------

namespace {

cc_pbils_params ToAbi(const PbilsParams& p) {
  cc_pbils_params a;
  cc_pbils_default_params(&a);
  a.k = p.k;
  a.population = p.population;
  a.tournament = p.tournament;
  a.iterations = p.iterations;
  a.early_stop = p.early_stop;
  a.threads = p.threads;
  a.perturbation = p.perturbation;
  a.gww = p.gww;
  a.seed = p.seed;
  a.time_limit_sec = p.time_limit_sec;
  a.verbose = p.verbose ? 1 : 0;
  a.ls_kernel = p.ls_kernel;
  a.profile = p.profile ? 1 : 0;
  return a;
}

PbilsResult FromAbi(const cc_pbils_stats& s, std::vector<int> labels) {
  PbilsResult r;
  r.objective = s.objective;
  r.labels = std::move(labels);
  r.iterations_done = s.iterations_done;
  r.clusters_used = s.clusters_used;
  r.seconds = s.seconds;
  r.local_searches = s.local_searches;
  r.accepted_moves = s.accepted_moves;
  r.descent = s.descent;
  r.descent_threads = s.threads;
  r.search_ms = s.search_ms;
  r.rebuild_ms = s.rebuild_ms;
  return r;
}

PbilsResult Neighborhood(const Graph& graph, const PbilsParams& params, bool local_search) {
  const cc_pbils_params a = ToAbi(params);
  std::vector<int> labels(graph.Size());
  cc_pbils_stats stats;
  char err[512] = {0};
  const int rc = cc_hip_solve_neighborhood(graph.Bits().data(), graph.Size(), graph.EdgeCount(), &a,
                                           local_search ? 1 : 0, labels.data(), &stats, err, sizeof(err));
  if (rc != CC_HIP_OK) throw std::runtime_error(err);
  return FromAbi(stats, std::move(labels));
}

}  // namespace

PbilsResult SolveGpu(const Graph& graph, const PbilsParams& params) {
  const cc_pbils_params a = ToAbi(params);
  std::vector<int> labels(graph.Size());
  cc_pbils_stats stats;
  char err[512] = {0};
  const int rc = cc_hip_solve_pbils(graph.Bits().data(), graph.Size(), graph.EdgeCount(), &a, labels.data(), &stats,
                                    nullptr, err, sizeof(err));
  if (rc != CC_HIP_OK) throw std::runtime_error(err);
  return FromAbi(stats, std::move(labels));
}

PbilsResult SolveNeighborhoodGpu(const Graph& graph, const PbilsParams& params) {
  return Neighborhood(graph, params, false);
}

std::string AutotuneGpu(const Graph& graph, const PbilsParams& params, int iterations) {
  static const char* const kNames[] = {"generic, global", "generic, LDS", "k = 2, registers", "k = 2, packed",
                                       "k = 3, packed",   "registers",    "mixed"};
  const cc_pbils_params a = ToAbi(params);
  char err[512] = {0};
  cc_hip_graph* g = nullptr;
  if (cc_hip_graph_upload(graph.Bits().data(), graph.Size(), graph.EdgeCount(), &g, err, sizeof(err)) != CC_HIP_OK) {
    throw std::runtime_error(err);
  }
  std::vector<cc_tune_candidate> cands(64);
  int32_t count = 0, chosen = -1;
  const int rc = cc_hip_autotune(g, &a, iterations, cands.data(), static_cast<int32_t>(cands.size()), &count, &chosen,
                                 err, sizeof(err));
  cc_hip_graph_free(g);
  if (rc != CC_HIP_OK) throw std::runtime_error(err);
  std::string report;
  char line[256];
  for (int i = 0; i < std::min<int>(count, static_cast<int>(cands.size())); ++i) {
    const cc_tune_candidate& c = cands[i];
    const char* name = c.descent >= 0 && c.descent <= 6 ? kNames[c.descent] : "?";
    std::snprintf(line, sizeof(line), "tune %c %-18s threads=%-5d vpt=%-3d blocks=%-4d lds=%-7lld %10.3f ms\n",
                  i == chosen ? '*' : ' ', name, c.threads, c.vertices_per_thread, c.blocks,
                  static_cast<long long>(c.lds_bytes), c.search_ms);
    report += line;
  }
  return report;
}

PbilsResult SolveNeighborhoodWithManyLocalSearchesGpu(const Graph& graph, const PbilsParams& params) {
  return Neighborhood(graph, params, true);
}

}  // namespace cc

namespace cc {
namespace {

// A host copy of a graph's bits for the caller (released with cc_host_free).
uint64_t* HandOut(const Graph& g) {
  const std::vector<Word>& bits = g.Bits();
  uint64_t* out = static_cast<uint64_t*>(std::malloc(std::max<size_t>(1, bits.size()) * sizeof(uint64_t)));
  if (!out) throw std::runtime_error("out of host memory for the bitset");
  std::copy(bits.begin(), bits.end(), out);
  return out;
}

int HostError(const std::exception& e, char* err, size_t err_len) {
  if (err && err_len) std::snprintf(err, err_len, "%s", e.what());
  return CC_HIP_EINVAL;
}

}  // namespace

}  // namespace cc

// ---- the generator / recount part of the C ABI ----------------------------------------------------------------

extern "C" {

uint32_t cc_words64(uint32_t n) { return cc::WordsFor(n); }

uint64_t cc_graph_erdos_renyi(uint32_t n, double p, uint64_t seed, uint64_t* bits, int32_t threads) {
  cc::GeneratorErdosRenyi(n, p, seed, bits, threads);
  return cc::CountEdges(bits, static_cast<size_t>(n) * cc::WordsFor(n));
}

uint64_t cc_graph_sbm(uint32_t n, int32_t blocks, double p_in, double p_out, uint64_t seed, uint64_t* bits,
                      int32_t threads) {
  cc::GeneratorSbm(n, blocks, p_in, p_out, seed, bits, threads);
  return cc::CountEdges(bits, static_cast<size_t>(n) * cc::WordsFor(n));
}

uint64_t cc_graph_chung_lu(uint32_t n, double gamma, double mean_degree, uint64_t seed, uint64_t* bits,
                           int32_t threads) {
  cc::GeneratorChungLu(n, gamma, mean_degree, seed, bits, threads);
  return cc::CountEdges(bits, static_cast<size_t>(n) * cc::WordsFor(n));
}

int cc_graph_load_tags(const char* path, const char* similarity, double threshold, uint32_t n, uint64_t seed,
                       uint64_t** bits_out, uint32_t* n_out, uint64_t* edges_out, char* err, size_t err_len) {
  *bits_out = nullptr;
  try {
    const cc::Graph g = cc::Graph::LoadTags(path, similarity, threshold, n, seed);
    *bits_out = cc::HandOut(g);
    *n_out = g.Size();
    *edges_out = g.EdgeCount();
    return CC_HIP_OK;
  } catch (const std::exception& e) {
    if (err && err_len) std::snprintf(err, err_len, "%s", e.what());
    return CC_HIP_EINVAL;
  }
}

void cc_host_free(void* p) { std::free(p); }

int cc_graph_load(const char* path, int32_t format, uint64_t** bits_out, uint32_t* n_out, uint64_t* edges_out,
                  char* err, size_t err_len) {
  *bits_out = nullptr;
  try {
    cc::Graph g;
    switch (format) {
      case CC_GRAPH_MATRIX: g = cc::Graph::LoadMatrix(path); break;
      case CC_GRAPH_BASELINE_JSON: g = cc::Graph::LoadBaselineJson(path); break;
      case CC_GRAPH_EDGE_LIST: g = cc::Graph::LoadEdgeList(path); break;
      default: throw std::runtime_error("unknown graph file format " + std::to_string(format));
    }
    *bits_out = cc::HandOut(g);
    *n_out = g.Size();
    *edges_out = g.EdgeCount();
    return CC_HIP_OK;
  } catch (const std::exception& e) {
    return cc::HostError(e, err, err_len);
  }
}

int cc_graph_save(const uint64_t* bits, uint32_t n, const char* path, int32_t format, char* err, size_t err_len) {
  try {
    const cc::Graph g = cc::Graph::FromBits(std::vector<cc::Word>(bits, bits + static_cast<size_t>(n) * cc::WordsFor(n)), n);
    switch (format) {
      case CC_GRAPH_MATRIX: g.SaveMatrix(path); break;
      case CC_GRAPH_BASELINE_JSON: g.SaveBaselineJson(path); break;
      case CC_GRAPH_EDGE_LIST: g.SaveEdgeList(path); break;
      default: throw std::runtime_error("unknown graph file format " + std::to_string(format));
    }
    return CC_HIP_OK;
  } catch (const std::exception& e) {
    return cc::HostError(e, err, err_len);
  }
}

void cc_sample_without_replacement(uint32_t total, uint32_t n, uint64_t seed, uint32_t* out) {
  const std::vector<unsigned> order = cc::SampleWithoutReplacement(total, n, seed);
  std::copy(order.begin(), order.end(), out);
}

int64_t cc_objective_direct(const uint64_t* bits, uint32_t n, const int32_t* labels) {
  const uint32_t words = cc::WordsFor(n);
  int64_t distance = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint64_t* row = bits + static_cast<size_t>(i) * words;
    for (uint32_t j = i + 1; j < n; ++j) {
      const bool joined = (row[j / 64] >> (j % 64)) & 1u;
      if ((labels[i] == labels[j]) != joined) ++distance;
    }
  }
  return distance;
}

int64_t cc_objective_bitwise(const uint64_t* bits, uint32_t n, const int32_t* labels) {
  const uint32_t words = cc::WordsFor(n);
  std::vector<cc::Word> copy(bits, bits + static_cast<size_t>(n) * words);
  const cc::Graph graph = cc::Graph::FromBits(std::move(copy), n);
  return cc::ObjectiveBitwise(graph, std::vector<int>(labels, labels + n));
}

}  // extern "C"
