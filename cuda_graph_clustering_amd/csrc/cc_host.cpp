// cc_host.cpp — host side of the library: the Graph type with its generators and file formats, the objective
// recounts, the GWW helpers, the C++ entry points of cc_gpu.hpp over the C ABI, and the generator/recount part of
// the C ABI.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "cc_gpu.hpp"
#include "cc_graph.hpp"
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

std::string ReadWholeFile(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open file: " + path);
  std::stringstream buffer;
  buffer << in.rdbuf();
  return buffer.str();
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
  FillPairs(n, WordsFor(n), StreamSeed(seed, kChungLuStream, 0), bits, threads, [&](unsigned i, unsigned j) {
    const double q = w[i] * w[j] * inv_total;
    return q > 1.0 ? 1.0 : q;
  });
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

Graph Graph::LoadMatrix(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open graph file: " + path);
  unsigned n = 0;
  in >> n;
  if (!in || n == 0) throw std::runtime_error("bad graph header in " + path);
  Graph g(n);
  for (unsigned i = 0; i < n; ++i) {
    for (unsigned j = 0; j < n; ++j) {
      int ch = in.get();
      while (ch != EOF && std::isspace(ch)) ch = in.get();
      if (ch == EOF) throw std::runtime_error("truncated matrix in " + path);
      if (ch == '1' && j > i) g.AddEdge(i, j);
    }
  }
  return g;
}

void Graph::SaveMatrix(const std::string& path) const {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write graph file: " + path);
  out << n_ << "\n";
  std::string line(n_, '0');
  for (unsigned i = 0; i < n_; ++i) {
    for (unsigned j = 0; j < n_; ++j) line[j] = IsJoined(i, j) ? '1' : '0';
    out << line << "\n";
  }
}

Graph Graph::LoadBaselineJson(const std::string& path) {
  const std::string text = ReadWholeFile(path);
  const size_t key = text.find("\"graph\"");
  if (key == std::string::npos) throw std::runtime_error("no \"graph\" key in " + path);
  size_t pos = text.find('[', key);
  if (pos == std::string::npos) throw std::runtime_error("malformed \"graph\" block in " + path);
  std::vector<std::vector<char>> rows;
  std::vector<char> current;
  int depth = 0;
  for (; pos < text.size(); ++pos) {
    const char ch = text[pos];
    if (ch == '[') {
      if (++depth == 2) current.clear();
    } else if (ch == ']') {
      if (depth == 2) rows.push_back(current);
      if (--depth == 0) break;
    } else if (depth == 2 && (ch == '0' || ch == '1')) {
      current.push_back(static_cast<char>(ch - '0'));
    }
  }
  if (rows.empty()) throw std::runtime_error("empty \"graph\" block in " + path);
  const unsigned n = static_cast<unsigned>(rows.size());
  Graph g(n);
  for (unsigned i = 0; i < n; ++i) {
    if (rows[i].size() != n) {
      throw std::runtime_error("row " + std::to_string(i) + " of \"graph\" is not n long in " + path);
    }
    for (unsigned j = i + 1; j < n; ++j) {
      if (rows[i][j]) g.AddEdge(i, j);
    }
  }
  return g;
}

void Graph::SaveBaselineJson(const std::string& path) const {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write json file: " + path);
  out << "{\n\"size\": " << n_ << ",\n\"density\": " << Density() << ",\n\"graph\": [\n";
  for (unsigned i = 0; i < n_; ++i) {
    out << "[";
    for (unsigned j = 0; j < n_; ++j) out << (IsJoined(i, j) ? 1 : 0) << (j + 1 != n_ ? "," : "");
    out << (i + 1 != n_ ? "],\n" : "]\n");
  }
  out << "]\n}\n";
}

Graph Graph::LoadEdgeList(const std::string& path) {
  const std::string text = ReadWholeFile(path);
  std::vector<uint64_t> ends;
  size_t pos = 0;
  auto read_id = [&](uint64_t& id) {
    while (pos < text.size() && (text[pos] == ' ' || text[pos] == '\t' || text[pos] == ',')) ++pos;
    if (pos >= text.size() || !std::isdigit(static_cast<unsigned char>(text[pos]))) return false;
    for (id = 0; pos < text.size() && std::isdigit(static_cast<unsigned char>(text[pos])); ++pos) {
      if (id > (UINT64_MAX - 9) / 10) throw std::runtime_error("vertex id does not fit 64 bits in " + path);
      id = id * 10 + static_cast<uint64_t>(text[pos] - '0');
    }
    return true;
  };
  while (pos < text.size()) {
    const size_t line_end = std::min(text.find('\n', pos), text.size());
    uint64_t u = 0, v = 0;
    if (text[pos] != '#' && text[pos] != '%' && read_id(u) && read_id(v)) {
      ends.push_back(u);
      ends.push_back(v);
    }
    pos = line_end + 1;
  }
  if (ends.empty()) throw std::runtime_error("no edges in " + path);
  std::vector<uint64_t> ids(ends);
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  if (ids.size() > UINT32_MAX) throw std::runtime_error("too many vertices in " + path);
  auto rank = [&](uint64_t id) {
    return static_cast<unsigned>(std::lower_bound(ids.begin(), ids.end(), id) - ids.begin());
  };
  Graph g(static_cast<unsigned>(ids.size()));
  for (size_t i = 0; i < ends.size(); i += 2) g.AddEdge(rank(ends[i]), rank(ends[i + 1]));
  return g;
}

void Graph::SaveEdgeList(const std::string& path) const {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write edge list: " + path);
  out << "# " << n_ << " vertices, " << edges_ << " edges\n";
  for (unsigned i = 0; i < n_; ++i) {
    for (unsigned j = i + 1; j < n_; ++j) {
      if (IsJoined(i, j)) out << i << ' ' << j << '\n';
    }
  }
}

// ---- recounts and GWW helpers -----------------------------------------------------------------------------------

long long ObjectiveDirect(const Graph& graph, const std::vector<int>& labels) {
  const unsigned n = graph.Size();
  if (labels.size() != n) throw std::runtime_error("labels length != graph size");
  long long distance = 0;
  for (unsigned i = 0; i < n; ++i) {
    for (unsigned j = i + 1; j < n; ++j) {
      if ((labels[i] == labels[j]) != graph.IsJoined(i, j)) ++distance;
    }
  }
  return distance;
}

long long ObjectiveBitwise(const Graph& graph, const std::vector<int>& labels) {
  const unsigned n = graph.Size();
  if (labels.size() != n) throw std::runtime_error("labels length != graph size");
  const unsigned words = graph.WordsPerRow();
  int k = 0;
  for (const int label : labels) {
    if (label < 0) throw std::runtime_error("negative label");
    k = std::max(k, label + 1);
  }
  std::vector<Word> masks(static_cast<size_t>(std::max(k, 1)) * words, 0);
  for (unsigned v = 0; v < n; ++v) masks[static_cast<size_t>(labels[v]) * words + v / kWordBits] |= Word(1) << (v % kWordBits);
  constexpr unsigned kRowsPerItem = 256;
  const int items = static_cast<int>((n + kRowsPerItem - 1) / kRowsPerItem);
  std::vector<long long> sums(static_cast<size_t>(std::max(items, 1)), 0);
  ParallelFor(items, 0, [&](int item) {
    long long acc = 0;
    for (unsigned v = item * kRowsPerItem; v < std::min(n, (item + 1) * kRowsPerItem); ++v) {
      const Word* row = graph.Row(v);
      const Word* mask = masks.data() + static_cast<size_t>(labels[v]) * words;
      for (unsigned w = 0; w < words; ++w) acc += __builtin_popcountll(row[w] ^ mask[w]);
      acc -= 1;  // the vertex's own bit: set in its mask, clear in its row
    }
    sums[item] = acc;
  });
  long long total = 0;
  for (const long long s : sums) total += s;
  return total / 2;
}

int CountClustersUsed(const std::vector<int>& labels, int k) {
  std::vector<char> seen(static_cast<size_t>(std::max(k, 1)), 0);
  for (const int label : labels) {
    if (label >= 0 && label < k) seen[label] = 1;
  }
  int used = 0;
  for (const char s : seen) used += s;
  return used;
}

int WinnerCopies(double gww, int population) { return static_cast<int>(gww * population); }

std::vector<int> RankSlots(const long long* f, int population) {
  std::vector<int> order(population);
  for (int slot = 0; slot < population; ++slot) order[slot] = slot;
  std::stable_sort(order.begin(), order.end(), [f](int a, int b) { return f[a] < f[b]; });
  return order;
}

void CheckParams(const PbilsParams& params) {
  if (params.k < 1) throw std::runtime_error("k must be >= 1");
  if (params.population < 1) throw std::runtime_error("population must be >= 1");
  if (params.gww < 0.0 || params.gww > kMaxGwwShare) throw std::runtime_error("gww must be in [0, 0.5]");
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
