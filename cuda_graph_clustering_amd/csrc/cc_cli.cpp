// cc_cli.cpp — the cc_gpu command line on the HIP library.
//
// What a harness of the reference sees is kept: the flag names and their lax number parsing (atoi/strtod: a
// malformed number reads as 0), the choice of instance (tags, then baseline JSON, then edge list, then matrix, then
// a generated graph), run r solved with seed + r, the stdout lines, the recount check (exit status 2 on a mismatch),
// the --out JSON keys and the --labels-out format (include/cc_cli.hpp:124-183, 238-277, 285-376 of the reference).
// The program around them is this file's: a typed flag table read by one dispatch, a result tally, and output built
// with stdio. Additions: --generator (er, sbm, chung-lu) with its parameters, --save-edges, --autotune and --profile.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "cc_gpu.hpp"
#include "cc_graph.hpp"
#include "cc_pbils.hpp"

namespace {

enum class Algo { kPbils, kNeighborhood, kNeighborhoodLocalSearch };

const char* AlgoName(Algo a) {
  switch (a) {
    case Algo::kNeighborhood: return "NeighborhoodCuda";
    case Algo::kNeighborhoodLocalSearch: return "NeighborhoodWithManyLocalSearchesCuda";
    case Algo::kPbils: break;
  }
  return "PBILS";
}

struct Options {
  // instance
  unsigned n = 500;
  double density = 0.5;
  uint64_t graph_seed = 1;
  std::string generator = "er";
  int blocks = 5;
  double p_in = 0.4, p_out = 0.05, gamma = 2.5, mean_degree = 64.0;
  std::string matrix_in, json_in, edges_in, tags_in;
  std::string similarity = "jaccard";
  double threshold = 0.5;
  std::string matrix_out, json_out, edges_out;
  // search
  cc::PbilsParams params;
  int algorithm = static_cast<int>(Algo::kPbils);
  // experiment
  int runs = 1;
  int autotune = 0;
  bool verify = true;
  std::string result_out, labels_out;
};

// ---- the flag table --------------------------------------------------------------------------------------------

enum class Kind { kInt, kUnsigned, kUnsigned64, kReal, kText, kSetTrue, kSetFalse, kChoice };

struct Named {
  const char* name;
  int value;
};

struct Flag {
  const char* name;
  const char* arg;  // the value's name in the help; nullptr for a switch
  const char* help;
  Kind kind;
  void* target;
  std::vector<Named> choices = {};
};

const std::vector<Named> kAlgorithms = {{"pbils", static_cast<int>(Algo::kPbils)},
                                        {"NeighborhoodCuda", static_cast<int>(Algo::kNeighborhood)},
                                        {"NeighborhoodWithManyLocalSearchesCuda",
                                         static_cast<int>(Algo::kNeighborhoodLocalSearch)}};
const std::vector<Named> kDescents = {{"auto", cc::PbilsParams::kLocalSearchAuto},
                                      {"shared", cc::PbilsParams::kLocalSearchShared},
                                      {"global", cc::PbilsParams::kLocalSearchGlobal},
                                      {"wide", cc::PbilsParams::kLocalSearchWide}};

struct Section {
  const char* title;
  std::vector<Flag> flags;
};

std::vector<Section> FlagTable(Options& o) {
  cc::PbilsParams& p = o.params;
  return {
      {"instance",
       {{"--n", "N", "vertices of a generated graph; objects sampled for --tags (0: all)", Kind::kUnsigned, &o.n},
        {"--density", "P", "edge probability of --generator er", Kind::kReal, &o.density},
        {"--graph-seed", "S", "generator seed: fixes the instance", Kind::kUnsigned64, &o.graph_seed},
        {"--generator", "G", "er, sbm or chung-lu", Kind::kText, &o.generator},
        {"--blocks", "B", "sbm: hidden clusters", Kind::kInt, &o.blocks},
        {"--p-in", "P", "sbm: edge probability inside a block", Kind::kReal, &o.p_in},
        {"--p-out", "P", "sbm: edge probability across blocks", Kind::kReal, &o.p_out},
        {"--gamma", "G", "chung-lu: degree exponent", Kind::kReal, &o.gamma},
        {"--mean-degree", "D", "chung-lu: mean degree", Kind::kReal, &o.mean_degree},
        {"--graph", "PATH", "read a 0/1 matrix text instead of generating", Kind::kText, &o.matrix_in},
        {"--graph-json", "PATH", "read the \"graph\" block of a baseline result file", Kind::kText, &o.json_in},
        {"--edges", "PATH", "read an edge list (SNAP, CSV); a vertex is the rank of its id", Kind::kText, &o.edges_in},
        {"--tags", "PATH", "tag-similarity graph of a {\"key\": [tags]} file over --n objects drawn by --graph-seed",
         Kind::kText, &o.tags_in},
        {"--similarity", "M", "--tags measure: jaccard, cosine, dice or overlap", Kind::kText, &o.similarity},
        {"--threshold", "T", "--tags: join objects whose similarity is at least T", Kind::kReal, &o.threshold},
        {"--save-graph", "PATH", "write the instance as a 0/1 matrix", Kind::kText, &o.matrix_out},
        {"--save-graph-json", "PATH", "write the instance as a baseline JSON", Kind::kText, &o.json_out},
        {"--save-edges", "PATH", "write the instance as an edge list", Kind::kText, &o.edges_out}}},
      {"algorithm",
       {{"--algorithm", "NAME", "pbils, NeighborhoodCuda or NeighborhoodWithManyLocalSearchesCuda (k = 2)",
         Kind::kChoice, &o.algorithm, kAlgorithms},
        {"--k", "K", "largest number of clusters", Kind::kInt, &p.k},
        {"--pop", "P", "population; for Neighborhood the start partitions of a batch", Kind::kInt, &p.population},
        {"--tournament", "T", "tournament size", Kind::kInt, &p.tournament},
        {"--iters", "I", "iteration limit", Kind::kInt, &p.iterations},
        {"--early-stop", "E", "stop after E iterations without a new record", Kind::kInt, &p.early_stop},
        {"--perturb", "Q", "probability that a perturbation relabels a vertex", Kind::kReal, &p.perturbation},
        {"--gww", "R", "share of the worst individuals replaced by copies of the best (0 to 0.5)", Kind::kReal,
         &p.gww},
        {"--seed", "S", "search seed (run r uses S + r)", Kind::kUnsigned64, &p.seed},
        {"--threads", "T", "accepted for the reference's CPU binary; the GPU ignores it", Kind::kInt, &p.threads},
        {"--time-limit", "T", "seconds, 0: none", Kind::kReal, &p.time_limit_sec},
        {"--ls-kernel", "WHERE", "descent kernel: auto, shared (G in LDS), global or wide (32-bit G)", Kind::kChoice,
         &p.ls_kernel, kDescents}}},
      {"experiment",
       {{"--runs", "R", "independent runs; min/avg/max are reported", Kind::kInt, &o.runs},
        {"--autotune", "I", "time every descent plan on I iterations first and keep the fastest", Kind::kInt,
         &o.autotune},
        {"--no-verify", nullptr, "do not recount the objective of the result", Kind::kSetFalse, &o.verify},
        {"--out", "PATH", "write the best result as JSON", Kind::kText, &o.result_out},
        {"--labels-out", "PATH", "write the best clustering's labels", Kind::kText, &o.labels_out},
        {"--verbose", nullptr, "report every iteration", Kind::kSetTrue, &p.verbose},
        {"--profile", nullptr, "time the descent and rebuild kernels with HIP events", Kind::kSetTrue, &p.profile}}},
  };
}

// Stores `text` into the flag's target; false if a choice does not name one of its values.
bool Assign(const Flag& f, const char* text) {
  switch (f.kind) {
    case Kind::kInt: *static_cast<int*>(f.target) = std::atoi(text); return true;
    case Kind::kUnsigned: *static_cast<unsigned*>(f.target) = static_cast<unsigned>(std::strtoul(text, nullptr, 10)); return true;
    case Kind::kUnsigned64: *static_cast<uint64_t*>(f.target) = std::strtoull(text, nullptr, 10); return true;
    case Kind::kReal: *static_cast<double*>(f.target) = std::strtod(text, nullptr); return true;
    case Kind::kText: *static_cast<std::string*>(f.target) = text; return true;
    case Kind::kSetTrue: *static_cast<bool*>(f.target) = true; return true;
    case Kind::kSetFalse: *static_cast<bool*>(f.target) = false; return true;
    case Kind::kChoice:
      for (const Named& c : f.choices) {
        if (std::strcmp(c.name, text) == 0) {
          *static_cast<int*>(f.target) = c.value;
          return true;
        }
      }
      return false;
  }
  return false;
}

std::string Default(const Flag& f) {
  char text[64] = "";
  switch (f.kind) {
    case Kind::kInt: std::snprintf(text, sizeof(text), "%d", *static_cast<int*>(f.target)); break;
    case Kind::kUnsigned: std::snprintf(text, sizeof(text), "%u", *static_cast<unsigned*>(f.target)); break;
    case Kind::kUnsigned64:
      std::snprintf(text, sizeof(text), "%llu", static_cast<unsigned long long>(*static_cast<uint64_t*>(f.target)));
      break;
    case Kind::kReal: std::snprintf(text, sizeof(text), "%g", *static_cast<double*>(f.target)); break;
    case Kind::kText: return *static_cast<std::string*>(f.target);
    case Kind::kChoice:
      for (const Named& c : f.choices) {
        if (c.value == *static_cast<int*>(f.target)) return c.name;
      }
      break;
    case Kind::kSetTrue:
    case Kind::kSetFalse: break;
  }
  return text;
}

void Usage(const char* program) {
  Options defaults;
  std::printf("usage: %s [options]\n", program);
  for (const Section& s : FlagTable(defaults)) {
    std::printf("\n%s:\n", s.title);
    for (const Flag& f : s.flags) {
      char head[48];
      std::snprintf(head, sizeof(head), "%s%s%s", f.name, f.arg ? " " : "", f.arg ? f.arg : "");
      const std::string fallback = Default(f);
      std::printf("  %-26s %s%s%s%s\n", head, f.help, fallback.empty() ? "" : " [", fallback.c_str(),
                  fallback.empty() ? "" : "]");
    }
  }
}

// 0: options read; otherwise the exit status (help or a bad command line: 1, as the reference).
int ReadCommandLine(int argc, char** argv, Options& o) {
  const std::vector<Section> table = FlagTable(o);
  for (int i = 1; i < argc; ++i) {
    const std::string name = argv[i];
    if (name == "--help" || name == "-h") {
      Usage(argv[0]);
      return 1;
    }
    const Flag* found = nullptr;
    for (const Section& s : table) {
      for (const Flag& f : s.flags) found = name == f.name ? &f : found;
    }
    if (!found) {
      std::printf("error: no option %s (--help lists them)\n", name.c_str());
      return 1;
    }
    const char* value = "";
    if (found->arg) {
      if (i + 1 == argc) {
        std::printf("error: option %s expects a value %s\n", name.c_str(), found->arg);
        return 1;
      }
      value = argv[++i];
    }
    if (!Assign(*found, value)) {
      std::printf("error: %s is not a value of %s (--help lists them)\n", value, name.c_str());
      return 1;
    }
  }
  return 0;
}

// ---- the instance ----------------------------------------------------------------------------------------------

cc::Graph Instance(const Options& o) {
  if (!o.tags_in.empty()) return cc::Graph::LoadTags(o.tags_in, o.similarity, o.threshold, o.n, o.graph_seed);
  if (!o.json_in.empty()) return cc::Graph::LoadBaselineJson(o.json_in);
  if (!o.edges_in.empty()) return cc::Graph::LoadEdgeList(o.edges_in);
  if (!o.matrix_in.empty()) return cc::Graph::LoadMatrix(o.matrix_in);
  if (o.generator == "er") return cc::Graph::ErdosRenyi(o.n, o.density, o.graph_seed);
  if (o.generator == "sbm") return cc::Graph::StochasticBlock(o.n, o.blocks, o.p_in, o.p_out, o.graph_seed);
  if (o.generator == "chung-lu") return cc::Graph::ChungLu(o.n, o.gamma, o.mean_degree, o.graph_seed);
  throw std::runtime_error("--generator takes er, sbm or chung-lu, not '" + o.generator + "'");
}

// ---- results ---------------------------------------------------------------------------------------------------

// The runs so far: the best result (smallest f, the first of equals), the largest f, and the sums for the means.
struct Tally {
  int count = 0;
  cc::PbilsResult best;
  long long worst = 0;
  double f_sum = 0.0, seconds_sum = 0.0;

  void Add(const cc::PbilsResult& r) {
    if (count == 0 || r.objective < best.objective) best = r;
    worst = count == 0 ? r.objective : std::max(worst, r.objective);
    f_sum += static_cast<double>(r.objective);
    seconds_sum += r.seconds;
    ++count;
  }
  double MeanF() const { return f_sum / count; }
  double MeanSeconds() const { return seconds_sum / count; }
};

// A double as the reference's std::ostream writes it (six significant digits).
std::string Num(double x) {
  char text[40];
  std::snprintf(text, sizeof(text), "%g", x);
  return text;
}

// The result file; keys and their order are the reference's (include/cc_cli.hpp:256-276), which keeps the
// baseline's.
bool WriteResult(const std::string& path, const cc::Graph& g, const Options& o, const Tally& t) {
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  const cc::PbilsParams& p = o.params;
  std::string s = "{\n";
  auto field = [&s](const char* key, const std::string& value) { s += std::string("  \"") + key + "\": " + value + ",\n"; };
  field("algorithm", std::string("\"") + AlgoName(static_cast<Algo>(o.algorithm)) + "\"");
  field("backend", "\"gpu\"");
  field("size", std::to_string(g.Size()));
  field("edges", std::to_string(g.EdgeCount()));
  field("density", Num(g.Density()));
  field("k", std::to_string(p.k));
  field("population", std::to_string(p.population));
  field("gww", Num(p.gww));
  field("runs", std::to_string(o.runs));
  field("objective function value", std::to_string(t.best.objective));
  field("objective average", Num(t.MeanF()));
  field("objective max", std::to_string(t.worst));
  field("computation time seconds", Num(t.best.seconds));
  field("computation time average", Num(t.MeanSeconds()));
  field("clusters used", std::to_string(t.best.clusters_used));
  s += "  \"clustering vector\": [";
  for (size_t v = 0; v < t.best.labels.size(); ++v) {
    if (v) s += ',';
    s += std::to_string(t.best.labels[v]);
  }
  s += "]\n}\n";
  const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
  return std::fclose(f) == 0 && ok;
}

bool WriteLabels(const std::string& path, const std::vector<int>& labels) {
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  std::string s;
  for (size_t v = 0; v < labels.size(); ++v) {
    s += std::to_string(labels[v]);
    s += v + 1 == labels.size() ? '\n' : ' ';
  }
  const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
  return std::fclose(f) == 0 && ok;
}

cc::PbilsResult Solve(Algo a, const cc::Graph& g, const cc::PbilsParams& p) {
  switch (a) {
    case Algo::kNeighborhood: return cc::SolveNeighborhoodGpu(g, p);
    case Algo::kNeighborhoodLocalSearch: return cc::SolveNeighborhoodWithManyLocalSearchesGpu(g, p);
    case Algo::kPbils: break;
  }
  return cc::SolveGpu(g, p);
}

// Graphs up to this size are recounted pair by pair; larger ones over the bit rows.
constexpr unsigned kPairwiseRecountMaxN = 32767;

int Run(const Options& o) {
  cc::Graph g;
  try {
    g = Instance(o);
    if (!o.matrix_out.empty()) g.SaveMatrix(o.matrix_out);
    if (!o.json_out.empty()) g.SaveBaselineJson(o.json_out);
    if (!o.edges_out.empty()) g.SaveEdgeList(o.edges_out);
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  const cc::PbilsParams& p = o.params;
  const Algo algo = static_cast<Algo>(o.algorithm);
  std::printf("backend      gpu\n");
  std::printf("instance     n=%u  edges=%llu  density=%.4f\n", g.Size(), static_cast<unsigned long long>(g.EdgeCount()),
              g.Density());
  if (algo == Algo::kPbils) {
    std::printf("algorithm    PBILS  k=%d  pop=%d  tournament=%d  iters=%d  early-stop=%d  perturb=%.2f  gww=%.2f\n",
                p.k, p.population, p.tournament, p.iterations, p.early_stop, p.perturbation, p.gww);
  } else {
    std::printf("algorithm    %s  k=%d  batch=%d\n", AlgoName(algo), p.k, p.population);
  }
  try {
    if (o.autotune > 0 && algo == Algo::kPbils) std::fputs(cc::AutotuneGpu(g, p, o.autotune).c_str(), stdout);
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }

  const bool pairwise = g.Size() <= kPairwiseRecountMaxN;
  Tally tally;
  for (int r = 0; r < o.runs; ++r) {
    cc::PbilsParams run_params = p;
    run_params.seed = p.seed + static_cast<uint64_t>(r);
    cc::PbilsResult res;
    try {
      res = Solve(algo, g, run_params);
    } catch (const std::exception& e) {
      std::printf("error: %s\n", e.what());
      return 1;
    }
    if (o.verify) {
      const long long recount = pairwise ? cc::ObjectiveDirect(g, res.labels) : cc::ObjectiveBitwise(g, res.labels);
      if (recount != res.objective) {
        std::printf("MISMATCH run %d: reported %lld, recount %lld\n", r, res.objective, recount);
        return 2;
      }
    }
    tally.Add(res);
    std::printf("run %-3d      f=%-10lld  clusters=%-3d  iters=%-4d  %.3fs\n", r, res.objective, res.clusters_used,
                res.iterations_done, res.seconds);
    if (p.verbose && res.local_searches > 0) {
      std::printf("moves        %lld accepted, %.0f per local search\n", res.accepted_moves,
                  static_cast<double>(res.accepted_moves) / static_cast<double>(res.local_searches));
    }
    if (p.profile) {
      const double rate = res.search_ms > 0 ? 1e3 * static_cast<double>(res.accepted_moves) / res.search_ms : 0.0;
      std::printf("device       descent %.3f ms, rebuild %.3f ms, %.3g moves/s\n", res.search_ms, res.rebuild_ms, rate);
    }
  }
  if (tally.count == 0) return 0;
  std::printf("---\n");
  std::printf("objective    min=%lld  avg=%.2f  max=%lld\n", tally.best.objective, tally.MeanF(), tally.worst);
  std::printf("time         best=%.3fs  avg=%.3fs\n", tally.best.seconds, tally.MeanSeconds());
  if (o.verify) {
    std::printf(pairwise ? "verified     reported value recounted pairwise in O(n^2)\n"
                         : "verified     reported value recounted over bit rows in O(n^2/32)\n");
  }
  if (!o.result_out.empty()) {
    if (WriteResult(o.result_out, g, o, tally)) {
      std::printf("written      %s\n", o.result_out.c_str());
    } else {
      std::printf("warning: cannot write %s\n", o.result_out.c_str());
    }
  }
  if (!o.labels_out.empty()) {
    if (WriteLabels(o.labels_out, tally.best.labels)) {
      std::printf("written      %s\n", o.labels_out.c_str());
    } else {
      std::printf("warning: cannot write %s\n", o.labels_out.c_str());
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  if (const int status = ReadCommandLine(argc, argv, o)) return status;
  return Run(o);
}
