// cc_cli.cpp — the cc_gpu command line: the flags and output lines of the reference CLI (include/cc_cli.hpp of the
// reference, flags at :124-183, output at :307-376), on the HIP library. Additions: --generator (er, sbm,
// chung-lu) with its parameters, --save-edges, and --profile.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <utility>
#include <vector>

#include "cc_gpu.hpp"
#include "cc_graph.hpp"
#include "cc_pbils.hpp"

namespace {

enum Algorithm { kPbils, kNeighborhood, kNeighborhoodWithManyLocalSearches };

const char* AlgorithmName(int algorithm) {
  switch (algorithm) {
    case kNeighborhood: return "NeighborhoodCuda";
    case kNeighborhoodWithManyLocalSearches: return "NeighborhoodWithManyLocalSearchesCuda";
    default: return "PBILS";
  }
}

struct Options {
  unsigned n = 500;
  double density = 0.5;
  uint64_t graph_seed = 1;
  std::string generator = "er";
  int blocks = 5;
  double p_in = 0.4;
  double p_out = 0.05;
  double gamma = 2.5;
  double mean_degree = 64.0;
  std::string graph_path, graph_json_path, edges_path, save_graph_path, save_graph_json_path, save_edges_path;
  std::string tags_path, similarity = "jaccard";
  double threshold = 0.5;
  cc::PbilsParams params;
  int algorithm = kPbils;
  int runs = 1;
  int autotune = 0;  // iterations per plan; 0: no tuning
  bool verify = true;
  std::string out_path, labels_out_path;
};

struct Flag {
  const char* name;
  const char* value;  // nullptr: a switch
  const char* help;
  std::function<bool(const char*)> apply;
};

template <typename T>
Flag Option(const char* name, const char* value, const char* help, T& field) {
  return {name, value, help, [&field](const char* text) {
            if constexpr (std::is_same_v<T, int>) field = std::atoi(text);
            else if constexpr (std::is_same_v<T, unsigned>) field = static_cast<unsigned>(std::strtoul(text, nullptr, 10));
            else if constexpr (std::is_same_v<T, uint64_t>) field = std::strtoull(text, nullptr, 10);
            else if constexpr (std::is_same_v<T, double>) field = std::strtod(text, nullptr);
            else field = text;
            return true;
          }};
}

Flag Switch(const char* name, const char* help, bool& field, bool on) {
  return {name, nullptr, help, [&field, on](const char*) {
            field = on;
            return true;
          }};
}

Flag Choice(const char* name, const char* value, const char* help, int& field,
            std::vector<std::pair<const char*, int>> choices) {
  return {name, value, help, [&field, choices](const char* text) {
            for (const auto& c : choices) {
              if (!std::strcmp(text, c.first)) {
                field = c.second;
                return true;
              }
            }
            return false;
          }};
}

std::vector<Flag> Flags(Options& o) {
  cc::PbilsParams& p = o.params;
  return {
      Option("--n", "N", "generate a graph on N vertices", o.n),
      Option("--density", "P", "edge probability of --generator er", o.density),
      Option("--graph-seed", "S", "generator seed, fixes the instance", o.graph_seed),
      Option("--generator", "G", "er, sbm or chung-lu", o.generator),
      Option("--blocks", "B", "sbm: hidden clusters", o.blocks),
      Option("--p-in", "P", "sbm: edge probability inside a block", o.p_in),
      Option("--p-out", "P", "sbm: edge probability across blocks", o.p_out),
      Option("--gamma", "G", "chung-lu: degree exponent", o.gamma),
      Option("--mean-degree", "D", "chung-lu: mean degree", o.mean_degree),
      Option("--graph", "PATH", "load a 0/1 matrix instead of generating", o.graph_path),
      Option("--graph-json", "PATH", "load the \"graph\" block of a baseline result file", o.graph_json_path),
      Option("--edges", "PATH", "load an edge list (SNAP, CSV): vertex = rank of its id", o.edges_path),
      Option("--tags", "PATH",
             "tag-similarity graph of a {\"key\": [tags]} JSON file: --n objects sampled by --graph-seed "
             "(0: all), joined when the similarity of their tags is >= --threshold",
             o.tags_path),
      Option("--similarity", "M", "measure for --tags: jaccard, cosine, dice or overlap", o.similarity),
      Option("--threshold", "T", "threshold for --tags", o.threshold),
      Option("--save-graph", "PATH", "save the instance as a 0/1 matrix", o.save_graph_path),
      Option("--save-graph-json", "PATH", "save the instance in the baseline JSON format", o.save_graph_json_path),
      Option("--save-edges", "PATH", "save the instance as an edge list", o.save_edges_path),
      Choice("--algorithm", "NAME", "pbils, NeighborhoodCuda or NeighborhoodWithManyLocalSearchesCuda", o.algorithm,
             {{"pbils", kPbils},
              {"NeighborhoodCuda", kNeighborhood},
              {"NeighborhoodWithManyLocalSearchesCuda", kNeighborhoodWithManyLocalSearches}}),
      Option("--k", "K", "upper bound on the number of clusters", p.k),
      Option("--pop", "P", "population; for Neighborhood the start partitions per batch", p.population),
      Option("--tournament", "T", "tournament size", p.tournament),
      Option("--iters", "I", "iteration limit", p.iterations),
      Option("--early-stop", "E", "stop after E iterations without a new record", p.early_stop),
      Option("--perturb", "Q", "probability of relabelling a vertex", p.perturbation),
      Option("--gww", "R", "GWW share of the worst replaced by copies of the best, up to 0.5; 0 disables", p.gww),
      Option("--seed", "S", "search seed", p.seed),
      Option("--threads", "T", "CPU only in the reference; accepted and ignored", p.threads),
      Option("--time-limit", "T", "time limit in seconds, 0 disables", p.time_limit_sec),
      Choice("--ls-kernel", "WHERE", "auto, shared (LDS), global or wide (32-bit G)", p.ls_kernel,
             {{"auto", cc::PbilsParams::kLocalSearchAuto},
              {"shared", cc::PbilsParams::kLocalSearchShared},
              {"global", cc::PbilsParams::kLocalSearchGlobal},
              {"wide", cc::PbilsParams::kLocalSearchWide}}),
      Option("--runs", "R", "independent runs (seed, seed+1, ...), prints min/avg/max", o.runs),
      Option("--autotune", "I", "before the runs, time every descent plan on I iterations and use the fastest",
             o.autotune),
      Switch("--no-verify", "skip the recount of the objective", o.verify, false),
      Option("--out", "PATH", "save the result as JSON", o.out_path),
      Option("--labels-out", "PATH", "save the best clustering as labels", o.labels_out_path),
      Switch("--verbose", "print progress per iteration", p.verbose, true),
      Switch("--profile", "time the descent and rebuild kernels with HIP events", p.profile, true),
  };
}

void PrintUsage(const char* program) {
  Options defaults;
  std::printf("usage: %s [options]\n", program);
  for (const Flag& f : Flags(defaults)) {
    std::string head = std::string("  ") + f.name + (f.value ? std::string(" ") + f.value : "");
    head.resize(std::max<size_t>(head.size() + 1, 26), ' ');
    std::printf("%s%s\n", head.c_str(), f.help);
  }
}

bool Parse(int argc, char** argv, Options& options) {
  const std::vector<Flag> flags = Flags(options);
  for (int i = 1; i < argc; ++i) {
    const char* name = argv[i];
    if (!std::strcmp(name, "--help") || !std::strcmp(name, "-h")) {
      PrintUsage(argv[0]);
      return false;
    }
    const Flag* flag = nullptr;
    for (const Flag& f : flags) {
      if (!std::strcmp(name, f.name)) flag = &f;
    }
    if (!flag) {
      std::printf("error: unknown option %s\n", name);
      PrintUsage(argv[0]);
      return false;
    }
    if (flag->value && i + 1 >= argc) {
      std::printf("error: %s needs a value\n", name);
      return false;
    }
    const char* value = flag->value ? argv[++i] : nullptr;
    if (!flag->apply(value)) {
      std::printf("error: %s does not accept %s, see --help\n", name, value);
      return false;
    }
  }
  return true;
}

cc::Graph LoadInstance(const Options& o) {
  if (!o.tags_path.empty()) return cc::Graph::LoadTags(o.tags_path, o.similarity, o.threshold, o.n, o.graph_seed);
  if (!o.graph_json_path.empty()) return cc::Graph::LoadBaselineJson(o.graph_json_path);
  if (!o.edges_path.empty()) return cc::Graph::LoadEdgeList(o.edges_path);
  if (!o.graph_path.empty()) return cc::Graph::LoadMatrix(o.graph_path);
  if (o.generator == "sbm") return cc::Graph::StochasticBlock(o.n, o.blocks, o.p_in, o.p_out, o.graph_seed);
  if (o.generator == "chung-lu") return cc::Graph::ChungLu(o.n, o.gamma, o.mean_degree, o.graph_seed);
  if (o.generator != "er") throw std::runtime_error("unknown generator '" + o.generator + "'");
  return cc::Graph::ErdosRenyi(o.n, o.density, o.graph_seed);
}

void WriteResultJson(const std::string& path, const cc::Graph& graph, const Options& o, const cc::PbilsResult& best,
                     double avg_objective, long long worst, double avg_seconds) {
  std::ofstream out(path);
  if (!out) {
    std::printf("warning: cannot write %s\n", path.c_str());
    return;
  }
  // key names of the reference (include/cc_cli.hpp:256-276), which keeps those of the baseline
  out << "{\n";
  out << "  \"algorithm\": \"" << AlgorithmName(o.algorithm) << "\",\n";
  out << "  \"backend\": \"gpu\",\n";
  out << "  \"size\": " << graph.Size() << ",\n";
  out << "  \"edges\": " << graph.EdgeCount() << ",\n";
  out << "  \"density\": " << graph.Density() << ",\n";
  out << "  \"k\": " << o.params.k << ",\n";
  out << "  \"population\": " << o.params.population << ",\n";
  out << "  \"gww\": " << o.params.gww << ",\n";
  out << "  \"runs\": " << o.runs << ",\n";
  out << "  \"objective function value\": " << best.objective << ",\n";
  out << "  \"objective average\": " << avg_objective << ",\n";
  out << "  \"objective max\": " << worst << ",\n";
  out << "  \"computation time seconds\": " << best.seconds << ",\n";
  out << "  \"computation time average\": " << avg_seconds << ",\n";
  out << "  \"clusters used\": " << best.clusters_used << ",\n";
  out << "  \"clustering vector\": [";
  for (size_t i = 0; i < best.labels.size(); ++i) out << best.labels[i] << (i + 1 != best.labels.size() ? "," : "");
  out << "]\n}\n";
}

constexpr unsigned kPairwiseVerifyLimit = 32767;

}  // namespace

int main(int argc, char** argv) {
  Options o;
  if (!Parse(argc, argv, o)) return 1;
  cc::Graph graph;
  try {
    graph = LoadInstance(o);
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  if (!o.save_graph_path.empty()) graph.SaveMatrix(o.save_graph_path);
  if (!o.save_graph_json_path.empty()) graph.SaveBaselineJson(o.save_graph_json_path);
  if (!o.save_edges_path.empty()) graph.SaveEdgeList(o.save_edges_path);

  const cc::PbilsParams& p = o.params;
  std::printf("backend      gpu\n");
  std::printf("instance     n=%u  edges=%llu  density=%.4f\n", graph.Size(),
              static_cast<unsigned long long>(graph.EdgeCount()), graph.Density());
  if (o.algorithm == kPbils) {
    std::printf("algorithm    PBILS  k=%d  pop=%d  tournament=%d  iters=%d  early-stop=%d  perturb=%.2f  gww=%.2f\n",
                p.k, p.population, p.tournament, p.iterations, p.early_stop, p.perturbation, p.gww);
  } else {
    std::printf("algorithm    %s  k=%d  batch=%d\n", AlgorithmName(o.algorithm), p.k, p.population);
  }

  if (o.autotune > 0 && o.algorithm == kPbils) std::fputs(cc::AutotuneGpu(graph, p, o.autotune).c_str(), stdout);
  const bool pairwise = graph.Size() <= kPairwiseVerifyLimit;
  cc::PbilsResult best;
  double objective_sum = 0.0, seconds_sum = 0.0;
  long long worst = 0;
  for (int run = 0; run < o.runs; ++run) {
    cc::PbilsParams params = p;
    params.seed = p.seed + static_cast<uint64_t>(run);
    cc::PbilsResult result;
    try {
      if (o.algorithm == kNeighborhood) result = cc::SolveNeighborhoodGpu(graph, params);
      else if (o.algorithm == kNeighborhoodWithManyLocalSearches) result = cc::SolveNeighborhoodWithManyLocalSearchesGpu(graph, params);
      else result = cc::SolveGpu(graph, params);
    } catch (const std::exception& e) {
      std::printf("error: %s\n", e.what());
      return 1;
    }
    if (o.verify) {
      const long long truth =
          pairwise ? cc::ObjectiveDirect(graph, result.labels) : cc::ObjectiveBitwise(graph, result.labels);
      if (truth != result.objective) {
        std::printf("MISMATCH run %d: reported %lld, recount %lld\n", run, result.objective, truth);
        return 2;
      }
    }
    objective_sum += static_cast<double>(result.objective);
    seconds_sum += result.seconds;
    if (run == 0 || result.objective < best.objective) best = result;
    if (run == 0 || result.objective > worst) worst = result.objective;
    std::printf("run %-3d      f=%-10lld  clusters=%-3d  iters=%-4d  %.3fs\n", run, result.objective,
                result.clusters_used, result.iterations_done, result.seconds);
    if (p.verbose && result.local_searches > 0) {
      std::printf("moves        %lld accepted, %.0f per local search\n", result.accepted_moves,
                  static_cast<double>(result.accepted_moves) / static_cast<double>(result.local_searches));
    }
    if (p.profile) {
      std::printf("device       descent %.3f ms, rebuild %.3f ms, %.3g moves/s\n", result.search_ms, result.rebuild_ms,
                  result.search_ms > 0 ? 1e3 * static_cast<double>(result.accepted_moves) / result.search_ms : 0.0);
    }
  }
  const double avg_objective = objective_sum / o.runs;
  const double avg_seconds = seconds_sum / o.runs;
  std::printf("---\n");
  std::printf("objective    min=%lld  avg=%.2f  max=%lld\n", best.objective, avg_objective, worst);
  std::printf("time         best=%.3fs  avg=%.3fs\n", best.seconds, avg_seconds);
  if (o.verify) {
    std::printf(pairwise ? "verified     reported value recounted pairwise in O(n^2)\n"
                         : "verified     reported value recounted over bit rows in O(n^2/64)\n");
  }
  if (!o.out_path.empty()) {
    WriteResultJson(o.out_path, graph, o, best, avg_objective, worst, avg_seconds);
    std::printf("written      %s\n", o.out_path.c_str());
  }
  if (!o.labels_out_path.empty()) {
    std::ofstream out(o.labels_out_path);
    for (size_t v = 0; v < best.labels.size(); ++v) out << best.labels[v] << (v + 1 == best.labels.size() ? "\n" : " ");
    std::printf("written      %s\n", o.labels_out_path.c_str());
  }
  return 0;
}
