// cc_gpu.hpp — the GPU entry points, same declarations as the reference (include/cc_gpu.hpp:10-20 of the
// reference), implemented over the C ABI of cc_hip.h on AMD Instinct (gfx950). Errors are thrown as
// std::runtime_error, as in the reference.
#pragma once

#include "cc_graph.hpp"
#include "cc_pbils.hpp"

namespace cc {

// Largest k the kernels hold.
constexpr int kMaxClusters = 64;

PbilsResult SolveGpu(const Graph& graph, const PbilsParams& params);

// NeighborhoodCuda and NeighborhoodWithManyLocalSearchesCuda, k = 2 only: for each v the first cluster is v and its
// neighbours, the second the rest; the second algorithm improves each start by local search. Equal f: the smaller v;
// equal improvements: the smaller moved vertex. population bounds the starts per batch, ls_kernel picks the
// descent; the other PBILS settings except verbose are unused.
PbilsResult SolveNeighborhoodGpu(const Graph& graph, const PbilsParams& params);
PbilsResult SolveNeighborhoodWithManyLocalSearchesGpu(const Graph& graph, const PbilsParams& params);

// Times every descent plan for this graph's shape on `iterations` PBILS iterations and records the fastest for
// ls_kernel = auto in this process (no counterpart in the reference, which fixes its tuning at compile time).
// Returns one printable line per plan tried, the chosen one marked.
std::string AutotuneGpu(const Graph& graph, const PbilsParams& params, int iterations);

}  // namespace cc
