// cc_graph.hpp — an undirected, unweighted, loop-free graph as a bitset adjacency with rows padded to whole 64-bit
// words (one wave64 lane per bit of a word; __popcll of row & mask is one row of the product A Z).
//
// Same interface as the reference cc::Graph (include/cc_graph.hpp:12-58 of the reference) except the word width;
// bit j of row i is bit j % 64 of word j / 64.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cc {

using Word = uint64_t;
constexpr unsigned kWordBits = 64;

class Graph {
 public:
  Graph() = default;
  explicit Graph(unsigned n);

  unsigned Size() const { return n_; }
  unsigned WordsPerRow() const { return words_per_row_; }
  uint64_t EdgeCount() const { return edges_; }
  const std::vector<Word>& Bits() const { return bits_; }
  const Word* Row(unsigned v) const { return bits_.data() + static_cast<size_t>(v) * words_per_row_; }
  bool IsJoined(unsigned i, unsigned j) const { return (Row(i)[j / kWordBits] >> (j % kWordBits)) & 1u; }

  void AddEdge(unsigned i, unsigned j);
  double Density() const;

  // Reference G(n, p) (src/cc_graph.cpp:168-177): pair (i, j), j < i, is draw number i(i-1)/2 + j of the stream
  // StreamSeed(seed, kErdosRenyiStream, 0); an edge iff RandFloat < p. Evaluated on all host cores.
  static Graph ErdosRenyi(unsigned n, double density, uint64_t seed, int threads = 0);
  // Planted partition: vertex v in block v * blocks / n; p_in inside a block, p_out across; ER's draw order on
  // the stream kSbmStream.
  static Graph StochasticBlock(unsigned n, int blocks, double p_in, double p_out, uint64_t seed, int threads = 0);
  // Chung-Lu: w_i = (i + 1)^(-1 / (gamma - 1)) scaled to mean mean_degree, p_ij = min(1, w_i w_j / (mean_degree n)),
  // each pair independently; drawn in O(n + m) by geometric skips along each row, row u on the stream
  // StreamSeed(seed, kChungLuStream, u + 1).
  static Graph ChungLu(unsigned n, double gamma, double mean_degree, uint64_t seed, int threads = 0);
  // Adopts a bitset (n rows of WordsPerRow words) and counts its edges.
  static Graph FromBits(std::vector<Word> bits, unsigned n);

  // Text: n, then n rows of n '0'/'1' (with or without spaces) (reference src/cc_graph.cpp:179-211).
  static Graph LoadMatrix(const std::string& path);
  void SaveMatrix(const std::string& path) const;
  // The "graph" block of the baseline's result JSON (reference src/cc_graph.cpp:213-268).
  static Graph LoadBaselineJson(const std::string& path);
  void SaveBaselineJson(const std::string& path) const;
  // SNAP / CSV edge list: vertex = rank of its id among all ids (reference src/cc_graph.cpp:270-306).
  static Graph LoadEdgeList(const std::string& path);
  void SaveEdgeList(const std::string& path) const;
  // Tag-similarity graph (reference src/cc_graph.cpp:311-371): a JSON object {"key": ["tag", ...], ...}; the
  // vertices are SampleWithoutReplacement(objects, n, seed) (n = 0: all objects, in file order) and u, v are
  // joined iff the similarity of their tag sets (jaccard, cosine, dice or overlap, in double precision) is at least
  // threshold. No loops.
  static Graph LoadTags(const std::string& path, const std::string& similarity, double threshold, unsigned n,
                        uint64_t seed);

 private:
  // Sets bit (j, i) for every bit (i, j) above the diagonal and counts those as the edges: the loaders fill the upper
  // triangle only.
  void MirrorUpperTriangle();

  unsigned n_ = 0;
  unsigned words_per_row_ = 1;
  uint64_t edges_ = 0;
  std::vector<Word> bits_;
};

// The first n of a Fisher-Yates shuffle of 0 .. total - 1 on the stream StreamSeed(seed, kSampleStream, 0), or all of
// them in order when n = 0 or n >= total (reference src/cc_graph.cpp:373-381).
std::vector<unsigned> SampleWithoutReplacement(unsigned total, unsigned n, uint64_t seed);

}  // namespace cc
