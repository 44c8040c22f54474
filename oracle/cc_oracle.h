/*
 * cc_oracle.h — CPU restatement of the reference PBILS correlation-clustering path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product (cuda_graph_clustering_amd) never links or calls it.
 *
 * Every function names the reference file:line whose behaviour it restates. The bitset layout differs on
 * purpose: rows are padded to 64-bit words (the reference pads to 32-bit words, include/cc_graph.hpp:12-58);
 * on a little-endian host bit j of a row is bit j%64 of word j/64 in both layouts.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   rng/math   — splitmix64 known outputs, closed forms, and the reference-held golden trajectories
 *                (docs/data/gpu.csv, gww.csv, gww_stop.csv) which exercise every random decision end to end;
 *                plus bit-exact agreement with the reference CPU built out of tree (oracle/_ref).
 *   objective  — pairwise definition vs closed form vs bitwise recount on word-boundary sizes.
 *   rebuild    — brute-force neighbour counting and the invariants sum_c G[v][c] = deg(v).
 *   search     — incremental == rebuild after random moves; local-optimum property; known optima.
 *   GWW        — RankSlots/WinnerCopies known answers (src/cc_verify.cpp:379-383).
 *   generators — ER identical to the reference generator; SBM and Chung-Lu are defined here (they do not
 *                exist in the reference) and pinned by edge-count statistics, the closed-form draw rule (SBM) and
 *                per-pair edge frequencies over many seeds against p_ij (Chung-Lu).
 */
#ifndef CC_ORACLE_H
#define CC_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG and move arithmetic (include/cc_math.hpp:56-145) ---- */
uint64_t orc_splitmix64(uint64_t* state);
uint32_t orc_next_u32(uint64_t* state);
uint32_t orc_rand_below(uint64_t* state, uint32_t bound);
float orc_rand_float(uint64_t* state);
uint64_t orc_stream_seed(uint64_t seed, uint64_t stream, uint64_t step);
uint32_t orc_sample_without_replacement(uint32_t total, uint32_t n, uint64_t seed, uint32_t* out);
int orc_initial_label(uint64_t base, uint64_t index, int k);
int orc_tournament(uint64_t base, int slot, int iteration, const int64_t* f, int population, int tournament);
int orc_perturbed_label(uint64_t base, uint64_t index, int iteration, int label, int k, float probability);
int64_t orc_same_pairs(const int32_t* sizes, int k);
int64_t orc_objective(int64_t edges, const int32_t* sizes, int k, int64_t intra_edges);
int orc_move_delta(int size_from, int size_to, int g_from, int g_to);

/* ---- graphs: row-major n x words uint64 bitset ---- */
uint32_t orc_words_per_row(uint32_t n);
/* Reference G(n,p) (src/cc_graph.cpp:168-177), evaluated pair-parallel with the counter form of splitmix64. */
uint64_t orc_graph_erdos_renyi(uint32_t n, double p, uint64_t seed, uint64_t* bits, int threads);
/* Planted partition: block(v) = v*blocks/n; same lower-triangle draw order as ER, own stream. */
uint64_t orc_graph_sbm(uint32_t n, int blocks, double p_in, double p_out, uint64_t seed, uint64_t* bits,
                       int threads);
/* Chung-Lu: w_i = (i+1)^(-1/(gamma-1)) scaled to mean degree d; p_ij = min(1, w_i w_j / (d n)), each pair
 * independently; drawn in O(n + m) by geometric skips along each row (Miller and Hagberg 2011), row u on the stream
 * StreamSeed(seed, 0xC1A9E5, u + 1). */
uint64_t orc_graph_chung_lu(uint32_t n, double gamma, double mean_degree, uint64_t seed, uint64_t* bits,
                            int threads);
uint64_t orc_edge_count(const uint64_t* bits, uint32_t n, uint32_t words);

/* ---- objective recounts (src/cc_pbils_cpu.cpp:12-53) ---- */
int64_t orc_objective_direct(const uint64_t* bits, uint32_t n, uint32_t words, const int32_t* labels);
int64_t orc_objective_bitwise(const uint64_t* bits, uint32_t n, uint32_t words, const int32_t* labels);
int orc_clusters_used(const int32_t* labels, uint32_t n, int k);

/* ---- solution state (src/cc_pbils_cpu.cpp:65-161) ---- */
typedef struct orc_state {
  int n, k;
  int32_t* labels; /* n */
  int32_t* g;      /* k*n, g[c*n+v] */
  int32_t* sizes;  /* k */
  int64_t f;
} orc_state;

void orc_state_init(orc_state* s, int n, int k);
void orc_state_free(orc_state* s);
void orc_state_copy(orc_state* dst, const orc_state* src);
void orc_rebuild(orc_state* s, const uint64_t* bits, uint32_t words, uint64_t edges);
int orc_best_move(const orc_state* s, int* out_v, int* out_to);
int orc_best_move_plain(const orc_state* s, int* out_v, int* out_to);
void orc_apply_move(orc_state* s, const uint64_t* bits, uint32_t words, int v, int to);
int64_t orc_local_search(orc_state* s, const uint64_t* bits, uint32_t words, int64_t* accepted_moves);
void orc_perturb(orc_state* s, const uint64_t* bits, uint32_t words, uint64_t edges, double probability,
                 uint64_t base, int slot, int iteration);

/* Convenience for tests: labels in, G/sizes/f out (g is k*n). */
int64_t orc_rebuild_labels(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, int k,
                           const int32_t* labels, int32_t* g_out, int32_t* sizes_out);
/* Local search from given labels; returns f, writes final labels and the number of accepted moves. */
int orc_best_move_labels(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, int k,
                         const int32_t* labels, int plain, int* v, int* to);
int64_t orc_local_search_labels(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, int k,
                                int32_t* labels, int64_t* moves);

/* ---- PBILS (include/cc_pbils.hpp:26-85, src/cc_pbils_cpu.cpp:163-252) ---- */
typedef struct orc_params {
  int32_t k, population, tournament, iterations, early_stop, threads;
  double perturbation, gww;
  uint64_t seed;
} orc_params;

typedef struct orc_result {
  int64_t objective;
  int32_t iterations_done, clusters_used;
  int64_t local_searches, accepted_moves;
  double seconds;
} orc_result;

void orc_default_params(orc_params* p);
int orc_winner_copies(double gww, int population);
void orc_rank_slots(const int64_t* f, int population, int32_t* order);
/* Returns 0 on success, -1 on invalid params (message in err). trace (nullable) receives the record after each
 * completed iteration (iterations entries); labels_out receives n labels. */
int orc_solve_pbils(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, const orc_params* params,
                    int32_t* labels_out, orc_result* result, int64_t* trace, char* err, size_t err_len);

/* Neighborhood enumeration for k = 2 (cuda/cc_verify_gpu.cu:21-35): for every v the start partition
 * N[v] u {v} -> 0, the rest -> 1; optional local search; best f with ties to the smaller v. */
int orc_neighborhood(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, int local_search,
                     int threads, int32_t* labels_out, orc_result* result);

#ifdef __cplusplus
}
#endif
#endif
