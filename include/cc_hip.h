/*
 * cc_hip.h — the C ABI of the MI355X PBILS correlation-clustering library (libcc_hip.so).
 *
 * Plain C: POD structs, raw pointers and sizes, integer status codes. The C++ entry points of the reference
 * (include/cc_gpu.hpp:12-20 of the reference) are re-declared in cc_gpu.hpp over this ABI; any other host language
 * binds these functions directly (a ctypes, cgo or JNI stub over these declarations; the Python package of this
 * library is such a ctypes binding).
 *
 * Graph layout everywhere: n rows of words64(n) = max(1, ceil(n/64)) uint64 words, row-major; bit j of row i is
 * bit (j % 64) of word j / 64. The matrix is symmetric with a zero diagonal and zero padding bits.
 *
 * Status codes: CC_HIP_OK (0) or one of the CC_HIP_E* codes; a message is written to err (if err_len > 0). The
 * reference throws std::runtime_error for the same conditions; cc_gpu.hpp rethrows them as such.
 */
#ifndef CC_HIP_H
#define CC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CC_HIP_OK 0
#define CC_HIP_EINVAL 1  /* invalid parameters or graph (reference: std::runtime_error from CheckParams etc.) */
#define CC_HIP_ENOMEM 2  /* device memory pre-flight failed (reference: cuda/cc_pbils_gpu.cu:198-207) */
#define CC_HIP_EDEVICE 3 /* HIP runtime error (reference: CC_CUDA_CHECK, cuda/cc_gpu_common.cuh:13-20) */
#define CC_HIP_ENODEV 4  /* no usable device */

/* Largest k the kernels hold (reference include/cc_gpu.hpp:10). */
#define CC_HIP_MAX_CLUSTERS 64
/* Largest n of the 16-bit-G narrow path and of any path (reference cuda/cc_gpu_common.cuh:29-32). */
#define CC_HIP_MAX_VERTICES_NARROW 32767
#define CC_HIP_MAX_VERTICES_WIDE (1 << 20)

/* PbilsParams::LocalSearchKernel (reference include/cc_pbils.hpp:30-35). */
#define CC_LS_AUTO 0
#define CC_LS_GLOBAL 1
#define CC_LS_SHARED 2
#define CC_LS_WIDE 3

/* Mirrors PbilsParams (reference include/cc_pbils.hpp:26-49); same defaults via cc_pbils_default_params. */
typedef struct cc_pbils_params {
  int32_t k;
  int32_t population;
  int32_t tournament;
  int32_t iterations;
  int32_t early_stop;
  int32_t threads; /* CPU-only in the reference; ignored here */
  double perturbation;
  double gww;
  uint64_t seed;
  double time_limit_sec; /* 0 disables; checked between iteration chunks */
  int32_t verbose;
  int32_t ls_kernel; /* CC_LS_* */
  int32_t profile;   /* 1: time the search and rebuild kernels with HIP events (fills the *_ms fields) */
  int32_t rebuild;   /* CC_REBUILD_*: how a Neighborhood batch computes G */
} cc_pbils_params;

/* The G rebuild of a Neighborhood batch (k = 2): popcounts of adjacency words against cluster masks, or an int8
 * matrix product on the matrix cores (n <= 32767; exact, the same G). PBILS runs always use the popcounts. */
#define CC_REBUILD_POPCOUNT 0
#define CC_REBUILD_MFMA 1

/* Mirrors PbilsResult (reference include/cc_pbils.hpp:51-59) plus the plan and device timings. */
typedef struct cc_pbils_stats {
  int64_t objective;
  int32_t iterations_done;
  int32_t clusters_used;
  int64_t local_searches;
  int64_t accepted_moves;
  double seconds; /* solver clock: from entry (allocation and upload included, as the reference) to exit */
  /* plan actually used */
  int32_t descent;      /* CC_DESCENT_* */
  int32_t threads;      /* threads per block of the descent kernel */
  int32_t blocks;       /* blocks of the descent kernel */
  int32_t vertices_per_thread;
  int64_t lds_bytes;    /* dynamic LDS of the descent kernel */
  /* device timings, filled when params.profile != 0 */
  double search_ms;     /* sum over descent launches */
  double rebuild_ms;    /* sum over G rebuilds */
  double iteration_ms;  /* device time of all PBILS iterations */
  int64_t search_launches;
  int64_t rebuild_launches;
} cc_pbils_stats;

#define CC_DESCENT_GENERIC_GLOBAL 0
#define CC_DESCENT_GENERIC_LDS 1
#define CC_DESCENT_K2_REGISTERS 2
#define CC_DESCENT_K2_WIDE 3
#define CC_DESCENT_K3_WIDE 4
#define CC_DESCENT_REGISTERS 5 /* k in [3, 32], n up to 16384 at k = 3 (fewer at larger k): state in VGPRs */
#define CC_DESCENT_MIXED 6     /* batched: one launch over runs of different register kernels */
#define CC_DESCENT_K3_CHUNK 7  /* k = 3 wide: packed words plus per-chunk move summaries in LDS */
#define CC_DESCENT_K3_PACKED 8 /* k = 3, n <= 32767: three 16-bit counts and the label per vertex in two VGPRs */

void cc_pbils_default_params(cc_pbils_params* params);

/* ---- one-shot solves on a host-resident graph ------------------------------------------------------------ */

/* Replaces cc::SolveGpu (reference include/cc_gpu.hpp:12; cuda/cc_pbils_gpu.cu:441-454). labels_out: n entries.
 * trace_out (nullable): params.iterations entries, the record after each completed iteration. */
int cc_hip_solve_pbils(const uint64_t* bits, uint32_t n, uint64_t edges, const cc_pbils_params* params,
                       int32_t* labels_out, cc_pbils_stats* stats, int64_t* trace_out, char* err, size_t err_len);

/* Replaces cc::SolveNeighborhoodGpu (local_search = 0) and cc::SolveNeighborhoodWithManyLocalSearchesGpu
 * (local_search = 1) (reference include/cc_gpu.hpp:19-20; cuda/cc_pbils_gpu.cu:389-436). params.population is the
 * batch of start partitions, params.ls_kernel picks the descent; k must be 2. */
int cc_hip_solve_neighborhood(const uint64_t* bits, uint32_t n, uint64_t edges, const cc_pbils_params* params,
                              int32_t local_search, int32_t* labels_out, cc_pbils_stats* stats, char* err,
                              size_t err_len);

/* ---- a device-resident graph: upload once, solve many times (no reference counterpart: the reference uploads
 *      the bits on every call, cuda/cc_pbils_gpu.cu:212) ----------------------------------------------------- */

typedef struct cc_hip_graph cc_hip_graph;

int cc_hip_graph_upload(const uint64_t* bits, uint32_t n, uint64_t edges, cc_hip_graph** out, char* err,
                        size_t err_len);
void cc_hip_graph_free(cc_hip_graph* graph);
int cc_hip_graph_solve_pbils(cc_hip_graph* graph, const cc_pbils_params* params, int32_t* labels_out,
                             cc_pbils_stats* stats, int64_t* trace_out, char* err, size_t err_len);
int cc_hip_graph_solve_neighborhood(cc_hip_graph* graph, const cc_pbils_params* params, int32_t local_search,
                                    int32_t* labels_out, cc_pbils_stats* stats, char* err, size_t err_len);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): cc_hip_graph_solve_pbils with every piece of device state
 * (the two populations, G, record, trace, GWW and rebuild scratch, launch tables) laid out in a static region of
 * CC_HIP_STATIC_SCRATCH_BYTES per GPU, held from the library's load until the process ends, instead of fresh
 * allocations. No hipMalloc, hipHostMalloc or HIP graph instantiation: iterations are launched one by one on
 * `stream` (a hipStream_t; NULL: the default stream) and params.profile must be 0 (CC_HIP_EINVAL otherwise). A run
 * whose state exceeds the region (cc_hip_footprint minus the graph, plus the launch tables) fails with
 * CC_HIP_ENOMEM before anything runs. Every call on one GPU uses the same region and leaves its contents behind:
 * calls that overlap on one GPU (other host threads, other streams) race and must not be made. Results are
 * bit-identical to cc_hip_graph_solve_pbils. */
#define CC_HIP_STATIC_SCRATCH_BYTES ((size_t)64 << 20)
int cc_hip_graph_solve_pbils_static(cc_hip_graph* graph, const cc_pbils_params* params, void* stream,
                                    int32_t* labels_out, cc_pbils_stats* stats, int64_t* trace_out, char* err,
                                    size_t err_len);

/* ---- batched solves: many independent PBILS runs (different graphs, n, k and parameters) advanced together by
 *      one launch per phase, descents claimed from persistent work queues, iterations replayed from a HIP graph.
 *      Each run's result is bit-identical to its own cc_hip_graph_solve_pbils. No reference counterpart: the
 *      reference solves one instance per call. --------------------------------------------------------------- */

typedef struct cc_batch_run {
  const cc_hip_graph* graph;
  cc_pbils_params params;
  int32_t* labels_out;  /* n entries */
  int64_t* trace_out;   /* nullable: params.iterations entries */
  cc_pbils_stats stats; /* filled; stats.seconds is the wall time of the whole batch */
} cc_batch_run;

int cc_hip_solve_batch(cc_batch_run* runs, int32_t count, int32_t use_graph, double* seconds, char* err,
                       size_t err_len);

/* ---- serving with continuous batching: requests are queued, admitted between iterations while their HBM
 *      footprint (n^2/8 bitset + per-vertex state of the population, cc_hip_footprint) fits the budget, advanced
 *      together with the other active runs of their lane, and retired as soon as they stop. Each result is
 *      bit-identical to a solo solve. No reference counterpart. ---------------------------------------------- */

#define CC_HIP_ETIMEOUT 5 /* cc_hip_server_wait: not done within the timeout */

typedef struct cc_serve_options {
  int64_t budget_bytes; /* device bytes the server may hold; <= 0: 85% of the free memory at creation */
  int32_t lanes;        /* streams with their own continuous batch; lane 0 takes the large requests */
  int32_t max_active;   /* active runs per lane */
  int64_t large_bytes;  /* requests with a footprint at or above this go to lane 0 */
} cc_serve_options;

typedef struct cc_hip_server cc_hip_server;

void cc_serve_default_options(cc_serve_options* options);
int cc_hip_server_create(const cc_serve_options* options, cc_hip_server** out, char* err, size_t err_len);
void cc_hip_server_destroy(cc_hip_server* server); /* finishes every queued request first */
/* Copies the graph; returns at once with the request id. */
int cc_hip_server_submit(cc_hip_server* server, const uint64_t* bits, uint32_t n, uint64_t edges,
                         const cc_pbils_params* params, uint64_t* id_out, char* err, size_t err_len);
/* Blocks until the request is done (timeout_s <= 0: no limit). latency_s: submit to result; queued_s: submit to
 * admission. trace_out (nullable): params.iterations entries. */
int cc_hip_server_wait(cc_hip_server* server, uint64_t id, double timeout_s, int32_t* labels_out, int64_t* trace_out,
                       cc_pbils_stats* stats, double* latency_s, double* queued_s, char* err, size_t err_len);
int64_t cc_hip_footprint(uint32_t n, const cc_pbils_params* params);
/* Requests admitted and completed, budget left now, and the peaks of active runs and of admitted bytes. */
void cc_hip_server_counters(cc_hip_server* server, int64_t* admitted, int64_t* completed, int64_t* budget_left,
                            int64_t* peak_active, int64_t* peak_bytes);

/* ---- descent autotuner: times every descent plan the kernels support for a graph's shape (n, k, population, G
 *      width) on a short PBILS run of that graph, checks that all follow one trajectory, and records the fastest
 *      in a process-wide table that ls_kernel = CC_LS_AUTO then uses for that shape. Replaces the reference's
 *      compile-time tuning for one GPU (cuda/cc_gpu_tuning.cuh:13-77, plan choice cuda/cc_pbils_gpu.cu:69-102). -- */

typedef struct cc_tune_candidate {
  int32_t descent; /* CC_DESCENT_* */
  int32_t threads; /* threads per workgroup */
  int32_t vertices_per_thread;
  int32_t rows_in_lds; /* k = 2 register kernel: adjacency staged in LDS */
  int32_t blocks;      /* persistent workgroups launched */
  int32_t reserved;
  int64_t lds_bytes; /* dynamic LDS (generic kernel: G and labels in LDS) */
  int64_t accepted_moves;
  double search_ms; /* device time of the descent launches */
} cc_tune_candidate;

/* iterations: PBILS iterations timed per candidate (>= 1). out receives up to capacity candidates; count the number
 * tried; chosen the index of the fastest (now recorded for this shape). */
int cc_hip_autotune(cc_hip_graph* graph, const cc_pbils_params* params, int32_t iterations, cc_tune_candidate* out,
                    int32_t capacity, int32_t* count, int32_t* chosen, char* err, size_t err_len);
void cc_hip_tune_clear(void);
/* Text table, one line per shape: n k population wide descent threads vpt rows_in_lds lds_bytes blocks. */
int cc_hip_tune_save(const char* path, char* err, size_t err_len);
int cc_hip_tune_load(const char* path, int32_t* entries, char* err, size_t err_len);

/* ---- host-side graph generators and recounts of the product (the reference's Graph::ErdosRenyi,
 *      src/cc_graph.cpp:168-177, and the recounts of src/cc_pbils_cpu.cpp:12-53) --------------------------- */

uint32_t cc_words64(uint32_t n);
/* Each returns the edge count and fills bits (n * cc_words64(n) words). threads = 0: all host cores. */
uint64_t cc_graph_erdos_renyi(uint32_t n, double p, uint64_t seed, uint64_t* bits, int32_t threads);
uint64_t cc_graph_sbm(uint32_t n, int32_t blocks, double p_in, double p_out, uint64_t seed, uint64_t* bits,
                      int32_t threads);
uint64_t cc_graph_chung_lu(uint32_t n, double gamma, double mean_degree, uint64_t seed, uint64_t* bits,
                           int32_t threads);
/* Tag-similarity graph, the reference's Graph::LoadTags (include/cc_graph.hpp:51, src/cc_graph.cpp:311-371): n
 * sampled objects of a {"key": ["tag", ...]} JSON file (n = 0: all), joined when the similarity ("jaccard",
 * "cosine", "dice" or "overlap") of their tag sets is >= threshold. On success *bits_out is a host buffer of
 * *n_out * cc_words64(*n_out) words that the caller releases with cc_host_free. CC_HIP_EINVAL for an unreadable or
 * malformed file or an unknown similarity. */
int cc_graph_load_tags(const char* path, const char* similarity, double threshold, uint32_t n, uint64_t seed,
                       uint64_t** bits_out, uint32_t* n_out, uint64_t* edges_out, char* err, size_t err_len);
void cc_host_free(void* p);
/* Instance files of the reference's command line (include/cc_graph.hpp:35-48, src/cc_graph.cpp:179-306): the 0/1
 * matrix text (n, then n rows of n cells '0'/'1', spaced or not), the "graph" block of a baseline result JSON, and a
 * SNAP/CSV edge list (two ids per line, a vertex is the rank of its id; loading only). On success *bits_out is a host
 * buffer of *n_out * cc_words64(*n_out) words that the caller releases with cc_host_free. CC_HIP_EINVAL for an
 * unreadable or malformed file. */
#define CC_GRAPH_MATRIX 0
#define CC_GRAPH_BASELINE_JSON 1
#define CC_GRAPH_EDGE_LIST 2
int cc_graph_load(const char* path, int32_t format, uint64_t** bits_out, uint32_t* n_out, uint64_t* edges_out,
                  char* err, size_t err_len);
/* Writes bits (n rows of cc_words64(n) words, symmetric, zero diagonal) as CC_GRAPH_MATRIX, CC_GRAPH_BASELINE_JSON
 * (byte for byte the reference's Graph::SaveMatrix / SaveBaselineJson) or CC_GRAPH_EDGE_LIST ("# n vertices, m
 * edges", then "i j" with i < j per line). */
int cc_graph_save(const uint64_t* bits, uint32_t n, const char* path, int32_t format, char* err, size_t err_len);
/* The reference's SampleWithoutReplacement (include/cc_graph.hpp:62): out receives the first min(n, total) entries
 * of a partial Fisher-Yates shuffle of 0 .. total - 1 on StreamSeed(seed, 0x7A65, 0), or 0 .. total - 1 when n = 0
 * or n >= total. */
void cc_sample_without_replacement(uint32_t total, uint32_t n, uint64_t seed, uint32_t* out);
int64_t cc_objective_direct(const uint64_t* bits, uint32_t n, const int32_t* labels);
int64_t cc_objective_bitwise(const uint64_t* bits, uint32_t n, const int32_t* labels);

/* ---- device information ---------------------------------------------------------------------------------- */

typedef struct cc_device_info {
  char name[64];
  char arch[32];
  int32_t compute_units;
  int32_t lds_per_block;
  int32_t l2_bytes;
  int32_t clock_khz;
  int64_t total_bytes;
  int64_t free_bytes;
} cc_device_info;

int cc_hip_device_info(cc_device_info* info, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif
#endif
