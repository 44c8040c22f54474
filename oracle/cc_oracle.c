/*
 * cc_oracle.c — CPU restatement of the reference PBILS path. TEST INFRASTRUCTURE ONLY (the checker).
 *
 * Written from the behaviour of the reference files cited per function; the adjacency is a 64-bit-word bitset.
 * Multithreading follows the reference's slot-parallel pool (include/cc_parallel.hpp:20-95): items are handed
 * out through an atomic counter, so the answer does not depend on the thread count.
 */
#include "cc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* ------------------------------------------------------------------------------------------------------------ */
/* RNG (include/cc_math.hpp:84-111)                                                                              */
/* ------------------------------------------------------------------------------------------------------------ */

#define ORC_GOLDEN 0x9E3779B97F4A7C15ull

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* include/cc_math.hpp:84-90 */
uint64_t orc_splitmix64(uint64_t* state) {
  *state += ORC_GOLDEN;
  return mix64(*state);
}

/* include/cc_math.hpp:92 */
uint32_t orc_next_u32(uint64_t* state) { return (uint32_t)(orc_splitmix64(state) >> 32); }

/* include/cc_math.hpp:95-98: (u32 * bound) >> 32, 0 for bound 0 */
uint32_t orc_rand_below(uint64_t* state, uint32_t bound) {
  if (bound == 0) return 0;
  return (uint32_t)(((uint64_t)orc_next_u32(state) * bound) >> 32);
}

/* include/cc_math.hpp:101-104: top 24 bits of the u32 scaled by 2^-24 (exact in float) */
float orc_rand_float(uint64_t* state) {
  return (float)(orc_next_u32(state) >> 8) * (1.0f / 16777216.0f);
}

/* include/cc_math.hpp:107-111: the advanced state, not the output */
uint64_t orc_stream_seed(uint64_t seed, uint64_t stream, uint64_t step) {
  uint64_t s = seed ^ (stream * 0xD1B54A32D192ED03ull) ^ (step * 0xA0761D6478BD642Full);
  orc_splitmix64(&s);
  return s;
}

/* Stream constants: include/cc_math.hpp:118-121, src/cc_graph.cpp:21 */
#define ORC_INIT_STREAM 0xA11CEull
#define ORC_SELECT_STREAM 0x5E1EC7ull
#define ORC_PERTURB_STREAM 0xBEEFull
#define ORC_PERTURB_FIRST_STEP 0x5000000ull
#define ORC_ER_STREAM 0xE4D05Full
/* Streams of the generators that the reference does not have (defined by this project). */
#define ORC_SBM_STREAM 0x5B3D10Cull
#define ORC_CHUNG_LU_STREAM 0xC1A9E5ull

/* src/cc_graph.cpp:22 */
#define ORC_SAMPLE_STREAM 0x7A65ull

/* src/cc_graph.cpp:373-381: the first n of a Fisher-Yates shuffle of 0 .. total - 1, swapping position i with
   i + RandBelow(total - i) on the stream StreamSeed(seed, kSampleStream, 0); 0 .. total - 1 in order when n = 0 or
   n >= total. Returns the number of entries written to out. */
uint32_t orc_sample_without_replacement(uint32_t total, uint32_t n, uint64_t seed, uint32_t* out) {
  for (uint32_t i = 0; i < total; ++i) out[i] = i;
  if (n == 0 || n >= total) return total;
  uint64_t state = orc_stream_seed(seed, ORC_SAMPLE_STREAM, 0);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t j = i + orc_rand_below(&state, total - i);
    const uint32_t t = out[i];
    out[i] = out[j];
    out[j] = t;
  }
  return n;
}

/* include/cc_math.hpp:123-126 */
int orc_initial_label(uint64_t base, uint64_t index, int k) {
  uint64_t rng = orc_stream_seed(base, index, 0);
  return (int)orc_rand_below(&rng, (uint32_t)k);
}

/* include/cc_math.hpp:129-138: strict < keeps the first drawn among equals */
int orc_tournament(uint64_t base, int slot, int iteration, const int64_t* f, int population, int tournament) {
  uint64_t rng = orc_stream_seed(base, (uint64_t)(int64_t)slot, (uint64_t)(int64_t)(iteration + 1));
  int chosen = (int)orc_rand_below(&rng, (uint32_t)population);
  for (int t = 1; t < tournament; ++t) {
    const int candidate = (int)orc_rand_below(&rng, (uint32_t)population);
    if (f[candidate] < f[chosen]) chosen = candidate;
  }
  return chosen;
}

/* include/cc_math.hpp:141-145 */
int orc_perturbed_label(uint64_t base, uint64_t index, int iteration, int label, int k, float probability) {
  uint64_t rng = orc_stream_seed(base, index, (uint64_t)(int64_t)iteration + ORC_PERTURB_FIRST_STEP);
  if (orc_rand_float(&rng) >= probability) return label;
  return (label + 1 + (int)orc_rand_below(&rng, (uint32_t)(k - 1))) % k;
}

/* include/cc_math.hpp:56-77 */
int64_t orc_same_pairs(const int32_t* sizes, int k) {
  int64_t acc = 0;
  for (int c = 0; c < k; ++c) {
    const int64_t s = sizes[c];
    acc += s * (s - 1) / 2;
  }
  return acc;
}

int64_t orc_objective(int64_t edges, const int32_t* sizes, int k, int64_t intra_edges) {
  return edges + orc_same_pairs(sizes, k) - 2 * intra_edges;
}

int orc_move_delta(int size_from, int size_to, int g_from, int g_to) {
  return (size_to - size_from + 1) - 2 * (g_to - g_from);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* A small parallel-for with dynamic hand-out (include/cc_parallel.hpp:20-95)                                     */
/* ------------------------------------------------------------------------------------------------------------ */

typedef void (*orc_body)(void* ctx, int item, int worker);

typedef struct {
  orc_body body;
  void* ctx;
  int count;
  atomic_int next;
} orc_job;

typedef struct {
  orc_job* job;
  int worker;
} orc_worker_arg;

static void* orc_worker_main(void* arg) {
  orc_worker_arg* w = (orc_worker_arg*)arg;
  for (int item = atomic_fetch_add(&w->job->next, 1); item < w->job->count;
       item = atomic_fetch_add(&w->job->next, 1)) {
    w->job->body(w->job->ctx, item, w->worker);
  }
  return NULL;
}

static int orc_hardware_threads(void) {
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}

static int orc_resolve_threads(int threads) { return threads > 0 ? threads : orc_hardware_threads(); }

static void orc_parallel_for(int count, int threads, orc_body body, void* ctx) {
  if (count <= 0) return;
  if (threads > count) threads = count;
  if (threads <= 1) {
    for (int i = 0; i < count; ++i) body(ctx, i, 0);
    return;
  }
  orc_job job;
  job.body = body;
  job.ctx = ctx;
  job.count = count;
  atomic_init(&job.next, 0);
  pthread_t* tids = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  orc_worker_arg* args = (orc_worker_arg*)malloc(sizeof(orc_worker_arg) * (size_t)threads);
  for (int t = 0; t < threads; ++t) {
    args[t].job = &job;
    args[t].worker = t;
  }
  for (int t = 1; t < threads; ++t) pthread_create(&tids[t], NULL, orc_worker_main, &args[t]);
  orc_worker_main(&args[0]);
  for (int t = 1; t < threads; ++t) pthread_join(tids[t], NULL);
  free(tids);
  free(args);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Graph generators                                                                                              */
/* ------------------------------------------------------------------------------------------------------------ */

uint32_t orc_words_per_row(uint32_t n) {
  uint32_t w = (n + 63u) / 64u;
  return w == 0 ? 1u : w; /* src/cc_graph.cpp:149-153: an empty graph still has one word per row */
}

/* The draw of pair (i, j), j < i, in the reference's lower-triangle order (src/cc_graph.cpp:170-175): it is the
 * (t+1)-th splitmix64 output of the stream, t = i(i-1)/2 + j, and splitmix64's state after m steps is
 * state0 + m * golden, so any draw is computed directly. Returns the RandFloat value. */
static inline float pair_draw(uint64_t state0, uint32_t i, uint32_t j) {
  const uint64_t t = (uint64_t)i * (i - 1) / 2 + j;
  const uint32_t u = (uint32_t)(mix64(state0 + (t + 1) * ORC_GOLDEN) >> 32);
  return (float)(u >> 8) * (1.0f / 16777216.0f);
}

typedef enum { GEN_ER, GEN_SBM } gen_kind;

typedef struct {
  gen_kind kind;
  uint32_t n, words;
  uint64_t state0;
  double p, p_in, p_out;
  int blocks;
  uint64_t* bits;
} gen_ctx;

static inline double gen_probability(const gen_ctx* g, uint32_t i, uint32_t j) {
  switch (g->kind) {
    case GEN_ER: return g->p;
    case GEN_SBM: {
      const uint32_t bi = (uint32_t)(((uint64_t)i * (uint64_t)g->blocks) / g->n);
      const uint32_t bj = (uint32_t)(((uint64_t)j * (uint64_t)g->blocks) / g->n);
      return bi == bj ? g->p_in : g->p_out;
    }
  }
  return 0.0;
}

/* Row i in full: draw(i, j) for j < i and draw(j, i) for j > i, so rows are independent work items. */
static void gen_row(void* ctx, int item, int worker) {
  (void)worker;
  const gen_ctx* g = (const gen_ctx*)ctx;
  const uint32_t i = (uint32_t)item;
  uint64_t* row = g->bits + (size_t)i * g->words;
  memset(row, 0, sizeof(uint64_t) * g->words);
  for (uint32_t j = 0; j < g->n; ++j) {
    if (j == i) continue;
    const uint32_t hi = i > j ? i : j;
    const uint32_t lo = i > j ? j : i;
    const float draw = pair_draw(g->state0, hi, lo);
    if ((double)draw < gen_probability(g, hi, lo)) row[j >> 6] |= 1ull << (j & 63);
  }
}

uint64_t orc_edge_count(const uint64_t* bits, uint32_t n, uint32_t words) {
  uint64_t ones = 0;
  for (size_t i = 0; i < (size_t)n * words; ++i) ones += (uint64_t)__builtin_popcountll(bits[i]);
  return ones / 2;
}

static uint64_t gen_run(gen_ctx* g, int threads) {
  g->words = orc_words_per_row(g->n);
  orc_parallel_for((int)g->n, orc_resolve_threads(threads), gen_row, g);
  return orc_edge_count(g->bits, g->n, g->words);
}

/* src/cc_graph.cpp:168-177 */
uint64_t orc_graph_erdos_renyi(uint32_t n, double p, uint64_t seed, uint64_t* bits, int threads) {
  gen_ctx g;
  memset(&g, 0, sizeof(g));
  g.kind = GEN_ER;
  g.n = n;
  g.p = p;
  g.state0 = orc_stream_seed(seed, ORC_ER_STREAM, 0);
  g.bits = bits;
  return gen_run(&g, threads);
}

uint64_t orc_graph_sbm(uint32_t n, int blocks, double p_in, double p_out, uint64_t seed, uint64_t* bits,
                       int threads) {
  gen_ctx g;
  memset(&g, 0, sizeof(g));
  g.kind = GEN_SBM;
  g.n = n;
  g.blocks = blocks < 1 ? 1 : blocks;
  g.p_in = p_in;
  g.p_out = p_out;
  g.state0 = orc_stream_seed(seed, ORC_SBM_STREAM, 0);
  g.bits = bits;
  return gen_run(&g, threads);
}

/* Chung-Lu in O(n + m), by the edge-skipping method of J. C. Miller and A. Hagberg, "Efficient generation of
 * networks with given expected degrees" (WAW 2011). The weights fall with the index, so along row u the pair
 * probability p(u, v) = min(1, w_u w_v / S) falls with v. From the current v with probability p the number of pairs
 * skipped is geometric, floor(log(r) / log(1 - p)) for r uniform in (0, 1]; the candidate reached is kept with
 * probability p(u, v) / p (a uniform in [0, 1) below it), and p becomes p(u, v). Every pair (u, v) is then an edge with
 * probability p(u, v), independently of the others. Row u draws 53-bit uniforms from the stream
 * StreamSeed(seed, ORC_CHUNG_LU_STREAM, u + 1), so rows are independent work items: each worker draws the pairs
 * above the diagonal of its rows, and the lower triangle is mirrored afterwards. */
typedef struct {
  uint32_t n, words;
  uint64_t seed;
  const double* weights;
  double inv_total;
  uint64_t* bits;
} cl_ctx;

static inline double cl_uniform_open0(uint64_t* state) { /* (0, 1] */
  return (double)((orc_splitmix64(state) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

static inline double cl_uniform_open1(uint64_t* state) { /* [0, 1) */
  return (double)(orc_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

static inline double cl_probability(const cl_ctx* c, uint32_t u, uint32_t v) {
  const double q = c->weights[u] * c->weights[v] * c->inv_total;
  return q > 1.0 ? 1.0 : q;
}

static void cl_row(void* ctx, int item, int worker) {
  (void)worker;
  const cl_ctx* c = (const cl_ctx*)ctx;
  const uint32_t u = (uint32_t)item;
  const uint32_t n = c->n;
  uint64_t* row = c->bits + (size_t)u * c->words;
  memset(row, 0, sizeof(uint64_t) * c->words);
  uint64_t state = orc_stream_seed(c->seed, ORC_CHUNG_LU_STREAM, (uint64_t)u + 1);
  uint32_t v = u + 1;
  if (v >= n) return;
  double p = cl_probability(c, u, v);
  while (v < n && p > 0.0) {
    if (p < 1.0) {
      const double skip = floor(log(cl_uniform_open0(&state)) / log1p(-p));
      if (skip >= (double)(n - v)) break;
      v += (uint32_t)skip;
    }
    const double q = cl_probability(c, u, v);
    if (cl_uniform_open1(&state) < q / p) row[v >> 6] |= 1ull << (v & 63);
    p = q;
    ++v;
  }
}

uint64_t orc_graph_chung_lu(uint32_t n, double gamma, double mean_degree, uint64_t seed, uint64_t* bits,
                            int threads) {
  double* w = (double*)malloc(sizeof(double) * (n ? n : 1));
  const double exponent = -1.0 / (gamma - 1.0);
  double sum = 0.0;
  for (uint32_t i = 0; i < n; ++i) {
    w[i] = pow((double)(i + 1), exponent);
    sum += w[i];
  }
  const double total = mean_degree * (double)n;
  const double scale = n ? total / sum : 0.0;
  for (uint32_t i = 0; i < n; ++i) w[i] *= scale;
  cl_ctx c;
  c.n = n;
  c.words = orc_words_per_row(n);
  c.seed = seed;
  c.weights = w;
  c.inv_total = n ? 1.0 / total : 0.0;
  c.bits = bits;
  orc_parallel_for((int)n, orc_resolve_threads(threads), cl_row, &c);
  /* mirror: every bit (u, v) with v > u sets (v, u) */
  for (uint32_t u = 0; u < n; ++u) {
    const uint64_t* row = bits + (size_t)u * c.words;
    for (uint32_t wd = (u + 1) >> 6; wd < c.words; ++wd) {
      uint64_t x = row[wd];
      while (x) {
        const uint32_t v = (wd << 6) + (uint32_t)__builtin_ctzll(x);
        x &= x - 1;
        if (v > u) bits[(size_t)v * c.words + (u >> 6)] |= 1ull << (u & 63);
      }
    }
  }
  free(w);
  return orc_edge_count(bits, n, c.words);
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Objective recounts (src/cc_pbils_cpu.cpp:12-53)                                                               */
/* ------------------------------------------------------------------------------------------------------------ */

static inline int joined(const uint64_t* bits, uint32_t words, uint32_t i, uint32_t j) {
  return (int)((bits[(size_t)i * words + (j >> 6)] >> (j & 63)) & 1ull);
}

int64_t orc_objective_direct(const uint64_t* bits, uint32_t n, uint32_t words, const int32_t* labels) {
  int64_t distance = 0;
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t j = i + 1; j < n; ++j) {
      const int same = labels[i] == labels[j];
      if (same != joined(bits, words, i, j)) ++distance;
    }
  }
  return distance;
}

int64_t orc_objective_bitwise(const uint64_t* bits, uint32_t n, uint32_t words, const int32_t* labels) {
  int k = 0;
  for (uint32_t v = 0; v < n; ++v) {
    if (labels[v] < 0) return -1;
    if (labels[v] + 1 > k) k = labels[v] + 1;
  }
  uint64_t* masks = (uint64_t*)calloc((size_t)(k ? k : 1) * words, sizeof(uint64_t));
  for (uint32_t v = 0; v < n; ++v) masks[(size_t)labels[v] * words + (v >> 6)] |= 1ull << (v & 63);
  int64_t acc = 0;
  for (uint32_t v = 0; v < n; ++v) {
    const uint64_t* row = bits + (size_t)v * words;
    const uint64_t* mask = masks + (size_t)labels[v] * words;
    for (uint32_t w = 0; w < words; ++w) acc += __builtin_popcountll(row[w] ^ mask[w]);
    acc -= 1; /* the vertex's own bit: set in its mask, clear in its row */
  }
  free(masks);
  return acc / 2;
}

/* src/cc_pbils_cpu.cpp:55-63 */
int orc_clusters_used(const int32_t* labels, uint32_t n, int k) {
  if (k <= 0) return 0;
  char* seen = (char*)calloc((size_t)k, 1);
  for (uint32_t v = 0; v < n; ++v) {
    if (labels[v] >= 0 && labels[v] < k) seen[labels[v]] = 1;
  }
  int used = 0;
  for (int c = 0; c < k; ++c) used += seen[c];
  free(seen);
  return used;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* State (src/cc_pbils_cpu.cpp:65-161)                                                                           */
/* ------------------------------------------------------------------------------------------------------------ */

void orc_state_init(orc_state* s, int n, int k) {
  s->n = n;
  s->k = k;
  s->labels = (int32_t*)calloc((size_t)(n ? n : 1), sizeof(int32_t));
  s->g = (int32_t*)calloc((size_t)k * (size_t)(n ? n : 1), sizeof(int32_t));
  s->sizes = (int32_t*)calloc((size_t)(k ? k : 1), sizeof(int32_t));
  s->f = 0;
}

void orc_state_free(orc_state* s) {
  free(s->labels);
  free(s->g);
  free(s->sizes);
  s->labels = NULL;
  s->g = NULL;
  s->sizes = NULL;
}

void orc_state_copy(orc_state* dst, const orc_state* src) {
  memcpy(dst->labels, src->labels, sizeof(int32_t) * (size_t)src->n);
  memcpy(dst->g, src->g, sizeof(int32_t) * (size_t)src->k * (size_t)src->n);
  memcpy(dst->sizes, src->sizes, sizeof(int32_t) * (size_t)src->k);
  dst->f = src->f;
}

/* src/cc_pbils_cpu.cpp:74-95: sizes, per-cluster masks, G[c][v] = popc(row_v & mask_c), f */
void orc_rebuild(orc_state* s, const uint64_t* bits, uint32_t words, uint64_t edges) {
  const int n = s->n;
  const int k = s->k;
  memset(s->sizes, 0, sizeof(int32_t) * (size_t)k);
  for (int v = 0; v < n; ++v) ++s->sizes[s->labels[v]];
  uint64_t* masks = (uint64_t*)calloc((size_t)k * words, sizeof(uint64_t));
  for (int v = 0; v < n; ++v) masks[(size_t)s->labels[v] * words + ((unsigned)v >> 6)] |= 1ull << (v & 63);
  int64_t intra_twice = 0;
  for (int v = 0; v < n; ++v) {
    const uint64_t* row = bits + (size_t)v * words;
    for (int c = 0; c < k; ++c) {
      const uint64_t* mask = masks + (size_t)c * words;
      int acc = 0;
      for (uint32_t w = 0; w < words; ++w) acc += __builtin_popcountll(row[w] & mask[w]);
      s->g[(size_t)c * n + v] = acc;
    }
    intra_twice += s->g[(size_t)s->labels[v] * n + v];
  }
  free(masks);
  s->f = orc_objective((int64_t)edges, s->sizes, k, intra_twice / 2);
}

/* src/cc_pbils_cpu.cpp:99-118, as written there: one scan over (v, target) in order, keeping a strictly smaller
 * delta; so the smallest delta < 0 wins, then the smallest v, then the smallest target. */
int orc_best_move_plain(const orc_state* s, int* out_v, int* out_to) {
  const int n = s->n;
  const int k = s->k;
  int best = 0;
  *out_v = -1;
  *out_to = -1;
  for (int v = 0; v < n; ++v) {
    const int from = s->labels[v];
    const int g_from = s->g[(size_t)from * n + v];
    const int size_from = s->sizes[from];
    for (int to = 0; to < k; ++to) {
      if (to == from) continue;
      const int delta = orc_move_delta(size_from, s->sizes[to], g_from, s->g[(size_t)to * n + v]);
      if (delta < best) {
        best = delta;
        *out_v = v;
        *out_to = to;
      }
    }
  }
  return best;
}

/* The same move as orc_best_move_plain, found in two steps so that the compiler can vectorise the scan (the
 * descents of n = 10^5 vertices take minutes with the plain scan): with q_c(v) = size_c - 2 G[c][v], the delta of
 * v: a -> c is q_c(v) - q_a(v) + 1 (the formula of orc_move_delta), and c = a gives +1, which never beats a
 * negative best, so every target can be scanned. Step 1 takes the minimum delta of each block of vertices and keeps
 * the first block with the smallest one; step 2 scans that block in the plain order for the first (v, target) that
 * attains it, which is exactly the move the plain scan keeps. */
enum { ORC_SCAN_BLOCK = 256 };

int orc_best_move(const orc_state* s, int* out_v, int* out_to) {
  const int n = s->n;
  const int k = s->k;
  int32_t qa[ORC_SCAN_BLOCK];
  int best = 0;
  int best_block = -1;
  *out_v = -1;
  *out_to = -1;
  for (int b0 = 0; b0 < n; b0 += ORC_SCAN_BLOCK) {
    const int len = n - b0 < ORC_SCAN_BLOCK ? n - b0 : ORC_SCAN_BLOCK;
    for (int i = 0; i < len; ++i) {
      const int a = s->labels[b0 + i];
      qa[i] = s->sizes[a] - 2 * s->g[(size_t)a * n + b0 + i];
    }
    int m = 1;
    for (int c = 0; c < k; ++c) {
      const int32_t* gc = s->g + (size_t)c * n + b0;
      const int sc = s->sizes[c] + 1;
      for (int i = 0; i < len; ++i) {
        const int d = sc - 2 * gc[i] - qa[i];
        m = d < m ? d : m;
      }
    }
    if (m < best) {
      best = m;
      best_block = b0;
    }
  }
  if (best_block < 0) return 0;
  const int end = n - best_block < ORC_SCAN_BLOCK ? n : best_block + ORC_SCAN_BLOCK;
  for (int v = best_block; v < end; ++v) {
    const int from = s->labels[v];
    for (int to = 0; to < k; ++to) {
      if (to == from) continue;
      if (orc_move_delta(s->sizes[from], s->sizes[to], s->g[(size_t)from * n + v], s->g[(size_t)to * n + v]) == best) {
        *out_v = v;
        *out_to = to;
        return best;
      }
    }
  }
  return best; /* not reached: the block holds a move with delta == best */
}

/* src/cc_pbils_cpu.cpp:120-139 */
void orc_apply_move(orc_state* s, const uint64_t* bits, uint32_t words, int v, int to) {
  const int n = s->n;
  const int from = s->labels[v];
  if (from == to) return;
  const int delta = orc_move_delta(s->sizes[from], s->sizes[to], s->g[(size_t)from * n + v], s->g[(size_t)to * n + v]);
  int32_t* g_from = s->g + (size_t)from * n;
  int32_t* g_to = s->g + (size_t)to * n;
  const uint64_t* row = bits + (size_t)v * words;
  for (uint32_t w = 0; w < words; ++w) {
    for (uint64_t b = row[w]; b != 0; b &= b - 1) {
      const unsigned u = w * 64u + (unsigned)__builtin_ctzll(b);
      --g_from[u];
      ++g_to[u];
    }
  }
  --s->sizes[from];
  ++s->sizes[to];
  s->labels[v] = to;
  s->f += delta;
}

/* src/cc_pbils_cpu.cpp:141-149 */
int64_t orc_local_search(orc_state* s, const uint64_t* bits, uint32_t words, int64_t* accepted_moves) {
  int v = -1;
  int to = -1;
  while (orc_best_move(s, &v, &to) < 0) {
    orc_apply_move(s, bits, words, v, to);
    if (accepted_moves) ++*accepted_moves;
  }
  return s->f;
}

/* src/cc_pbils_cpu.cpp:151-161: label index slot*n + v, rebuild only if a label changed */
void orc_perturb(orc_state* s, const uint64_t* bits, uint32_t words, uint64_t edges, double probability,
                 uint64_t base, int slot, int iteration) {
  if (s->k < 2) return;
  int touched = 0;
  for (int v = 0; v < s->n; ++v) {
    const uint64_t index = (uint64_t)slot * (uint64_t)s->n + (uint64_t)v;
    const int label = orc_perturbed_label(base, index, iteration, s->labels[v], s->k, (float)probability);
    touched |= label != s->labels[v];
    s->labels[v] = label;
  }
  if (touched) orc_rebuild(s, bits, words, edges);
}

int64_t orc_rebuild_labels(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, int k,
                           const int32_t* labels, int32_t* g_out, int32_t* sizes_out) {
  orc_state s;
  orc_state_init(&s, (int)n, k);
  memcpy(s.labels, labels, sizeof(int32_t) * n);
  orc_rebuild(&s, bits, words, edges);
  if (g_out) memcpy(g_out, s.g, sizeof(int32_t) * (size_t)k * n);
  if (sizes_out) memcpy(sizes_out, s.sizes, sizeof(int32_t) * (size_t)k);
  const int64_t f = s.f;
  orc_state_free(&s);
  return f;
}

int64_t orc_local_search_labels(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, int k,
                                int32_t* labels, int64_t* moves) {
  orc_state s;
  orc_state_init(&s, (int)n, k);
  memcpy(s.labels, labels, sizeof(int32_t) * n);
  orc_rebuild(&s, bits, words, edges);
  int64_t applied = 0;
  orc_local_search(&s, bits, words, &applied);
  memcpy(labels, s.labels, sizeof(int32_t) * n);
  if (moves) *moves = applied;
  const int64_t f = s.f;
  orc_state_free(&s);
  return f;
}

/* The best move of the state that `labels` define (plain = 1: the literal scan), for the test that pins the blocked
 * scan to the plain one. Returns the delta (0: no improving move) and the move in *v, *to (-1 when none). */
int orc_best_move_labels(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, int k,
                         const int32_t* labels, int plain, int* v, int* to) {
  orc_state s;
  orc_state_init(&s, (int)n, k);
  memcpy(s.labels, labels, sizeof(int32_t) * n);
  orc_rebuild(&s, bits, words, edges);
  const int delta = plain ? orc_best_move_plain(&s, v, to) : orc_best_move(&s, v, to);
  orc_state_free(&s);
  return delta;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* PBILS (include/cc_pbils.hpp:26-85, src/cc_pbils_cpu.cpp:163-252)                                             */
/* ------------------------------------------------------------------------------------------------------------ */

/* include/cc_pbils.hpp:37-48 */
void orc_default_params(orc_params* p) {
  p->k = 2;
  p->population = 128;
  p->tournament = 5;
  p->iterations = 100;
  p->early_stop = 6;
  p->perturbation = 0.4;
  p->gww = 0.1;
  p->threads = 1;
  p->seed = 1;
}

/* src/cc_pbils_cpu.cpp:168 */
int orc_winner_copies(double gww, int population) { return (int)(gww * population); }

/* src/cc_pbils_cpu.cpp:170-175: by f, equal f by slot (insertion sort keeps it dependency-free) */
void orc_rank_slots(const int64_t* f, int population, int32_t* order) {
  for (int i = 0; i < population; ++i) order[i] = i;
  for (int i = 1; i < population; ++i) {
    const int32_t x = order[i];
    int j = i - 1;
    while (j >= 0 && (f[order[j]] > f[x] || (f[order[j]] == f[x] && order[j] > x))) {
      order[j + 1] = order[j];
      --j;
    }
    order[j + 1] = x;
  }
}

typedef struct {
  const uint64_t* bits;
  uint32_t n, words;
  uint64_t edges;
  const orc_params* params;
  orc_state* population;
  orc_state* next;
  int64_t* fitness;
  int64_t* moves;
  int32_t* order;
  uint64_t init_stream, select_stream, perturb_stream;
  int iteration, copies;
} pbils_ctx;

static void phase_init(void* vctx, int slot, int worker) {
  (void)worker;
  pbils_ctx* c = (pbils_ctx*)vctx;
  orc_state* s = &c->population[slot];
  for (uint32_t v = 0; v < c->n; ++v) {
    s->labels[v] = orc_initial_label(c->init_stream, (uint64_t)slot * c->n + v, c->params->k);
  }
  orc_rebuild(s, c->bits, c->words, c->edges);
}

static void phase_select_search(void* vctx, int slot, int worker) {
  (void)worker;
  pbils_ctx* c = (pbils_ctx*)vctx;
  const int chosen = orc_tournament(c->select_stream, slot, c->iteration, c->fitness, c->params->population,
                                    c->params->tournament);
  orc_state_copy(&c->next[slot], &c->population[chosen]);
  int64_t applied = 0;
  orc_local_search(&c->next[slot], c->bits, c->words, &applied);
  c->moves[slot] = applied;
}

static void phase_copy_winners(void* vctx, int i, int worker) {
  (void)worker;
  pbils_ctx* c = (pbils_ctx*)vctx;
  const int p = c->params->population;
  orc_state_copy(&c->next[c->order[p - 1 - i]], &c->next[c->order[i]]);
}

static void phase_perturb(void* vctx, int slot, int worker) {
  (void)worker;
  pbils_ctx* c = (pbils_ctx*)vctx;
  orc_perturb(&c->next[slot], c->bits, c->words, c->edges, c->params->perturbation, c->perturb_stream, slot,
              c->iteration);
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Record = first minimum over slots, taken only when strictly better (src/cc_pbils_cpu.cpp:208-214). As in the
 * reference, the "have a record" test is "record labels are non-empty", so with n = 0 every offer improves. */
static int offer_record(const orc_state* states, int population, uint32_t n, int* have, int64_t* objective,
                        int32_t* labels) {
  int best = 0;
  for (int s = 1; s < population; ++s) {
    if (states[s].f < states[best].f) best = s;
  }
  if (*have && n > 0 && states[best].f >= *objective) return 0;
  *objective = states[best].f;
  memcpy(labels, states[best].labels, sizeof(int32_t) * n);
  *have = 1;
  return 1;
}

int orc_solve_pbils(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, const orc_params* params,
                    int32_t* labels_out, orc_result* result, int64_t* trace, char* err, size_t err_len) {
  /* src/cc_pbils_cpu.cpp:177-181 */
  if (params->k < 1) {
    snprintf(err, err_len, "k must be >= 1");
    return -1;
  }
  if (params->population < 1) {
    snprintf(err, err_len, "population must be >= 1");
    return -1;
  }
  if (params->gww < 0.0 || params->gww > 0.5) {
    snprintf(err, err_len, "gww must be in [0, 0.5]");
    return -1;
  }
  const double started = now_seconds();
  const int size = params->population;
  pbils_ctx c;
  memset(&c, 0, sizeof(c));
  c.bits = bits;
  c.n = n;
  c.words = words;
  c.edges = edges;
  c.params = params;
  c.init_stream = orc_stream_seed(params->seed, ORC_INIT_STREAM, 0);
  c.select_stream = orc_stream_seed(params->seed, ORC_SELECT_STREAM, 0);
  c.perturb_stream = orc_stream_seed(params->seed, ORC_PERTURB_STREAM, 0);
  c.population = (orc_state*)calloc((size_t)size, sizeof(orc_state));
  c.next = (orc_state*)calloc((size_t)size, sizeof(orc_state));
  c.fitness = (int64_t*)calloc((size_t)size, sizeof(int64_t));
  c.moves = (int64_t*)calloc((size_t)size, sizeof(int64_t));
  c.order = (int32_t*)calloc((size_t)size, sizeof(int32_t));
  for (int s = 0; s < size; ++s) {
    orc_state_init(&c.population[s], (int)n, params->k);
    orc_state_init(&c.next[s], (int)n, params->k);
  }
  int threads = orc_resolve_threads(params->threads);
  if (threads > size) threads = size; /* src/cc_pbils_cpu.cpp:163-166 */

  orc_parallel_for(size, threads, phase_init, &c);

  memset(result, 0, sizeof(*result));
  result->objective = -1;
  int have = 0;
  offer_record(c.population, size, n, &have, &result->objective, labels_out);
  int stall = 0;

  for (int iteration = 0; iteration < params->iterations; ++iteration) {
    c.iteration = iteration;
    for (int s = 0; s < size; ++s) c.fitness[s] = c.population[s].f;
    orc_parallel_for(size, threads, phase_select_search, &c);

    const int improved = offer_record(c.next, size, n, &have, &result->objective, labels_out);
    for (int s = 0; s < size; ++s) result->accepted_moves += c.moves[s];
    /* Progress::Next, include/cc_pbils.hpp:69-78 */
    ++result->iterations_done;
    result->local_searches += size;
    stall = improved ? 0 : stall + 1;
    if (trace) trace[iteration] = result->objective;
    if (stall >= params->early_stop) break;

    c.copies = orc_winner_copies(params->gww, size);
    if (c.copies > 0) {
      for (int s = 0; s < size; ++s) c.fitness[s] = c.next[s].f;
      orc_rank_slots(c.fitness, size, c.order);
      orc_parallel_for(c.copies, threads, phase_copy_winners, &c);
    }
    if (params->k >= 2 && params->perturbation > 0.0) orc_parallel_for(size, threads, phase_perturb, &c);
    orc_state* tmp = c.population;
    c.population = c.next;
    c.next = tmp;
  }

  result->seconds = now_seconds() - started;
  result->clusters_used = orc_clusters_used(labels_out, n, params->k);
  for (int s = 0; s < size; ++s) {
    orc_state_free(&c.population[s]);
    orc_state_free(&c.next[s]);
  }
  free(c.population);
  free(c.next);
  free(c.fitness);
  free(c.moves);
  free(c.order);
  return 0;
}

/* ------------------------------------------------------------------------------------------------------------ */
/* Neighborhood (cuda/cc_verify_gpu.cu:21-35, semantics of cuda/cc_pbils_gpu.cu:389-436)                         */
/* ------------------------------------------------------------------------------------------------------------ */

typedef struct {
  const uint64_t* bits;
  uint32_t n, words;
  uint64_t edges;
  int local_search;
  int64_t* f;
  int64_t* moves;
  orc_state* scratch; /* one per worker */
} nbh_ctx;

static void neighborhood_start(const nbh_ctx* c, orc_state* s, uint32_t v, int64_t* moves) {
  for (uint32_t u = 0; u < c->n; ++u) s->labels[u] = (u == v || joined(c->bits, c->words, v, u)) ? 0 : 1;
  orc_rebuild(s, c->bits, c->words, c->edges);
  if (c->local_search) orc_local_search(s, c->bits, c->words, moves);
}

static void phase_neighborhood(void* vctx, int item, int worker) {
  nbh_ctx* c = (nbh_ctx*)vctx;
  int64_t applied = 0;
  neighborhood_start(c, &c->scratch[worker], (uint32_t)item, &applied);
  c->f[item] = c->scratch[worker].f;
  c->moves[item] = applied;
}

int orc_neighborhood(const uint64_t* bits, uint32_t n, uint32_t words, uint64_t edges, int local_search,
                     int threads, int32_t* labels_out, orc_result* result) {
  memset(result, 0, sizeof(*result));
  result->objective = -1;
  if (n == 0) return -1;
  const double started = now_seconds();
  threads = orc_resolve_threads(threads);
  if (threads > (int)n) threads = (int)n;
  nbh_ctx c;
  c.bits = bits;
  c.n = n;
  c.words = words;
  c.edges = edges;
  c.local_search = local_search;
  c.f = (int64_t*)calloc(n, sizeof(int64_t));
  c.moves = (int64_t*)calloc(n, sizeof(int64_t));
  c.scratch = (orc_state*)calloc((size_t)threads, sizeof(orc_state));
  for (int t = 0; t < threads; ++t) orc_state_init(&c.scratch[t], (int)n, 2);
  orc_parallel_for((int)n, threads, phase_neighborhood, &c);
  uint32_t best = 0;
  for (uint32_t v = 0; v < n; ++v) {
    result->accepted_moves += c.moves[v];
    if (c.f[v] < c.f[best]) best = v;
  }
  int64_t ignored = 0;
  neighborhood_start(&c, &c.scratch[0], best, &ignored);
  memcpy(labels_out, c.scratch[0].labels, sizeof(int32_t) * n);
  result->objective = c.f[best];
  result->local_searches = local_search ? (int64_t)n : 0;
  result->clusters_used = orc_clusters_used(labels_out, n, 2);
  result->seconds = now_seconds() - started;
  for (int t = 0; t < threads; ++t) orc_state_free(&c.scratch[t]);
  free(c.scratch);
  free(c.f);
  free(c.moves);
  return 0;
}
