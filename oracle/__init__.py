"""ctypes bindings to the checker libraries. TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py import this package. It wraps
  * liboracle.so      — the C restatement of the reference path (cc_oracle.c), and
  * _ref/libccref.so  — the reference CPU solver compiled out of tree with ref_shim.cpp (optional: present only
                        where the reference sources were available at build time).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libccref.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)


class Result(ctypes.Structure):
    _fields_ = [
        ("objective", ctypes.c_int64),
        ("iterations_done", ctypes.c_int32),
        ("clusters_used", ctypes.c_int32),
        ("local_searches", ctypes.c_int64),
        ("accepted_moves", ctypes.c_int64),
        ("seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Params(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32),
        ("population", ctypes.c_int32),
        ("tournament", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("early_stop", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("perturbation", ctypes.c_double),
        ("gww", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


def _ptr(array, kind):
    return array.ctypes.data_as(kind)


def words_per_row(n):
    return max(1, (n + 63) // 64)


class Graph:
    """A 64-bit-word bitset adjacency: bits[v, w] holds neighbours 64w..64w+63 of v."""

    def __init__(self, bits, n, edges):
        self.bits = np.ascontiguousarray(bits, dtype=np.uint64)
        self.n = int(n)
        self.words = words_per_row(n)
        self.edges = int(edges)
        assert self.bits.shape == (self.n, self.words) or (self.n == 0 and self.bits.size == 0)

    def neighbours(self, v):
        row = self.bits[v]
        bits = np.unpackbits(row.view(np.uint8), bitorder="little")[: self.n]
        return np.nonzero(bits)[0]

    def dense(self):
        if self.n == 0:
            return np.zeros((0, 0), dtype=np.uint8)
        return np.unpackbits(self.bits.view(np.uint8), axis=1, bitorder="little")[:, : self.n]

    @staticmethod
    def from_dense(adj):
        adj = np.asarray(adj, dtype=np.uint8)
        n = adj.shape[0]
        words = words_per_row(n)
        padded = np.zeros((n, words * 64), dtype=np.uint8)
        padded[:, :n] = adj
        bits = np.packbits(padded, axis=1, bitorder="little").view(np.uint64).reshape(n, words)
        return Graph(bits.copy(), n, int(adj.sum()) // 2)


class Oracle:
    """The C restatement (cc_oracle.c)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run `make -C oracle`")
        lib = ctypes.CDLL(path)
        self.lib = lib
        lib.orc_splitmix64.argtypes = [_u64p]
        lib.orc_splitmix64.restype = ctypes.c_uint64
        lib.orc_stream_seed.argtypes = [ctypes.c_uint64] * 3
        lib.orc_stream_seed.restype = ctypes.c_uint64
        lib.orc_initial_label.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
        lib.orc_initial_label.restype = ctypes.c_int
        lib.orc_tournament.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_int, ctypes.c_int]
        lib.orc_tournament.restype = ctypes.c_int
        lib.orc_perturbed_label.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_float]
        lib.orc_perturbed_label.restype = ctypes.c_int
        lib.orc_rand_float.argtypes = [_u64p]
        lib.orc_rand_float.restype = ctypes.c_float
        lib.orc_rand_below.argtypes = [_u64p, ctypes.c_uint32]
        lib.orc_rand_below.restype = ctypes.c_uint32
        lib.orc_sample_without_replacement.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                                       ctypes.POINTER(ctypes.c_uint32)]
        lib.orc_sample_without_replacement.restype = ctypes.c_uint32
        for name in ("orc_graph_erdos_renyi",):
            getattr(lib, name).argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, _u64p, ctypes.c_int]
            getattr(lib, name).restype = ctypes.c_uint64
        lib.orc_graph_sbm.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_uint64, _u64p, ctypes.c_int]
        lib.orc_graph_sbm.restype = ctypes.c_uint64
        lib.orc_graph_chung_lu.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                           _u64p, ctypes.c_int]
        lib.orc_graph_chung_lu.restype = ctypes.c_uint64
        for name in ("orc_objective_direct", "orc_objective_bitwise"):
            getattr(lib, name).argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, _i32p]
            getattr(lib, name).restype = ctypes.c_int64
        lib.orc_rebuild_labels.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                           _i32p, _i32p, _i32p]
        lib.orc_rebuild_labels.restype = ctypes.c_int64
        lib.orc_local_search_labels.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                                ctypes.c_int, _i32p, _i64p]
        lib.orc_local_search_labels.restype = ctypes.c_int64
        lib.orc_best_move_labels.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                             _i32p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]
        lib.orc_best_move_labels.restype = ctypes.c_int
        lib.orc_rank_slots.argtypes = [_i64p, ctypes.c_int, _i32p]
        lib.orc_winner_copies.argtypes = [ctypes.c_double, ctypes.c_int]
        lib.orc_winner_copies.restype = ctypes.c_int
        lib.orc_solve_pbils.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.POINTER(Params), _i32p, ctypes.POINTER(Result), _i64p,
                                        ctypes.c_char_p, ctypes.c_size_t]
        lib.orc_solve_pbils.restype = ctypes.c_int
        lib.orc_neighborhood.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                         ctypes.c_int, _i32p, ctypes.POINTER(Result)]
        lib.orc_neighborhood.restype = ctypes.c_int
        lib.orc_clusters_used.argtypes = [_i32p, ctypes.c_uint32, ctypes.c_int]
        lib.orc_clusters_used.restype = ctypes.c_int

    # -- rng --
    def splitmix_outputs(self, seed, count):
        state = ctypes.c_uint64(seed)
        return [self.lib.orc_splitmix64(ctypes.byref(state)) for _ in range(count)]

    def stream_seed(self, seed, stream, step):
        return self.lib.orc_stream_seed(seed, stream, step)

    def sample_without_replacement(self, total, n, seed):
        out = np.zeros(max(total, 1), dtype=np.uint32)
        count = self.lib.orc_sample_without_replacement(total, n, seed, _ptr(out, ctypes.POINTER(ctypes.c_uint32)))
        return out[:count]

    def initial_label(self, base, index, k):
        return self.lib.orc_initial_label(base, index, k)

    def tournament(self, base, slot, iteration, f, tournament):
        f = np.ascontiguousarray(f, dtype=np.int64)
        return self.lib.orc_tournament(base, slot, iteration, _ptr(f, _i64p), len(f), tournament)

    def perturbed_label(self, base, index, iteration, label, k, p):
        return self.lib.orc_perturbed_label(base, index, iteration, label, k, p)

    # -- graphs --
    def erdos_renyi(self, n, p, seed, threads=0):
        bits = np.zeros((n, words_per_row(n)), dtype=np.uint64)
        edges = self.lib.orc_graph_erdos_renyi(n, p, seed, _ptr(bits, _u64p), threads)
        return Graph(bits, n, edges)

    def sbm(self, n, blocks, p_in, p_out, seed, threads=0):
        bits = np.zeros((n, words_per_row(n)), dtype=np.uint64)
        edges = self.lib.orc_graph_sbm(n, blocks, p_in, p_out, seed, _ptr(bits, _u64p), threads)
        return Graph(bits, n, edges)

    def chung_lu(self, n, gamma, mean_degree, seed, threads=0):
        bits = np.zeros((n, words_per_row(n)), dtype=np.uint64)
        edges = self.lib.orc_graph_chung_lu(n, gamma, mean_degree, seed, _ptr(bits, _u64p), threads)
        return Graph(bits, n, edges)

    # -- objective / state --
    def objective_direct(self, g, labels):
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        return self.lib.orc_objective_direct(_ptr(g.bits, _u64p), g.n, g.words, _ptr(labels, _i32p))

    def objective_bitwise(self, g, labels):
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        return self.lib.orc_objective_bitwise(_ptr(g.bits, _u64p), g.n, g.words, _ptr(labels, _i32p))

    def rebuild(self, g, k, labels):
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        gg = np.zeros((k, max(g.n, 1)), dtype=np.int32)
        sizes = np.zeros(k, dtype=np.int32)
        f = self.lib.orc_rebuild_labels(_ptr(g.bits, _u64p), g.n, g.words, g.edges, k, _ptr(labels, _i32p),
                                        _ptr(gg, _i32p), _ptr(sizes, _i32p))
        return f, gg[:, : g.n], sizes

    def local_search(self, g, k, labels):
        labels = np.array(labels, dtype=np.int32)
        moves = ctypes.c_int64(0)
        f = self.lib.orc_local_search_labels(_ptr(g.bits, _u64p), g.n, g.words, g.edges, k, _ptr(labels, _i32p),
                                             ctypes.byref(moves))
        return f, labels, moves.value

    def best_move(self, g, k, labels, plain=False):
        """(delta, v, to) of the steepest move from `labels`; plain=True uses the literal reference scan."""
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        v, to = ctypes.c_int(0), ctypes.c_int(0)
        d = self.lib.orc_best_move_labels(_ptr(g.bits, _u64p), g.n, g.words, g.edges, k, _ptr(labels, _i32p),
                                          int(plain), ctypes.byref(v), ctypes.byref(to))
        return d, v.value, to.value

    def rank_slots(self, f):
        f = np.ascontiguousarray(f, dtype=np.int64)
        order = np.zeros(len(f), dtype=np.int32)
        self.lib.orc_rank_slots(_ptr(f, _i64p), len(f), _ptr(order, _i32p))
        return order

    def winner_copies(self, gww, population):
        return self.lib.orc_winner_copies(gww, population)

    def solve_pbils(self, g, k=2, population=128, tournament=5, iterations=100, early_stop=6, perturbation=0.4,
                    gww=0.1, seed=1, threads=0, trace=False):
        p = Params(k, population, tournament, iterations, early_stop, threads, perturbation, gww, seed)
        labels = np.zeros(max(g.n, 1), dtype=np.int32)
        res = Result()
        tr = np.full(max(iterations, 1), -1, dtype=np.int64)
        err = ctypes.create_string_buffer(256)
        rc = self.lib.orc_solve_pbils(_ptr(g.bits, _u64p), g.n, g.words, g.edges, ctypes.byref(p),
                                      _ptr(labels, _i32p), ctypes.byref(res), _ptr(tr, _i64p), err, 256)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        out = res.as_dict()
        out["labels"] = labels[: g.n]
        if trace:
            out["trace"] = tr[: res.iterations_done]
        return out

    def neighborhood(self, g, local_search, threads=0):
        labels = np.zeros(max(g.n, 1), dtype=np.int32)
        res = Result()
        rc = self.lib.orc_neighborhood(_ptr(g.bits, _u64p), g.n, g.words, g.edges, int(local_search), threads,
                                       _ptr(labels, _i32p), ctypes.byref(res))
        if rc != 0:
            raise RuntimeError("Neighborhood requires n >= 1")
        out = res.as_dict()
        out["labels"] = labels[: g.n]
        return out


class Reference:
    """The reference CPU solver built out of tree (oracle/_ref/libccref.so)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run `make -C oracle ref` where the reference exists")
        lib = ctypes.CDLL(path)
        self.lib = lib
        lib.ref_erdos_renyi.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, _u64p, ctypes.c_uint32]
        lib.ref_erdos_renyi.restype = ctypes.c_uint64
        lib.ref_solve_cpu.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32] + [ctypes.c_int32] * 5 + [
            ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_uint64, _i32p, ctypes.POINTER(Result),
            ctypes.c_char_p, ctypes.c_size_t]
        lib.ref_solve_cpu.restype = ctypes.c_int
        lib.ref_neighborhood.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, _i32p,
                                         ctypes.POINTER(Result)]
        lib.ref_rebuild.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, _i32p, _i32p, _i32p]
        lib.ref_rebuild.restype = ctypes.c_int64
        lib.ref_local_search.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, _i32p, _i64p]
        lib.ref_local_search.restype = ctypes.c_int64
        for name in ("ref_objective_direct", "ref_objective_bitwise"):
            getattr(lib, name).argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, _i32p]
            getattr(lib, name).restype = ctypes.c_int64
        lib.ref_rank_slots.argtypes = [_i64p, ctypes.c_int32, _i32p]
        lib.ref_winner_copies.argtypes = [ctypes.c_double, ctypes.c_int32]
        lib.ref_winner_copies.restype = ctypes.c_int32
        lib.ref_stream_seed.argtypes = [ctypes.c_uint64] * 3
        lib.ref_stream_seed.restype = ctypes.c_uint64
        lib.ref_initial_label.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int32]
        lib.ref_initial_label.restype = ctypes.c_int32
        lib.ref_perturbed_label.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_float]
        lib.ref_perturbed_label.restype = ctypes.c_int32
        lib.ref_tournament.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, _i64p, ctypes.c_int32,
                                       ctypes.c_int32]
        lib.ref_tournament.restype = ctypes.c_int32
        lib.ref_load_tags.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_uint32,
                                      ctypes.c_uint64, _u64p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.c_char_p, ctypes.c_size_t]
        lib.ref_load_tags.restype = ctypes.c_int64
        lib.ref_load_graph.argtypes = [ctypes.c_char_p, ctypes.c_int32, _u64p, ctypes.c_uint64,
                                       ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, ctypes.c_size_t]
        lib.ref_load_graph.restype = ctypes.c_int64
        lib.ref_save_graph.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_int32]
        lib.ref_save_graph.restype = ctypes.c_int
        lib.ref_sample_without_replacement.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                                       ctypes.POINTER(ctypes.c_uint32)]

    def erdos_renyi(self, n, p, seed):
        words = words_per_row(n)
        bits = np.zeros((n, words), dtype=np.uint64)
        edges = self.lib.ref_erdos_renyi(n, p, seed, _ptr(bits, _u64p), words)
        return Graph(bits, n, edges)

    def solve_pbils(self, g, k=2, population=128, tournament=5, iterations=100, early_stop=6, perturbation=0.4,
                    gww=0.1, seed=1, threads=0):
        labels = np.zeros(max(g.n, 1), dtype=np.int32)
        res = Result()
        err = ctypes.create_string_buffer(256)
        rc = self.lib.ref_solve_cpu(_ptr(g.bits, _u64p), g.n, g.words, k, population, tournament, iterations,
                                    early_stop, perturbation, gww, threads, seed, _ptr(labels, _i32p),
                                    ctypes.byref(res), err, 256)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        out = res.as_dict()
        out["labels"] = labels[: g.n]
        return out

    def neighborhood(self, g, local_search):
        labels = np.zeros(max(g.n, 1), dtype=np.int32)
        res = Result()
        self.lib.ref_neighborhood(_ptr(g.bits, _u64p), g.n, g.words, int(local_search), _ptr(labels, _i32p),
                                  ctypes.byref(res))
        out = res.as_dict()
        out["labels"] = labels[: g.n]
        return out

    def rebuild(self, g, k, labels):
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        gg = np.zeros((k, max(g.n, 1)), dtype=np.int32)
        sizes = np.zeros(k, dtype=np.int32)
        f = self.lib.ref_rebuild(_ptr(g.bits, _u64p), g.n, g.words, k, _ptr(labels, _i32p), _ptr(gg, _i32p),
                                 _ptr(sizes, _i32p))
        return f, gg[:, : g.n], sizes

    def local_search(self, g, k, labels):
        labels = np.array(labels, dtype=np.int32)
        moves = ctypes.c_int64(0)
        f = self.lib.ref_local_search(_ptr(g.bits, _u64p), g.n, g.words, k, _ptr(labels, _i32p), ctypes.byref(moves))
        return f, labels, moves.value

    def objective_direct(self, g, labels):
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        return self.lib.ref_objective_direct(_ptr(g.bits, _u64p), g.n, g.words, _ptr(labels, _i32p))

    def rank_slots(self, f):
        f = np.ascontiguousarray(f, dtype=np.int64)
        order = np.zeros(len(f), dtype=np.int32)
        self.lib.ref_rank_slots(_ptr(f, _i64p), len(f), _ptr(order, _i32p))
        return order

    def load_tags(self, path, similarity, threshold, n, seed):
        """Graph::LoadTags of the reference; raises RuntimeError with its message."""
        edges = ctypes.c_uint64(0)
        err = ctypes.create_string_buffer(512)
        args = (os.fsencode(path), similarity.encode(), threshold, n, seed)
        size = self.lib.ref_load_tags(*args, None, 0, ctypes.byref(edges), err, 512)
        if size < 0:
            raise RuntimeError(err.value.decode())
        bits = np.zeros((size, words_per_row(size)), dtype=np.uint64)
        self.lib.ref_load_tags(*args, _ptr(bits, _u64p), bits.size, ctypes.byref(edges), err, 512)
        return Graph(bits, size, edges.value)

    GRAPH_FORMATS = {"matrix": 0, "json": 1, "edges": 2}

    def load_graph(self, path, fmt):
        """Graph::LoadMatrix / LoadBaselineJson / LoadEdgeList of the reference; raises RuntimeError with its
        message."""
        edges = ctypes.c_uint64(0)
        err = ctypes.create_string_buffer(512)
        code = self.GRAPH_FORMATS[fmt]
        size = self.lib.ref_load_graph(os.fsencode(path), code, None, 0, ctypes.byref(edges), err, 512)
        if size < 0:
            raise RuntimeError(err.value.decode())
        bits = np.zeros((size, words_per_row(size)), dtype=np.uint64)
        self.lib.ref_load_graph(os.fsencode(path), code, _ptr(bits, _u64p), bits.size, ctypes.byref(edges), err, 512)
        return Graph(bits, size, edges.value)

    def save_graph(self, g, path, fmt):
        """Graph::SaveMatrix / SaveBaselineJson of the reference."""
        bits = np.ascontiguousarray(g.bits, dtype=np.uint64)
        if self.lib.ref_save_graph(_ptr(bits, _u64p), g.n, g.words, os.fsencode(path), self.GRAPH_FORMATS[fmt]):
            raise RuntimeError("reference save failed")

    def sample_without_replacement(self, total, n, seed):
        out = np.zeros(max(total, 1), dtype=np.uint32)
        self.lib.ref_sample_without_replacement(total, n, seed, _ptr(out, ctypes.POINTER(ctypes.c_uint32)))
        return out[: (total if n == 0 or n >= total else n)]


def reference_available():
    return os.path.exists(REF_SO)
