"""Python binding of the MI355X PBILS correlation-clustering library (lib/libcc_hip.so) over its C ABI.

The library is the product: every solve runs the HIP kernels of csrc/ on the GPU. Importing this module raises if
the library has not been built (`make -C cuda_graph_clustering_amd`); there is no CPU fallback.

Mirrors the reference's PbilsParams / PbilsResult (include/cc_pbils.hpp:26-59 of the reference) and its GPU entry
points SolveGpu, SolveNeighborhoodGpu and SolveNeighborhoodWithManyLocalSearchesGpu (include/cc_gpu.hpp:12-20).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcc_hip.so")
CLI_PATH = os.path.join(HERE, "bin", "cc_gpu")

LS_AUTO, LS_GLOBAL, LS_SHARED, LS_WIDE = 0, 1, 2, 3
LS_KERNELS = {"auto": LS_AUTO, "global": LS_GLOBAL, "shared": LS_SHARED, "wide": LS_WIDE}
REBUILDS = {"popcount": 0, "mfma": 1}
DESCENTS = {0: "generic-global", 1: "generic-lds", 2: "k2-registers", 3: "k2-wide", 4: "k3-wide", 5: "registers",
            6: "mixed", 7: "k3-chunk", 8: "k3-packed"}

OK, EINVAL, ENOMEM, EDEVICE, ENODEV = 0, 1, 2, 3, 4


class CcError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


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
        ("time_limit_sec", ctypes.c_double),
        ("verbose", ctypes.c_int32),
        ("ls_kernel", ctypes.c_int32),
        ("profile", ctypes.c_int32),
        ("rebuild", ctypes.c_int32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("objective", ctypes.c_int64),
        ("iterations_done", ctypes.c_int32),
        ("clusters_used", ctypes.c_int32),
        ("local_searches", ctypes.c_int64),
        ("accepted_moves", ctypes.c_int64),
        ("seconds", ctypes.c_double),
        ("descent", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("blocks", ctypes.c_int32),
        ("vertices_per_thread", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int64),
        ("search_ms", ctypes.c_double),
        ("rebuild_ms", ctypes.c_double),
        ("iteration_ms", ctypes.c_double),
        ("search_launches", ctypes.c_int64),
        ("rebuild_launches", ctypes.c_int64),
    ]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["descent_name"] = DESCENTS.get(self.descent, "?")
        return d


class BatchRun(ctypes.Structure):
    _fields_ = [
        ("graph", ctypes.c_void_p),
        ("params", Params),
        ("labels_out", ctypes.POINTER(ctypes.c_int32)),
        ("trace_out", ctypes.POINTER(ctypes.c_int64)),
        ("stats", Stats),
    ]


class ServeOptions(ctypes.Structure):
    _fields_ = [
        ("budget_bytes", ctypes.c_int64),
        ("lanes", ctypes.c_int32),
        ("max_active", ctypes.c_int32),
        ("large_bytes", ctypes.c_int64),
    ]


class TuneCandidate(ctypes.Structure):
    _fields_ = [("descent", ctypes.c_int32), ("threads", ctypes.c_int32), ("vertices_per_thread", ctypes.c_int32),
                ("rows_in_lds", ctypes.c_int32), ("blocks", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("lds_bytes", ctypes.c_int64), ("accepted_moves", ctypes.c_int64), ("search_ms", ctypes.c_double)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("arch", ctypes.c_char * 32),
        ("compute_units", ctypes.c_int32),
        ("lds_per_block", ctypes.c_int32),
        ("l2_bytes", ctypes.c_int32),
        ("clock_khz", ctypes.c_int32),
        ("total_bytes", ctypes.c_int64),
        ("free_bytes", ctypes.c_int64),
    ]


_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)

# Every symbol the C ABI (include/cc_hip.h) exports.
ABI_SYMBOLS = [
    "cc_pbils_default_params",
    "cc_hip_solve_pbils",
    "cc_hip_solve_neighborhood",
    "cc_hip_graph_upload",
    "cc_hip_graph_free",
    "cc_hip_graph_solve_pbils",
    "cc_hip_graph_solve_pbils_static",
    "cc_hip_graph_solve_neighborhood",
    "cc_hip_solve_batch",
    "cc_serve_default_options",
    "cc_hip_server_create",
    "cc_hip_server_destroy",
    "cc_hip_server_submit",
    "cc_hip_server_wait",
    "cc_hip_footprint",
    "cc_hip_server_counters",
    "cc_words64",
    "cc_graph_erdos_renyi",
    "cc_graph_sbm",
    "cc_graph_chung_lu",
    "cc_objective_direct",
    "cc_objective_bitwise",
    "cc_graph_load_tags",
    "cc_graph_load",
    "cc_graph_save",
    "cc_host_free",
    "cc_sample_without_replacement",
    "cc_hip_device_info",
    "cc_hip_autotune",
    "cc_hip_tune_clear",
    "cc_hip_tune_save",
    "cc_hip_tune_load",
]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `make -C cuda_graph_clustering_amd` (there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.cc_pbils_default_params.argtypes = [ctypes.POINTER(Params)]
    lib.cc_hip_solve_pbils.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(Params), _i32p,
                                       ctypes.POINTER(Stats), _i64p, ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_solve_neighborhood.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(Params),
                                              ctypes.c_int32, _i32p, ctypes.POINTER(Stats), ctypes.c_char_p,
                                              ctypes.c_size_t]
    lib.cc_hip_graph_upload.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p),
                                        ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_graph_free.argtypes = [ctypes.c_void_p]
    lib.cc_hip_graph_solve_pbils.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), _i32p, ctypes.POINTER(Stats),
                                             _i64p, ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_graph_solve_pbils_static.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_void_p, _i32p,
                                                    ctypes.POINTER(Stats), _i64p, ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_graph_solve_pbils_static.restype = ctypes.c_int
    lib.cc_hip_graph_solve_neighborhood.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int32, _i32p,
                                                    ctypes.POINTER(Stats), ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_solve_batch.argtypes = [ctypes.POINTER(BatchRun), ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_serve_default_options.argtypes = [ctypes.POINTER(ServeOptions)]
    lib.cc_hip_server_create.argtypes = [ctypes.POINTER(ServeOptions), ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_server_destroy.argtypes = [ctypes.c_void_p]
    lib.cc_hip_server_submit.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint32, ctypes.c_uint64,
                                         ctypes.POINTER(Params), ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p,
                                         ctypes.c_size_t]
    lib.cc_hip_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double, _i32p, _i64p,
                                       ctypes.POINTER(Stats), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_footprint.argtypes = [ctypes.c_uint32, ctypes.POINTER(Params)]
    lib.cc_hip_footprint.restype = ctypes.c_int64
    lib.cc_hip_server_counters.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int64)] * 5
    lib.cc_words64.argtypes = [ctypes.c_uint32]
    lib.cc_words64.restype = ctypes.c_uint32
    lib.cc_graph_erdos_renyi.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, _u64p, ctypes.c_int32]
    lib.cc_graph_erdos_renyi.restype = ctypes.c_uint64
    lib.cc_graph_sbm.argtypes = [ctypes.c_uint32, ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                 _u64p, ctypes.c_int32]
    lib.cc_graph_sbm.restype = ctypes.c_uint64
    lib.cc_graph_chung_lu.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _u64p,
                                      ctypes.c_int32]
    lib.cc_graph_chung_lu.restype = ctypes.c_uint64
    for name in ("cc_objective_direct", "cc_objective_bitwise"):
        getattr(lib, name).argtypes = [_u64p, ctypes.c_uint32, _i32p]
        getattr(lib, name).restype = ctypes.c_int64
    lib.cc_graph_load_tags.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_uint32,
                                       ctypes.c_uint64, ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_uint32),
                                       ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_graph_load_tags.restype = ctypes.c_int
    lib.cc_graph_load.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(_u64p),
                                  ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p,
                                  ctypes.c_size_t]
    lib.cc_graph_load.restype = ctypes.c_int
    lib.cc_graph_save.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p,
                                  ctypes.c_size_t]
    lib.cc_graph_save.restype = ctypes.c_int
    lib.cc_host_free.argtypes = [ctypes.c_void_p]
    lib.cc_sample_without_replacement.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                                  ctypes.POINTER(ctypes.c_uint32)]
    lib.cc_hip_device_info.argtypes = [ctypes.POINTER(DeviceInfo), ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_autotune.argtypes = [ctypes.c_void_p, ctypes.POINTER(Params), ctypes.c_int32,
                                    ctypes.POINTER(TuneCandidate), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_tune_clear.argtypes = []
    lib.cc_hip_tune_save.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
    lib.cc_hip_tune_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p,
                                     ctypes.c_size_t]
    return lib


lib = _load()


def words64(n):
    return max(1, (n + 63) // 64)


class Graph:
    """Host bitset: n rows of words64(n) uint64 words."""

    def __init__(self, bits, n, edges):
        self.n = int(n)
        self.words = words64(self.n)
        self.bits = np.ascontiguousarray(bits, dtype=np.uint64).reshape(self.n, self.words)
        self.edges = int(edges)

    def ptr(self):
        return self.bits.ctypes.data_as(_u64p)


def erdos_renyi(n, p, seed, threads=0):
    bits = np.zeros((n, words64(n)), dtype=np.uint64)
    edges = lib.cc_graph_erdos_renyi(n, p, seed, bits.ctypes.data_as(_u64p), threads)
    return Graph(bits, n, edges)


def sbm(n, blocks, p_in, p_out, seed, threads=0):
    bits = np.zeros((n, words64(n)), dtype=np.uint64)
    edges = lib.cc_graph_sbm(n, blocks, p_in, p_out, seed, bits.ctypes.data_as(_u64p), threads)
    return Graph(bits, n, edges)


def chung_lu(n, gamma, mean_degree, seed, threads=0):
    bits = np.zeros((n, words64(n)), dtype=np.uint64)
    edges = lib.cc_graph_chung_lu(n, gamma, mean_degree, seed, bits.ctypes.data_as(_u64p), threads)
    return Graph(bits, n, edges)


def from_bits(bits, n):
    bits = np.ascontiguousarray(bits, dtype=np.uint64).reshape(n, words64(n))
    edges = int(np.unpackbits(bits.view(np.uint8)).sum(dtype=np.int64)) // 2
    return Graph(bits, n, edges)


def load_tags(path, similarity="jaccard", threshold=0.5, n=0, seed=1):
    """Graph::LoadTags: the tag-similarity graph of a {"key": ["tag", ...]} JSON file over n objects sampled by
    seed (0: all), an edge when the similarity of two tag sets is >= threshold."""
    size = ctypes.c_uint32(0)
    edges = ctypes.c_uint64(0)
    buf = _u64p()
    err = ctypes.create_string_buffer(512)
    _check(lib.cc_graph_load_tags(os.fsencode(path), similarity.encode(), float(threshold), int(n), int(seed),
                                  ctypes.byref(buf), ctypes.byref(size), ctypes.byref(edges), err, 512), err)
    try:
        count = size.value * words64(size.value)
        bits = np.ctypeslib.as_array(buf, shape=(count,)).copy() if count else np.zeros(0, dtype=np.uint64)
    finally:
        lib.cc_host_free(buf)
    return Graph(bits, size.value, edges.value)


GRAPH_FORMATS = {"matrix": 0, "json": 1, "edges": 2}


def _take_bits(buf, n):
    try:
        count = n * words64(n)
        return np.ctypeslib.as_array(buf, shape=(count,)).copy() if count else np.zeros(0, dtype=np.uint64)
    finally:
        lib.cc_host_free(buf)


def load_graph(path, fmt="matrix"):
    """An instance file of the command line: "matrix" (Graph::LoadMatrix), "json" (the "graph" block of a baseline
    result, Graph::LoadBaselineJson) or "edges" (SNAP/CSV edge list, Graph::LoadEdgeList)."""
    size = ctypes.c_uint32(0)
    edges = ctypes.c_uint64(0)
    buf = _u64p()
    err = ctypes.create_string_buffer(512)
    _check(lib.cc_graph_load(os.fsencode(path), GRAPH_FORMATS[fmt], ctypes.byref(buf), ctypes.byref(size),
                             ctypes.byref(edges), err, 512), err)
    return Graph(_take_bits(buf, size.value), size.value, edges.value)


def save_graph(graph, path, fmt="matrix"):
    """Writes a graph as "matrix" (Graph::SaveMatrix), "json" (Graph::SaveBaselineJson) or "edges"."""
    err = ctypes.create_string_buffer(512)
    _check(lib.cc_graph_save(graph.ptr(), graph.n, os.fsencode(path), GRAPH_FORMATS[fmt], err, 512), err)


def sample_without_replacement(total, n, seed):
    """SampleWithoutReplacement: the objects LoadTags takes, in vertex order."""
    out = np.zeros(max(total, 1), dtype=np.uint32)
    lib.cc_sample_without_replacement(total, n, seed, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out[: (total if n == 0 or n >= total else n)]


def objective_bitwise(graph, labels):
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    return lib.cc_objective_bitwise(graph.ptr(), graph.n, labels.ctypes.data_as(_i32p))


def objective_direct(graph, labels):
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    return lib.cc_objective_direct(graph.ptr(), graph.n, labels.ctypes.data_as(_i32p))


def make_params(k=2, population=128, tournament=5, iterations=100, early_stop=6, perturbation=0.4, gww=0.1, seed=1,
                time_limit_sec=0.0, verbose=False, ls_kernel="auto", profile=False, rebuild="popcount"):
    p = Params()
    lib.cc_pbils_default_params(ctypes.byref(p))
    p.k, p.population, p.tournament, p.iterations, p.early_stop = k, population, tournament, iterations, early_stop
    p.perturbation, p.gww, p.seed, p.time_limit_sec = perturbation, gww, seed, time_limit_sec
    p.verbose = int(bool(verbose))
    p.ls_kernel = LS_KERNELS[ls_kernel] if isinstance(ls_kernel, str) else int(ls_kernel)
    p.profile = int(bool(profile))
    p.rebuild = REBUILDS[rebuild] if isinstance(rebuild, str) else int(rebuild)
    return p


def _result(stats, labels, n, trace=None):
    out = stats.as_dict()
    out["labels"] = labels[:n].copy()
    if trace is not None:
        out["trace"] = trace[: stats.iterations_done].copy()
    return out


def _check(rc, err):
    if rc != OK:
        raise CcError(rc, err.value.decode(errors="replace"))


def solve_pbils(graph, **kw):
    """SolveGpu: PBILS on the GPU; returns the PbilsResult fields plus the plan, device timings and the trace."""
    p = make_params(**kw)
    labels = np.zeros(max(graph.n, 1), dtype=np.int32)
    trace = np.zeros(max(p.iterations, 1), dtype=np.int64)
    stats = Stats()
    err = ctypes.create_string_buffer(512)
    rc = lib.cc_hip_solve_pbils(graph.ptr(), graph.n, graph.edges, ctypes.byref(p), labels.ctypes.data_as(_i32p),
                                ctypes.byref(stats), trace.ctypes.data_as(_i64p), err, 512)
    _check(rc, err)
    return _result(stats, labels, graph.n, trace)


def solve_neighborhood(graph, local_search, **kw):
    """SolveNeighborhoodGpu (local_search False) / SolveNeighborhoodWithManyLocalSearchesGpu (True)."""
    p = make_params(**kw)
    labels = np.zeros(max(graph.n, 1), dtype=np.int32)
    stats = Stats()
    err = ctypes.create_string_buffer(512)
    rc = lib.cc_hip_solve_neighborhood(graph.ptr(), graph.n, graph.edges, ctypes.byref(p), int(bool(local_search)),
                                       labels.ctypes.data_as(_i32p), ctypes.byref(stats), err, 512)
    _check(rc, err)
    return _result(stats, labels, graph.n)


class DeviceGraph:
    """A graph uploaded once to the GPU; solves on it start with the bitset already in HBM."""

    def __init__(self, graph):
        self.graph = graph  # the host bitset it was uploaded from
        self.n = graph.n
        self.edges = graph.edges
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib.cc_hip_graph_upload(graph.ptr(), graph.n, graph.edges, ctypes.byref(handle), err, 512)
        _check(rc, err)
        self.handle = handle

    def close(self):
        if self.handle:
            lib.cc_hip_graph_free(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def solve_pbils(self, **kw):
        p = make_params(**kw)
        labels = np.zeros(max(self.n, 1), dtype=np.int32)
        trace = np.zeros(max(p.iterations, 1), dtype=np.int64)
        stats = Stats()
        err = ctypes.create_string_buffer(512)
        rc = lib.cc_hip_graph_solve_pbils(self.handle, ctypes.byref(p), labels.ctypes.data_as(_i32p),
                                          ctypes.byref(stats), trace.ctypes.data_as(_i64p), err, 512)
        _check(rc, err)
        return _result(stats, labels, self.n, trace)

    def solve_pbils_static(self, stream=None, **kw):
        """solve_pbils without any device allocation: the run's state lives in the library's static scratch region
        (cc_hip_graph_solve_pbils_static); `stream` is a hipStream_t handle as an int, or None for the default
        stream. Calls that overlap on one GPU race."""
        p = make_params(**kw)
        labels = np.zeros(max(self.n, 1), dtype=np.int32)
        trace = np.zeros(max(p.iterations, 1), dtype=np.int64)
        stats = Stats()
        err = ctypes.create_string_buffer(512)
        rc = lib.cc_hip_graph_solve_pbils_static(self.handle, ctypes.byref(p), stream, labels.ctypes.data_as(_i32p),
                                                 ctypes.byref(stats), trace.ctypes.data_as(_i64p), err, 512)
        _check(rc, err)
        return _result(stats, labels, self.n, trace)

    def solve_neighborhood(self, local_search, **kw):
        p = make_params(**kw)
        labels = np.zeros(max(self.n, 1), dtype=np.int32)
        stats = Stats()
        err = ctypes.create_string_buffer(512)
        rc = lib.cc_hip_graph_solve_neighborhood(self.handle, ctypes.byref(p), int(bool(local_search)),
                                                 labels.ctypes.data_as(_i32p), ctypes.byref(stats), err, 512)
        _check(rc, err)
        return _result(stats, labels, self.n)


def solve_batch(device_graphs, params, use_graph=True):
    """Many independent PBILS runs advanced together (cc_hip_solve_batch). device_graphs: DeviceGraph per run;
    params: one dict of make_params() keywords per run. Returns (results, seconds); each result is the same dict
    solve_pbils returns, bit-identical to the run's own solve."""
    count = len(device_graphs)
    runs = (BatchRun * count)()
    keep = []
    for i, (dg, kw) in enumerate(zip(device_graphs, params)):
        p = make_params(**kw)
        labels = np.zeros(max(dg.n, 1), dtype=np.int32)
        trace = np.zeros(max(p.iterations, 1), dtype=np.int64)
        keep.append((labels, trace, dg.n))
        runs[i].graph = dg.handle
        runs[i].params = p
        runs[i].labels_out = labels.ctypes.data_as(_i32p)
        runs[i].trace_out = trace.ctypes.data_as(_i64p)
    seconds = ctypes.c_double(0.0)
    err = ctypes.create_string_buffer(512)
    rc = lib.cc_hip_solve_batch(runs, count, int(bool(use_graph)), ctypes.byref(seconds), err, 512)
    _check(rc, err)
    results = [_result(runs[i].stats, keep[i][0], keep[i][2], keep[i][1]) for i in range(count)]
    return results, seconds.value


def autotune(device_graph, tune_iterations=2, **kw):
    """Times every descent plan for this graph's shape on `tune_iterations` PBILS iterations (cc_hip_autotune) and
    records the fastest for ls_kernel = auto. Returns (candidates, chosen): one dict per plan, with its device time."""
    p = make_params(**kw)
    capacity = 64
    out = (TuneCandidate * capacity)()
    count = ctypes.c_int32(0)
    chosen = ctypes.c_int32(-1)
    err = ctypes.create_string_buffer(512)
    rc = lib.cc_hip_autotune(device_graph.handle, ctypes.byref(p), tune_iterations, out, capacity, ctypes.byref(count),
                             ctypes.byref(chosen), err, 512)
    _check(rc, err)
    cands = []
    for i in range(min(count.value, capacity)):
        c = out[i]
        cands.append({"descent": DESCENTS.get(c.descent, str(c.descent)), "threads": c.threads,
                      "vertices_per_thread": c.vertices_per_thread, "rows_in_lds": bool(c.rows_in_lds),
                      "blocks": c.blocks, "lds_bytes": c.lds_bytes, "accepted_moves": c.accepted_moves,
                      "search_ms": c.search_ms})
    return cands, chosen.value


def tune_clear():
    """Forget every tuned plan."""
    lib.cc_hip_tune_clear()


def tune_save(path):
    err = ctypes.create_string_buffer(512)
    _check(lib.cc_hip_tune_save(os.fsencode(path), err, 512), err)


def tune_load(path):
    """Adds the plans of a saved table; returns how many were read."""
    entries = ctypes.c_int32(0)
    err = ctypes.create_string_buffer(512)
    _check(lib.cc_hip_tune_load(os.fsencode(path), ctypes.byref(entries), err, 512), err)
    return entries.value


def footprint(n, **kw):
    """Device bytes a served solve of an n-vertex graph holds: bitset and degrees plus the run's state."""
    p = make_params(**kw)
    return lib.cc_hip_footprint(n, ctypes.byref(p))


class Server:
    """Serving with continuous batching (cc_hip_server_*): submit() returns at once; wait() returns the result with
    its latency (submit to result) and queueing time (submit to admission)."""

    def __init__(self, budget_bytes=0, lanes=2, max_active=4096, large_bytes=1 << 30):
        o = ServeOptions()
        lib.cc_serve_default_options(ctypes.byref(o))
        o.budget_bytes, o.lanes, o.max_active, o.large_bytes = budget_bytes, lanes, max_active, large_bytes
        handle = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        _check(lib.cc_hip_server_create(ctypes.byref(o), ctypes.byref(handle), err, 512), err)
        self.handle = handle
        self.shapes = {}

    def submit(self, graph, **kw):
        p = make_params(**kw)
        rid = ctypes.c_uint64()
        err = ctypes.create_string_buffer(512)
        _check(lib.cc_hip_server_submit(self.handle, graph.ptr(), graph.n, graph.edges, ctypes.byref(p),
                                        ctypes.byref(rid), err, 512), err)
        self.shapes[rid.value] = (graph.n, p.iterations)
        return rid.value

    def wait(self, rid, timeout=0.0):
        n, iterations = self.shapes[rid]
        labels = np.zeros(max(n, 1), dtype=np.int32)
        trace = np.zeros(max(iterations, 1), dtype=np.int64)
        stats = Stats()
        latency, queued = ctypes.c_double(), ctypes.c_double()
        err = ctypes.create_string_buffer(512)
        rc = lib.cc_hip_server_wait(self.handle, rid, timeout, labels.ctypes.data_as(_i32p),
                                    trace.ctypes.data_as(_i64p), ctypes.byref(stats), ctypes.byref(latency),
                                    ctypes.byref(queued), err, 512)
        _check(rc, err)
        del self.shapes[rid]
        out = _result(stats, labels, n, trace)
        out["latency"], out["queued"] = latency.value, queued.value
        return out

    def counters(self):
        vals = [ctypes.c_int64() for _ in range(5)]
        lib.cc_hip_server_counters(self.handle, *[ctypes.byref(v) for v in vals])
        return dict(zip(("admitted", "completed", "budget_left", "peak_active", "peak_bytes"),
                        (v.value for v in vals)))

    def close(self):
        if self.handle:
            lib.cc_hip_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def device_info():
    info = DeviceInfo()
    err = ctypes.create_string_buffer(512)
    rc = lib.cc_hip_device_info(ctypes.byref(info), err, 512)
    _check(rc, err)
    return {name: (getattr(info, name).decode() if isinstance(getattr(info, name), bytes) else getattr(info, name))
            for name, _ in info._fields_}
