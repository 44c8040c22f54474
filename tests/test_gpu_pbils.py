"""PBILS on the GPU against the oracle: the same graph and --seed must give the same trajectory — the same record
after every iteration, the same final labels, objective, iterations done and accepted-move count — for every
descent kernel (reference claim README.md:62-64; the check the reference only made in docs/plot_benchmarks.py:145).
All results are integers and compared exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("objective", "iterations_done", "clusters_used", "local_searches", "accepted_moves")


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


def as_oracle_graph(g):
    import oracle

    return oracle.Graph(g.bits, g.n, g.edges)


def check_same(got, want, context):
    for key in KEYS:
        assert got[key] == want[key], f"{context}: {key} {got[key]} != {want[key]}"
    assert (got["labels"] == want["labels"]).all(), f"{context}: labels"
    if "trace" in want:
        assert (got["trace"] == want["trace"]).all(), f"{context}: record trace"


def run_both(cga, orc, g, cpu_threads, **kw):
    got = cga.solve_pbils(g, **kw)
    params = {k: v for k, v in kw.items() if k != "ls_kernel"}
    want = orc.solve_pbils(as_oracle_graph(g), threads=cpu_threads, trace=True, **params)
    return got, want


CASES = [
    # n, generator, k, pop, iterations, early_stop, perturbation, gww, seed
    (1, "er", 2, 4, 3, 3, 0.4, 0.1, 1),
    (2, "er", 2, 3, 4, 4, 0.4, 0.0, 2),
    (31, "er", 2, 16, 8, 8, 0.4, 0.1, 3),
    (64, "er", 3, 16, 8, 8, 0.4, 0.25, 4),
    (65, "sbm", 5, 12, 10, 10, 0.3, 0.1, 5),
    (129, "er", 1, 8, 5, 5, 0.4, 0.1, 6),
    (200, "sbm", 8, 20, 6, 6, 0.0, 0.1, 7),
    (150, "cl", 4, 9, 6, 2, 0.5, 0.5, 8),
    (97, "er", 64, 5, 4, 4, 0.4, 0.2, 9),
    (300, "sbm", 2, 33, 12, 3, 0.4, 0.1, 10),
    (257, "er", 3, 1, 6, 6, 0.4, 0.0, 11),
    (700, "sbm", 12, 10, 5, 5, 0.4, 0.1, 12),
    (1500, "er", 16, 8, 4, 4, 0.4, 0.1, 13),
    (300, "cl", 9, 6, 4, 4, 0.4, 0.2, 14),
    (400, "er", 17, 8, 4, 4, 0.4, 0.1, 15),
    (900, "sbm", 20, 6, 4, 4, 0.4, 0.1, 16),
    (2000, "er", 24, 6, 3, 3, 0.4, 0.1, 17),
    (1100, "cl", 30, 5, 3, 3, 0.4, 0.2, 18),
    (300, "sbm", 28, 6, 4, 4, 0.4, 0.1, 19),
]


def make_graph(cga, kind, n, seed):
    if kind == "er":
        return cga.erdos_renyi(n, 0.4, seed)
    if kind == "sbm":
        return cga.sbm(n, 4, 0.4, 0.05, seed)
    return cga.chung_lu(n, 2.5, 8.0, seed)


@pytest.mark.parametrize("ls_kernel", ["auto", "global", "shared", "wide"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}-{c[1]}-k{c[2]}-p{c[3]}")
def test_trajectory_matches_oracle(cga, orc, cpu_threads, case, ls_kernel):
    n, kind, k, pop, iters, es, pert, gww, seed = case
    g = make_graph(cga, kind, n, seed)
    got, want = run_both(cga, orc, g, cpu_threads, k=k, population=pop, iterations=iters, early_stop=es,
                         perturbation=pert, gww=gww, seed=seed, ls_kernel=ls_kernel)
    check_same(got, want, f"n={n} k={k} kernel={ls_kernel}")
    assert got["objective"] == cga.objective_direct(g, got["labels"])


def test_c1_full_config(cga, orc, cpu_threads):
    """ER G(1000, 0.5) graph-seed 1, k = 2, pop 128, 100 iterations without early stop, gww 0.1, seed 42."""
    g = cga.erdos_renyi(1000, 0.5, 1)
    got, want = run_both(cga, orc, g, cpu_threads, k=2, population=128, iterations=100, early_stop=100, gww=0.1,
                         seed=42)
    check_same(got, want, "C1")
    assert got["descent_name"] == "k2-registers"


@pytest.fixture(scope="module")
def c2_graph(cga):
    """SBM n=10000 in 5 equal blocks, p_in 0.4, p_out 0.05, graph-seed 2."""
    return cga.sbm(10000, 5, 0.4, 0.05, 2)


@pytest.mark.parametrize("ls_kernel", ["auto", "shared"])
def test_c2_full_size_trajectory(cga, orc, cpu_threads, c2_graph, ls_kernel):
    """The C2 graph at full size (n=10000, k=5) bit-exact against the oracle; population 16 and 2 iterations keep
    the CPU side to seconds."""
    got, want = run_both(cga, orc, c2_graph, cpu_threads, k=5, population=16, iterations=2, early_stop=2, gww=0.1,
                         seed=42, ls_kernel=ls_kernel)
    check_same(got, want, f"C2 pop16 {ls_kernel}")


@pytest.fixture(scope="module")
def wide_graph(cga):
    """Chung-Lu n=40000 (> 32767: the wide path), degree exponent 2.5, mean degree 16."""
    return cga.chung_lu(40000, 2.5, 16.0, 9)


@pytest.mark.parametrize("k,ls_kernel", [(2, "auto"), (3, "auto"), (3, "global"), (4, "auto"), (13, "auto")])
def test_wide_path_full_size_trajectory(cga, orc, cpu_threads, wide_graph, k, ls_kernel):
    """n above the 16-bit range: int32 G, the packed k = 2 kernel and the k = 3 chunk-summary kernel (auto) or the
    generic global kernel, whose wide scan loads the counts eight clusters at a time (k = 13: a full and a partial
    group). (The streaming k = 3 kernel runs beside the others in the autotuner test.)"""
    got, want = run_both(cga, orc, wide_graph, cpu_threads, k=k, population=4, iterations=2, early_stop=2,
                         gww=0.25, seed=3, ls_kernel=ls_kernel)
    check_same(got, want, f"wide k={k} {ls_kernel}")
    expected = {(2, "auto"): "k2-wide", (3, "auto"): "k3-chunk", (3, "global"): "generic-global",
                (4, "auto"): "generic-global", (13, "auto"): "generic-global"}[(k, ls_kernel)]
    assert got["descent_name"] == expected


def test_c5_graph_trajectory(cga, orc, c5):
    """The wide-path workload's own graph (Chung-Lu n=100000, degree exponent 2.5, mean degree 64, graph-seed 4)
    at k = 3 through the default wide plan (the chunk-summary kernel): pop 4 and 2 iterations follow the oracle's
    trajectory bit for bit (labels, record trace, f, moves)."""
    got = cga.solve_pbils(c5.graph, **c5.KW)
    check_same(got, c5.want(orc), "C5 graph pop 4")
    assert got["descent_name"] == "k3-chunk"


def test_c5_full_config_properties(cga, orc, c5):
    """The wide-path workload at full size (pop 128, 20 iterations, early stop off): the reported f is the
    oracle's bitwise recount of the labels, the labels are a local optimum (the oracle's steepest-move scan finds
    no improving move), the record never increases and ends at f, and the counters are those of 20 full
    iterations."""
    kw = dict(k=3, population=128, iterations=20, early_stop=20, gww=0.1, seed=42)
    got = cga.solve_pbils(c5.graph, **kw)
    og = c5.oracle_graph()
    assert got["iterations_done"] == 20 and got["local_searches"] == 20 * 128
    assert got["objective"] == orc.objective_bitwise(og, got["labels"])
    assert orc.best_move(og, 3, got["labels"]) == (0, -1, -1)
    assert (np.diff(got["trace"]) <= 0).all() and got["trace"][-1] == got["objective"]


def local_optimum_violations(graph, labels, k):
    """Number of vertices with an improving single move, from G recomputed with numpy (no oracle code)."""
    adj = np.unpackbits(graph.bits.view(np.uint8), axis=1, bitorder="little")[:, : graph.n].astype(np.float32)
    z = np.zeros((graph.n, k), dtype=np.float32)
    z[np.arange(graph.n), labels] = 1.0
    g = (adj @ z).astype(np.int64)  # exact: counts < 2^24
    sizes = np.bincount(labels, minlength=k).astype(np.int64)
    own = g[np.arange(graph.n), labels]
    delta = (sizes[None, :] - sizes[labels][:, None] + 1) - 2 * (g - own[:, None])
    delta[np.arange(graph.n), labels] = 1
    return int((delta < 0).any(axis=1).sum()), g, sizes


def test_c2_full_config_properties(cga, c2_graph):
    """C2 at its full size (pop 128, 200 iterations): properties that hold for the exact trajectory at any size —
    the reported f is the recount of the labels, the labels are a local optimum, the record never increases, and
    a second run (a fresh allocation) reproduces the first bit for bit."""
    kw = dict(k=5, population=128, iterations=200, early_stop=200, gww=0.1, seed=42)
    a = cga.solve_pbils(c2_graph, **kw)
    b = cga.solve_pbils(c2_graph, **kw)
    assert a["iterations_done"] == 200 and a["local_searches"] == 200 * 128
    assert a["objective"] == cga.objective_bitwise(c2_graph, a["labels"])
    violations, g, sizes = local_optimum_violations(c2_graph, a["labels"], 5)
    assert violations == 0
    assert (np.diff(a["trace"]) <= 0).all()
    for key in KEYS:
        assert a[key] == b[key]
    assert (a["labels"] == b["labels"]).all() and (a["trace"] == b["trace"]).all()


@pytest.mark.parametrize("ls_kernel", ["auto", "global", "shared"])
def test_golden_gpu_csv_rows(cga, cpu_threads, ls_kernel):
    """docs/data/gpu.csv of the reference: n=1000, k=3, pop 128, 20 iterations, gww 0, graph-seed 7, seeds 6-8."""
    import csv
    import os

    rows = [r for r in csv.DictReader(open(os.path.join(os.path.dirname(__file__), "golden", "gpu_trajectories.csv")))
            if r["series"] == "n128" and r["n"] == "1000"]
    g = cga.erdos_renyi(1000, 0.33, 7)
    for r in rows:
        got = cga.solve_pbils(g, k=3, population=128, iterations=20, early_stop=1000, gww=0.0, seed=int(r["seed"]),
                              ls_kernel=ls_kernel)
        assert got["objective"] == int(r["f"])


def test_golden_gpu_csv_cluster_series(cga):
    """The reference's published k sweep (docs/benchmarks.md, ER G(2000, 0.33) graph-seed 7, pop 128, 20
    iterations, gww 0, seeds 6-8, k from 2 to 32): every final f of the default plan equals the reference's."""
    import csv
    import os

    rows = [r for r in csv.DictReader(open(os.path.join(os.path.dirname(__file__), "golden", "gpu_trajectories.csv")))
            if r["series"] == "k128"]
    assert len({r["k"] for r in rows}) >= 9
    g = cga.erdos_renyi(2000, 0.33, 7)
    for r in rows:
        got = cga.solve_pbils(g, k=int(r["k"]), population=128, iterations=20, early_stop=1000, gww=0.0,
                              seed=int(r["seed"]))
        assert got["objective"] == int(r["f"]), (r["k"], r["seed"])


def test_golden_gww_curves(cga):
    """docs/data/gww.csv of the reference: the record after iterations 1..50 at n=1000, k=5, pop 128."""
    import csv
    import os

    rows = list(csv.DictReader(open(os.path.join(os.path.dirname(__file__), "golden", "gww_records.csv"))))
    curves = {}
    for r in rows:
        if r["n"] != "1000":
            continue
        curves.setdefault((int(r["graph_seed"]), float(r["gww"]), int(r["seed"])), {})[int(r["iter"])] = int(r["record"])
    graphs = {s: cga.erdos_renyi(1000, 0.33, s) for s in (7, 8)}
    checked = 0
    for (gs, gww, seed), points in sorted(curves.items()):
        if seed not in (2, 3):
            continue
        last = max(points)
        got = cga.solve_pbils(graphs[gs], k=5, population=128, iterations=last, early_stop=last, gww=gww, seed=seed)
        for it, record in points.items():
            assert got["trace"][it - 1] == record, (gs, gww, seed, it)
            checked += 1
    assert checked > 50


def test_golden_early_stop(cga):
    """docs/data/gww_stop.csv of the reference: default settings with early stop 6: final f and iterations done."""
    import csv
    import os

    rows = [r for r in csv.DictReader(open(os.path.join(os.path.dirname(__file__), "golden", "gww_stop.csv")))
            if r["n"] in ("1000", "2000") and int(r["seed"]) <= 4]
    graphs = {}
    for r in rows:
        key = (int(r["n"]), int(r["graph_seed"]))
        if key not in graphs:
            graphs[key] = cga.erdos_renyi(key[0], 0.33, key[1])
        got = cga.solve_pbils(graphs[key], k=int(r["k"]), population=int(r["pop"]), gww=float(r["gww"]),
                              seed=int(r["seed"]))
        assert (got["objective"], got["iterations_done"]) == (int(r["f"]), int(r["iters"])), r


def test_invalid_input_rejected(cga):
    g = cga.erdos_renyi(10, 0.5, 1)
    for kw in ({"k": 0}, {"population": 0}, {"gww": 0.6}, {"gww": -0.1}, {"k": 65}):
        with pytest.raises(cga.CcError) as e:
            cga.solve_pbils(g, **kw)
        assert e.value.code == cga.EINVAL
    with pytest.raises(cga.CcError):
        cga.solve_pbils(cga.erdos_renyi(0, 0.5, 1))
    # the library stays usable after rejected calls
    assert cga.solve_pbils(g, k=3, iterations=1)["iterations_done"] == 1


def test_shared_kernel_rejects_state_larger_than_lds(cga):
    g = cga.erdos_renyi(3000, 0.1, 1)
    with pytest.raises(cga.CcError) as e:
        cga.solve_pbils(g, k=64, population=2, iterations=1, ls_kernel="shared")
    assert e.value.code == cga.EINVAL


def test_device_graph_repeated_solves_are_identical(cga):
    g = cga.sbm(500, 4, 0.4, 0.05, 3)
    with cga.DeviceGraph(g) as dg:
        a = dg.solve_pbils(k=4, population=32, iterations=10, early_stop=10, seed=5)
        b = dg.solve_pbils(k=4, population=32, iterations=10, early_stop=10, seed=5)
    c = cga.solve_pbils(g, k=4, population=32, iterations=10, early_stop=10, seed=5)
    for key in KEYS:
        assert a[key] == b[key] == c[key]
    assert (a["labels"] == b["labels"]).all() and (a["labels"] == c["labels"]).all()


def run_with_plan(cga, orc, g, cpu_threads, tmp_path, plan, **kw):
    """Solve with one tuned-plan entry forced (the table line: n k pop wide descent threads vpt rows_in_lds lds
    blocks) and with the oracle."""
    table = tmp_path / "plan.txt"
    table.write_text(plan + "\n")
    cga.tune_clear()
    try:
        assert cga.tune_load(table) == 1
        return run_both(cga, orc, g, cpu_threads, **kw)
    finally:
        cga.tune_clear()


@pytest.fixture(scope="module")
def dense_wide_graph(cga):
    """ER n=33001, p=0.12: a dense graph on the wide path (mean degree about 3960, every 64-vertex chunk touched by
    every move), n not a multiple of the streaming kernels' tiles (1024 vertices at k = 2, 512 at k = 3)."""
    return cga.erdos_renyi(33001, 0.12, 21)


# (k, threads, blocks, tiles in LDS): None = the default plan's (every whole tile that fits, at most three per
# wave); k = 2 at 1024 threads keeps 32 of the 33 tiles of n = 33001 (the partial last one is
# streamed), k = 3 keeps 48 of 65; 16 tiles of LDS at 256 threads are capped at 12 (three per wave); 5 tiles leave
# waves 5-15 without one; 0 streams every tile.
@pytest.mark.parametrize("k,threads,blocks,tiles", [(2, 1024, 0, None), (3, 1024, 0, None), (2, 256, 0, 0),
                                                    (3, 256, 2, 16), (2, 1024, 0, 5), (3, 1024, 0, 0)])
def test_dense_wide_path_streaming_kernels(cga, orc, cpu_threads, tmp_path, dense_wide_graph, k, threads, blocks,
                                           tiles):
    """The streaming wide kernels (per lane a group of 12 bytes: four 24-bit vertex fields at k = 2, two 48-bit words
    at k = 3; one row dword per group) on a dense graph, where auto takes them for k = 3 too, with the slot's first
    3 KiB tiles kept in LDS; each workgroup size, fewer resident workgroups than slots, and fewer or no tiles in LDS
    are forced through a tuned-plan entry. The trajectory must be the oracle's."""
    g = dense_wide_graph
    tile_bytes = 3072
    kw = dict(k=k, population=4, iterations=2, early_stop=2, seed=5, gww=0.25)
    if tiles is None:
        got, want = run_both(cga, orc, g, cpu_threads, **kw)
        tiles = min(g.n // (1024 if k == 2 else 512), (160 * 1024 - 2048) // tile_bytes,
                    threads // 64 * 3)
    else:
        descent = 3 if k == 2 else 4
        got, want = run_with_plan(cga, orc, g, cpu_threads, tmp_path,
                                  f"{g.n} {k} 4 1 {descent} {threads} 0 0 {tiles * tile_bytes} {blocks}", **kw)
    check_same(got, want, f"dense wide k={k} threads={threads} blocks={blocks} tiles in LDS={tiles}")
    assert (got["descent_name"], got["threads"]) == ("k2-wide" if k == 2 else "k3-wide", threads)
    assert got["lds_bytes"] == tiles * tile_bytes


def test_k3_streaming_words_follow_the_largest_degree(cga, orc, cpu_threads):
    """The k = 3 streaming kernel's 48-bit words hold the three counts (15 bits each) while every degree is below
    2^15, and otherwise the label and the two differences G_c - G_a (21 bits each). An ER graph (n = 40000, p = 0.05:
    dense enough for the streaming kernel) takes the counts; the same graph with vertex 0 joined to every other
    vertex (degree 39999) takes the differences. Both must follow the oracle."""
    base = cga.erdos_renyi(40000, 0.05, 13)
    n = base.n
    bits = base.bits.copy()
    row0 = np.full(bits.shape[1], np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    row0[0] &= ~np.uint64(1)
    if n % 64:
        row0[-1] &= np.uint64((1 << (n % 64)) - 1)
    bits[0] = row0
    bits[1:, 0] |= np.uint64(1)
    hub = cga.from_bits(bits, n)
    assert hub.edges == base.edges + (n - 1) - int(np.unpackbits(base.bits[0].view(np.uint8)).sum())
    kw = dict(k=3, population=4, iterations=2, early_stop=2, seed=9, gww=0.25)
    for name, g in (("max degree < 2^15", base), ("hub of degree n - 1", hub)):
        got, want = run_both(cga, orc, g, cpu_threads, **kw)
        check_same(got, want, f"k3 streaming, {name}")
        assert got["descent_name"] == "k3-wide", name


@pytest.mark.parametrize("n,p", [(40, 0.3), (64, 0.4), (65, 0.2), (257, 0.4), (1000, 0.05)])
def test_chunk_kernel_small_graphs(cga, orc, cpu_threads, tmp_path, n, p):
    """The k = 3 chunk-summary kernel below and at one chunk (n = 40, 64), one vertex into a second chunk (65) and on
    graphs too dense for auto to pick it, forced through a tuned wide-path plan (ls_kernel = wide)."""
    g = cga.erdos_renyi(n, p, n)
    lds = (n + 63) // 64 * 32
    got, want = run_with_plan(cga, orc, g, cpu_threads, tmp_path, f"{n} 3 6 1 7 1024 0 0 {lds} 0", k=3,
                              population=6, iterations=4, early_stop=4, seed=7, gww=0.1, ls_kernel="wide")
    check_same(got, want, f"chunk kernel n={n}")
    assert got["descent_name"] == "k3-chunk"


def test_wide_kernel_choice_follows_density(cga):
    """auto on the wide path at k = 3: the chunk-summary kernel on a sparse graph, the streaming kernel once twice the
    mean degree exceeds the number of 64-vertex chunks."""
    sparse = cga.solve_pbils(cga.chung_lu(40000, 2.5, 16.0, 9), k=3, population=2, iterations=1, early_stop=1)
    dense = cga.solve_pbils(cga.erdos_renyi(40000, 0.02, 9), k=3, population=2, iterations=1, early_stop=1)
    assert (sparse["descent_name"], dense["descent_name"]) == ("k3-chunk", "k3-wide")


@pytest.mark.parametrize("threads", [256, 1024])
def test_wide_chunk_kernel_with_several_chunks_per_thread(cga, orc, cpu_threads, tmp_path, threads):
    """The k = 3 chunk-summary kernel when a thread owns several 64-vertex chunks (n = 70000: 1094 chunks): across
    chunks only the gain may decide, the earlier chunk winning a tie, or vertex order breaks. Both workgroup sizes
    are forced through a tuned-plan entry and must follow the oracle."""
    g = cga.chung_lu(70000, 2.5, 8.0, 3)
    kw = dict(k=3, population=4, iterations=2, early_stop=2, seed=3, gww=0.1)
    lds = (70000 + 63) // 64 * 32
    table = tmp_path / "plan.txt"
    table.write_text(f"70000 3 4 1 7 {threads} 0 0 {lds} 0\n")
    cga.tune_clear()
    try:
        assert cga.tune_load(table) == 1
        got, want = run_both(cga, orc, g, cpu_threads, **kw)
    finally:
        cga.tune_clear()
    check_same(got, want, f"chunk kernel {threads} threads")
    assert (got["descent_name"], got["threads"]) == ("k3-chunk", threads)


def test_register_kernel_descent_longer_than_padding_reset(cga, orc, cpu_threads):
    """Descents of more than 4096 moves in the register kernel with padding slots (n = 12000 on 12 x 1024 slots):
    padding lanes take unmasked row words and are reset every 4096 moves, so a first descent from random labels
    must still follow the oracle exactly."""
    g = cga.sbm(12000, 3, 0.4, 0.05, 21)
    kw = dict(k=3, population=2, iterations=1, early_stop=1, seed=4, gww=0.1)
    got, want = run_both(cga, orc, g, cpu_threads, **kw)
    check_same(got, want, "register kernel, long descent")
    assert got["descent_name"] == "registers"
    assert got["accepted_moves"] > 2 * 4096, "the descents must pass the padding reset"


def test_register_kernel_sixteen_slots_per_thread(cga, orc, cpu_threads):
    """n = 16000 at k = 3 takes the register kernel with 16 vertex slots per thread (the widest layout: slot index
    in 4 bits, vertex in 15); it must follow the oracle exactly."""
    g = cga.sbm(16000, 3, 0.3, 0.05, 23)
    kw = dict(k=3, population=2, iterations=2, early_stop=2, seed=6, gww=0.1)
    got, want = run_both(cga, orc, g, cpu_threads, **kw)
    check_same(got, want, "register kernel, 16 slots")
    assert (got["descent_name"], got["threads"]) == ("registers", 1024)


@pytest.mark.parametrize("n,k,threads", [(400, 17, 256), (900, 20, 256), (2000, 24, 1024), (600, 32, 1024),
                                         (300, 28, 256), (2048, 31, 1024)])
def test_register_kernel_above_sixteen_clusters(cga, orc, cpu_threads, n, k, threads):
    """k in [17, 32] runs the register kernel with five-bit label fields and k rounded up to a multiple of four
    registers per slot (the extra clusters stay empty); it must follow the oracle exactly, and the empty clusters
    must never appear among the labels."""
    g = cga.erdos_renyi(n, 0.33, k)
    kw = dict(k=k, population=4, iterations=3, early_stop=3, seed=k + 1, gww=0.1)
    got, want = run_both(cga, orc, g, cpu_threads, **kw)
    check_same(got, want, f"register kernel, n={n} k={k}")
    assert (got["descent_name"], got["threads"]) == ("registers", threads)
    assert got["labels"].max() < k


@pytest.mark.parametrize("n,vpt", [(16385, 20), (24000, 24), (32767, 32)])
def test_packed_k3_kernel(cga, orc, cpu_threads, n, vpt):
    """k = 3 beyond the register kernel's 16384 vertices runs the packed kernel (three 16-bit counts and the label
    in two registers per vertex, consecutive vertices per thread) up to the narrow path's last n; it must follow the
    oracle exactly, with the smallest slot count that covers n (16385: almost a whole slot of padding)."""
    g = cga.sbm(n, 3, 0.3, 0.05, 29)
    kw = dict(k=3, population=2, iterations=2, early_stop=2, seed=7, gww=0.1)
    got, want = run_both(cga, orc, g, cpu_threads, **kw)
    check_same(got, want, f"packed k = 3, n={n}")
    assert (got["descent_name"], got["threads"], got["vertices_per_thread"]) == ("k3-packed", 1024, vpt)


@pytest.mark.parametrize("n,k,vpt", [(10000, 5, 20), (12000, 5, 24), (9000, 7, 20), (6000, 8, 12), (16000, 3, 32),
                                     (14000, 4, 28), (4000, 16, 8)])
def test_register_kernel_with_512_threads(cga, orc, cpu_threads, tmp_path, n, k, vpt):
    """The 512-thread register descent (eight waves, up to 32 vertex slots per thread; from five slot bits on, the
    keys carry q from bit 13 instead of 12) forced through a tuned-plan entry: it must follow the oracle exactly."""
    g = cga.sbm(n, 4, 0.3, 0.05, 31 + k)
    kw = dict(k=k, population=4, iterations=2, early_stop=2, seed=k, gww=0.1)
    table = tmp_path / "plan.txt"
    table.write_text(f"{n} {k} 4 0 5 512 {vpt} 0 0 0\n")
    cga.tune_clear()
    try:
        assert cga.tune_load(table) == 1
        got, want = run_both(cga, orc, g, cpu_threads, **kw)
    finally:
        cga.tune_clear()
    check_same(got, want, f"512-thread register kernel n={n} k={k} vpt={vpt}")
    assert (got["descent_name"], got["threads"], got["vertices_per_thread"]) == ("registers", 512, vpt)
