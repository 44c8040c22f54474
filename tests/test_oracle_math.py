"""The oracle pinned by what the mathematics fixes: known RNG outputs, closed forms, brute force, invariants and
the known answers of the reference's own CPU tests (src/cc_verify.cpp)."""
import itertools

import numpy as np
import pytest

import oracle


def test_splitmix64_known_outputs(orc):
    # Published splitmix64 reference outputs for state 0 (Steele, Lea, Flood 2014; Vigna's splitmix64.c).
    assert orc.splitmix_outputs(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_rand_float_is_24bit_grid(orc):
    state = np.uint64(12345)
    import ctypes

    s = ctypes.c_uint64(12345)
    for _ in range(1000):
        x = orc.lib.orc_rand_float(ctypes.byref(s))
        assert 0.0 <= x < 1.0
        assert float(x) * 16777216.0 == int(float(x) * 16777216.0)
    del state


def test_rand_below_bounds(orc):
    import ctypes

    s = ctypes.c_uint64(7)
    for bound in (1, 2, 3, 5, 128, 1000, 4096):
        values = [orc.lib.orc_rand_below(ctypes.byref(s), bound) for _ in range(2000)]
        assert min(values) >= 0 and max(values) < bound
        if bound <= 5:
            assert set(values) == set(range(bound))
    assert orc.lib.orc_rand_below(ctypes.byref(s), 0) == 0


def test_perturbed_label_never_keeps_when_certain(orc):
    base = orc.stream_seed(3, 0xBEEF, 0)
    for k in (2, 3, 7):
        for idx in range(200):
            label = idx % k
            assert orc.perturbed_label(base, idx, 4, label, k, 1.0) != label
            assert orc.perturbed_label(base, idx, 4, label, k, 0.0) == label


def test_tournament_picks_minimum_of_its_draws(orc):
    f = np.arange(64, dtype=np.int64)[::-1].copy()  # slot 63 is best
    base = orc.stream_seed(9, 0x5E1EC7, 0)
    for slot in range(20):
        one = orc.tournament(base, slot, 3, f, 1)
        five = orc.tournament(base, slot, 3, f, 5)
        assert f[five] <= f[one]  # the first draw is among the five


def brute_g(adj, labels, k):
    n = adj.shape[0]
    g = np.zeros((k, n), dtype=np.int64)
    for v in range(n):
        for u in range(n):
            if adj[v, u]:
                g[labels[u], v] += 1
    return g


def pairwise(adj, labels):
    n = adj.shape[0]
    d = 0
    for i in range(n):
        for j in range(i + 1, n):
            d += int((labels[i] == labels[j]) != bool(adj[i, j]))
    return d


@pytest.mark.parametrize("n", [1, 2, 3, 7, 31, 32, 33, 63, 64, 65, 137])
def test_objective_closed_form_bitwise_and_rebuild_agree(orc, n):
    rng = np.random.default_rng(n)
    for p in (0.0, 0.15, 0.5, 1.0):
        g = orc.erdos_renyi(n, p, 100 + n, threads=1)
        adj = g.dense()
        assert (adj == adj.T).all() and not adj.diagonal().any()
        assert g.edges == adj.sum() // 2
        for k in (1, 2, 3, 5, 9):
            labels = rng.integers(0, k, n).astype(np.int32)
            truth = pairwise(adj, labels)
            assert orc.objective_direct(g, labels) == truth
            assert orc.objective_bitwise(g, labels) == truth
            f, gg, sizes = orc.rebuild(g, k, labels)
            assert f == truth
            assert (gg == brute_g(adj, labels, k)).all()
            assert (sizes == np.bincount(labels, minlength=k)).all()
            assert (gg.sum(axis=0) == adj.sum(axis=1)).all()  # sum_c G[v][c] = deg(v)


def test_local_search_reaches_local_optimum_and_counts_moves(orc):
    rng = np.random.default_rng(777)
    for n in (4, 23, 80, 210):
        for k in (2, 3, 4):
            g = orc.erdos_renyi(n, 0.35, int(rng.integers(1 << 30)), threads=1)
            labels = rng.integers(0, k, n).astype(np.int32)
            before = pairwise(g.dense(), labels)
            f, out, moves = orc.local_search(g, k, labels)
            assert f <= before
            assert f == orc.objective_direct(g, out)
            # every move strictly improves, so moves <= before - f (each at least 1)
            assert moves <= before - f
            # a local optimum: no single move improves
            fo, gg, sizes = orc.rebuild(g, k, out)
            for v in range(n):
                a = out[v]
                for c in range(k):
                    if c != a:
                        delta = (sizes[c] - sizes[a] + 1) - 2 * (gg[c, v] - gg[a, v])
                        assert delta >= 0


def brute_best_move(adj, labels, k):
    """The steepest move by enumeration in numpy: the smallest delta < 0, then the smallest v, then the smallest
    target (src/cc_pbils_cpu.cpp:99-118)."""
    n = len(labels)
    z = np.zeros((n, k), dtype=np.int64)
    z[np.arange(n), labels] = 1
    g = adj.astype(np.int64) @ z
    sizes = z.sum(axis=0)
    own = g[np.arange(n), labels]
    delta = (sizes[None, :] - sizes[labels][:, None] + 1) - 2 * (g - own[:, None])
    delta[np.arange(n), labels] = 1
    best = int(delta.min())
    if best >= 0:
        return 0, -1, -1
    v, to = np.argwhere(delta == best)[0]  # row-major: smallest v, then smallest target
    return best, int(v), int(to)


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 513, 700])
def test_blocked_best_move_scan_equals_plain_scan(orc, n):
    """The oracle's vectorisable two-step scan returns exactly the move of the reference's single in-order scan,
    ties included: random labels (many improving moves), sparse graphs with few distinct deltas (many ties, also
    across the 256-vertex blocks), and local optima (no move)."""
    rng = np.random.default_rng(n)
    for density in (0.0, 0.02, 0.5):
        g = orc.erdos_renyi(n, density, int(rng.integers(1 << 30)), threads=1)
        adj = g.dense()
        for k in (1, 2, 3, 5, 9):
            for trial in range(3):
                labels = rng.integers(0, k, n).astype(np.int32)
                if trial == 2:
                    _, labels, _ = orc.local_search(g, k, labels)
                want = brute_best_move(adj, labels, k)
                assert orc.best_move(g, k, labels, plain=True) == want
                assert orc.best_move(g, k, labels) == want


def test_known_optima(orc):
    # complete graphs reach 0 with one cluster (src/cc_verify.cpp:131-143)
    for n in (2, 6, 33):
        g = orc.erdos_renyi(n, 1.0, 1, threads=1)
        labels = np.random.default_rng(5).integers(0, 3, n).astype(np.int32)
        f, out, _ = orc.local_search(g, 3, labels)
        assert f == 0 and len(set(out.tolist())) == 1
    # edgeless 4-vertex graph: all together f = 6, LS at k = 2 reaches 2 (src/cc_verify.cpp:145-155)
    g = oracle.Graph.from_dense(np.zeros((4, 4), dtype=np.uint8))
    assert orc.rebuild(g, 2, np.zeros(4, np.int32))[0] == 6
    assert orc.local_search(g, 2, np.zeros(4, np.int32))[0] == 2
    # two disjoint triangles split for free (src/cc_verify.cpp:157-168)
    adj = np.zeros((6, 6), dtype=np.uint8)
    for a, b in [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5)]:
        adj[a, b] = adj[b, a] = 1
    g = oracle.Graph.from_dense(adj)
    assert orc.local_search(g, 2, np.array([0, 1, 0, 1, 0, 1], np.int32))[0] == 0
    # n = 1
    g = oracle.Graph.from_dense(np.zeros((1, 1), dtype=np.uint8))
    assert orc.rebuild(g, 2, np.zeros(1, np.int32))[0] == 0


@pytest.mark.parametrize("n,k", [(6, 2), (6, 3), (9, 2), (9, 4), (12, 3), (12, 4)])
def test_edgeless_graph_reaches_balanced_split(orc, n, k):
    # closed form of src/cc_verify.cpp:236-257: the balanced partition's sum of C(s, 2)
    g = oracle.Graph.from_dense(np.zeros((n, n), dtype=np.uint8))
    labels = np.random.default_rng(n * 31 + k).integers(0, k, n).astype(np.int32)
    f, _, _ = orc.local_search(g, k, labels)
    base, rem = divmod(n, k)
    want = sum((base + (c < rem)) * (base + (c < rem) - 1) // 2 for c in range(k))
    assert f == want


def test_rank_slots_and_winner_copies_known_answers(orc):
    # src/cc_verify.cpp:379-383
    assert orc.rank_slots([5, 3, 5, 1, 3]).tolist() == [3, 1, 4, 0, 2]
    assert orc.winner_copies(0.1, 128) == 12
    assert orc.winner_copies(0.5, 7) == 3


def test_solver_reports_truth_and_is_thread_independent(orc):
    g = orc.erdos_renyi(120, 0.5, 4242, threads=1)
    for k in (2, 3, 6):
        for gww in (0.0, 0.25):
            r1 = orc.solve_pbils(g, k=k, population=16, iterations=12, early_stop=4, seed=7, gww=gww, threads=1)
            r4 = orc.solve_pbils(g, k=k, population=16, iterations=12, early_stop=4, seed=7, gww=gww, threads=4)
            assert r1["objective"] == orc.objective_direct(g, r1["labels"])
            assert 1 <= r1["clusters_used"] <= k
            for key in ("objective", "iterations_done", "accepted_moves", "local_searches"):
                assert r1[key] == r4[key]
            assert (r1["labels"] == r4["labels"]).all()


def test_trace_is_a_prefix_property(orc):
    # with early stop off the first I records of a longer run equal the final record of a run of I iterations
    g = orc.erdos_renyi(150, 0.4, 99, threads=1)
    long = orc.solve_pbils(g, k=3, population=24, iterations=12, early_stop=12, gww=0.25, trace=True)
    for it in (1, 4, 9):
        short = orc.solve_pbils(g, k=3, population=24, iterations=it, early_stop=12, gww=0.25)
        assert short["objective"] == long["trace"][it - 1]


def test_generators_are_symmetric_loop_free_and_thread_independent(orc):
    for make in (lambda t: orc.erdos_renyi(300, 0.2, 5, threads=t),
                 lambda t: orc.sbm(300, 4, 0.4, 0.05, 5, threads=t),
                 lambda t: orc.chung_lu(300, 2.5, 12.0, 5, threads=t)):
        a, b = make(1), make(4)
        assert (a.bits == b.bits).all() and a.edges == b.edges
        adj = a.dense()
        assert (adj == adj.T).all() and not adj.diagonal().any()
        assert a.edges == adj.sum() // 2


def test_sbm_edge_densities_match_the_plan(orc):
    n, blocks = 2000, 5
    g = orc.sbm(n, blocks, 0.4, 0.05, 2)
    adj = g.dense().astype(np.int64)
    block = (np.arange(n) * blocks) // n
    same = block[:, None] == block[None, :]
    np.fill_diagonal(same, False)
    inside = adj[same].sum() / same.sum()
    outside = adj[~same & ~np.eye(n, dtype=bool)].sum() / (~same & ~np.eye(n, dtype=bool)).sum()
    assert abs(inside - 0.4) < 0.01 and abs(outside - 0.05) < 0.005


def test_chung_lu_mean_degree(orc):
    n = 4000
    g = orc.chung_lu(n, 2.5, 16.0, 4)
    mean = 2 * g.edges / n
    # capping p at 1 for the heaviest pairs removes a little mass; the mean stays within a few percent
    assert 0.85 * 16 < mean < 1.02 * 16
    deg = g.dense().sum(axis=1)
    assert deg[:10].mean() > 10 * deg[-1000:].mean()  # heavy head, light tail


@pytest.mark.parametrize("n,mean_degree", [(9, 2.0), (7, 4.5)])
def test_chung_lu_pairs_follow_their_probabilities(orc, n, mean_degree):
    # The O(n + m) sampler (geometric skips along each row) must make every pair an edge with its own probability
    # p_ij = min(1, w_i w_j / (d n)), independently. Over 12000 seeds each pair's frequency lies within 5 standard
    # errors of p_ij (pairs with p_ij = 1 always joined), and two pairs of one row, which share the row's skips, are
    # joined together as often as p_ij p_ik says. At mean degree 4.5 on 7 vertices the heaviest pairs are capped at 1.
    trials = 12000
    w = np.arange(1, n + 1, dtype=np.float64) ** (-1.0 / 1.5)
    w *= mean_degree * n / w.sum()
    p = np.minimum(1.0, np.outer(w, w) / (mean_degree * n))
    count = np.zeros((n, n))
    both = 0
    for seed in range(trials):
        adj = orc.chung_lu(n, 2.5, mean_degree, seed, threads=1).dense()
        count += adj
        both += int(adj[1, 2] and adj[1, 4])
    freq = count / trials
    for i in range(n):
        for j in range(i + 1, n):
            se = np.sqrt(p[i, j] * (1 - p[i, j]) / trials)
            assert abs(freq[i, j] - p[i, j]) <= 5 * se + 1e-12, (i, j, freq[i, j], p[i, j])
    pb = p[1, 2] * p[1, 4]
    assert abs(both / trials - pb) <= 5 * np.sqrt(pb * (1 - pb) / trials) + 1e-12
    if mean_degree > 4:
        assert p[0, 1] == 1.0 and freq[0, 1] == 1.0


def test_chung_lu_large_graph_is_cheap_and_has_its_mean_degree(orc):
    # O(n + m): the wide-path workload's graph (n = 100000, mean degree 64) in well under a second of one core's work
    # would be 5e9 pair draws in the pairwise order; its mean degree is the target less what capping removes.
    import time

    t0 = time.perf_counter()
    g = orc.chung_lu(100000, 2.5, 64.0, 4, threads=1)
    assert time.perf_counter() - t0 < 20.0
    assert 0.95 * 64 < 2 * g.edges / g.n < 64.0


def test_er_draw_is_the_reference_sequence(orc):
    # ER is defined by the reference's sequential draw order (src/cc_graph.cpp:168-177); the oracle evaluates
    # pair (i, j) as draw number i(i-1)/2 + j of the stream. Check that against a sequential walk of the stream.
    import ctypes

    n, p, seed = 70, 0.3, 11
    g = orc.erdos_renyi(n, p, seed, threads=1)
    s = ctypes.c_uint64(orc.stream_seed(seed, 0xE4D05F, 0))
    adj = np.zeros((n, n), dtype=np.uint8)
    for i in range(n):
        for j in range(i):
            if float(orc.lib.orc_rand_float(ctypes.byref(s))) < p:
                adj[i, j] = adj[j, i] = 1
    assert (g.dense() == adj).all()


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_neighborhood_exhaustive_small_graphs_vs_brute_force(orc, n):
    pairs = list(itertools.combinations(range(n), 2))
    for mask in range(1 << len(pairs)):
        adj = np.zeros((n, n), dtype=np.uint8)
        for b, (i, j) in enumerate(pairs):
            if (mask >> b) & 1:
                adj[i, j] = adj[j, i] = 1
        g = oracle.Graph.from_dense(adj)
        for ls in (False, True):
            got = orc.neighborhood(g, ls, threads=1)
            best, best_labels, moves = None, None, 0
            for v in range(n):
                labels = np.array([0 if (u == v or adj[v, u]) else 1 for u in range(n)], np.int32)
                if ls:
                    f, labels, m = orc.local_search(g, 2, labels)
                    moves += m
                else:
                    f = pairwise(adj, labels)
                if best is None or f < best:
                    best, best_labels = f, labels
            assert got["objective"] == best
            assert (got["labels"] == best_labels).all()
            assert got["accepted_moves"] == moves
            assert got["local_searches"] == (n if ls else 0)
