"""The restatement (oracle/cc_oracle.c) against the reference CPU itself, built out of tree into oracle/_ref.
Skipped where the reference build is absent (e.g. on a host that only received the repository)."""
import numpy as np
import pytest

import oracle


def test_rng_primitives_match(orc, ref):
    rng = np.random.default_rng(1)
    for _ in range(300):
        seed, stream, step = (int(x) for x in rng.integers(0, 2**63, 3, dtype=np.uint64))
        assert orc.stream_seed(seed, stream, step) == ref.lib.ref_stream_seed(seed, stream, step)
        base = orc.stream_seed(seed, 0xA11CE, 0)
        idx = int(rng.integers(0, 2**40))
        k = int(rng.integers(1, 65))
        assert orc.initial_label(base, idx, k) == ref.lib.ref_initial_label(base, idx, k)
        label = int(rng.integers(0, k))
        it = int(rng.integers(0, 1000))
        for p in (0.0, 0.1, 0.4, 0.999, 1.0):
            assert orc.perturbed_label(base, idx, it, label, k, p) == ref.lib.ref_perturbed_label(
                base, idx, it, label, k, p)
    f = rng.integers(0, 50, 97).astype(np.int64)  # many ties
    import ctypes

    for slot in range(97):
        for t in (1, 2, 5, 9):
            got = orc.tournament(12345, slot, 17, f, t)
            want = ref.lib.ref_tournament(12345, slot, 17, f.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 97, t)
            assert got == want


@pytest.mark.parametrize("n", [0, 1, 2, 31, 32, 33, 63, 64, 65, 127, 129, 257, 600])
def test_erdos_renyi_bit_exact(orc, ref, n):
    for p, seed in ((0.0, 1), (0.33, 7), (0.5, 1), (1.0, 3), (0.1, 19 + n)):
        a = orc.erdos_renyi(n, p, seed)
        b = ref.erdos_renyi(n, p, seed)
        assert a.edges == b.edges
        assert (a.bits == b.bits).all()


def test_rebuild_and_local_search_match(orc, ref):
    rng = np.random.default_rng(3)
    for n in (1, 5, 33, 96, 200):
        for k in (1, 2, 3, 7):
            g = orc.sbm(n, 3, 0.5, 0.1, n + k, threads=1)
            labels = rng.integers(0, k, n).astype(np.int32)
            fo, go, so = orc.rebuild(g, k, labels)
            fr, gr, sr = ref.rebuild(g, k, labels)
            assert fo == fr and (go == gr).all() and (so == sr).all()
            fo, lo, mo = orc.local_search(g, k, labels)
            fr, lr, mr = ref.local_search(g, k, labels)
            assert (fo, mo) == (fr, mr) and (lo == lr).all()


def test_rank_slots_match(orc, ref):
    rng = np.random.default_rng(5)
    for p in (1, 2, 7, 64, 128, 300):
        f = rng.integers(0, 10, p).astype(np.int64)
        assert (orc.rank_slots(f) == ref.rank_slots(f)).all()
    for gww in (0.0, 0.05, 0.1, 0.25, 0.3, 0.5):
        for p in (1, 3, 7, 64, 128, 1000):
            assert orc.winner_copies(gww, p) == ref.lib.ref_winner_copies(gww, p)


CASES = [
    # n, graph, k, pop, iterations, early_stop, perturbation, gww, seed
    (60, "er", 2, 16, 8, 8, 0.4, 0.1, 1),
    (120, "er", 3, 16, 12, 4, 0.4, 0.25, 7),
    (97, "sbm", 5, 12, 10, 10, 0.3, 0.0, 42),
    (150, "er", 1, 8, 5, 5, 0.4, 0.1, 3),
    (70, "cl", 4, 9, 6, 2, 0.5, 0.5, 11),
    (200, "sbm", 8, 20, 6, 6, 0.0, 0.1, 2),
    (64, "er", 2, 1, 5, 5, 0.4, 0.0, 5),
    (33, "er", 64, 5, 4, 4, 0.4, 0.2, 6),
    (1, "er", 2, 4, 3, 3, 0.4, 0.1, 1),
]


def make_graph(orc, kind, n, seed):
    if kind == "er":
        return orc.erdos_renyi(n, 0.4, seed)
    if kind == "sbm":
        return orc.sbm(n, 4, 0.4, 0.05, seed)
    return orc.chung_lu(n, 2.5, 8.0, seed)


@pytest.mark.parametrize("case", CASES)
def test_pbils_trajectory_matches(orc, ref, case):
    n, kind, k, pop, iters, es, pert, gww, seed = case
    g = make_graph(orc, kind, n, seed)
    a = orc.solve_pbils(g, k=k, population=pop, iterations=iters, early_stop=es, perturbation=pert, gww=gww,
                        seed=seed, threads=4)
    b = ref.solve_pbils(g, k=k, population=pop, iterations=iters, early_stop=es, perturbation=pert, gww=gww,
                        seed=seed, threads=4)
    for key in ("objective", "iterations_done", "clusters_used", "local_searches", "accepted_moves"):
        assert a[key] == b[key], key
    assert (a["labels"] == b["labels"]).all()


@pytest.mark.parametrize("n", [31, 32, 33, 127])
def test_neighborhood_matches(orc, ref, n):
    for p in (0.0, 0.2, 0.5, 1.0):
        g = orc.erdos_renyi(n, p, 19 + n)
        for ls in (False, True):
            a = orc.neighborhood(g, ls)
            b = ref.neighborhood(g, ls)
            for key in ("objective", "local_searches", "accepted_moves", "clusters_used"):
                assert a[key] == b[key], key
            assert (a["labels"] == b["labels"]).all()
