"""NeighborhoodCuda / NeighborhoodWithManyLocalSearchesCuda on the GPU against the oracle's sequential enumeration,
with the reference GPU test's coverage (cuda/cc_verify_gpu.cu:79-112): every graph with n <= 4, word- and
block-boundary sizes, every descent kernel, a partial last batch, a batch of one and a batch larger than n. Both G
rebuilds (popcounts and the int8 matrix product on the matrix cores) must give the same results."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


def compare(cga, orc, g, batch, kernel, rebuild="popcount"):
    import oracle

    og = oracle.Graph(g.bits, g.n, g.edges)
    for ls in (False, True):
        want = orc.neighborhood(og, ls, threads=1)
        got = cga.solve_neighborhood(g, ls, population=batch, ls_kernel=kernel, rebuild=rebuild)
        ctx = f"n={g.n} batch={batch} kernel={kernel} ls={ls} rebuild={rebuild}"
        assert got["objective"] == want["objective"], ctx
        assert (got["labels"] == want["labels"]).all(), ctx + " labels and tie-breaking"
        assert got["objective"] == cga.objective_direct(g, got["labels"]), ctx
        assert got["accepted_moves"] == want["accepted_moves"], ctx + " moves"
        assert got["local_searches"] == (g.n if ls else 0), ctx
        assert got["clusters_used"] == want["clusters_used"], ctx


def test_all_graphs_up_to_four_vertices(cga, orc):
    for n in range(1, 5):
        pairs = list(itertools.combinations(range(n), 2))
        for mask in range(1 << len(pairs)):
            adj = np.zeros((n, n), dtype=np.uint8)
            for b, (i, j) in enumerate(pairs):
                if (mask >> b) & 1:
                    adj[i, j] = adj[j, i] = 1
            bits = np.packbits(np.pad(adj, ((0, 0), (0, 64 - n))), axis=1, bitorder="little").view(np.uint64)
            for rebuild in ("popcount", "mfma"):
                compare(cga, orc, cga.from_bits(bits, n), 3, "auto", rebuild)


@pytest.mark.parametrize("n", [31, 32, 33, 127, 129, 257])
def test_boundaries_and_kernels(cga, orc, n):
    for density in (0.0, 0.2, 0.5, 1.0):
        g = cga.erdos_renyi(n, density, 19 + n)
        for kernel in ("auto", "global", "shared", "wide"):
            compare(cga, orc, g, 17, kernel)
        compare(cga, orc, g, 1, "auto")
        compare(cga, orc, g, n + 3, "auto")
        for batch in (1, 17, n + 3):
            compare(cga, orc, g, batch, "auto", "mfma")


def test_matrix_core_rebuild_at_full_size(cga, orc, cpu_threads):
    """ER n=4096, p=0.1, graph-seed 3, batches of 256: the matrix-core rebuild gives the popcount results with and
    without local search, and the enumeration without local search matches the oracle."""
    import oracle

    g = cga.erdos_renyi(4096, 0.1, 3)
    with cga.DeviceGraph(g) as dg:
        for ls in (False, True):
            a = dg.solve_neighborhood(ls, k=2, population=256)
            b = dg.solve_neighborhood(ls, k=2, population=256, rebuild="mfma")
            for key in ("objective", "accepted_moves", "clusters_used", "local_searches"):
                assert a[key] == b[key], (ls, key)
            assert (a["labels"] == b["labels"]).all()
            if not ls:
                want = orc.neighborhood(oracle.Graph(g.bits, g.n, g.edges), False, threads=cpu_threads)
                assert b["objective"] == want["objective"] and (b["labels"] == want["labels"]).all()


def test_invalid_input_rejected(cga):
    for kw, n in (({}, 0), ({"k": 3}, 1), ({"population": 0}, 1), ({"rebuild": 7}, 5)):
        g = cga.erdos_renyi(n, 0.5, 1)
        for ls in (False, True):
            with pytest.raises(cga.CcError) as e:
                cga.solve_neighborhood(g, ls, **kw)
            assert e.value.code == cga.EINVAL


def test_c3_with_local_search_matches_oracle(cga, orc, cpu_threads):
    """NeighborhoodWithManyLocalSearchesCuda at its workload size (ER n=4096, p=0.1, graph-seed 3, batches of 256):
    the best f, its labels (ties to the smaller start vertex) and the moves summed over all 4096 descents equal the
    oracle's sequential enumeration, for every descent kernel and both G rebuilds. The reference's own check
    (cuda/cc_verify_gpu.cu:37-54, 102-112) stops at n = 257."""
    import oracle

    g = cga.erdos_renyi(4096, 0.1, 3)
    want = orc.neighborhood(oracle.Graph(g.bits, g.n, g.edges), True, threads=cpu_threads)
    legs = [("auto", "popcount"), ("global", "popcount"), ("shared", "popcount"), ("wide", "popcount"),
            ("auto", "mfma")]
    with cga.DeviceGraph(g) as dg:
        for kernel, rebuild in legs:
            got = dg.solve_neighborhood(True, k=2, population=256, ls_kernel=kernel, rebuild=rebuild)
            ctx = f"C3 kernel={kernel} rebuild={rebuild}"
            assert got["objective"] == want["objective"], ctx
            assert (got["labels"] == want["labels"]).all(), ctx + " labels"
            assert got["accepted_moves"] == want["accepted_moves"], ctx + " moves"
            assert got["local_searches"] == 4096, ctx
