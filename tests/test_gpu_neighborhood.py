"""NeighborhoodCuda / NeighborhoodWithManyLocalSearchesCuda on the GPU against the oracle's sequential enumeration,
with the reference GPU test's coverage (cuda/cc_verify_gpu.cu:79-112): every graph with n <= 4, word- and
block-boundary sizes, every descent kernel, a partial last batch, a batch of one and a batch larger than n."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


def compare(cga, orc, g, batch, kernel):
    import oracle

    og = oracle.Graph(g.bits, g.n, g.edges)
    for ls in (False, True):
        want = orc.neighborhood(og, ls, threads=1)
        got = cga.solve_neighborhood(g, ls, population=batch, ls_kernel=kernel)
        ctx = f"n={g.n} batch={batch} kernel={kernel} ls={ls}"
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
            compare(cga, orc, cga.from_bits(bits, n), 3, "auto")


@pytest.mark.parametrize("n", [31, 32, 33, 127, 129, 257])
def test_boundaries_and_kernels(cga, orc, n):
    for density in (0.0, 0.2, 0.5, 1.0):
        g = cga.erdos_renyi(n, density, 19 + n)
        for kernel in ("auto", "global", "shared", "wide"):
            compare(cga, orc, g, 17, kernel)
        compare(cga, orc, g, 1, "auto")
        compare(cga, orc, g, n + 3, "auto")


def test_invalid_input_rejected(cga):
    for kw, n in (({}, 0), ({"k": 3}, 1), ({"population": 0}, 1)):
        g = cga.erdos_renyi(n, 0.5, 1)
        for ls in (False, True):
            with pytest.raises(cga.CcError) as e:
                cga.solve_neighborhood(g, ls, **kw)
            assert e.value.code == cga.EINVAL
