import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950); runs the HIP path")


@pytest.fixture(scope="session")
def orc():
    import oracle

    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle

    if not oracle.reference_available():
        pytest.skip("reference CPU build (oracle/_ref) not present")
    return oracle.Reference()


@pytest.fixture(scope="session")
def cpu_threads():
    return min(16, os.cpu_count() or 1)


class _C5:
    """The wide-path workload graph (Chung-Lu n=100000, degree exponent 2.5, mean degree 64, graph-seed 4), built
    once per session, with the oracle's trajectory of its short check run (k=3, pop 4, 2 iterations, seed 42)
    computed on first use."""

    KW = dict(k=3, population=4, iterations=2, early_stop=2, gww=0.1, seed=42)

    def __init__(self, threads):
        import cuda_graph_clustering_amd as cga

        self.threads = threads
        self.graph = cga.chung_lu(100000, 2.5, 64.0, 4)
        self._want = None

    def oracle_graph(self):
        import oracle

        return oracle.Graph(self.graph.bits, self.graph.n, self.graph.edges)

    def want(self, orc):
        if self._want is None:
            self._want = orc.solve_pbils(self.oracle_graph(), threads=self.threads, trace=True, **self.KW)
        return self._want


@pytest.fixture(scope="session")
def c5(cpu_threads):
    return _C5(cpu_threads)
