"""The solve that allocates nothing (cc_hip_graph_solve_pbils_static): called through the C ABI, it must return
exactly what cc_hip_graph_solve_pbils returns (labels, record trace, objective, iterations, moves), on two calls
in a row with the second on a smaller input (the static scratch region keeps the first call's contents), and must
reject a run whose state exceeds the region before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("objective", "iterations_done", "clusters_used", "local_searches", "accepted_moves")


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


def solve_static(cga, dg, **kw):
    """One call of the entry point through the C ABI, on the default stream."""
    p = cga.make_params(**kw)
    labels = np.zeros(max(dg.n, 1), dtype=np.int32)
    trace = np.zeros(max(p.iterations, 1), dtype=np.int64)
    stats = cga.Stats()
    err = ctypes.create_string_buffer(512)
    rc = cga.lib.cc_hip_graph_solve_pbils_static(dg.handle, ctypes.byref(p), None,
                                                 labels.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                 ctypes.byref(stats),
                                                 trace.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), err, 512)
    return rc, err.value.decode(), stats, labels[: dg.n], trace[: stats.iterations_done]


def check(cga, dg, **kw):
    rc, msg, stats, labels, trace = solve_static(cga, dg, **kw)
    assert rc == cga.OK, msg
    want = dg.solve_pbils(**kw)
    for key in KEYS:
        assert getattr(stats, key) == want[key], key
    assert np.array_equal(labels, want["labels"])
    assert np.array_equal(trace, want["trace"])
    return stats


def test_static_solve_matches_the_allocating_solve_twice(cga):
    big = cga.sbm(3000, 5, 0.4, 0.05, 2)
    small = cga.erdos_renyi(700, 0.3, 9)
    with cga.DeviceGraph(big) as a, cga.DeviceGraph(small) as b:
        first = check(cga, a, k=5, population=32, iterations=6, early_stop=6, gww=0.1, seed=42)
        second = check(cga, b, k=3, population=16, iterations=5, early_stop=5, gww=0.25, seed=7)
        assert first.iterations_done == 6 and second.iterations_done == 5


def test_static_solve_rejects_a_run_beyond_the_region(cga):
    g = cga.erdos_renyi(6000, 0.01, 4)
    with cga.DeviceGraph(g) as dg:
        # two populations of 512 x 64 x 6000 int16 counts: far beyond 64 MiB
        rc, msg, _, _, _ = solve_static(cga, dg, k=64, population=512, iterations=1, early_stop=1)
        assert rc == cga.ENOMEM and "static scratch region" in msg
        rc, msg, _, _, _ = solve_static(cga, dg, k=3, population=8, iterations=1, early_stop=1, profile=True)
        assert rc == cga.EINVAL
        # the region is still usable afterwards
        check(cga, dg, k=3, population=8, iterations=2, early_stop=2, seed=3)
