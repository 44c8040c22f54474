"""Batched solves: many runs of different n, k and parameters advanced together must each reproduce their own
single-instance solve bit for bit (record trace, labels, objective, iterations, accepted moves), with and without
the HIP-graph replay of iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("objective", "iterations_done", "clusters_used", "local_searches", "accepted_moves")


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


def mixed(cga, count, n_hi, population, iterations):
    from cuda_graph_clustering_amd import workloads

    insts = workloads.mixed_instances(count, master=7, n_hi=n_hi)
    graphs = [cga.DeviceGraph(workloads.build_graph(cga, inst)) for inst in insts]
    params = [dict(k=inst["k"], population=population, iterations=iterations, early_stop=iterations, gww=0.1,
                   seed=inst["seed"]) for inst in insts]
    return insts, graphs, params


def same(a, b):
    for key in KEYS:
        assert a[key] == b[key], key
    assert (a["labels"] == b["labels"]).all()
    assert (a["trace"] == b["trace"]).all()


@pytest.mark.parametrize("use_graph", [True, False])
def test_mixed_batch_matches_solo_runs(cga, use_graph):
    insts, graphs, params = mixed(cga, 24, 2000, 16, 8)
    results, seconds = cga.solve_batch(graphs, params, use_graph=use_graph)
    assert seconds > 0
    for g, kw, res in zip(graphs, params, results):
        same(res, g.solve_pbils(**kw))


def test_batch_with_per_run_parameters_and_early_stops(cga):
    insts, graphs, params = mixed(cga, 12, 1500, 12, 15)
    variants = [dict(early_stop=2), dict(gww=0.0), dict(gww=0.5), dict(perturbation=0.0), dict(population=5),
                dict(iterations=1), dict(tournament=1), dict(early_stop=15, iterations=15), dict(population=33),
                dict(k=1), dict(ls_kernel="global"), dict(ls_kernel="shared")]
    for kw, v in zip(params, variants):
        kw.update(v)
    results, _ = cga.solve_batch(graphs, params)
    for g, kw, res in zip(graphs, params, results):
        same(res, g.solve_pbils(**kw))


def test_batch_of_one_equals_solve(cga):
    g = cga.sbm(3000, 5, 0.4, 0.05, 2)
    with cga.DeviceGraph(g) as dg:
        kw = dict(k=5, population=32, iterations=6, early_stop=6, gww=0.1, seed=42)
        (res,), _ = cga.solve_batch([dg], [kw])
        same(res, dg.solve_pbils(**kw))


def test_mixed_batch_spans_all_block_sizes(cga):
    """Register runs up to 2048 vertices take the 256-thread mixed descent (four workgroups per CU), up to 4096 the
    512-thread one, larger ones the 1024-thread one; all three in one batch, each run equal to its solo solve."""
    shapes = [(1500, 2), (4096, 5), (4097, 8), (4097, 2), (7000, 3), (2500, 7)]
    graphs, params = [], []
    for i, (n, k) in enumerate(shapes):
        graphs.append(cga.DeviceGraph(cga.sbm(n, 4, 0.3, 0.05, 30 + i)))
        params.append(dict(k=k, population=12, iterations=4, early_stop=4, gww=0.1, seed=100 + i))
    try:
        results, _ = cga.solve_batch(graphs, params)
        for (n, _), g, kw, res in zip(shapes, graphs, params, results):
            same(res, g.solve_pbils(**kw))
            assert res["descent_name"] == "mixed"
            assert res["threads"] == (256 if n <= 2048 else 512 if n <= 4096 else 1024), n
    finally:
        for g in graphs:
            g.close()
