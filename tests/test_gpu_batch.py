"""Batched solves: many runs of different n, k and parameters advanced together must each follow the oracle's
trajectory (the reference CPU algorithm) and reproduce their own single-instance solve bit for bit (record trace,
labels, objective, iterations, accepted moves), with and without the HIP-graph replay of iterations. The full
512-instance mixed stream is checked through properties on every instance and against the oracle on a sample."""
import numpy as np
import pytest

import oracle

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


def same(a, b, context=""):
    for key in KEYS:
        assert a[key] == b[key], f"{context} {key}: {a[key]} != {b[key]}"
    assert (a["labels"] == b["labels"]).all(), f"{context} labels"
    assert (a["trace"] == b["trace"]).all(), f"{context} record trace"


def oracle_graph(g):
    return oracle.Graph(g.bits, g.n, g.edges)


def oracle_solve(orc, g, kw, threads):
    params = {k: v for k, v in kw.items() if k != "ls_kernel"}
    return orc.solve_pbils(oracle_graph(g), threads=threads, trace=True, **params)


@pytest.mark.parametrize("use_graph", [True, False])
def test_mixed_batch_matches_oracle_and_solo_runs(cga, orc, cpu_threads, use_graph):
    insts, graphs, params = mixed(cga, 24, 2000, 16, 8)
    results, seconds = cga.solve_batch(graphs, params, use_graph=use_graph)
    assert seconds > 0
    for inst, g, kw, res in zip(insts, graphs, params, results):
        same(res, oracle_solve(orc, g.graph, kw, cpu_threads), f"instance {inst['index']} vs oracle")
        same(res, g.solve_pbils(**kw), f"instance {inst['index']} vs solo")


def test_batch_with_per_run_parameters_and_early_stops(cga, orc, cpu_threads):
    insts, graphs, params = mixed(cga, 12, 1500, 12, 15)
    variants = [dict(early_stop=2), dict(gww=0.0), dict(gww=0.5), dict(perturbation=0.0), dict(population=5),
                dict(iterations=1), dict(tournament=1), dict(early_stop=15, iterations=15), dict(population=33),
                dict(k=1), dict(ls_kernel="global"), dict(ls_kernel="shared")]
    for kw, v in zip(params, variants):
        kw.update(v)
    results, _ = cga.solve_batch(graphs, params)
    for g, kw, res in zip(graphs, params, results):
        same(res, oracle_solve(orc, g.graph, kw, cpu_threads), f"{kw} vs oracle")
        same(res, g.solve_pbils(**kw), f"{kw} vs solo")


def test_batch_of_one_equals_solve(cga):
    g = cga.sbm(3000, 5, 0.4, 0.05, 2)
    with cga.DeviceGraph(g) as dg:
        kw = dict(k=5, population=32, iterations=6, early_stop=6, gww=0.1, seed=42)
        (res,), _ = cga.solve_batch([dg], [kw])
        same(res, dg.solve_pbils(**kw))


def test_mixed_batch_spans_all_block_sizes(cga, orc, cpu_threads):
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
            same(res, oracle_solve(orc, g.graph, kw, cpu_threads), f"n={n} vs oracle")
            same(res, g.solve_pbils(**kw), f"n={n} vs solo")
            assert res["descent_name"] == "mixed"
            assert res["threads"] == (256 if n <= 2048 else 512 if n <= 4096 else 1024), n
    finally:
        for g in graphs:
            g.close()


def test_full_mixed_stream_of_512_instances(cga, orc, cpu_threads):
    """The batched workload at its full size: 512 instances (n log-uniform in [200, 8000], ER or SBM, k in 2..8,
    pop 64, 50 iterations, gww 0.1, seeds from master seed 7) in one batched solve with HIP-graph replay. Every
    instance: the reported f is the oracle's recount of its labels, the labels are a local optimum (the oracle's
    steepest-move scan finds no improving move), the record trace never increases and ends at f, and the counters
    are those of 50 full iterations. The 24 smallest instances and 8 more drawn with a fixed seed among those up to
    1200 vertices follow the oracle's trajectory bit for bit."""
    insts, graphs, params = mixed(cga, 512, 8000, 64, 50)
    try:
        results, _ = cga.solve_batch(graphs, params)
        for inst, g, kw, res in zip(insts, graphs, params, results):
            og = oracle_graph(g.graph)
            ctx = f"instance {inst['index']} n={g.graph.n} k={kw['k']}"
            assert res["iterations_done"] == 50 and res["local_searches"] == 50 * 64, ctx
            assert res["objective"] == orc.objective_bitwise(og, res["labels"]), ctx
            assert orc.best_move(og, kw["k"], res["labels"]) == (0, -1, -1), ctx + " not a local optimum"
            assert (np.diff(res["trace"]) <= 0).all() and res["trace"][-1] == res["objective"], ctx
            assert res["descent_name"] == "mixed", ctx
        by_size = sorted(range(len(insts)), key=lambda i: graphs[i].graph.n)
        rest = [i for i in by_size[24:] if graphs[i].graph.n <= 1200]
        picked = by_size[:24] + list(np.random.default_rng(7).choice(rest, size=min(8, len(rest)), replace=False))
        for i in picked:
            same(results[i], oracle_solve(orc, graphs[i].graph, params[i], cpu_threads),
                 f"instance {i} n={graphs[i].graph.n} vs oracle")
    finally:
        for g in graphs:
            g.close()
