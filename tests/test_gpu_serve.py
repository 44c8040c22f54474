"""Serving with continuous batching: requests submitted while others run are admitted as memory frees up, run
together on the device, and each served result is bit-identical to the same solve run alone."""
import time

import pytest

pytestmark = pytest.mark.gpu

KEYS = ("objective", "iterations_done", "local_searches", "accepted_moves", "clusters_used")


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


def stream(cga, count, n_hi=1500, population=16, iterations=8):
    from cuda_graph_clustering_amd import workloads

    out = []
    for inst in workloads.mixed_instances(count, master=11, n_hi=n_hi):
        g = workloads.build_graph(cga, inst)
        out.append((g, dict(k=inst["k"], population=population, iterations=iterations, early_stop=iterations,
                            gww=0.1, seed=inst["seed"])))
    return out


def alone(cga, g, kw):
    with cga.DeviceGraph(g) as dg:
        return dg.solve_pbils(**kw)


def check(served, solo):
    for key in KEYS:
        assert served[key] == solo[key], key
    assert (served["labels"] == solo["labels"]).all()
    assert (served["trace"] == solo["trace"]).all()


def test_served_results_equal_solo_results(cga):
    reqs = stream(cga, 20)
    with cga.Server(lanes=2) as server:
        ids = []
        for i, (g, kw) in enumerate(reqs):
            ids.append(server.submit(g, **kw))
            if i % 5 == 4:
                time.sleep(0.01)  # later requests arrive while earlier ones are on the device
        results = [server.wait(rid, timeout=120) for rid in ids]
        counters = server.counters()
    assert counters["completed"] == len(reqs) and counters["admitted"] == len(reqs)
    for (g, kw), res in zip(reqs, results):
        check(res, alone(cga, g, kw))
        assert res["latency"] >= res["queued"] >= 0


def test_admission_by_footprint(cga):
    reqs = stream(cga, 8, n_hi=1200)
    sizes = [cga.footprint(g.n, **kw) for g, kw in reqs]
    budget = max(sizes) + min(sizes)
    with cga.Server(budget_bytes=budget, lanes=1) as server:
        ids = [server.submit(g, **kw) for g, kw in reqs]
        results = [server.wait(rid, timeout=120) for rid in ids]
        counters = server.counters()
    assert counters["peak_bytes"] <= budget  # admitted footprints never exceed the budget
    assert counters["peak_active"] < len(reqs)  # so not everything could run at once
    assert counters["budget_left"] == budget  # every byte returned
    for (g, kw), res in zip(reqs, results):
        check(res, alone(cga, g, kw))


def test_request_over_budget_is_rejected(cga):
    g = cga.erdos_renyi(2000, 0.3, 1)
    with cga.Server(budget_bytes=1 << 20, lanes=1) as server:
        with pytest.raises(cga.CcError) as e:
            server.submit(g, k=3, population=32, iterations=2)
        assert e.value.code == cga.ENOMEM


def test_large_lane_runs_beside_small_requests(cga):
    big = cga.chung_lu(40000, 2.5, 16.0, 9)
    small = stream(cga, 10, n_hi=1000)
    with cga.Server(lanes=2, large_bytes=64 << 20) as server:
        bid = server.submit(big, k=3, population=8, iterations=2, early_stop=2, seed=5)
        ids = [server.submit(g, **kw) for g, kw in small]
        results = [server.wait(rid, timeout=120) for rid in ids]
        bres = server.wait(bid, timeout=300)
    for (g, kw), res in zip(small, results):
        check(res, alone(cga, g, kw))
    check(bres, alone(cga, big, dict(k=3, population=8, iterations=2, early_stop=2, seed=5)))
