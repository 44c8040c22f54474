"""Serving with continuous batching: requests submitted while others run are admitted as memory frees up, run
together on the device, and each served result follows the oracle's trajectory (the reference CPU algorithm) and is
bit-identical to the same solve run alone."""
import time

import pytest

import oracle

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


def check(served, solo, context=""):
    for key in KEYS:
        assert served[key] == solo[key], f"{context} {key}: {served[key]} != {solo[key]}"
    assert (served["labels"] == solo["labels"]).all(), f"{context} labels"
    assert (served["trace"] == solo["trace"]).all(), f"{context} record trace"


def by_oracle(orc, g, kw, threads):
    return orc.solve_pbils(oracle.Graph(g.bits, g.n, g.edges), threads=threads, trace=True, **kw)


def test_served_results_equal_oracle_and_solo_results(cga, orc, cpu_threads):
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
    for i, ((g, kw), res) in enumerate(zip(reqs, results)):
        check(res, by_oracle(orc, g, kw, cpu_threads), f"request {i} vs oracle")
        check(res, alone(cga, g, kw), f"request {i} vs solo")
        assert res["latency"] >= res["queued"] >= 0


def test_admission_by_footprint(cga, orc, cpu_threads):
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
    for i, ((g, kw), res) in enumerate(zip(reqs, results)):
        check(res, by_oracle(orc, g, kw, cpu_threads), f"request {i} vs oracle")
        check(res, alone(cga, g, kw), f"request {i} vs solo")


def test_request_over_budget_is_rejected(cga):
    g = cga.erdos_renyi(2000, 0.3, 1)
    with cga.Server(budget_bytes=1 << 20, lanes=1) as server:
        with pytest.raises(cga.CcError) as e:
            server.submit(g, k=3, population=32, iterations=2)
        assert e.value.code == cga.ENOMEM


def test_large_lane_runs_beside_small_requests(cga, orc, cpu_threads):
    big = cga.chung_lu(40000, 2.5, 16.0, 9)
    big_kw = dict(k=3, population=8, iterations=2, early_stop=2, seed=5)
    small = stream(cga, 10, n_hi=1000)
    with cga.Server(lanes=2, large_bytes=64 << 20) as server:
        bid = server.submit(big, **big_kw)
        ids = [server.submit(g, **kw) for g, kw in small]
        results = [server.wait(rid, timeout=120) for rid in ids]
        bres = server.wait(bid, timeout=300)
    for i, ((g, kw), res) in enumerate(zip(small, results)):
        check(res, by_oracle(orc, g, kw, cpu_threads), f"small request {i} vs oracle")
        check(res, alone(cga, g, kw), f"small request {i} vs solo")
    check(bres, by_oracle(orc, big, big_kw, cpu_threads), "large request vs oracle")
    check(bres, alone(cga, big, big_kw), "large request vs solo")


def test_wide_path_workload_served_beside_a_request_stream(cga, orc, c5):
    """The wide-path workload's graph (Chung-Lu n=100000, degree exponent 2.5, mean degree 64, graph-seed 4; k=3)
    on the large lane, with requests of the mixed-batch distribution arriving while it runs: the large run follows
    the oracle's trajectory, and every small request equals the oracle's and its own solo solve."""
    small = stream(cga, 12, n_hi=1500)
    with cga.Server(lanes=2, large_bytes=256 << 20) as server:
        bid = server.submit(c5.graph, **c5.KW)
        ids = []
        for i, (g, kw) in enumerate(small):
            ids.append(server.submit(g, **kw))
            time.sleep(0.02)  # arrivals spread over the large run
        results = [server.wait(rid, timeout=120) for rid in ids]
        bres = server.wait(bid, timeout=300)
        counters = server.counters()
    assert counters["completed"] == len(small) + 1
    check(bres, c5.want(orc), "wide-path run vs oracle")
    for i, ((g, kw), res) in enumerate(zip(small, results)):
        check(res, by_oracle(orc, g, kw, c5.threads), f"request {i} vs oracle")
        check(res, alone(cga, g, kw), f"request {i} vs solo")
