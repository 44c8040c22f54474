"""The oracle against the reference-held golden trajectories (tests/golden/*.csv, copied as data from the
reference's docs/data). All use G(n, 0.33) with the reference ER generator, tournament 5, perturbation 0.4."""
import csv
import os

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name), newline="") as f:
        return list(csv.DictReader(f))


@pytest.fixture(scope="module")
def graphs(orc):
    cache = {}

    def get(n, seed):
        if (n, seed) not in cache:
            cache[(n, seed)] = orc.erdos_renyi(n, 0.33, seed)
        return cache[(n, seed)]

    return get


# gpu.csv: --graph-seed 7 --iters 20 --early-stop 1000 --gww 0 (docs/gpu_sweep.sh:13)
GPU_ROWS = [r for r in load("gpu_trajectories.csv") if (r["series"], r["n"], r["pop"]) in
            {("n128", "1000", "128"), ("k128", "2000", "128")} and (r["n"] == "1000" or r["k"] == "2")]


@pytest.mark.parametrize("row", GPU_ROWS, ids=lambda r: f"{r['series']}-n{r['n']}-k{r['k']}-s{r['seed']}")
def test_gpu_csv_final_objective(orc, graphs, cpu_threads, row):
    g = graphs(int(row["n"]), 7)
    r = orc.solve_pbils(g, k=int(row["k"]), population=int(row["pop"]), iterations=20, early_stop=1000, gww=0.0,
                        seed=int(row["seed"]), threads=cpu_threads)
    assert r["objective"] == int(row["f"])
    assert r["objective"] == orc.objective_bitwise(g, r["labels"])


RECORD_ROWS = load("gww_records.csv")
CURVES = sorted({(r["n"], r["graph_seed"], r["k"], r["pop"], r["gww"], r["seed"]) for r in RECORD_ROWS
                 if r["n"] == "1000" and r["seed"] in ("2", "5")})


@pytest.mark.parametrize("curve", CURVES, ids=lambda c: "-".join(c))
def test_gww_csv_record_curve(orc, graphs, cpu_threads, curve):
    n, gs, k, pop, gww, seed = curve
    points = {int(r["iter"]): int(r["record"]) for r in RECORD_ROWS
              if (r["n"], r["graph_seed"], r["k"], r["pop"], r["gww"], r["seed"]) == curve and int(r["iter"]) <= 20}
    g = graphs(int(n), int(gs))
    last = max(points)
    r = orc.solve_pbils(g, k=int(k), population=int(pop), iterations=last, early_stop=last, gww=float(gww),
                        seed=int(seed), threads=cpu_threads, trace=True)
    for it, record in points.items():
        assert r["trace"][it - 1] == record, f"iteration {it}"


STOP_ROWS = [r for r in load("gww_stop.csv") if r["n"] == "1000" and r["seed"] == "3" and
             (r["graph_seed"], r["gww"]) in {("7", "0.1"), ("8", "0")}]


@pytest.mark.parametrize("row", STOP_ROWS, ids=lambda r: f"n{r['n']}-g{r['graph_seed']}-gww{r['gww']}-s{r['seed']}")
def test_gww_stop_csv_early_stop(orc, graphs, cpu_threads, row):
    g = graphs(int(row["n"]), int(row["graph_seed"]))
    r = orc.solve_pbils(g, k=int(row["k"]), population=int(row["pop"]), gww=float(row["gww"]), seed=int(row["seed"]),
                        threads=cpu_threads)
    assert r["objective"] == int(row["f"])
    assert r["iterations_done"] == int(row["iters"])
