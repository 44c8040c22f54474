"""The cc_gpu command line and the instance formats around the path: a solve through the CLI matches the oracle;
the matrix text, baseline JSON and edge-list files it writes load back to the same graph and the same result; the
--out JSON keys and --labels-out format are the reference's (include/cc_cli.hpp:248-277, 372-376); invalid input
ends with `error:` and exit status 1 (include/cc_cli.hpp:329-335)."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cuda_graph_clustering_amd", "bin", "cc_gpu")

PBILS = ["--k", "3", "--pop", "16", "--iters", "5", "--early-stop", "5", "--seed", "7", "--gww", "0.1"]


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


def run(args, tmp_path, expect_ok=True):
    proc = subprocess.run([CLI] + [str(a) for a in args], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    if expect_ok:
        assert proc.returncode == 0, proc.stdout + proc.stderr
    return proc


def objective_line(stdout):
    for line in stdout.splitlines():
        if line.startswith("run 0"):
            return int(line.split("f=")[1].split()[0])
    raise AssertionError(stdout)


def test_pbils_through_cli_matches_oracle_and_formats_round_trip(cga, orc, cpu_threads, tmp_path):
    import oracle

    g = cga.erdos_renyi(300, 0.3, 5)
    want = orc.solve_pbils(oracle.Graph(g.bits, g.n, g.edges), threads=cpu_threads, k=3, population=16,
                           iterations=5, early_stop=5, seed=7, gww=0.1)
    run(["--n", 300, "--density", 0.3, "--graph-seed", 5, *PBILS, "--save-graph", "a.txt", "--save-graph-json",
         "a.json", "--save-edges", "a.edges", "--out", "r.json", "--labels-out", "l.txt"], tmp_path)
    result = json.loads((tmp_path / "r.json").read_text())
    assert result["objective function value"] == want["objective"]
    assert result["size"] == 300 and result["edges"] == g.edges and result["k"] == 3
    assert result["clustering vector"] == want["labels"].tolist()
    labels = np.array([int(x) for x in (tmp_path / "l.txt").read_text().split()])
    assert (labels == want["labels"]).all()
    for flag, path in (("--graph", "a.txt"), ("--graph-json", "a.json"), ("--edges", "a.edges")):
        out = run([flag, path, *PBILS], tmp_path).stdout
        assert objective_line(out) == want["objective"], flag


def test_neighborhood_algorithms_through_cli(cga, orc, tmp_path):
    import oracle

    g = cga.erdos_renyi(120, 0.25, 3)
    og = oracle.Graph(g.bits, g.n, g.edges)
    for algorithm, local_search in (("NeighborhoodCuda", 0), ("NeighborhoodWithManyLocalSearchesCuda", 1)):
        out = run(["--algorithm", algorithm, "--n", 120, "--density", 0.25, "--graph-seed", 3, "--k", 2, "--pop", 32],
                  tmp_path).stdout
        want = orc.neighborhood(og, local_search=local_search)
        assert objective_line(out) == want["objective"], algorithm


def test_autotune_flag_reports_plans(tmp_path):
    out = run(["--n", 400, "--density", 0.3, "--graph-seed", 2, *PBILS, "--autotune", 1], tmp_path).stdout
    tuned = [line for line in out.splitlines() if line.startswith("tune")]
    assert len(tuned) >= 3 and sum(line.startswith("tune *") for line in tuned) == 1


def test_invalid_input_is_an_error(tmp_path):
    proc = run(["--n", 50, "--k", 65], tmp_path, expect_ok=False)
    assert proc.returncode == 1 and "error:" in proc.stdout
    proc = run(["--graph", "missing.txt"], tmp_path, expect_ok=False)
    assert proc.returncode == 1 and "error:" in proc.stdout


def test_tags_graph_through_cli(cga, orc, cpu_threads, tmp_path):
    """--tags builds the tag-similarity graph of the reference's CLI (include/cc_cli.hpp:138-143, 239-241): --n
    sampled objects by --graph-seed, --similarity and --threshold; the graph saved by the CLI is the library's
    load_tags graph and the solve on it matches the oracle."""
    import random

    import oracle

    rng = random.Random(3)
    vocabulary = [f"t{i}" for i in range(15)]
    objects = {str(i): [rng.choice(vocabulary) for _ in range(rng.randint(0, 4))] for i in range(260)}
    (tmp_path / "tags.json").write_text(json.dumps(objects))
    g = cga.load_tags(tmp_path / "tags.json", "dice", 0.6, 200, 11)
    want = orc.solve_pbils(oracle.Graph(g.bits, g.n, g.edges), threads=cpu_threads, k=3, population=16,
                           iterations=5, early_stop=5, seed=7, gww=0.1)
    out = run(["--tags", "tags.json", "--similarity", "dice", "--threshold", 0.6, "--n", 200, "--graph-seed", 11,
               *PBILS, "--save-edges", "t.edges"], tmp_path).stdout
    assert objective_line(out) == want["objective"]
    saved = np.zeros((g.n, g.n), dtype=bool)
    for line in (tmp_path / "t.edges").read_text().splitlines()[1:]:
        i, j = map(int, line.split())
        saved[i, j] = saved[j, i] = True
    dense = np.unpackbits(g.bits.view(np.uint8), axis=1, bitorder="little")[:, : g.n].astype(bool)
    assert (saved == dense).all()
    proc = run(["--tags", "tags.json", "--similarity", "euclid", *PBILS], tmp_path, expect_ok=False)
    assert proc.returncode == 1 and "error:" in proc.stdout


REF_CLI = os.path.join(ROOT, "oracle", "_ref", "cc_gpu_ref")


def without_times(stdout):
    """stdout with the clock readings and the written paths masked: what must be identical between two runs."""
    import re

    return [re.sub(r"\d+\.\d+s", "T", line) for line in stdout.splitlines() if not line.startswith("written")]


@pytest.mark.skipif(not os.path.exists(REF_CLI), reason="the reference's command line is built only where its sources are")
def test_reference_command_line_runs_on_this_library(cga, orc, cpu_threads, tmp_path):
    """The reference's own cc_gpu (cuda/main_gpu.cu with include/cc_cli.hpp, compiled unchanged against include/
    and linked to libcc_hip.so) solves through this library: its result line and --out JSON match the oracle, its
    --labels-out is the oracle's labels, and its stdout equals this package's cc_gpu line for line, times aside."""
    import oracle

    g = cga.erdos_renyi(300, 0.3, 5)
    want = orc.solve_pbils(oracle.Graph(g.bits, g.n, g.edges), threads=cpu_threads, k=3, population=16,
                           iterations=5, early_stop=5, seed=7, gww=0.1)
    args = ["--n", 300, "--density", 0.3, "--graph-seed", 5, *PBILS, "--runs", 2]
    proc = subprocess.run([REF_CLI] + [str(a) for a in args] + ["--out", "r.json", "--labels-out", "l.txt"],
                          cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert objective_line(proc.stdout) == want["objective"]
    result = json.loads((tmp_path / "r.json").read_text())
    assert result["backend"] == "gpu" and result["runs"] == 2
    assert result["objective function value"] <= want["objective"]  # the better of seeds 7 and 8
    labels = np.array([int(x) for x in (tmp_path / "l.txt").read_text().split()])
    assert result["clustering vector"] == labels.tolist()
    assert cga.objective_direct(g, labels) == result["objective function value"]
    ours = run(args, tmp_path).stdout
    assert without_times(ours) == without_times(proc.stdout)
    small = cga.erdos_renyi(120, 0.25, 3)
    small_oracle = oracle.Graph(small.bits, small.n, small.edges)
    for algorithm, local_search in (("NeighborhoodCuda", 0), ("NeighborhoodWithManyLocalSearchesCuda", 1)):
        out = subprocess.run([REF_CLI, "--algorithm", algorithm, "--n", "120", "--density", "0.25", "--graph-seed",
                              "3", "--k", "2", "--pop", "32"], cwd=tmp_path, capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stdout
        assert objective_line(out.stdout) == orc.neighborhood(small_oracle, local_search=local_search)["objective"]
