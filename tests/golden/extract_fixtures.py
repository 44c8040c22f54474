"""Copy small subsets of the reference's published trajectories into tests/golden/ as data fixtures.

Run once in a container that has the reference tree (it reads nothing at test time):
    python tests/golden/extract_fixtures.py [/root/reference]

Sources (reference docs/data/*.csv) and the flags that produced them:
  gpu.csv       docs/gpu_sweep.sh:13-16  --density 0.33 --graph-seed 7 --iters 20 --early-stop 1000 --gww 0, row
                `run=j` used --seed 5+j; gpu_f == cpu_f on every row. Columns kept: series,n,k,pop,seed,f, plus
                the reference's measured seconds of the run (gpu_s: cc_gpu --ls-kernel global on its RTX 4070 Ti
                SUPER, cpu_s: cc_cpu --threads 16 on its Ryzen 7 3700X; docs/benchmarks.md:127-142) as
                published timings for the sweep comparison in bench.py.
  gww.csv       docs/gww_sweep.py:34-86  --density 0.33 --graph-seed 7|8 --seed 1 --runs 9 --iters I --early-stop I;
                the record after iteration `iter`; the seed column is the actual --seed. Random graphs only, iter <= 50.
  gww_stop.csv  docs/gww_sweep.py:88-92  defaults (100 iterations, early stop 6); final f and iterations done.
"""
import csv
import os
import sys

ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
DATA = os.path.join(ROOT, "docs", "data")
OUT = os.path.dirname(os.path.abspath(__file__))


def rows(name):
    with open(os.path.join(DATA, name), newline="") as f:
        return list(csv.DictReader(f))


def write(name, fields, data):
    with open(os.path.join(OUT, name), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        w.writerows(data)
    print(name, len(data), "rows")


def main():
    gpu = []
    for r in rows("gpu.csv"):
        assert r["gpu_f"] == r["cpu_f"], r
        gpu.append(dict(series=r["series"], n=r["n"], k=r["k"], pop=r["pop"], seed=5 + int(r["run"]), f=r["gpu_f"],
                        gpu_s=r["gpu_s"], cpu_s=r["cpu_s"]))
    write("gpu_trajectories.csv", ["series", "n", "k", "pop", "seed", "f", "gpu_s", "cpu_s"], gpu)

    curves = []
    for r in rows("gww.csv"):
        if r["set"] != "random" or int(r["iter"]) > 50:
            continue
        graph_seed = int(r["graph"].split("_")[1])
        curves.append(dict(n=r["n"], graph_seed=graph_seed, k=r["k"], pop=r["pop"], gww=r["gww"], seed=r["seed"],
                           iter=r["iter"], record=r["record"]))
    write("gww_records.csv", ["n", "graph_seed", "k", "pop", "gww", "seed", "iter", "record"], curves)

    stops = []
    for r in rows("gww_stop.csv"):
        if r["set"] != "random":
            continue
        graph_seed = int(r["graph"].split("_")[1])
        stops.append(dict(n=r["n"], graph_seed=graph_seed, k=r["k"], pop=r["pop"], gww=r["gww"], seed=r["seed"],
                          f=r["f"], iters=r["iters"]))
    write("gww_stop.csv", ["n", "graph_seed", "k", "pop", "gww", "seed", "f", "iters"], stops)


if __name__ == "__main__":
    main()
