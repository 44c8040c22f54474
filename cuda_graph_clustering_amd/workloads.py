"""Synthetic workloads of the benchmark configurations, defined once and shared by bench.py, the serving loop and
the tests.

The mixed batch: instance i of master seed m draws its shape from the splitmix64 stream
StreamSeed(m, kBatchShapeStream, i) (the RNG of include/cc_math.hpp): n = floor(lo * (hi/lo)^u) (log-uniform),
the generator (ER or SBM, 50/50), ER density U[0.05, 0.5) or SBM blocks in {2..8} (p_in 0.4, p_out 0.05), and
k in {2..8}; its graph seed is StreamSeed(m, kBatchGraphStream, i) and its search seed
StreamSeed(m, kBatchSearchStream, i).
"""
import math

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
BATCH_GRAPH_STREAM = 0xBA7C4
BATCH_SEARCH_STREAM = 0xBA7C5
BATCH_SHAPE_STREAM = 0xBA7C6


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def stream_seed(seed, stream, step):
    s = (seed ^ ((stream * 0xD1B54A32D192ED03) & MASK) ^ ((step * 0xA0761D6478BD642F) & MASK)) & MASK
    return (s + GOLDEN) & MASK


class Rng:
    def __init__(self, state):
        self.state = state

    def u32(self):
        self.state = (self.state + GOLDEN) & MASK
        return _mix(self.state) >> 32

    def below(self, bound):
        return (self.u32() * bound) >> 32

    def unit(self):
        return (self.u32() >> 8) / 16777216.0


def mixed_instance(master, i, n_lo=200, n_hi=8000):
    rng = Rng(stream_seed(master, BATCH_SHAPE_STREAM, i))
    n = int(math.floor(n_lo * (n_hi / n_lo) ** rng.unit()))
    n = max(n_lo, min(n_hi, n))
    if rng.below(2) == 0:
        gen = ("er", n, 0.05 + 0.45 * rng.unit())
    else:
        gen = ("sbm", n, 2 + rng.below(7), 0.4, 0.05)
    k = 2 + rng.below(7)
    return {"index": i, "gen": gen, "k": k, "graph_seed": stream_seed(master, BATCH_GRAPH_STREAM, i),
            "seed": stream_seed(master, BATCH_SEARCH_STREAM, i)}


def mixed_instances(count, master=7, n_lo=200, n_hi=8000):
    return [mixed_instance(master, i, n_lo, n_hi) for i in range(count)]


def build_graph(cga, inst):
    gen = inst["gen"]
    if gen[0] == "er":
        return cga.erdos_renyi(gen[1], gen[2], inst["graph_seed"])
    return cga.sbm(gen[1], gen[2], gen[3], gen[4], inst["graph_seed"])


def describe(inst):
    gen = inst["gen"]
    if gen[0] == "er":
        return f"ER(n={gen[1]}, p={gen[2]:.3f}) k={inst['k']}"
    return f"SBM(n={gen[1]}, blocks={gen[2]}) k={inst['k']}"
