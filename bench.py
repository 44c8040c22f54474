"""PBILS throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2] [--no-cpu]

A step is one PBILS iteration (select + steepest descent of every individual + record + GWW + perturb + G
rebuild) of the planted-partition workload: SBM n=10000 in 5 equal hidden clusters, p_in 0.4, p_out 0.05,
graph-seed 2; k=5, pop 128, gww 0.1, seed 42, early stop off. The graph is uploaded to HBM before the timed region;
the timed region is one solve of K iterations (its allocation and initial rebuild included), with the iterations
replayed from the library's captured HIP graph as any caller's solve is. Prints one JSON line.

value       PBILS iterations per second (whole job, 1 GPU)
roofline    the descent kernel (95 % of the device time), its launch time taken with HIP events on the library's
            stream in a second solve of the same trajectory: VALU issue from the instruction counts per move that
            rocprofv3 measured for this plan (MEASURED; profiles/), with global-memory bytes and the algorithmic
            lane operations beside it
cpu_baseline  the reference CPU solver (built out of tree into oracle/_ref) on this host's cores: the median time
            per iteration over three pairs of solves after a warm-up; bit_exact reports that the GPU reproduces its
            result on those solves.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Hardware counters of the descent kernels per accepted move, from rocprofv3 --pmc runs of this script (the commands,
# raw per-dispatch values and their means are in the profiles/ file named beside each entry). Keyed by (config,
# descent, threads per workgroup, vertices per thread): another plan issues other counts and is reported without
# them. valu: wave64 VALU instructions (SQ_INSTS_VALU); lds: LDS instructions (SQ_INSTS_LDS); lds_conflict: cycles
# of LDS bank conflicts (SQ_LDS_BANK_CONFLICT); fabric: bytes moved beyond L2, FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
# (KiB -> bytes; the factors from a 1 GiB device copy, profiles/copy_calibration_1gib.txt).
MEASURED = {
    ("c2", "registers", 1024, 10): dict(valu=2064.2, lds=224.0, lds_conflict=0.0, fabric=877.3,
                                        profile="profiles/descent_registers_sbm_n10000_k5.txt"),
    ("c5", "k3-chunk", 1024, 0): dict(valu=7466.4, lds=202.9, lds_conflict=32.2, fabric=29893.0,
                                      profile="profiles/descent_k3chunk_chunglu_n100000_k3.txt"),
}

METRIC = "PBILS iter/s + LS moves/s (1 GPU); batched graphs/s & p99 vs plain, bit-exact"
HBM_PEAK_GBS = 8000.0
# MI355X: 256 CUs at 2.4 GHz; a CU issues at most one wave64 VALU instruction per cycle (four SIMDs, four cycles
# each), i.e. 64 int32 lane operations per CU-cycle.
CUS = 256
CLOCK_HZ = 2.4e9
VALU_PEAK_LANE_OPS = CUS * 64 * CLOCK_HZ

CONFIGS = {
    "c1": dict(workload="ER G(n=1000, p=0.5) graph-seed 1; k=2, pop=128, gww=0.1, seed=42, early stop off",
               gen=("er", 1000, 0.5, 1), k=2, population=128, gww=0.1, seed=42, iterations=100, cpu_iterations=20),
    "c2": dict(workload="SBM n=10000, 5 equal blocks, p_in=0.4, p_out=0.05, graph-seed 2; k=5, pop=128, gww=0.1, "
                        "seed=42, early stop off",
               gen=("sbm", 10000, 5, 0.4, 0.05, 2), k=5, population=128, gww=0.1, seed=42, iterations=200,
               cpu_iterations=1),
    "c5": dict(workload="Chung-Lu n=100000, degree exponent 2.5, mean degree 64, graph-seed 4; k=3, pop=128, "
                        "gww=0.1, seed=42, early stop off (wide path, int32 G, packed k=3 descent)",
               gen=("cl", 100000, 2.5, 64.0, 4), k=3, population=128, gww=0.1, seed=42, iterations=20,
               cpu_iterations=1),
}


def note(msg):
    """A progress line on stderr as each phase of the bench ends (stdout carries only the JSON result)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_graph(cga, gen):
    if gen[0] == "er":
        return cga.erdos_renyi(gen[1], gen[2], gen[3])
    if gen[0] == "cl":
        return cga.chung_lu(gen[1], gen[2], gen[3], gen[4])
    return cga.sbm(gen[1], gen[2], gen[3], gen[4], gen[5])


REGISTER_DESCENTS = ("registers", "k2-registers", "k3-packed", "generic-lds")


def streamed_state_bytes(n, k, lds_bytes):
    """State bytes a streaming wide descent passes over per move: one packed word per vertex (k = 2: 3 bytes; k = 3:
    6 bytes) less the whole 3 KiB tiles it keeps in LDS (1024 vertices at k = 2, 512 at k = 3; at most three per
    wave of 1024 threads)."""
    if k == 2:
        return 3 * n - 3072 * min(lds_bytes // 3072, n // 1024, 48)
    return 6 * n - 3072 * min(lds_bytes // 3072, n // 512, 48)


def descent_bytes_per_move(n, k, descent, mean_degree, lds_bytes=0):
    """Algorithmic global-memory bytes of one accepted move: the adjacency row of the moved vertex, plus what the
    kernel keeps in memory rather than on chip. The register kernels (and the generic kernel with G in LDS) hold
    the individual's state on chip, so a move reads only the row. The generic global kernel scans G and the labels
    (k n |G| + n) in memory every move. The wide streaming kernels stream one packed word per vertex except the first
    tiles they keep in LDS (streamed_state_bytes; `lds_bytes` is the plan's dynamic LDS), and write back the
    neighbours' words; the chunk-summary kernel re-reads only the 64-vertex chunks that hold the moved vertex and its
    neighbours (at most mean degree + 1 chunks of 512 bytes, written back)."""
    words = max(1, (n + 63) // 64)
    row = 8 * words
    if descent in REGISTER_DESCENTS:
        return row
    if descent in ("k2-wide", "k3-wide"):
        state = streamed_state_bytes(n, k, lds_bytes)
        return state + row + state / n * mean_degree
    if descent == "k3-chunk":
        return row + 2 * 512 * min(mean_degree + 1, words)
    gbytes = 2 if n <= 32767 else 4
    return k * n * gbytes + n + row


def descent_state_bytes(n, k):
    """Bytes of one individual's state that a descent reads at its start and writes at its end: G (k n |G|) and
    the labels (n), each way."""
    gbytes = 2 if n <= 32767 else 4
    return 2 * (k * n * gbytes + n)


def roofline(config, cfg, graph, prof, mean_launch_s, moves_per_launch):
    """The descent kernel (95 % of the device time) against the roof that binds it. The register and chunk kernels
    keep an individual's state on chip and issue VALU instructions back to back, so with counters measured for the
    plan the bound is VALU issue: SQ_INSTS_VALU per move x moves per launch x 64 lanes over the launch time, against
    256 CUs x 64 lanes x 2.4 GHz; one workgroup holds one individual and takes one CU, so with fewer individuals than
    CUs only cus_used CUs can issue at all (frac_of_used_cus). Global-memory bytes (algorithmic, and the measured
    fabric traffic) and the survey's algorithmic lane operations per move, 7 (k - 1) n, are given beside it. Without
    counters for the plan the line falls back to the algorithmic bytes against HBM."""
    n, k, pop = graph.n, cfg["k"], cfg["population"]
    descent = prof["descent_name"]
    key = (config, descent, prof["threads"], prof["vertices_per_thread"])
    meas = MEASURED.get(key)
    bytes_per_launch = (moves_per_launch * descent_bytes_per_move(n, k, descent, 2.0 * graph.edges / n) +
                        pop * descent_state_bytes(n, k))
    gbs = bytes_per_launch / mean_launch_s / 1e9 if mean_launch_s > 0 else 0.0
    hbm = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
           "algorithmic_bytes_per_launch": round(bytes_per_launch)}
    cus_used = min(pop, CUS)
    out = {"kernel": f"descent ({descent}, {prof['threads']} threads, {prof['vertices_per_thread']} vertices per "
                     f"thread)",
           "mean_launch_ms": round(1e3 * mean_launch_s, 4), "moves_per_launch": round(moves_per_launch),
           "cus_used": cus_used}
    if meas is None or mean_launch_s <= 0:
        out.update(bound="hbm", achieved=hbm["achieved"], peak=HBM_PEAK_GBS, unit="GB/s", frac=hbm["frac"],
                   traffic=None, bytes_per_launch=hbm["algorithmic_bytes_per_launch"],
                   note="no counters measured for this plan")
        return out
    lane_ops = meas["valu"] * moves_per_launch * 64 / mean_launch_s
    frac = lane_ops / VALU_PEAK_LANE_OPS
    algo_ops = 7 * (k - 1) * n
    algo = moves_per_launch * algo_ops / mean_launch_s
    out.update(bound="valu", achieved=round(lane_ops / 1e12, 3), peak=round(VALU_PEAK_LANE_OPS / 1e12, 2),
               unit="Tlane-op/s", frac=round(frac, 4), frac_of_used_cus=round(frac * CUS / cus_used, 4),
               traffic=round(meas["fabric"] * moves_per_launch),
               valu_instructions_per_move=meas["valu"], lds_instructions_per_move=meas["lds"],
               lds_bank_conflict_cycles_per_move=meas["lds_conflict"],
               hbm=dict(hbm, fabric_gbs=round(meas["fabric"] * moves_per_launch / mean_launch_s / 1e9, 1)),
               algorithmic={"lane_ops_per_move": algo_ops, "achieved": round(algo / 1e12, 3),
                            "frac": round(algo / VALU_PEAK_LANE_OPS, 4)},
               counters=meas["profile"])
    return out


def reference_solver():
    import oracle

    if not oracle.reference_available():
        return None
    return oracle.Reference()


def same_result(a, b):
    return (a["objective"] == b["objective"] and a["accepted_moves"] == b["accepted_moves"] and
            a["iterations_done"] == b["iterations_done"] and bool((a["labels"] == b["labels"]).all()))


def host_cpu():
    """nproc and the CPU model name of this host (lscpu, else /proc/cpuinfo)."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_sample(cfg, graph, dg, threads, repeats=3):
    """The steady-state iteration on the CPU: after one warm-up solve, `repeats` pairs of reference solves of m and
    m + 1 iterations (m = cpu_iterations), each pair differenced on the solver's own clock; the median difference is
    the CPU's time per iteration. The GPU solves the same two prefixes and must match them bit for bit."""
    import statistics

    import oracle

    ref = reference_solver()
    if ref is None:
        return None
    og = oracle.Graph(graph.bits, graph.n, graph.edges)
    common = dict(k=cfg["k"], population=cfg["population"], gww=cfg["gww"], seed=cfg["seed"])
    m = cfg["cpu_iterations"]
    ref.solve_pbils(og, iterations=m, early_stop=m, threads=threads, **common)  # warm-up: pages, caches, threads
    out = {"iterations": m, "bit_exact": True, "pairs": []}
    for rep in range(repeats):
        secs = {}
        for it in (m, m + 1):
            cpu = ref.solve_pbils(og, iterations=it, early_stop=it, threads=threads, **common)
            secs[it] = cpu["seconds"]
            if rep == 0:
                gpu = dg.solve_pbils(iterations=it, early_stop=it, **common)
                out[f"gpu_{it}"] = gpu["seconds"]
                out["bit_exact"] &= same_result(cpu, gpu)
        out["pairs"].append([round(secs[m], 4), round(secs[m + 1], 4)])
        note(f"cpu sample {rep + 1}/{repeats}: {m} and {m + 1} iterations, {secs[m]:.2f} s and {secs[m + 1]:.2f} s")
    out["cpu_iteration_s"] = statistics.median(b - a for a, b in out["pairs"])
    out["gpu_iteration_s"] = out[f"gpu_{m + 1}"] - out[f"gpu_{m}"]
    return out


def percentile(values, q):
    import numpy as np

    return float(np.percentile(np.asarray(values, dtype=np.float64), q))


def bench_batch(cga, count, population=64, iterations=50, master=7, plain_leg=True):
    """The mixed stream: `count` independent instances (n log-uniform in [200, 8000], ER or SBM, k in 2..8), pop 64,
    50 iterations, gww 0.1, seeds from master seed 7; solved plain (one instance per solve call, one after the
    other) and batched (one cc_hip_solve_batch), with every batched result compared to its plain result."""
    from cuda_graph_clustering_amd import workloads

    insts = workloads.mixed_instances(count, master=master)
    graphs = [cga.DeviceGraph(workloads.build_graph(cga, inst)) for inst in insts]
    params = [dict(k=inst["k"], population=population, iterations=iterations, early_stop=iterations, gww=0.1,
                   seed=inst["seed"]) for inst in insts]
    cga.solve_batch(graphs[:4], params[:4])  # warm-up: kernels loaded, graph path exercised
    plain, latencies = [], []
    t0 = time.perf_counter()
    for g, kw in zip(graphs if plain_leg else [], params):
        t1 = time.perf_counter()
        plain.append(g.solve_pbils(**kw))
        latencies.append(time.perf_counter() - t1)
        if len(plain) % 128 == 0:
            note(f"mixed stream: {len(plain)}/{count} plain solves")
    plain_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    batched, _ = cga.solve_batch(graphs, params)
    batch_s = time.perf_counter() - t0
    note(f"mixed stream: batched solve of {count} in {batch_s:.2f} s")
    exact = all(a["objective"] == b["objective"] and a["accepted_moves"] == b["accepted_moves"] and
                a["iterations_done"] == b["iterations_done"] and bool((a["labels"] == b["labels"]).all())
                for a, b in zip(plain, batched))
    for g in graphs:
        g.close()
    if not plain_leg:
        return {"batched_graphs_per_s": round(count / batch_s, 3), "batched_seconds": round(batch_s, 3),
                "moves": int(sum(r["accepted_moves"] for r in batched))}
    return {
        "workload": f"{count} instances, n log-uniform [200, 8000], ER p~U[0.05,0.5) or SBM 2-8 blocks "
                    f"(0.4/0.05) 50/50, k~U{{2..8}}, pop {population}, {iterations} iterations, gww 0.1, "
                    f"master seed {master}",
        "batched_graphs_per_s": round(count / batch_s, 3),
        "plain_graphs_per_s": round(count / plain_s, 3),
        "speedup": round(plain_s / batch_s, 2),
        "batched_seconds": round(batch_s, 3),
        "plain_seconds": round(plain_s, 3),
        "plain_latency_p50_s": round(percentile(latencies, 50), 4),
        "plain_latency_p99_s": round(percentile(latencies, 99), 4),
        "batched_latency_s": round(batch_s, 3),
        "bit_exact": bool(exact),
        "moves": int(sum(r["accepted_moves"] for r in batched)),
    }


def bench_serve(cga, rate=50.0, pool=96, big_iterations=20, check=16, master=7):
    """Serving with continuous batching: the wide-path instance (Chung-Lu n=100000, gamma 2.5, mean degree 64,
    graph-seed 4; k=3, pop 128, 20 iterations) on the large lane, beside a Poisson stream (rate req/s) of requests
    drawn from the mixed-batch distribution (pop 64, 50 iterations) submitted while the large run is on the device.
    The stream cycles over `pool` pre-generated graphs, each request with its own search seed."""
    import threading

    import numpy as np
    from cuda_graph_clustering_amd import workloads

    insts = workloads.mixed_instances(pool, master=master)
    graphs = [workloads.build_graph(cga, inst) for inst in insts]
    big = cga.chung_lu(100000, 2.5, 64.0, 4)
    big_kw = dict(k=3, population=128, iterations=big_iterations, early_stop=big_iterations, gww=0.1, seed=42)
    arrivals = np.random.default_rng(master).exponential(1.0 / rate, size=100000)
    submitted = []
    stop = threading.Event()
    with cga.Server(lanes=2, large_bytes=256 << 20) as server:
        t0 = time.perf_counter()
        big_id = server.submit(big, **big_kw)

        def feed():
            when = 0.0
            for i, gap in enumerate(arrivals):
                when += gap
                delay = t0 + when - time.perf_counter()
                if delay > 0 and stop.wait(delay):
                    return
                if stop.is_set():
                    return
                inst = insts[i % pool]
                kw = dict(k=inst["k"], population=64, iterations=50, early_stop=50, gww=0.1,
                          seed=workloads.stream_seed(master, workloads.BATCH_SEARCH_STREAM, i))
                submitted.append((server.submit(graphs[i % pool], **kw), i % pool, kw))

        feeder = threading.Thread(target=feed)
        feeder.start()
        big_res = server.wait(big_id)
        note(f"serving: wide run done after {time.perf_counter() - t0:.2f} s, {len(submitted)} requests submitted")
        stop.set()
        feeder.join()
        stream_s = time.perf_counter() - t0
        results = [(server.wait(rid), gi, kw) for rid, gi, kw in submitted]
        total_s = time.perf_counter() - t0
        counters = server.counters()
    exact = True
    for res, gi, kw in results[:check]:
        with cga.DeviceGraph(graphs[gi]) as dg:
            solo = dg.solve_pbils(**kw)
        exact &= (res["objective"] == solo["objective"] and res["accepted_moves"] == solo["accepted_moves"] and
                  bool((res["labels"] == solo["labels"]).all()))
    lat = [r["latency"] for r, _, _ in results]
    queued = [r["queued"] for r, _, _ in results]
    return {
        "workload": f"wide-path run Chung-Lu n=100000 (gamma 2.5, mean degree 64, graph-seed 4) k=3 pop 128 "
                    f"{big_iterations} iterations on lane 0, beside a Poisson stream at {rate:g} req/s of mixed-batch "
                    f"requests (pop 64, 50 iterations) over {pool} graphs, submitted while the wide run is active",
        "requests": len(results),
        "graphs_per_s": round(len(results) / total_s, 3) if results else 0.0,
        "latency_p50_s": round(percentile(lat, 50), 3) if lat else None,
        "latency_p99_s": round(percentile(lat, 99), 3) if lat else None,
        "queued_p50_s": round(percentile(queued, 50), 3) if queued else None,
        "wide_run_latency_s": round(big_res["latency"], 3),
        "wide_run_objective": big_res["objective"],
        "stream_seconds": round(stream_s, 3),
        "drain_seconds": round(total_s - stream_s, 3),
        "peak_active": counters["peak_active"],
        "peak_bytes": counters["peak_bytes"],
        "bit_exact_sample": bool(exact),
        "checked": min(check, len(results)),
    }


def copy_calibration(mib, reps=4):
    """Device-to-device copies of `mib` MiB, for scaling hardware byte counters: each copy reads and writes exactly
    mib MiB of HBM."""
    import torch

    src = torch.full((mib << 20,), 1, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "copies": reps}))


def bench_neighborhood(cga, args):
    """The Neighborhood baselines on ER n=4096, p=0.1, graph-seed 3: all 4096 start partitions in batches of 256,
    without and with local search, for each local-search kernel and both G rebuilds (popcount, int8 matrix cores);
    the reference CPU enumeration (oracle/_ref) beside it. One JSON line."""
    graph = cga.erdos_renyi(4096, 0.1, 3)
    rows = []
    legs = [(False, "auto", "popcount"), (False, "auto", "mfma"), (True, "auto", "popcount"), (True, "auto", "mfma"),
            (True, "global", "popcount"), (True, "shared", "popcount"), (True, "wide", "popcount")]
    with cga.DeviceGraph(graph) as dg:
        dg.solve_neighborhood(False, k=2, population=256)  # warm-up
        dg.solve_neighborhood(False, k=2, population=256, rebuild="mfma")
        first = {}
        for local_search, ls, rebuild in legs:
            t0 = time.perf_counter()
            res = dg.solve_neighborhood(local_search, k=2, population=256, ls_kernel=ls, profile=True,
                                        rebuild=rebuild)
            wall = time.perf_counter() - t0
            base = first.setdefault(local_search, res)
            same = (res["objective"] == base["objective"] and res["accepted_moves"] == base["accepted_moves"] and
                    bool((res["labels"] == base["labels"]).all()))
            rows.append({"local_search": local_search, "ls_kernel": ls, "rebuild": rebuild,
                         "descent": res["descent_name"], "seconds": round(wall, 4),
                         "starts_per_s": round(graph.n / wall, 1), "objective": res["objective"],
                         "accepted_moves": res["accepted_moves"], "rebuild_ms": round(res["rebuild_ms"], 3),
                         "search_ms": round(res["search_ms"], 3), "rebuild_launches": res["rebuild_launches"],
                         "same_as_first": bool(same)})
            note(f"neighborhood local_search={local_search} {ls} {rebuild}: {wall:.3f} s")
    best = min(r["seconds"] for r in rows if r["local_search"])
    out = {"metric": "Neighborhood start partitions per second (with local search, best kernel)",
           "value": round(graph.n / best, 1), "unit": "starts/s", "n_gpus": 1, "higher_is_better": True,
           "data": "synthetic", "dtype": "int16",
           "config": {"workload": "ER n=4096, p=0.1, graph-seed 3; k=2; 4096 starts in batches of 256"},
           "edges": graph.edges, "runs": rows}
    ref = None if args.no_cpu else reference_solver()
    if ref is not None:
        import oracle

        og = oracle.Graph(graph.bits, graph.n, graph.edges)
        t0 = time.perf_counter()
        cpu = ref.neighborhood(og, local_search=False)
        out["cpu_baseline"] = {"value": round(graph.n / (time.perf_counter() - t0), 1), "unit": "starts/s",
                               "cores": 1, "kind": "reference",
                               "sample": "all 4096 starts without local search, the reference's sequential "
                                         "enumeration (cuda/cc_verify_gpu.cu:21-35)",
                               "bit_exact": cpu["objective"] == rows[0]["objective"]}
    print(json.dumps(out))


def bench_tags(cga, args):
    """The tag-similarity graph builder (Graph::LoadTags) on a synthetic tag file shaped like a photo-tag dataset:
    objects with 0-8 tags drawn Zipf-like from a vocabulary, jaccard >= 0.5 over every object, beside the reference's
    own LoadTags (oracle/_ref) on the same file; both graphs must be identical. Host code, all cores. One JSON
    line."""
    import random
    import tempfile

    import numpy as np

    objects, vocabulary = args.tags_objects, 2000
    rng = random.Random(17)
    weights = [1.0 / (r + 1) for r in range(vocabulary)]
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "tags.json")
        with open(path, "w") as out:
            out.write("{\n")
            for i in range(objects):
                tags = rng.choices(range(vocabulary), weights, k=rng.randint(0, 8))
                out.write(f'  "{i}": [' + ", ".join(f'"tag{t}"' for t in tags) + "]" +
                          (",\n" if i + 1 < objects else "\n"))
            out.write("}\n")
        t0 = time.perf_counter()
        g = cga.load_tags(path, "jaccard", 0.5, 0, 1)
        ours = time.perf_counter() - t0
        out = {"metric": "tag-similarity graph build (Graph::LoadTags)", "value": round(ours, 3), "unit": "s",
               "higher_is_better": False, "data": "synthetic",
               "config": {"workload": f"{objects} objects, 0-8 tags each from {vocabulary} (weights 1/rank), "
                                      "jaccard >= 0.5, all objects", "n": g.n, "edges": g.edges},
               "threads": os.cpu_count()}
        ref = None if args.no_cpu else reference_solver()
        if ref is not None:
            t0 = time.perf_counter()
            want = ref.load_tags(path, "jaccard", 0.5, 0, 1)
            theirs = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": round(theirs, 3), "unit": "s", "cores": os.cpu_count(),
                                   "kind": "reference", "sample": "the same file through the reference's LoadTags",
                                   "identical": bool(want.n == g.n and want.edges == g.edges and
                                                     np.array_equal(want.bits, g.bits))}
    print(json.dumps(out))


SWEEP_SERIES = {"n128": "graph size, k = 3, pop 128", "pop6000": "population, n = 6000, k = 3",
                "k128": "clusters, n = 2000, pop 128"}


def bench_sweep(cga, args):
    """The reference's published GPU sweeps (docs/benchmarks.md:127-202, docs/gpu_sweep.sh): ER G(n, 0.33),
    graph-seed 7, 20 iterations, early stop 1000, gww 0, tournament 5, perturbation 0.4; each point solved with
    seeds 6, 7 and 8 after a warm-up with seed 5, host graph in (so upload and allocation are inside the time, as
    on the reference's clock), with our default plan and with ls_kernel = global, the kernel family the reference
    timed. Every final f must equal the reference's (tests/golden/gpu_trajectories.csv, gpu_f == cpu_f there);
    the reference's seconds are its RTX 4070 Ti SUPER (CUDA) and Ryzen 7 3700X 16-thread runs. One JSON line."""
    import csv
    import statistics

    rows = list(csv.DictReader(open(os.path.join(ROOT, "tests", "golden", "gpu_trajectories.csv"))))
    points = {}
    for r in rows:
        if r["series"] in SWEEP_SERIES:
            key = (r["series"], int(r["n"]), int(r["k"]), int(r["pop"]))
            points.setdefault(key, []).append(r)
    out_points, exact = [], True
    graphs = {}
    for (series, n, k, pop), refs in sorted(points.items(), key=lambda kv: (kv[0][0], kv[0][1], kv[0][2], kv[0][3])):
        if n > args.sweep_max_n:
            continue
        if n not in graphs:
            graphs = {n: cga.erdos_renyi(n, 0.33, 7)}
        g = graphs[n]
        common = dict(k=k, population=pop, iterations=20, early_stop=1000, gww=0.0, tournament=5, perturbation=0.4)
        point = {"series": series, "n": n, "k": k, "pop": pop}
        for label, ls in (("auto", "auto"), ("global", "global")):
            cga.solve_pbils(g, seed=5, ls_kernel=ls, **common)
            times = []
            for r in sorted(refs, key=lambda r: int(r["seed"])):
                t0 = time.perf_counter()
                res = cga.solve_pbils(g, seed=int(r["seed"]), ls_kernel=ls, **common)
                times.append(time.perf_counter() - t0)
                exact = exact and res["objective"] == int(r["f"])
            point[f"ours_{label}_s"] = round(statistics.median(times), 4)
            if label == "auto":
                point["descent"] = res["descent_name"]
        point["reference_gpu_s"] = statistics.median(float(r["gpu_s"]) for r in refs)
        point["reference_cpu16_s"] = statistics.median(float(r["cpu_s"]) for r in refs)
        point["vs_reference_gpu"] = round(point["reference_gpu_s"] / point["ours_auto_s"], 1)
        point["vs_reference_gpu_same_kernel"] = round(point["reference_gpu_s"] / point["ours_global_s"], 1)
        out_points.append(point)
        note(f"sweep {series} n={n} k={k} pop={pop}: ours {point['ours_auto_s']} s (global "
             f"{point['ours_global_s']} s), reference GPU {point['reference_gpu_s']} s")
    print(json.dumps({"metric": "seconds per 20 PBILS iterations on the reference's published GPU sweeps",
                      "unit": "s", "higher_is_better": False, "data": "synthetic",
                      "config": {"workload": "ER G(n, 0.33) graph-seed 7, 20 iterations, gww 0; series: " +
                                 "; ".join(f"{k}: {v}" for k, v in SWEEP_SERIES.items())},
                      "bit_exact_vs_reference_f": exact, "points": out_points}))


# The reference's published wide-path runs on its two real tag graphs (docs/benchmarks.md:306-331; RTX 4070 Ti SUPER,
# pop 128, tournament 5, perturbation 0.4, early stop 6, --seed 5): (n, edge density, k) -> (seconds per iteration
# after the first, first iteration, accepted moves per descent).
REFERENCE_WIDE = {
    (193734, 0.126, 2): (10.2, 14.9, 81451),
    (193734, 0.126, 3): (23.3, 41.7, 87592),
    (330000, 0.128, 2): (41.6, 55.1, 137757),
    (330000, 0.128, 3): (105.0, 181.9, 145887),
}


def bench_scale(cga, args):
    """The wide path at the reference's published scale: synthetic ER graphs with the vertex counts and edge densities
    of its two tag graphs (NUS-WIDE, all 193 734 objects, 12.6 %; Stack Overflow sample, 330 000 objects, 12.8 %; the
    graphs themselves need downloaded data), k = 2 and 3, pop 128 and the reference's protocol (tournament 5,
    perturbation 0.4, gww 0.1, seed 5). Graph resident in HBM; the first iteration (descents from random labels) is
    timed by a solve of one iteration, the later ones by the difference to a solve of three. Each result's f must
    equal the library's own recount of its labels. The per-move figure divides an iteration by the accepted moves of
    one descent (all slots descend at once), which compares graphs whose descents differ in length. One JSON line."""
    import numpy as np

    sizes = [(int(a), float(b)) for a, b in (s.split(":") for s in args.scale_sizes.split(","))]
    common = dict(population=128, tournament=5, perturbation=0.4, gww=0.1, seed=5)
    points = []
    for n, density in sizes:
        t0 = time.perf_counter()
        g = cga.erdos_renyi(n, density, 7)
        max_degree = int(np.bitwise_count(g.bits).sum(axis=1, dtype=np.int64).max())
        note(f"scale: ER n={n} p={density} built in {time.perf_counter() - t0:.1f} s ({g.edges} edges, largest "
             f"degree {max_degree})")
        with cga.DeviceGraph(g) as dg:
            for k in (int(x) for x in args.scale_ks.split(",")):
                runs = {}
                for iters in (1, 3):
                    t0 = time.perf_counter()
                    res = dg.solve_pbils(k=k, iterations=iters, early_stop=1000, **common)
                    runs[iters] = (time.perf_counter() - t0, res)
                    note(f"scale n={n} k={k}: {iters} iteration(s) in {runs[iters][0]:.2f} s")
                    if cga.objective_bitwise(g, res["labels"]) != res["objective"]:
                        raise SystemExit(f"scale n={n} k={k}: reported f differs from the recount")
                first_s, first = runs[1]
                later_s = (runs[3][0] - first_s) / 2
                moves = (runs[3][1]["accepted_moves"] - first["accepted_moves"]) / (2 * common["population"])
                # all slots descend at once: algorithmic bytes of one move of every slot per move time, against HBM
                per_move = descent_bytes_per_move(n, k, first["descent_name"], 2.0 * g.edges / n, first["lds_bytes"])
                gbs = common["population"] * per_move / (later_s / moves) / 1e9
                point = {"n": n, "density": density, "edges": g.edges, "k": k, "descent": first["descent_name"],
                         "iteration_s": round(later_s, 3), "first_iteration_s": round(first_s, 3),
                         "moves_per_descent": round(moves), "us_per_move": round(1e6 * later_s / moves, 2),
                         "lds_bytes": first["lds_bytes"], "largest_degree": max_degree,
                         "hbm": {"bytes_per_move": round(per_move), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                 "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 3)},
                         "f_after_3": runs[3][1]["objective"], "f_recounted": True}
                ref = REFERENCE_WIDE.get((n, density, k))
                if ref:
                    point["reference"] = {"iteration_s": ref[0], "first_iteration_s": ref[1],
                                          "moves_per_descent": ref[2],
                                          "us_per_move": round(1e6 * ref[0] / ref[2], 2)}
                    point["per_move_vs_reference"] = round(point["reference"]["us_per_move"] / point["us_per_move"], 1)
                points.append(point)
                note(f"scale n={n} k={k} {point['descent']}: iteration {later_s:.2f} s, first {first_s:.2f} s, "
                     f"{moves:.0f} moves per descent")
    print(json.dumps({"metric": "wide-path PBILS iteration at the reference's published graph sizes",
                      "unit": "s", "higher_is_better": False, "data": "synthetic",
                      "config": {"workload": "ER graphs with the reference's tag-graph sizes and densities; k 2 and 3, "
                                             "pop 128, tournament 5, perturbation 0.4, gww 0.1, seed 5"},
                      "points": points}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS) + ["c3", "tags", "sweep", "scale"])
    ap.add_argument("--sweep-max-n", type=int, default=32000, help="largest n of --config sweep")
    ap.add_argument("--tags-objects", type=int, default=60000, help="objects of the --config tags file")
    ap.add_argument("--scale-sizes", default="193734:0.126,330000:0.128",
                    help="--config scale: comma-separated n:density pairs")
    ap.add_argument("--scale-ks", default="2,3", help="--config scale: comma-separated cluster counts")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-count", type=int, default=512, help="instances of the mixed stream; 0 skips it")
    ap.add_argument("--batch-only", action="store_true", help="mixed stream: skip the plain one-by-one leg")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="Poisson request rate; 0 skips serving")
    ap.add_argument("--serve-iterations", type=int, default=20, help="iterations of the wide-path served run")
    ap.add_argument("--serve-low-rate", type=float, default=15.0,
                    help="second serving leg at this rate, below capacity; 0 skips it")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--autotune", type=int, default=0, metavar="I",
                    help="before timing, let the autotuner time every descent plan on I iterations")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (to scale byte counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import cuda_graph_clustering_amd as cga

    if args.config == "tags":
        bench_tags(cga, args)
        return
    if args.config == "sweep":
        bench_sweep(cga, args)
        return
    if args.config == "scale":
        bench_scale(cga, args)
        return
    if args.config == "c3":
        bench_neighborhood(cga, args)
        return
    cfg = CONFIGS[args.config]
    steps = args.steps if args.steps else cfg["iterations"]
    graph = make_graph(cga, cfg["gen"])
    common = dict(k=cfg["k"], population=cfg["population"], gww=cfg["gww"], seed=cfg["seed"])

    tuned = None
    with cga.DeviceGraph(graph) as dg:
        if args.autotune > 0:  # before the warm-up, outside the timed region
            cands, chosen = cga.autotune(dg, tune_iterations=args.autotune, iterations=steps, early_stop=steps,
                                         **common)
            tuned = {"chosen": chosen, "candidates": [{k: c[k] for k in ("descent", "threads", "vertices_per_thread",
                                                                           "rows_in_lds", "blocks", "search_ms")}
                                                      for c in cands]}
            note(f"autotune: {len(cands)} plans, chose {cands[chosen]['descent']} {cands[chosen]['threads']} threads")
        if args.warmup > 0:
            dg.solve_pbils(iterations=args.warmup, early_stop=args.warmup, **common)
        # The timed solve runs as a user's would: iterations replayed from the captured HIP graph, no event timing.
        t0 = time.perf_counter()
        res = dg.solve_pbils(iterations=steps, early_stop=steps, **common)
        wall = time.perf_counter() - t0
        note(f"{args.config}: {res['iterations_done']} iterations in {wall:.3f} s")
        # The kernel times come from a second solve of the same trajectory with HIP events around every launch
        # (launched one by one on the library's stream); it must reproduce the timed solve bit for bit.
        prof = dg.solve_pbils(iterations=steps, early_stop=steps, profile=True, **common)
        if not same_result(res, prof):
            raise SystemExit("the profiled solve did not reproduce the timed one")
        sample = None if args.no_cpu else cpu_sample(cfg, graph, dg, args.cpu_threads)

    n, k = graph.n, cfg["k"]
    iters = res["iterations_done"]
    value = iters / wall
    launches = max(1, prof["search_launches"])
    mean_launch_s = prof["search_ms"] / 1e3 / launches
    moves_per_launch = prof["accepted_moves"] / launches
    descent = prof["descent_name"]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / max(iters, 1), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16" if n <= 32767 else "int32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "n": n, "edges": graph.edges, "k": k,
                   "population": cfg["population"], "iterations": steps, "descent": descent,
                   "descent_threads": prof["threads"], "vertices_per_thread": prof["vertices_per_thread"]},
        "moves_per_s": round(res["accepted_moves"] / wall, 1),
        "accepted_moves": res["accepted_moves"],
        "objective": res["objective"],
        "device_ms": {"descent": round(prof["search_ms"], 3), "rebuild": round(prof["rebuild_ms"], 3),
                      "iterations": round(prof["iteration_ms"], 3), "descent_launches": launches},
        "roofline": roofline(args.config, cfg, graph, prof, mean_launch_s, moves_per_launch),
    }
    if sample is not None:
        m = sample["iterations"]
        cpu_rate = 1.0 / sample["cpu_iteration_s"]
        host = host_cpu()
        out["cpu_baseline"] = {
            "value": round(cpu_rate, 5),
            "unit": "iter/s",
            "cores": args.cpu_threads,
            "kind": "reference",
            "sample": f"iteration {m + 1} of the same workload (tournament, {cfg['population']} descents from "
                      f"perturbed optima, record, GWW, perturb, rebuild): reference SolveCpu with "
                      f"{args.cpu_threads} threads; after a warm-up solve, the median over {len(sample['pairs'])} "
                      f"pairs of solves of {m} and {m + 1} iterations differenced on its own clock",
            "pairs_s": sample["pairs"],
            "host": host,
            "gpu_same_iteration_s": round(sample["gpu_iteration_s"], 5),
            "gpu_over_cpu": round(value / cpu_rate, 1),
            "bit_exact": sample["bit_exact"],
        }
    else:
        out["cpu_baseline"] = None
    out["autotune"] = tuned
    out["batched"] = bench_batch(cga, args.batch_count, plain_leg=not args.batch_only) if args.batch_count > 0 else None
    out["serving"] = (bench_serve(cga, rate=args.serve_rate, big_iterations=args.serve_iterations)
                      if args.serve_rate > 0 else None)
    # The same scenario at a rate below the device's batched capacity, where latency is service time, not queueing.
    out["serving_below_capacity"] = (bench_serve(cga, rate=args.serve_low_rate, big_iterations=args.serve_iterations)
                                     if args.serve_rate > 0 and args.serve_low_rate > 0 else None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
