"""The C ABI library without a GPU: it loads, exports every symbol include/cc_hip.h declares, and its host-side
generators and recounts agree bit-exactly with the oracle."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cc_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import cuda_graph_clustering_amd as cga

    declared = declared_functions()
    assert declared, "no declarations parsed"
    assert sorted(cga.ABI_SYMBOLS) == declared
    for name in declared:
        assert hasattr(cga.lib, name), name


def test_default_params_match_reference_defaults():
    import cuda_graph_clustering_amd as cga

    p = cga.make_params()
    # include/cc_pbils.hpp:37-48 of the reference
    assert (p.k, p.population, p.tournament, p.iterations, p.early_stop) == (2, 128, 5, 100, 6)
    assert (p.perturbation, p.gww, p.seed, p.ls_kernel) == (0.4, 0.1, 1, 0)


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 200, 517])
def test_product_generators_match_oracle(orc, n):
    import cuda_graph_clustering_amd as cga

    for seed in (1, 7):
        a, b = cga.erdos_renyi(n, 0.33, seed), orc.erdos_renyi(n, 0.33, seed)
        assert a.edges == b.edges and (a.bits == b.bits).all()
        a, b = cga.sbm(n, 5, 0.4, 0.05, seed), orc.sbm(n, 5, 0.4, 0.05, seed)
        assert a.edges == b.edges and (a.bits == b.bits).all()
        a, b = cga.chung_lu(n, 2.5, 16.0, seed), orc.chung_lu(n, 2.5, 16.0, seed)
        assert a.edges == b.edges and (a.bits == b.bits).all()


def test_product_recounts_match_oracle(orc):
    import cuda_graph_clustering_amd as cga

    rng = np.random.default_rng(0)
    for n in (1, 31, 64, 65, 300):
        g = cga.sbm(n, 3, 0.5, 0.1, n)
        og = orc.sbm(n, 3, 0.5, 0.1, n)
        for k in (1, 2, 5):
            labels = rng.integers(0, k, n).astype(np.int32)
            want = orc.objective_direct(og, labels)
            assert cga.objective_direct(g, labels) == want
            assert cga.objective_bitwise(g, labels) == want


def test_workload_streams_match_the_rng_contract(orc):
    from cuda_graph_clustering_amd import workloads

    rng = np.random.default_rng(3)
    for _ in range(200):
        seed, stream, step = (int(x) for x in rng.integers(0, 2**63, 3, dtype=np.uint64))
        assert workloads.stream_seed(seed, stream, step) == orc.stream_seed(seed, stream, step)
    a = workloads.mixed_instances(64)
    assert a == workloads.mixed_instances(64)
    ns = [x["gen"][1] for x in a]
    assert min(ns) >= 200 and max(ns) <= 8000
    assert {x["k"] for x in a} <= set(range(2, 9)) and len({x["gen"][0] for x in a}) == 2


def test_cli_binary_is_built():
    import cuda_graph_clustering_amd as cga

    assert os.access(cga.CLI_PATH, os.X_OK)


def test_tune_table_load_validates_entries(tmp_path):
    import cuda_graph_clustering_amd as cga

    good = tmp_path / "good.txt"
    good.write_text("2000 5 16 0 5 1024 2 0 0 0\n")
    bad = tmp_path / "bad.txt"
    bad.write_text("2000 5 16 0 99 1024 2 0 0 0\n")
    garbled = tmp_path / "garbled.txt"
    garbled.write_text("2000 5 16 0 5 1024 two\n")
    # the k = 3 chunk kernel at n = 1000 (16 chunks) needs 32 bytes of LDS per chunk: an entry with less is refused
    chunk_ok = tmp_path / "chunk_ok.txt"
    chunk_ok.write_text("1000 3 6 1 7 1024 0 0 512 0\n")
    chunk_small = tmp_path / "chunk_small.txt"
    chunk_small.write_text("1000 3 6 1 7 1024 0 0 511 0\n")
    try:
        assert cga.tune_load(good) == 1
        assert cga.tune_load(chunk_ok) == 1
        for path in (bad, garbled, chunk_small):
            with pytest.raises(cga.CcError) as e:
                cga.tune_load(path)
            assert e.value.code == cga.EINVAL
    finally:
        cga.tune_clear()


def test_reference_command_line_compiles_against_this_include():
    """The reference's GPU command line (cuda/main_gpu.cu with its include/cc_cli.hpp) compiles unchanged against
    include/ of this library: every declaration it uses (SolveGpu and the Neighborhood entry points, Graph and its
    loaders, PbilsParams, PbilsResult, ResolveThreads, ObjectiveDirect, ObjectiveBitwise) is there with its
    signature. Syntax check only; oracle/Makefile links the same source into oracle/_ref/cc_gpu_ref."""
    import shutil
    import subprocess
    import tempfile

    reference = "/root/reference"
    if not os.path.exists(os.path.join(reference, "cuda", "main_gpu.cu")):
        pytest.skip("the reference sources are not on this machine")
    with tempfile.TemporaryDirectory() as tmp:
        os.symlink(os.path.join(reference, "include", "cc_cli.hpp"), os.path.join(tmp, "cc_cli.hpp"))
        proc = subprocess.run([shutil.which("g++") or "g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-iquote", tmp,
                               "-I", os.path.join(ROOT, "include"), os.path.join(reference, "cuda", "main_gpu.cu")],
                              capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
