"""The instance files of the command line (matrix text, the "graph" block of a baseline result JSON, SNAP/CSV edge
lists, tag files) against the reference's own loaders and writers, compiled from its sources into oracle/_ref:
the same graph (bits and edge count) from every file either accepts, the same rejections, and the same bytes from
the writers (src/cc_graph.cpp:35-103, 179-306). Hand-written cases cover each rule of the formats; seeded random
byte strings over each format's alphabet cover what the cases miss. Host code only."""
import json
import os

import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


def same_graph(a, b):
    return a.n == b.n and a.edges == b.edges and np.array_equal(np.asarray(a.bits).reshape(a.n, -1),
                                                                  np.asarray(b.bits).reshape(b.n, -1))


def load_both(cga, ref, path, fmt):
    """(ours, reference): a graph each, or None where the loader rejected the file."""
    try:
        ours = cga.load_graph(path, fmt)
    except cga.CcError as e:
        assert e.code == cga.EINVAL
        ours = None
    try:
        theirs = ref.load_graph(path, fmt)
    except RuntimeError:
        theirs = None
    return ours, theirs


def agree(cga, ref, path, fmt):
    ours, theirs = load_both(cga, ref, path, fmt)
    assert (ours is None) == (theirs is None), (fmt, open(path, "rb").read()[:200])
    if ours is not None:
        assert same_graph(ours, theirs), (fmt, open(path, "rb").read()[:200])
    return ours


@pytest.mark.parametrize("n", [1, 2, 5, 63, 64, 65, 130])
def test_writers_are_byte_identical_and_round_trip(cga, ref, tmp_path, n):
    g = cga.erdos_renyi(n, 0.3, n)
    og = oracle.Graph(g.bits, g.n, g.edges)
    for fmt in ("matrix", "json"):
        cga.save_graph(g, tmp_path / f"ours.{fmt}", fmt)
        ref.save_graph(og, tmp_path / f"ref.{fmt}", fmt)
        assert (tmp_path / f"ours.{fmt}").read_bytes() == (tmp_path / f"ref.{fmt}").read_bytes(), fmt
        assert same_graph(agree(cga, ref, tmp_path / f"ours.{fmt}", fmt), g)
    cga.save_graph(g, tmp_path / "g.edges", "edges")
    back = agree(cga, ref, tmp_path / "g.edges", "edges")
    if g.edges:  # an edge list names only the vertices that have an edge
        degree = np.unpackbits(g.bits.view(np.uint8), axis=1, bitorder="little").sum(axis=1)
        assert back.n == int((degree > 0).sum()) and back.edges == g.edges


MATRIX_CASES = [
    "3\n010\n101\n010\n",
    "3\n0 1 0\n1 0 1\n0 1 0\n",
    "  3\r\n0 1 1\r\n1 0 0\r\n1 0 0\r\n",
    "3\n011\n000\n000\n",          # only the cells above the diagonal count
    "3\n000\n100\n110\n",          # the lower triangle alone gives no edge
    "3\n111\n111\n111\n",          # the diagonal never makes a loop
    "3\n0x1\n2 0 1\n0 9 0\n",      # any byte but '1' is a non-edge
    "2\n01\n10\ntrailing words\n",
    "+2\n01\n10\n",
    "1\n0\n",
    "1\n1",
    "2 0110",
    "",                            # no header
    "0\n",                         # n = 0
    "abc\n",
    "3\n010\n101\n01",             # truncated
]


@pytest.mark.parametrize("text", MATRIX_CASES)
def test_matrix_cases(cga, ref, tmp_path, text):
    path = tmp_path / "m.txt"
    path.write_text(text)
    agree(cga, ref, path, "matrix")


JSON_CASES = [
    '{"graph": [[0,1],[1,0]]}',
    '{"size": 3, "x": [[1,1,1]], "graph": [ [0, 1, 0], [1, 0, 1], [0, 1, 0] ], "after": [[1]]}',
    '{"graph": [[0,1,1],[0,0,0],[0,0,0]]}',
    '{"graph": [[0,1,[1,1,1]],[1,0]]}',   # deeper lists are skipped
    '{"graph": [[1]]}',
    '{"graph": [[0,1],[1,0]',              # the file ends inside the block: the rows read so far
    '{"graph": [[0,1],[1,0],[1',
    '{"graph": [[01],[10]]}',              # every '0' or '1' byte is a cell
    '{"graph": [[0,1],[1]]}',              # a short row
    '{"graph": []}',
    '{"graph": 5}',
    '{"grap": [[0]]}',
    '',
]


@pytest.mark.parametrize("text", JSON_CASES)
def test_baseline_json_cases(cga, ref, tmp_path, text):
    path = tmp_path / "g.json"
    path.write_text(text)
    agree(cga, ref, path, "json")


EDGE_CASES = [
    "1 2\n2 3\n",
    "# comment\n% other\nsrc,dst,weight\n10,20,0.5\n20,30,1\n",
    "5\t7\n7 5\n5 5\n5 7\n",                       # repeats, reversed and loops collapse
    "  3 4\n\n\n4 9 extra words\r\n",
    "1 2\n# 3 4\n 5 6\n,7,8\n",
    "9 8\n1\n2 x\n3 4",
    "18446744073709551609 1\n2 3\n",
    "99999999999999999999999 1\n",                # an id beyond 64 bits
    "# nothing\n",
    "",
    "0 0\n",
]


@pytest.mark.parametrize("text", EDGE_CASES)
def test_edge_list_cases(cga, ref, tmp_path, text):
    path = tmp_path / "g.edges"
    path.write_text(text)
    agree(cga, ref, path, "edges")


def test_missing_files(cga, ref, tmp_path):
    for fmt in ("matrix", "json", "edges"):
        ours, theirs = load_both(cga, ref, tmp_path / "missing", fmt)
        assert ours is None and theirs is None


def random_text(rng, alphabet, length):
    return "".join(alphabet[i] for i in rng.integers(0, len(alphabet), length))


@pytest.mark.parametrize("fmt,alphabet", [
    ("matrix", ["0", "1", "1", " ", "\n", "2", "3", "\t"]),
    ("json", ["0", "1", ",", "[", "]", " ", '"graph"', "\n", "{", "}"]),
    ("edges", ["1", "2", "3", "9", "0", " ", "\t", ",", "\n", "#", "%", "x", "12345678901234567890"]),
])
def test_random_files_agree(cga, ref, tmp_path, fmt, alphabet):
    rng = np.random.default_rng(len(fmt))
    path = tmp_path / "random"
    accepted = 0
    for trial in range(400):
        body = random_text(rng, alphabet, int(rng.integers(0, 120)))
        if fmt == "matrix":  # a one-digit vertex count, then random cells
            body = str(int(rng.integers(0, 7))) + random_text(rng, [" ", "\n"], 1) + body
        elif fmt == "json" and trial % 2:  # rows of about n cells, then one random byte changed
            n = int(rng.integers(1, 6))
            rows = ["[" + ",".join(random_text(rng, ["0", "1"], 1) for _ in range(n + int(rng.integers(-1, 2)) * (
                rng.random() < 0.2))) + "]" for _ in range(n)]
            body = '{"x": 1, "graph": [' + ", ".join(rows) + "]}"
            at = int(rng.integers(0, len(body)))
            if rng.random() < 0.5:
                body = body[:at] + random_text(rng, alphabet, 1) + body[at + 1:]
        path.write_text(body)
        accepted += agree(cga, ref, path, fmt) is not None
    assert accepted >= 40, "the random files must also exercise accepted inputs"


def test_random_tag_files_agree(cga, ref, tmp_path):
    """Tag files of random bytes over the tokens of the format: the tag reader accepts exactly what the
    reference's ReadTagSets accepts and builds the same graph."""
    rng = np.random.default_rng(5)
    tokens = ["{", "}", "[", "]", ":", ",", '"a"', '"b"', '"c\\"d"', '"k"', " ", "\n", "x"]
    path = tmp_path / "tags.json"
    accepted = 0
    for trial in range(600):
        if trial % 3 == 0:
            objects = {f"o{i}": [f"t{int(t)}" for t in rng.integers(0, 4, int(rng.integers(0, 4)))]
                       for i in range(int(rng.integers(0, 6)))}
            text = json.dumps(objects)
            if trial % 2:
                text = text[: int(rng.integers(0, len(text) + 1))]
        else:
            text = "{" + random_text(rng, tokens, int(rng.integers(0, 40)))
        path.write_text(text)
        try:
            ours = cga.load_tags(path, "jaccard", 0.5, 0, 1)
        except cga.CcError:
            ours = None
        try:
            theirs = ref.load_tags(path, "jaccard", 0.5, 0, 1)
        except RuntimeError:
            theirs = None
        assert (ours is None) == (theirs is None), text
        if ours is not None:
            accepted += 1
            assert same_graph(ours, theirs), text
    assert accepted >= 60
