"""The tag-similarity graph (Graph::LoadTags, reference src/cc_graph.cpp:311-371) and the object sample it draws
(SampleWithoutReplacement, src/cc_graph.cpp:373-381). Host code of the product, no GPU.

Checked three ways: against a pairwise restatement over Python sets with the baseline's formulas (the construction
the reference's own test compares with, src/cc_verify.cpp:277-343), on the objects of that test; against the
reference's LoadTags compiled out of tree (oracle/_ref) on random tag files, where that build exists; and the sample
against the oracle's restatement and the reference's function."""
import math
import random

import numpy as np
import pytest


@pytest.fixture(scope="module")
def cga():
    import cuda_graph_clustering_amd as m

    return m


# The objects of the reference's tags test (src/cc_verify.cpp:301-303): tags separated by spaces, with empty sets,
# repeated tags and a quote inside a tag.
REF_OBJECTS = ["a", "a b", "b a a", "c", "", "a b c", "d", "a", "b c", 'q"x', "a c d e", "e", "a b", "", "c d",
               'q"x a']
KINDS = ("jaccard", "cosine", "dice", "overlap")


def write_tags(path, objects):
    """{"1000": ["a", ...], "1007": [...], ...} with '"' and '\\' escaped, one object per line."""
    lines = []
    for i, tags in enumerate(objects):
        quoted = ", ".join('"' + t.replace("\\", "\\\\").replace('"', '\\"') + '"' for t in tags)
        lines.append(f'  "{1000 + 7 * i}": [{quoted}]')
    path.write_text("{\n" + ",\n".join(lines) + "\n}\n")
    return path


def similarity(kind, a, b):
    """The baseline's measure of two tag sets (as the reference test restates it, src/cc_verify.cpp:277-294); NaN
    for 0/0."""
    inter, union = len(a & b), len(a | b)
    num, den = {
        "jaccard": (inter, union),
        "cosine": (inter, math.sqrt(len(a) * len(b))),
        "dice": (2.0 * inter, len(a) + len(b)),
        "overlap": (inter, min(len(a), len(b))),
    }[kind]
    if den == 0:
        return math.nan if num == 0 else math.inf
    return num / den


def pairwise(objects, chosen, kind, threshold):
    sets = [set(objects[i]) for i in chosen]
    n = len(sets)
    adj = np.zeros((n, n), dtype=bool)
    for i in range(n):
        for j in range(i + 1, n):
            adj[i, j] = adj[j, i] = similarity(kind, sets[i], sets[j]) >= threshold
    return adj


def dense(graph):
    bits = np.unpackbits(graph.bits.view(np.uint8), axis=1, bitorder="little")
    return bits[:, : graph.n].astype(bool)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("threshold", [0.25, 0.5, 0.75, 1.0])
@pytest.mark.parametrize("n", [0, 9])
def test_reference_objects_match_pairwise_construction(cga, orc, tmp_path, kind, threshold, n):
    objects = [s.split() for s in REF_OBJECTS]
    path = write_tags(tmp_path / "tags.json", objects)
    g = cga.load_tags(path, kind, threshold, n, 42)
    chosen = orc.sample_without_replacement(len(objects), n, 42)
    want = pairwise(objects, chosen, kind, threshold)
    assert g.n == len(chosen)
    assert (dense(g) == want).all()
    assert g.edges == int(want.sum()) // 2


def random_objects(rng, count, vocabulary, max_tags):
    tags = [f"t{i}" for i in range(vocabulary)] + ['q"x', "back\\slash"]
    return [[rng.choice(tags) for _ in range(rng.randint(0, max_tags))] for _ in range(count)]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("threshold", [-1.0, 0.0, 0.3, 0.5, 1.0])
def test_random_objects_match_pairwise_construction(cga, orc, tmp_path, kind, threshold):
    """Thresholds <= 0 join sets without a common tag too (the measure of an empty intersection is 0, or NaN for
    two empty sets), which the product decides per pair of set sizes instead of per pair of objects."""
    objects = random_objects(random.Random(5), 150, 12, 5)
    path = write_tags(tmp_path / "tags.json", objects)
    for n in (0, 70):
        g = cga.load_tags(path, kind, threshold, n, 9)
        want = pairwise(objects, orc.sample_without_replacement(len(objects), n, 9), kind, threshold)
        assert (dense(g) == want).all(), (kind, threshold, n)
        assert g.edges == int(want.sum()) // 2


@pytest.mark.parametrize("kind", KINDS)
def test_random_tag_files_match_reference(cga, ref, tmp_path, kind):
    rng = random.Random(11)
    for trial, (count, vocabulary, max_tags) in enumerate([(600, 20, 6), (900, 60, 3), (300, 5, 8)]):
        objects = random_objects(rng, count, vocabulary, max_tags)
        path = write_tags(tmp_path / f"tags{trial}.json", objects)
        for threshold in (-0.5, 0.0, 0.2, 0.5, 0.8, 1.0, 1.5):
            for n, seed in ((0, 1), (250, 3), (count + 50, 4)):
                got = cga.load_tags(path, kind, threshold, n, seed)
                want = ref.load_tags(path, kind, threshold, n, seed)
                assert (got.n, got.edges) == (want.n, want.edges), (trial, threshold, n)
                assert np.array_equal(got.bits, want.bits), (trial, threshold, n)


def test_sample_matches_oracle_and_reference(cga, orc, request):
    reference = None
    import oracle

    if oracle.reference_available():
        reference = oracle.Reference()
    for total, n, seed in [(1, 0, 1), (1, 1, 1), (16, 9, 42), (16, 0, 42), (16, 16, 42), (16, 40, 42), (1000, 37, 7),
                           (100000, 5000, 123), (5, 4, 0)]:
        got = cga.sample_without_replacement(total, n, seed)
        assert np.array_equal(got, orc.sample_without_replacement(total, n, seed))
        if reference is not None:
            assert np.array_equal(got, reference.sample_without_replacement(total, n, seed))
        # a prefix of a permutation; everything in order when nothing is drawn
        assert len(got) == (total if n == 0 or n >= total else n)
        assert len(set(got.tolist())) == len(got) and (got < total).all()
        if n == 0 or n >= total:
            assert np.array_equal(got, np.arange(total))


def test_quote_escape_and_key_order(cga, tmp_path):
    """Keys are ignored (vertex = position in the file); an escaped quote is part of the tag."""
    path = tmp_path / "t.json"
    path.write_text('{"z": ["a\\"b", "c"], "a": ["a\\"b"], "m": ["c"], "b": []}')
    g = cga.load_tags(path, "overlap", 1.0, 0, 1)
    want = np.array([[0, 1, 1, 0], [1, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0]], dtype=bool)
    assert (dense(g) == want).all() and g.edges == 2


@pytest.mark.parametrize("text", ['', '[]', '{"a": ["x"]', '{"a": "x"}', '{"a": ["x" "y"]}', '{"a": ["x",]}',
                                  '{"a" ["x"]}', '{"a": ["x], "b": []}', '{}'])
def test_malformed_files_rejected(cga, tmp_path, text):
    path = tmp_path / "bad.json"
    path.write_text(text)
    with pytest.raises(cga.CcError) as e:
        cga.load_tags(path, "jaccard", 0.5, 0, 1)
    assert e.value.code == cga.EINVAL


def test_unknown_similarity_and_missing_file_rejected(cga, tmp_path):
    path = write_tags(tmp_path / "tags.json", [["a"], ["a", "b"]])
    with pytest.raises(cga.CcError) as e:
        cga.load_tags(path, "euclid", 0.5, 0, 1)
    assert e.value.code == cga.EINVAL and "similarity" in str(e.value)
    with pytest.raises(cga.CcError) as e:
        cga.load_tags(tmp_path / "none.json", "jaccard", 0.5, 0, 1)
    assert e.value.code == cga.EINVAL
