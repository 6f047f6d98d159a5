"""Input formats on the host: graph-tool .gt files and delimited edge lists,
against the facts the reference's own tests assert (crates/gt-io/tests/
u27_gt.rs, u29_csv.rs).  No GPU needed."""
from pathlib import Path

import numpy as np
import pytest

import gt_hip

GOLD = Path(__file__).resolve().parent / "golden"


def _degree(doc, v):
    s, t = doc.edges()
    return int((s == v).sum() + (t == v).sum())


def test_karate():
    doc = gt_hip.read_gt(GOLD / "karate.gt")
    assert (doc.n, doc.m, doc.directed) == (34, 78, False)
    assert "34 vertices, 78 edges, undirected" in doc.comment
    props = {p["name"]: p for p in doc.properties()}
    assert len(props) == 3
    assert props["description"]["kind"] == "string" and props["description"]["domain"] == "graph"
    assert props["readme"]["domain"] == "graph"
    assert props["pos"]["kind"] == "vector<double>" and props["pos"]["domain"] == "vertex"
    assert len(props["pos"]["values"]) == 34 and len(props["pos"]["values"][0]) == 2
    assert _degree(doc, 0) == 16  # the instructor


def test_lesmis_and_its_edge_weights():
    doc = gt_hip.read_gt(GOLD / "lesmis.gt")
    assert (doc.n, doc.m, doc.directed) == (77, 254, False)
    props = doc.properties()
    assert [(p["name"], p["domain"], p["kind"]) for p in props] == [
        ("description", "graph", "string"), ("readme", "graph", "string"),
        ("label", "vertex", "string"), ("pos", "vertex", "vector<double>"),
        ("value", "edge", "double")]
    labels, w = props[2]["values"], props[4]["values"]
    assert labels[0] == "Myriel" and "Valjean" in labels
    assert len(w) == 254
    s, t = doc.edges()
    # the first adjacency entry is Napoleon(1) -> Myriel(0), weight 1
    assert (s[0], t[0], w[0]) == (1, 0, 1.0)
    assert int((s == 0).sum()) == 0 and int((t == 0).sum()) == 10
    heaviest = int(np.argmax(w))
    assert w[heaviest] == 31.0 and (s[heaviest], t[heaviest]) == (26, 11)


def test_truncated_and_corrupt_gt_files_are_errors():
    b = (GOLD / "karate.gt").read_bytes()
    for cut in list(range(0, 80)) + list(range(80, len(b), 37)):
        with pytest.raises(ValueError):
            gt_hip.read_gt(b[:cut])
    with pytest.raises(ValueError, match="magic"):
        gt_hip.read_gt(b"not a graph")
    bad = bytearray(b)
    bad[6] = 2  # version
    with pytest.raises(ValueError, match="version"):
        gt_hip.read_gt(bytes(bad))


def test_adjlist_replay_matches_the_file_order():
    doc = gt_hip.read_gt(GOLD / "lesmis.gt")
    g = doc.adjlist()
    assert (g.num_vertices(), g.num_edges()) == (77, 254)
    s, t = g.edges()
    ds, dt = doc.edges()
    assert np.array_equal(s, ds) and np.array_equal(t, dt)


def _el(text, **kw):
    return gt_hip.read_edge_list(text.encode() if isinstance(text, str) else text, **kw)


def test_symbolic_endpoints_intern_in_first_seen_order():
    d = _el("b,a\na,c\nc,b\n")
    assert d.names() == ["b", "a", "c"]
    s, t = d.edges()
    assert list(zip(s, t)) == [(0, 1), (1, 2), (2, 0)]
    d = _el("x,x\n")  # a self-loop interns one vertex
    assert d.names() == ["x"] and d.n == 1


def test_interning_is_by_stripped_unquoted_value():
    d = _el('" a ",b\na,"b"\n')
    assert d.names() == ["a", "b"]
    assert d.n == 2


def test_unicode_whitespace_is_stripped_like_str_trim():
    d = _el("\u00a0a\u3000,b\u2028\n")
    assert d.names() == ["a", "b"]
    d = _el("+3,0\n", symbolic=False)
    assert d.n == 4


def test_empty_string_is_a_name_and_header_is_skipped():
    d = _el('"",x\n')
    assert d.names() == ["", "x"]
    d = _el("src,dst\nu,v\n", header=True)
    assert d.names() == ["u", "v"] and d.m == 1


def test_index_edge_list_grows_to_the_largest_endpoint():
    d = _el("0,5\n2,1\n", symbolic=False)
    assert (d.n, d.m) == (6, 2)
    with pytest.raises(ValueError, match="line 2"):
        _el("0,1\nx,2\n", symbolic=False)
    with pytest.raises(ValueError, match="line 2"):
        _el("0,1\n-1,2\n", symbolic=False)


def test_records_blank_lines_bom_and_extra_columns():
    d = _el("a,b\n\n   \nb,c")  # blank lines skipped, trailing record kept
    assert d.m == 2
    d = _el(b"\xef\xbb\xbfa,b\n")
    assert d.names() == ["a", "b"]
    d = _el('a,b,"x\ny",z\nc,d,1\n')  # a quoted third field may hold a newline
    assert d.m == 2 and d.names() == ["a", "b", "c", "d"]
    d = _el('a;b;c\n', delimiter=";")
    assert d.names() == ["a", "b"]


def test_malformed_edge_lists_name_their_line():
    with pytest.raises(ValueError, match="line 3"):
        _el("a,b\nc,d\nlonely\n")
    with pytest.raises(ValueError, match="unterminated"):
        _el('a,"b\n')
    with pytest.raises(ValueError, match="UTF-8"):
        _el(b"a,\xff\xfe\n")


def test_record_order_fixes_the_edge_ids():
    rng = np.random.default_rng(3)
    pairs = rng.integers(0, 500, size=(5000, 2))
    text = "\n".join(f"{a},{b}" for a, b in pairs)
    d = _el(text, symbolic=False)
    s, t = d.edges()
    assert np.array_equal(s, pairs[:, 0]) and np.array_equal(t, pairs[:, 1])
