"""Labeled graphs on the host (CPU only): <prefix>.vlabel.bin / .elabel.bin
loading with the checks graph.cc:37-82 asserts, the missing-file fallback,
the writer's round trip, the label frequency of Graph::computeLabelsFrequency
(graph.cc:440-457) and the neighbourhood label frequency of Graph::BuildNLF
(graph.cc:483-495).  The reference ships no labeled dataset or label test,
so the expected values come from numpy on the same arrays and from small
graphs worked by hand."""
from collections import Counter
from pathlib import Path

import numpy as np
import pytest

import mercury_amd as ma
from conftest import csr_from_edges

ERR_IO, ERR_FORMAT, ERR_INVALID = -1, -2, -3


def _meta(prefix: Path) -> list[int]:
    return [int(x) for x in Path(f"{prefix}.meta.txt").read_text().split()]


def _set_meta(prefix: Path, nvc: int, nec: int) -> None:
    f = _meta(prefix)
    f[8], f[9] = nvc, nec
    Path(f"{prefix}.meta.txt").write_text(" ".join(map(str, f)) + "\n")


@pytest.fixture
def rmat(tmp_path):
    """RMAT-9 with labels 1..6 on vertices and 1..4 on CSR entries, written
    to disk; returns (prefix, graph, vlabels, elabels)."""
    g = ma.Graph.rmat(9, 8, seed=3)
    rng = np.random.default_rng(5)
    v = rng.integers(1, 7, size=g.n).astype(np.uint8)
    e = rng.integers(1, 5, size=g.nnz).astype(np.uint16)
    g.set_labels(v, 6, e, 4)
    prefix = tmp_path / "graph"
    g.write(str(prefix))
    return prefix, g, v, e


def test_round_trip(rmat):
    prefix, g, v, e = rmat
    f = _meta(prefix)
    assert f[8:] == [6, 4]
    assert Path(f"{prefix}.vlabel.bin").stat().st_size == g.n
    assert Path(f"{prefix}.elabel.bin").stat().st_size == 2 * g.nnz
    h = ma.Graph.load(str(prefix), vlabels=True, elabels=True)
    hv, nv, he, ne = h.labels()
    assert (nv, ne) == (6, 4)
    assert np.array_equal(hv, v) and np.array_equal(he, e)
    # the CSR is unchanged by the labels
    for a, b in zip(h.csr(), g.csr()):
        assert np.array_equal(a, b)
    # one kind only, or none
    assert ma.Graph.load(str(prefix)).labels() == (None, 0, None, 0)
    hv, nv, he, ne = ma.Graph.load(str(prefix), vlabels=True).labels()
    assert nv == 6 and he is None and ne == 0
    hv, nv, he, ne = ma.Graph.load(str(prefix), elabels=True).labels()
    assert hv is None and nv == 0 and np.array_equal(he, e)


def test_unlabeled_write_keeps_zero_classes(tmp_path):
    g = ma.Graph.rmat(6, 4, seed=1)
    g.write(str(tmp_path / "g"))
    assert _meta(tmp_path / "g")[7:] == [0, 0, 0]
    assert not (tmp_path / "g.vlabel.bin").exists()
    # asking for labels of a graph whose meta has no classes is an error
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.load(str(tmp_path / "g"), vlabels=True)
    assert ei.value.code == ERR_FORMAT


@pytest.mark.parametrize("nvc", [0, -1, 255, 300])
def test_vertex_classes_out_of_range(rmat, nvc):
    prefix = rmat[0]
    _set_meta(prefix, nvc, 4)
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.load(str(prefix), vlabels=True)
    assert ei.value.code == ERR_FORMAT
    # edge labels alone still load
    assert ma.Graph.load(str(prefix), elabels=True).labels()[3] == 4


@pytest.mark.parametrize("nec", [0, 65535, 70000])
def test_edge_classes_out_of_range(rmat, nec):
    prefix = rmat[0]
    _set_meta(prefix, 6, nec)
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.load(str(prefix), elabels=True)
    assert ei.value.code == ERR_FORMAT


def test_vertex_label_count_must_equal_classes(rmat):
    # graph.cc:46-47: the file holds exactly num_vertex_classes labels
    prefix = rmat[0]
    for nvc in (5, 7, 254):
        _set_meta(prefix, nvc, 4)
        with pytest.raises(ma.MgpmError) as ei:
            ma.Graph.load(str(prefix), vlabels=True)
        assert ei.value.code == ERR_FORMAT, nvc


def test_edge_label_count_at_most_classes(rmat):
    # graph.cc:68-69: at most num_edge_classes distinct edge labels
    prefix = rmat[0]
    _set_meta(prefix, 6, 3)
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.load(str(prefix), elabels=True)
    assert ei.value.code == ERR_FORMAT
    for nec in (4, 5, 1000):
        _set_meta(prefix, 6, nec)
        assert ma.Graph.load(str(prefix), elabels=True).labels()[3] == nec


def test_short_label_files(rmat):
    prefix, g = rmat[0], rmat[1]
    vb = Path(f"{prefix}.vlabel.bin")
    vb.write_bytes(vb.read_bytes()[:-1])
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.load(str(prefix), vlabels=True)
    assert ei.value.code == ERR_IO
    eb = Path(f"{prefix}.elabel.bin")
    eb.write_bytes(eb.read_bytes()[:-2])
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.load(str(prefix), elabels=True)
    assert ei.value.code == ERR_IO


def test_missing_label_files_get_generated_labels(rmat):
    # graph.cc:49-55, 71-77: labels in 1..classes when the file is absent;
    # generated labels are not checked against the class count
    prefix, g = rmat[0], rmat[1]
    Path(f"{prefix}.vlabel.bin").unlink()
    Path(f"{prefix}.elabel.bin").unlink()
    _set_meta(prefix, 200, 60000)
    a = ma.Graph.load(str(prefix), vlabels=True, elabels=True).labels()
    b = ma.Graph.load(str(prefix), vlabels=True, elabels=True).labels()
    v, nv, e, ne = a
    assert (nv, ne) == (200, 60000)
    assert v.min() >= 1 and v.max() <= 200 and len(v) == g.n
    assert e.min() >= 1 and e.max() <= 60000 and len(e) == g.nnz
    assert np.array_equal(v, b[0]) and np.array_equal(e, b[2])
    # spread over the classes, not constant
    assert len(np.unique(v)) > 100 and len(np.unique(e)) > 1000
    _set_meta(prefix, 1, 1)
    v, _, e, _ = ma.Graph.load(str(prefix), vlabels=True,
                               elabels=True).labels()
    assert set(v.tolist()) == {1} and set(e.tolist()) == {1}


def test_set_labels_checks(tmp_path):
    g = ma.Graph.rmat(6, 4, seed=2)
    v = (np.arange(g.n) % 3 + 1).astype(np.uint8)
    with pytest.raises(ma.MgpmError) as ei:
        g.set_labels(v, 4)  # three distinct labels, four classes
    assert ei.value.code == ERR_FORMAT
    with pytest.raises(ma.MgpmError) as ei:
        g.set_labels(v, 255)
    assert ei.value.code == ERR_FORMAT
    assert g.labels() == (None, 0, None, 0)  # a failed set changes nothing
    g.set_labels(v, 3)
    assert g.labels()[1] == 3
    e = np.full(g.nnz, 9, np.uint16)
    with pytest.raises(ma.MgpmError) as ei:
        g.set_labels(v, 3, e, 0)
    assert ei.value.code == ERR_FORMAT
    g.set_labels(v, 3, e, 1)
    assert g.labels()[3] == 1
    g.set_labels()  # removes both kinds
    assert g.labels() == (None, 0, None, 0)
    with pytest.raises(ValueError):
        g.set_labels(v[:-1], 3)


def test_label_frequency(rmat):
    g, v = rmat[1], rmat[2]
    f = g.label_frequency()
    assert len(f) == 7
    assert np.array_equal(f, np.bincount(v, minlength=7))
    assert f[0] == 0 and f.sum() == g.n
    # a label above num_vertex_classes: the reference asserts
    h = ma.Graph.rmat(5, 4, seed=1)
    h.set_labels(np.where(np.arange(h.n) % 2, 7, 9).astype(np.uint8), 2)
    with pytest.raises(ma.MgpmError) as ei:
        h.label_frequency()
    assert ei.value.code == ERR_FORMAT
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.rmat(5, 4, seed=1).label_frequency()
    assert ei.value.code == ERR_INVALID


def test_nlf_small_by_hand():
    # path 0-1-2-3 plus 1-3, labels 1 2 1 2:
    #   N(0)={1}: {2:1}; N(1)={0,2,3}: {1:2, 2:1}; N(2)={1,3}: {2:2};
    #   N(3)={1,2}: {1:1, 2:1}
    rp, ci = csr_from_edges(4, [(0, 1), (1, 2), (2, 3), (1, 3)])
    g = ma.Graph.from_csr(rp, ci)
    g.set_labels(np.array([1, 2, 1, 2], np.uint8), 2)
    off, lab, cnt = g.nlf()
    got = [dict(zip(lab[off[v]:off[v + 1]].tolist(),
                    cnt[off[v]:off[v + 1]].tolist())) for v in range(4)]
    assert got == [{2: 1}, {1: 2, 2: 1}, {2: 2}, {1: 1, 2: 1}]
    # isolated vertices have an empty NLF
    rp, ci = csr_from_edges(3, [(0, 1)])
    g = ma.Graph.from_csr(rp, ci)
    g.set_labels(np.array([1, 1, 1], np.uint8), 1)
    off, lab, cnt = g.nlf()
    assert off.tolist() == [0, 1, 2, 2] and cnt.tolist() == [1, 1]


def test_nlf_rmat(rmat):
    g, v = rmat[1], rmat[2]
    rp, ci = g.csr()
    off, lab, cnt = g.nlf()
    assert off[0] == 0 and len(off) == g.n + 1
    deg = np.diff(rp)
    for u in range(g.n):
        want = Counter(v[ci[rp[u]:rp[u + 1]]].tolist())
        ls = lab[off[u]:off[u + 1]].tolist()
        assert ls == sorted(want)
        assert dict(zip(ls, cnt[off[u]:off[u + 1]].tolist())) == want
    # every neighbour is counted once
    per_v = np.add.reduceat(cnt, off[:-1][deg > 0]) if len(cnt) else []
    assert np.array_equal(per_v, deg[deg > 0])
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.rmat(5, 4, seed=1).nlf()
    assert ei.value.code == ERR_INVALID
