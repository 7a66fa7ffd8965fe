"""The oracle of the incident-exclusion pairs (oracle/network_np.py) against the reference's own test rule and
against brute force, on the networks of the reference's tests (tests/test_network_numba.py:18-95): wavy ring
rows of 2x3 to 400x37 nodes, a coincident column of nodes, signed zeros, NaN and infinite coordinates, first
at 0 and at n, self-loop segments.  The ring-edge aggregation (pack_ring_edges) against a hand-derived lattice,
the closure of every ring (the circulation into each node equals the circulation out) and lone rings, on the
strengths of the same tests.  CPU only."""
import numpy as np
import pytest

from oracle import network_np as NP
from openfawe_amd import network as NET


def wavy_rows(rng, depth, width):
    """Rows of a wavy wake sheet, (depth, width, 3) (the shape of tests/test_network_numba.py:30-34)."""
    x = 0.7 * np.arange(depth)[:, None] * np.ones((1, width))
    y = np.ones((depth, 1)) * np.linspace(-3.0, 3.0, width)[None, :]
    z = 0.2 * np.sin(x) * np.cos(y) + 0.01 * rng.normal(size=x.shape)
    return np.stack((x, y, z), axis=-1)


def ring_network(rows):
    """Nodes and edges of ring rows: spanwise segments in every row, trailing segments between rows."""
    depth, width = rows.shape[:2]
    nodes = rows.reshape(-1, 3).copy()
    idx = np.arange(depth * width).reshape(depth, width)
    span = np.stack([idx[:, :-1], idx[:, 1:]], axis=-1).reshape(-1, 2)
    trail = np.stack([idx[:-1], idx[1:]], axis=-1).reshape(-1, 2)
    edges = np.concatenate([span, trail])
    return nodes, edges


def tuple_key_pairs(nodes, edges, first):
    """The rule of the reference's test (tests/test_network_numba.py:18-27): points grouped by their coordinate
    tuples.  Tuples compare by value, so -0.0 and 0.0 unite here; the cases compared with it have no signed
    zeros.  NaN never matches (a fresh float object is unequal to every other)."""
    groups = {}
    for index, point in enumerate(nodes[first:]):
        groups.setdefault(tuple(point), []).append(index)
    out = [(i, s) for s, ends in enumerate(edges) for node in ends for i in groups.get(tuple(nodes[node]), ())]
    return np.asarray(out, dtype=np.int64).reshape(-1, 2)


def brute_pairs(nodes, edges, first):
    """Every (point, segment) pair by direct comparison of the coordinates' bytes."""
    key = [nodes[i].tobytes() for i in range(len(nodes))]
    nan = np.isnan(nodes).any(axis=1)
    out = []
    for s, ends in enumerate(edges):
        for node in ends:
            if nan[node]:
                continue
            out.extend((i - first, s) for i in range(first, len(nodes)) if not nan[i] and key[i] == key[node])
    return np.asarray(out, dtype=np.int64).reshape(-1, 2)


def all_forms(nodes, edges, first):
    a = NP.incident_pairs_loop(nodes, edges, first)
    b = NP.incident_pairs(nodes, edges, first)
    np.testing.assert_array_equal(a, b)
    assert b.dtype == np.int64 and b.shape[1] == 2
    return b


@pytest.mark.parametrize("depth,width", [(2, 3), (3, 5), (12, 9), (41, 17), (400, 37)])
def test_ring_rows_match_the_reference_rule(depth, width):
    rng = np.random.default_rng(7 + depth)
    nodes, edges = ring_network(wavy_rows(rng, depth, width))
    first = width                      # the bound row is not a target
    pairs = all_forms(nodes, edges, first)
    np.testing.assert_array_equal(pairs, tuple_key_pairs(nodes, edges, first))
    # distinct nodes: every free end of every edge is one pair (tests/test_network_numba.py:93-95)
    assert len(pairs) == int(np.sum(edges >= width))
    assert np.all(np.diff(pairs[:, 1]) >= 0)


def test_coincident_column_and_scrambled_edges():
    rng = np.random.default_rng(3)
    rows = wavy_rows(rng, 9, 8)
    rows[:, 4, :] = rows[:, 3, :]               # node 4 coincides with node 3 in every row
    nodes, edges = ring_network(rows)
    edges = edges[rng.permutation(len(edges))]
    edges[::3] = edges[::3, ::-1]               # reversed segments
    pairs = all_forms(nodes, edges, 8)
    np.testing.assert_array_equal(pairs, tuple_key_pairs(nodes, edges, 8))
    np.testing.assert_array_equal(pairs, brute_pairs(nodes, edges, 8))
    assert len(pairs) > int(np.sum(edges >= 8))  # the coincident nodes add pairs


def test_signed_zeros_nan_and_infinity():
    nodes = np.array([[0.0, 1.0, 2.0],      # 0
                      [-0.0, 1.0, 2.0],     # 1: differs from node 0 in the sign bit only
                      [0.0, 1.0, 2.0],      # 2: same bits as node 0
                      [np.nan, 0.0, 0.0],   # 3
                      [np.nan, 0.0, 0.0],   # 4: the same NaN bits as node 3, still never matches
                      [np.inf, -np.inf, 1.0],   # 5
                      [np.inf, -np.inf, 1.0],   # 6: infinities match by bits
                      [5.0, 5.0, 5.0]])     # 7
    edges = np.array([[0, 1], [3, 4], [5, 7], [2, 2], [6, 3], [1, 0]])
    pairs = all_forms(nodes, edges, 0)
    np.testing.assert_array_equal(pairs, brute_pairs(nodes, edges, 0))
    expected = [(0, 0), (2, 0), (1, 0),              # segment 0: start 0 -> {0, 2}, end 1 -> {1}
                (5, 2), (6, 2), (7, 2),              # segment 2: start 5 -> {5, 6}, end 7 -> {7}
                (0, 3), (2, 3), (0, 3), (2, 3),      # segment 3 is a self-loop: the group twice
                (5, 4), (6, 4),                      # segment 4: NaN end gives nothing
                (1, 5), (0, 5), (2, 5)]
    np.testing.assert_array_equal(pairs, np.array(expected))
    # the tuple rule unites the signed zeros, the bit rule does not
    assert len(tuple_key_pairs(nodes, edges, 0)) > len(pairs)
    # points before `first` are never listed, indices count from `first`
    tail = all_forms(nodes, edges, 2)
    np.testing.assert_array_equal(tail, brute_pairs(nodes, edges, 2))
    assert tail[:, 0].min() >= 0 and tail[:, 0].max() < len(nodes) - 2


def test_empty_cases():
    rng = np.random.default_rng(5)
    nodes, edges = ring_network(wavy_rows(rng, 4, 5))
    assert all_forms(nodes, edges, len(nodes)).shape == (0, 2)
    assert all_forms(nodes, edges[:0], 0).shape == (0, 2)
    assert all_forms(nodes[:0], np.zeros((0, 2), np.int64), 0).shape == (0, 2)


@pytest.mark.parametrize("seed", range(6))
def test_heavy_coincidence_against_brute_force(seed):
    rng = np.random.default_rng(seed)
    palette = np.array([[0.0, 0.0, 0.0], [-0.0, 0.0, 0.0], [1.0, 2.0, 3.0], [np.nan, 1.0, 1.0],
                        [np.inf, 0.0, 1.0], [1.0, 2.0, 3.0 + 2 ** -50]])
    nodes = palette[rng.integers(0, len(palette), 80)]
    edges = rng.integers(0, 80, size=(60, 2))
    for first in (0, 1, 37, 80):
        pairs = all_forms(nodes, edges, first)
        np.testing.assert_array_equal(pairs, brute_pairs(nodes, edges, first))


def test_host_validation_before_any_device_work():
    nodes = np.zeros((4, 3))
    edges = np.array([[0, 1]])
    for args in ((np.zeros((4, 2)), edges, 0), (nodes, np.array([0, 1]), 0), (nodes, edges, 5),
                 (nodes, edges, -1), (nodes, np.array([[0, 4]]), 0), (nodes, np.array([[-1, 0]]), 0)):
        with pytest.raises(ValueError):
            NET.incident_pairs(*args)
        with pytest.raises(ValueError):
            NP.incident_pairs(*args)


# ------------------------------------------------------------------ ring-edge aggregation (pack_ring_edges)
def ring_cases():
    """The strengths of the reference's pack_rings tests (tests/test_network_numba.py:52-95): repeated and
    cancelling strengths, an exact zero, a scrambled panel list with a reversed panel, a 400-row decaying history
    with zero rows."""
    rng = np.random.default_rng(7)
    for depth, width in ((2, 3), (3, 5), (12, 9), (41, 17)):
        strengths = rng.normal(size=(depth - 1, width - 1))
        strengths[1:, :] = strengths[:-1, :] * rng.choice([1.0, -1.0, 0.5], size=(depth - 2, width - 1))
        strengths[0, 0] = 0.0
        panels = np.column_stack([np.arange(width - 1), np.arange(1, width)])
        yield f"random-{depth}x{width}", depth, width, panels, strengths
    panels = np.array([[0, 1], [2, 1], [2, 3], [4, 5], [6, 5], [6, 7]])
    yield "scrambled", 9, 8, panels, np.random.default_rng(3).normal(size=(8, len(panels)))
    base = np.random.default_rng(11).normal(size=36)
    strengths = base[None, :] * np.exp(-0.01 * np.arange(399))[:, None]
    strengths[::7] = 0.0
    yield "history-400x37", 400, 37, np.column_stack([np.arange(36), np.arange(1, 37)]), strengths


def both_ring_forms(depth, width, panels, strengths):
    e1, g1 = NP.pack_ring_edges_loop(depth, width, panels, strengths)
    e2, g2 = NP.pack_ring_edges(depth, width, panels, strengths)
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(g1.view(np.int64), g2.view(np.int64))
    assert e2.dtype == np.int64 and g2.dtype == np.float64
    return e2, g2


def test_ring_edges_known_answer():
    # 2 x 2 rings on 3 x 3 nodes (ids row * 3 + column); strengths a=1, b=2 (row 0), c=1, d=0.5 (row 1):
    # the edge (3,4) between rows 0 and 1 of panel 0 gets -a + c = 0 and is dropped.
    edges, gamma = both_ring_forms(3, 3, np.array([[0, 1], [1, 2]]), np.array([[1.0, 2.0], [1.0, 0.5]]))
    assert edges.tolist() == [[0, 1], [1, 4], [0, 3], [1, 2], [2, 5], [4, 5], [4, 7], [6, 7], [3, 6], [5, 8],
                              [7, 8]]
    assert gamma.tolist() == [1.0, -1.0, -1.0, 2.0, 2.0, -1.5, 0.5, -1.0, -1.0, 0.5, -0.5]


@pytest.mark.parametrize("case", list(ring_cases()), ids=lambda c: c[0])
def test_ring_edges_forms_agree_and_rings_close(case):
    _, depth, width, panels, strengths = case
    edges, gamma = both_ring_forms(depth, width, panels, strengths)
    assert np.all(edges[:, 0] < edges[:, 1]) and np.all(gamma != 0.0)
    assert len(np.unique(edges, axis=0)) == len(edges)
    # every ring is closed, so the circulation entering each node equals the circulation leaving it
    balance = np.zeros(depth * width)
    np.add.at(balance, edges[:, 0], gamma)
    np.add.at(balance, edges[:, 1], -gamma)
    assert np.abs(balance).max() <= 1e-12 * max(1.0, np.abs(strengths).max())
    # a lone ring (one nonzero strength) gives exactly its four edges with +-strength
    one = np.zeros_like(strengths)
    one[-1, -1] = 1.25
    e1, g1 = both_ring_forms(depth, width, panels, one)
    assert len(e1) == 4 and sorted(np.abs(g1).tolist()) == [1.25] * 4


def test_ring_edges_host_validation():
    rows = np.zeros((3, 4, 3))
    rows[..., 0] = np.arange(4)
    good = np.ones((2, 3))
    for args in ((rows[:1], good[:0]), (rows[..., :2], good), (rows, good[:, :2]), (rows, good, [[0, 4]]),
                 (rows, good[:, :1], [[1, 1]]), (rows, good[:, :1], [[0.0, 1.0]]), (rows, np.full((2, 3), np.inf))):
        with pytest.raises(ValueError):
            NET.pack_rings(*args)
