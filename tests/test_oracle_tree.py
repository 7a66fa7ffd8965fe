"""The oracle's tree code pinned by the reference's own checks (tests/test_vortex_tree.py,
tests/test_cuda_build.py) and by what the mathematics fixes.

* build: the level-parallel build with renumbering equals the depth-first serial build element for
  element (the reference's T1 check, vortex_tree.py:208-214), a pure-Python restatement of the serial
  build, and structural invariants;
* moments: the exact line integrals of one segment (quadrature), the parallel-axis identity between a
  node and its children, closed rings have no monopole;
* far field: contracted form equals the term-by-term form; ring dipole limits;
* traversal: convergence in theta and order, theta -> 0 equals the direct sum, no excluded segment
  leaks into a node expansion, thread-count independence.
"""
import numpy as np
import pytest

from oracle import oracle as O
from openfawe_amd import workloads
from tests.helpers import midpoint_cases, oracle_direct, oracle_tree, prepare


# ------------------------------------------------------------------ build
def python_serial_build(mid, leaf):
    """vortex_tree.build_tree_serial restated with Python's stable sort (small inputs only)."""
    n = len(mid)
    order = list(range(n))
    lo, hi, left, right = [0], [n], [-1], [-1]
    stack = [0]
    while stack:
        node = stack.pop()
        a, b = lo[node], hi[node]
        if b - a <= leaf:
            continue
        pts = [mid[order[k]] for k in range(a, b)]
        low = [min(p[c] for p in pts) for c in range(3)]
        high = [max(p[c] for p in pts) for c in range(3)]
        axis, extent = 0, high[0] - low[0]
        for c in (1, 2):
            if high[c] - low[c] > extent:
                axis, extent = c, high[c] - low[c]
        if not extent > 0.0:
            continue
        order[a:b] = sorted(order[a:b], key=lambda s: mid[s][axis])
        middle = (a + b) // 2
        first = len(lo)
        lo += [a, middle]
        hi += [middle, b]
        left += [-1, -1]
        right += [-1, -1]
        left[node], right[node] = first, first + 1
        stack += [first + 1, first]
    return np.array(order), np.array(lo), np.array(hi), np.array(left), np.array(right), len(lo)


def check_structure(mid, leaf, tree):
    order, lo, hi, left, right, count = tree
    m = len(mid)
    assert sorted(order.tolist()) == list(range(m))
    assert lo[0] == 0 and hi[0] == m
    internal = 0
    stack = [0]
    while stack:                                   # pre-order, left first
        node = stack.pop()
        a, b = lo[node], hi[node]
        if left[node] < 0:
            assert right[node] < 0
            ext = mid[order[a:b]].max(axis=0) - mid[order[a:b]].min(axis=0) if b > a else np.zeros(3)
            assert b - a <= leaf or not ext.max() > 0
            continue
        # numbering rule of the serial build: the j-th split node in pre-order has children 2j+1, 2j+2
        assert (left[node], right[node]) == (2 * internal + 1, 2 * internal + 2)
        internal += 1
        middle = (a + b) // 2
        assert (lo[left[node]], hi[left[node]], lo[right[node]], hi[right[node]]) == (a, middle, middle, b)
        ext = mid[order[a:b]].max(axis=0) - mid[order[a:b]].min(axis=0)
        axis = int(np.argmax(ext))                 # first maximum: the lower axis on ties
        assert b - a > leaf and ext[axis] > 0
        assert mid[order[a:middle], axis].max() <= mid[order[middle:b], axis].min()
        stack += [right[node], left[node]]
    assert count == 2 * internal + 1


@pytest.mark.parametrize("case", midpoint_cases(), ids=lambda c: c[0])
def test_build_serial_equals_level_equals_python(case):
    name, mid, leaf = case
    level = O.build_tree(mid, leaf)
    serial = O.build_tree_serial(mid, leaf)
    for a, b in zip(level[:5], serial[:5]):
        assert np.array_equal(a, b)
    assert level[5] == serial[5]
    check_structure(mid, leaf, level)
    if len(mid) <= 2100:
        py = python_serial_build(mid, leaf)
        for a, b in zip(level[:5], py[:5]):
            assert np.array_equal(a, b)
        assert py[5] == level[5]


def test_build_negative_zero_ties():
    """np.argsort treats -0.0 == 0.0: such keys are ties kept in input order."""
    mid = np.zeros((9, 3))
    mid[:, 0] = [0.0, -0.0, 1.0, -0.0, 0.0, 2.0, -1.0, 0.0, -0.0]
    tree = O.build_tree(mid, 2)
    py = python_serial_build(mid, 2)
    for a, b in zip(tree[:5], py[:5]):
        assert np.array_equal(a, b)


def test_build_sheet_and_loop_thread_independent():
    case = workloads.sheet(60)
    _, _, _, mid = O.segment_arrays(case["starts"], case["directions"], case["lengths"], case["gamma"])
    one = O.build_tree(mid, 32, threads=1)
    many = O.build_tree(mid, 32, threads=8)
    serial = O.build_tree_serial(mid, 32)
    for a, b, c in zip(one[:5], many[:5], serial[:5]):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    check_structure(mid, 32, one)


# ------------------------------------------------------------------ moments
def moments_of(starts, directions, lengths, gamma, radius, leaf):
    lv, ends, alpha, mid = O.segment_arrays(starts, directions, lengths, gamma)
    order, lo, hi, left, right, _ = O.build_tree(mid, leaf)
    return (order, lo, hi, left, right), O.node_moments(order, lo, hi, starts, ends, alpha, lv, radius, gamma)


def test_single_segment_moments_are_exact_line_integrals():
    s = np.array([[0.3, -1.2, 0.5]])
    d = np.array([[1.0, 2.0, -0.5]])
    d /= np.linalg.norm(d)
    length, gamma = np.array([0.7]), np.array([1.9])
    _, (centre, bound, core, m0, m1, m2) = moments_of(s, d, length, gamma, np.array([0.05]), 8)
    L = d[0] * length[0]
    e = s[0] + L
    np.testing.assert_array_equal(centre[0], 0.5 * (np.minimum(s[0], e) + np.maximum(s[0], e)))
    # Gauss-Legendre line integrals of Gamma * L_i (x - c)_j (x - c)_k over the segment
    x, w = np.polynomial.legendre.leggauss(8)
    pts = s[0] + ((x + 1) / 2)[:, None] * L - centre[0]
    q0 = gamma[0] * L
    q1 = gamma[0] * np.einsum("i,qj,q->ij", L, pts, w / 2)
    q2 = gamma[0] * np.einsum("i,qj,qk,q->ijk", L, pts, pts, w / 2)
    np.testing.assert_allclose(m0[0], q0, rtol=1e-14)
    np.testing.assert_allclose(m1[0], q1, atol=1e-15 * np.abs(q0).max())
    np.testing.assert_allclose(m2[0], q2, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(bound[0], 0.5 * length[0] * (1 + 1e-12), rtol=1e-15)
    assert core[0] == 0.05


def test_parallel_axis_identity_and_symmetry():
    starts, directions, lengths, gamma, radius = (workloads.sheet(12)[k] for k in
                                                  ("starts", "directions", "lengths", "gamma", "radius"))
    gamma = gamma * np.cos(np.arange(len(gamma)))
    gamma[::5] = 0.0
    (order, lo, hi, left, right), (centre, bound, core, m0, m1, m2) = moments_of(
        starts, directions, lengths, gamma, radius, 16)
    assert np.array_equal(m2, m2.transpose(0, 1, 3, 2))        # exactly symmetric in the last two indices
    for node in np.flatnonzero(left >= 0):
        kids = (left[node], right[node])
        d = [centre[k] - centre[node] for k in kids]
        s0 = sum(m0[k] for k in kids)
        s1 = sum(m1[k] + np.outer(m0[k], dk) for k, dk in zip(kids, d))
        s2 = sum(m2[k] + np.einsum("ij,k->ijk", m1[k], dk) + np.einsum("ik,j->ijk", m1[k], dk)
                 + np.einsum("i,j,k->ijk", m0[k], dk, dk) for k, dk in zip(kids, d))
        scale = max(np.abs(m0[node]).max(), 1e-300)
        np.testing.assert_allclose(s0, m0[node], atol=1e-12 * scale)
        np.testing.assert_allclose(s1, m1[node], atol=1e-12 * max(np.abs(m1[node]).max(), scale))
        np.testing.assert_allclose(s2, m2[node], atol=1e-11 * max(np.abs(m2[node]).max(), scale))
        assert core[node] == max(core[k] for k in kids)
        # every end point of the node lies in its bounding sphere
        seg = order[lo[node]:hi[node]]
        ends = starts[seg] + directions[seg] * lengths[seg, None]
        for p in (starts[seg], ends):
            assert np.all(np.linalg.norm(p - centre[node], axis=1) <= bound[node])


def test_closed_ring_has_no_monopole():
    nodes, edges = workloads.ring(count=64)
    s, d, L = prepare(nodes, edges)
    _, (_, _, _, m0, m1, _) = moments_of(s, d, L, np.ones(64), np.zeros(64), 200)
    assert np.abs(m0[0]).max() < 1e-14
    assert np.abs(m1[0]).max() > 0.1


# ------------------------------------------------------------------ far field
def test_contracted_expansion_equals_term_by_term_form():
    rng = np.random.default_rng(5)
    m0 = rng.normal(size=(7, 3))
    m1 = rng.normal(size=(7, 3, 3))
    m2 = rng.normal(size=(7, 3, 3, 3))
    m2 = .5 * (m2 + m2.transpose(0, 1, 3, 2))
    w1, t2, w2 = O.node_contractions(m1, m2)
    for node in range(7):
        for order in (0, 1, 2):
            r = rng.normal(size=3) * 3.
            expected = O.multipole_explicit(r, m0[node], m1[node], m2[node], order)
            actual = O.far_velocity(r, m0[node], m1[node], m2[node], w1[node], t2[node], w2[node], order)
            np.testing.assert_allclose(actual, expected, rtol=1e-12, atol=1e-14 * np.abs(expected).max())


def test_ring_dipole_far_field():
    nodes, edges = workloads.ring(count=128)
    far = np.array([[6., 1., 4.], [0., 0., 9.], [-7., 3., -2.]])
    direct, _ = oracle_direct(far, nodes, edges, 1.0)
    scale = np.linalg.norm(direct, axis=1)
    errors = []
    for order in (0, 1, 2):
        out, _, _, _ = oracle_tree(far, nodes, edges, 1.0, theta=.5, order=order, leaf_size=200)
        errors.append(np.linalg.norm(out - direct, axis=1) / scale)
    np.testing.assert_allclose(errors[0], 1., atol=1e-12)
    assert np.all(errors[1] < 2. / np.linalg.norm(far, axis=1) ** 2)
    np.testing.assert_allclose(errors[2], errors[1], rtol=1e-6)
    angle = np.linspace(0., 2 * np.pi, 96, endpoint=False)
    loop = np.column_stack((1.4 * np.cos(angle) + .3, .8 * np.sin(angle), .4 * np.sin(2 * angle) + .2 * np.cos(angle)))
    edges = np.column_stack((np.arange(96), (np.arange(96) + 1) % 96))
    direct, _ = oracle_direct(far, loop, edges, 1.0)
    first, second = [np.max(np.linalg.norm(oracle_tree(far, loop, edges, 1.0, theta=.5, order=o, leaf_size=200)[0]
                                           - direct, axis=1) / np.linalg.norm(direct, axis=1)) for o in (1, 2)]
    assert second < first / 3


# ------------------------------------------------------------------ traversal
def test_converges_to_direct_with_theta_and_order():
    nodes, edges, gamma = workloads.lattice()
    rng = np.random.default_rng(2)
    points = nodes + .07 * rng.normal(size=nodes.shape)
    direct, _ = oracle_direct(points, nodes, edges, gamma, 2, .1)
    scale = np.max(np.linalg.norm(direct, axis=1))
    errors = {}
    for order in (0, 1, 2):
        for theta in (.6, .4, .25, .15):
            out = oracle_tree(points, nodes, edges, gamma, 2, .1, theta=theta, order=order, leaf_size=8)[0]
            errors[order, theta] = np.max(np.linalg.norm(out - direct, axis=1)) / scale
    for order in (0, 1, 2):
        seq = [errors[order, t] for t in (.6, .4, .25, .15)]
        assert all(b < a for a, b in zip(seq, seq[1:])), (order, seq)
    for theta in (.4, .25, .15):
        assert errors[2, theta] < errors[1, theta] < errors[0, theta]
    assert errors[2, .15] < 2e-4


def test_tiny_theta_is_the_direct_sum():
    nodes, edges, gamma = workloads.lattice(rows=10, span=9)
    points = nodes + [0., 0., .2]
    direct, _ = oracle_direct(points, nodes, edges, gamma, 1, .05)
    out, _, far, _ = oracle_tree(points, nodes, edges, gamma, 1, .05, theta=1e-9, order=2, leaf_size=4)
    scale = np.max(np.linalg.norm(direct, axis=1))
    np.testing.assert_allclose(out, direct, rtol=0, atol=1e-13 * scale)
    assert far.sum() == 0


def test_exclusions_never_leak():
    nodes, edges, gamma = workloads.lattice(rows=12, span=9)
    incident = np.asarray([(n, s) for s, (a, b) in enumerate(edges) for n in (a, b)], dtype=np.int64)
    direct, _ = oracle_direct(nodes, nodes, edges, gamma, 2, .1, incident)
    exact = oracle_tree(nodes, nodes, edges, gamma, 2, .1, incident, theta=1e-9, order=2, leaf_size=4)[0]
    scale = np.max(np.linalg.norm(direct, axis=1))
    np.testing.assert_allclose(exact, direct, rtol=0, atol=1e-13 * scale)
    approx = oracle_tree(nodes, nodes, edges, gamma, 2, .1, incident, theta=.3, order=2, leaf_size=4)[0]
    assert np.max(np.linalg.norm(approx - direct, axis=1)) / scale < 1e-2
    loud, silent = gamma.copy(), gamma.copy()
    loud[-1], silent[-1] = 1e6, 0.0
    far_pair = np.array([[0, len(edges) - 1]], dtype=np.int64)
    kw = dict(singularity="zero", theta=.6, order=2, leaf_size=4)
    tree = oracle_tree(nodes[:1], nodes, edges, loud, exclusions=far_pair, **kw)[0]
    quiet = oracle_tree(nodes[:1], nodes, edges, silent, exclusions=far_pair, **kw)[0]
    assert np.array_equal(tree, quiet)
    leaked = oracle_tree(nodes[:1], nodes, edges, loud, **kw)[0]
    assert np.linalg.norm(leaked - quiet) > 1e3 * np.linalg.norm(quiet)


def test_tree_thread_count_bitwise():
    nodes, edges, gamma = workloads.lattice()
    points = nodes + [0., 0., .3]
    s, d, L = prepare(nodes, edges)
    radius = np.full(len(s), .1)
    runs = [O.tree_sum(points, s, d, L, gamma, radius, 2, None, "raise", .35, 2, 16, 3.0, threads=t)[0]
            for t in (1, 8)]
    assert np.array_equal(runs[0].view(np.uint64), runs[1].view(np.uint64))


def test_singular_point_flags_in_tree():
    nodes, edges, gamma = workloads.lattice(rows=6, span=5)
    point = .5 * (nodes[edges[3, 0]] + nodes[edges[3, 1]])[None]
    _, flags, _, _ = oracle_tree(point, nodes, edges, gamma, theta=.4, order=2, leaf_size=2)
    assert flags[0] == O.FLAG_SINGULAR
