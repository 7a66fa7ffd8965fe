# SPDX-License-Identifier: Apache-2.0
# Test oracle: a restatement of OpenFAWE's src/openfawe/aero/network_numba.py and wake.pack_rings (Copyright 2026 Yinong Tian, Apache License 2.0, itself
# adapted from OpenFAST/OLAF; see NOTICE).
"""NumPy restatement of the reference's incident-exclusion pairs.  TEST INFRASTRUCTURE ONLY.

``incident_pairs`` (network_numba.py:242-253, with ``_incident_pairs_sorted`` :177-239 and the dictionary
form ``_incident_pairs`` :121-167 that it reproduces element for element) lists, for the free points
``nodes[first:]`` of a wake network, every (point, segment) pair in which the point coincides with an end of
the segment.  Coordinates match by bit pattern (-0.0 and 0.0 differ, network_numba.py:16-21), a node with a
NaN coordinate never matches, and the pairs come in segment order, the start end before the end end, and for
one end the coinciding points in increasing order (:229-239).  Point indices count from ``first``.

``incident_pairs_loop`` follows the reference's loops with a dictionary keyed by the bit patterns;
``incident_pairs`` is a vectorised form of the same rule, checked against the loop form by the tests.

``pack_ring_edges`` (network_numba.py:36-117, the compiled form of the loop of wake.pack_rings, wake.py:300-307)
aggregates the directed edges of a lattice of vortex rings: ring (i, j) of panel j = (left, right) runs
(i, left) -> (i, right) -> (i+1, right) -> (i+1, left) over node ids row * width + column; each edge is kept once as
(lo, hi) with the sum, from 0.0 in ring order, of +strength per ring that runs it lo -> hi and -strength otherwise;
edges come in order of first appearance and exact-zero sums are dropped.  ``pack_ring_edges_loop`` is the
dictionary form; ``pack_ring_edges`` the vectorised one.
"""
import numpy as np


def _check(nodes, edges, first):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    edges = np.ascontiguousarray(edges, dtype=np.int64)
    if nodes.ndim != 2 or nodes.shape[1] != 3 or edges.ndim != 2 or edges.shape[1] != 2:
        raise ValueError("nodes must be (n, 3) and edges (e, 2)")
    if not 0 <= int(first) <= len(nodes) or (edges.size and (edges.min() < 0 or edges.max() >= len(nodes))):
        raise ValueError("first or an edge index is out of range")
    return nodes, edges, int(first)


def incident_pairs_loop(nodes, edges, first):
    """network_numba.py:121-167 (dictionary form), keyed by the coordinates' bit patterns."""
    nodes, edges, first = _check(nodes, edges, first)
    bits = nodes.view(np.int64)
    nan = np.isnan(nodes).any(axis=1)
    groups = {}
    for point in range(first, len(nodes)):
        if not nan[point]:
            groups.setdefault(tuple(bits[point]), []).append(point - first)
    pairs = []
    for segment, (a, b) in enumerate(edges):
        for node in (a, b):
            if not nan[node]:
                pairs.extend((member, segment) for member in groups.get(tuple(bits[node]), ()))
    return np.asarray(pairs, dtype=np.int64).reshape(-1, 2)


def incident_pairs(nodes, edges, first):
    """The same pairs, vectorised: groups of equal bit patterns by a sort, segment ends matched by search."""
    nodes, edges, first = _check(nodes, edges, first)
    bits = np.ascontiguousarray(nodes.view(np.int64))
    nan = np.isnan(nodes).any(axis=1)
    free = np.arange(first, len(nodes))
    free = free[~nan[free]]
    if len(free) == 0 or len(edges) == 0:
        return np.zeros((0, 2), dtype=np.int64)
    key = bits.view([("x", "<i8"), ("y", "<i8"), ("z", "<i8")]).reshape(-1)
    uniq, group = np.unique(key[free], return_inverse=True)
    # members of each group in increasing point order (a stable sort keeps the order of `free`)
    by_group = np.argsort(group, kind="stable")
    members = free[by_group] - first
    start = np.searchsorted(group[by_group], np.arange(len(uniq)))
    size = np.bincount(group, minlength=len(uniq))
    ends = edges.reshape(-1)                                   # segment order, start end first
    pos = np.searchsorted(uniq, key[ends])
    hit = (pos < len(uniq)) & ~nan[ends]
    hit[hit] = uniq[pos[hit]] == key[ends[hit]]
    g = np.where(hit, pos, 0)
    counts = np.where(hit, size[g], 0)
    total = int(counts.sum())
    seg = np.repeat(np.arange(len(ends)) // 2, counts)
    offset = np.repeat(np.cumsum(counts) - counts, counts)
    first_member = np.repeat(start[g], counts)
    point = members[first_member + (np.arange(total) - offset)]
    return np.stack([point, seg], axis=1).astype(np.int64).reshape(-1, 2)


def _ring_check(depth, width, panel_edges, strengths):
    depth, width = int(depth), int(width)
    panel_edges = np.ascontiguousarray(panel_edges, dtype=np.int64).reshape(-1, 2)
    strengths = np.ascontiguousarray(strengths, dtype=np.float64)
    if depth < 2 or width < 2:
        raise ValueError("rows must have shape (depth+1,span+1,3)")
    if (np.any(panel_edges < 0) or np.any(panel_edges >= width)
            or np.any(panel_edges[:, 0] == panel_edges[:, 1])):
        raise ValueError("invalid spanwise panel connectivity")
    if strengths.shape != (depth - 1, len(panel_edges)):
        raise ValueError("ring strength shape mismatch")
    return depth, width, panel_edges, strengths


def ring_contributions(depth, width, panel_edges, strengths):
    """(lo, hi, value) of every ring edge in ring order (network_numba.py:36-64): 4 per ring."""
    depth, width, panel_edges, strengths = _ring_check(depth, width, panel_edges, strengths)
    i = np.repeat(np.arange(depth - 1), len(panel_edges))
    left = np.tile(panel_edges[:, 0], depth - 1)
    right = np.tile(panel_edges[:, 1], depth - 1)
    vertices = np.stack([i * width + left, i * width + right, (i + 1) * width + right, (i + 1) * width + left],
                        axis=1)
    a, b = vertices.reshape(-1), np.roll(vertices, -1, axis=1).reshape(-1)
    g = np.repeat(strengths.reshape(-1), 4)
    forward = a < b
    return np.where(forward, a, b), np.where(forward, b, a), np.where(forward, g, -g)


def pack_ring_edges_loop(depth, width, panel_edges, strengths):
    """The dictionary form (wake.py:300-307): (edges (e,2) int64, gamma (e,) float64)."""
    lo, hi, value = ring_contributions(depth, width, panel_edges, strengths)
    sums = {}
    for key in zip(lo.tolist(), hi.tolist(), value.tolist()):
        sums[key[:2]] = sums.get(key[:2], 0.0) + key[2]
    kept = [key for key, total in sums.items() if total != 0.0]
    return (np.asarray(kept, dtype=np.int64).reshape(-1, 2),
            np.asarray([sums[key] for key in kept], dtype=np.float64))


def pack_ring_edges(depth, width, panel_edges, strengths):
    """The same edges, vectorised: a stable sort by key, the group sums added member by member in ring order."""
    lo, hi, value = ring_contributions(depth, width, panel_edges, strengths)
    if len(lo) == 0:
        return np.zeros((0, 2), dtype=np.int64), np.zeros(0)
    key = lo * (int(depth) * int(width)) + hi
    order = np.argsort(key, kind="stable")
    starts = np.flatnonzero(np.r_[True, key[order][1:] != key[order][:-1]])
    sizes = np.diff(np.r_[starts, len(order)])
    total = np.zeros(len(starts))
    for r in range(int(sizes.max())):           # member r of every group that has one, in ring order
        has = sizes > r
        total[has] = total[has] + value[order[starts[has] + r]]
    first = order[starts]
    keep = np.argsort(first, kind="stable")
    keep = keep[total[keep] != 0.0]
    return np.stack([lo[first[keep]], hi[first[keep]]], axis=1).astype(np.int64), total[keep]
