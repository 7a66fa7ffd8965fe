"""Shared helpers of the tests: segment preparation and the oracle run as an engine."""
import numpy as np

from oracle import oracle as O


def prepare(nodes, edges):
    """starts, directions, lengths as the reference's segments_velocity prepares them (vortex.py:460-465)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    edges = np.asarray(edges, dtype=np.int64)
    starts = nodes[edges[:, 0]]
    delta = nodes[edges[:, 1]] - starts
    lengths = np.hypot(np.hypot(delta[:, 0], delta[:, 1]), delta[:, 2])
    return starts, delta / lengths[:, None], lengths


def oracle_direct(points, nodes, edges, gamma, model=0, core=0.0, exclusions=None, singularity="raise"):
    starts, directions, lengths = prepare(nodes, edges)
    m = len(starts)
    gamma = np.broadcast_to(np.asarray(gamma, float), (m,))
    radius = np.broadcast_to(np.asarray(core, float), (m,))
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    csr = O.exclusion_csr(len(points), exclusions if exclusions is not None else np.zeros((0, 2), np.int64))
    return O.direct_sum(points, starts, directions, lengths, gamma, radius, model, csr, singularity)


def oracle_tree(points, nodes, edges, gamma, model=0, core=0.0, exclusions=None, singularity="raise", theta=0.5,
                order=2, leaf_size=32, clearance=3.0):
    starts, directions, lengths = prepare(nodes, edges)
    m = len(starts)
    gamma = np.broadcast_to(np.asarray(gamma, float), (m,))
    radius = np.broadcast_to(np.asarray(core, float), (m,))
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    csr = O.exclusion_csr(len(points), exclusions if exclusions is not None else np.zeros((0, 2), np.int64))
    return O.tree_sum(points, starts, directions, lengths, gamma, radius, model, csr, singularity, theta, order,
                      leaf_size, clearance)


def seg1(point, model=0, core=0.5, gamma=1.0, a=(0., 0., -1.), b=(0., 0., 1.), singularity="raise"):
    """Single segment velocity through the oracle's pair kernel (segment_velocity of vortex.py:500)."""
    nodes = np.array([a, b], dtype=np.float64)
    out, flags = oracle_direct([point], nodes, [[0, 1]], gamma, model, core, singularity=singularity)
    return out[0], int(flags[0])


def midpoint_cases():
    """The reference's GPU-build cases (tests/test_cuda_build.py:20-28): random, a grid with many equal
    coordinates, leaf 1, coincident midpoints, fewer segments than a leaf, longer than one tile."""
    rng = np.random.default_rng(11)
    grid = np.array([[i * 0.5, j * 0.25, 0.0] for i in range(30) for j in range(20)])
    return [("random", rng.random((600, 3)) * [20.0, 5.0, 2.0], 8),
            ("grid", grid, 32),
            ("grid-leaf-1", grid[:200], 1),
            ("coincident", np.ones((40, 3)), 4),
            ("fewer-than-a-leaf", rng.random((5, 3)), 32),
            ("longer-than-a-tile", rng.random((2100, 3)) * [30.0, 3.0, 1.0], 16)]
