"""The oracle's pair kernel pinned by closed forms, the recorded upstream fixture, quadrature,
invariances and the A1 bound against an independent NumPy restatement.

Cases follow the reference's tests/test_vortex.py (K01-K09) and tests/test_vortex_numba.py (A1).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from openfawe_amd import workloads
from tests.helpers import oracle_direct, prepare, seg1

A = np.array([0., 0., -1.])
B = np.array([0., 0., 1.])


def close(actual, expected, rtol=1e-12, atol=1e-14):
    np.testing.assert_allclose(actual, expected, rtol=rtol, atol=atol)


def test_k01_closed_form():
    u, f = seg1([1, 0, 0], 0, 0.0)
    assert f == 0
    close(u, [0, 1 / (2 * np.pi * np.sqrt(2)), 0])


@pytest.mark.parametrize("h", [0.05, 1.0, 10.0])
@pytest.mark.parametrize("model", range(6))
def test_k02_symmetric_all_cores(h, model):
    ideal = 1 / (2 * np.pi * h * np.sqrt(1 + h * h))
    q = (h / .5) ** 2
    factors = [1, min(q, 1), 1 if -1.25643 * q < -10 else -np.expm1(-1.25643 * q),
               q / np.sqrt(1 + q * q),
               (2 * h * h * (1 + h * h)) / (2 * h * h * (1 + h * h) + 4 * .5 ** 2),
               h * h * np.sqrt(1 + h * h) / ((h * h + .5 ** 2) * np.sqrt(1 + h * h + .5 ** 2))]
    u, _ = seg1([h, 0, 0], model, 0.5)
    close(u, [0, ideal * factors[model], 0])


@pytest.mark.parametrize("exponent", [9.99, 10.01])
def test_k02_lamb_cutoff(exponent):
    h = .5 * np.sqrt(exponent / 1.25643)
    ideal = 1 / (2 * np.pi * h * np.sqrt(1 + h * h))
    full = ideal * -np.expm1(-exponent)
    v = seg1([h, 0, 0], 2, 0.5)[0][1]
    close(v, full if exponent < 10 else ideal)
    if exponent > 10:
        assert abs(v - full) > 1e-7


def test_k02_recorded_upstream_fixture():
    data = json.loads((Path(__file__).parent / "golden/upstream_kernel_check.json").read_text())
    assert len(data["cases"]) >= 10
    for case in data["cases"]:
        u, _ = seg1(case["point"], case["core_model_id"], case["core_radius_m"], data["circulation_m2_per_s"],
                    data["segment_start"], data["segment_end"])
        close(u, case["velocity_m_per_s"])


def test_independent_quadrature():
    start, end, point = np.array([-.7, .3, -.8]), np.array([1.2, -.4, 1.]), np.array([.5, 1.2, .4])
    x, w = np.polynomial.legendre.leggauss(96)
    delta = end - start
    r = point - (start + ((x + 1) / 2)[:, None] * delta)
    integral = np.sum((w / 2)[:, None] * np.cross(delta, r) / np.linalg.norm(r, axis=1)[:, None] ** 3, axis=0)
    u, _ = seg1(point, 0, 0.0, 1.7, start, end)
    close(u, 1.7 / (4 * np.pi) * integral)


def test_near_interior_stable_denominator():
    h = 1e-9
    close(seg1([h, 0, 0], 0, 0.0)[0][1], 1 / (2 * np.pi * h * np.sqrt(1 + h * h)))


def test_square_ring_centre_and_orientation():
    corners = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], dtype=float)
    edges = [[0, 1], [1, 2], [2, 3], [3, 0]]
    out, _ = oracle_direct([[0, 0, 0]], corners, edges, 1.0)
    close(out[0], [0, 0, np.sqrt(2) / np.pi])
    rev, _ = oracle_direct([[0, 0, 0]], corners[::-1], edges, 1.0)
    close(rev[0], -out[0])


@pytest.mark.parametrize("model", range(6))
def test_k05_orientation_linearity(model):
    u = seg1([1, 0, .2], model, .5, 2.0)[0]
    close(seg1([1, 0, .2], model, .5, 2.0, B, A)[0], -u)
    close(seg1([1, 0, .2], model, .5, -2.0)[0], -u)
    close(seg1([1, 0, .2], model, .5, 4.0)[0], 2 * u)


@pytest.mark.parametrize("model", range(6))
def test_k06_rotation_translation(model):
    axis = np.array([1., 2., 3.])
    axis /= np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    rot = np.eye(3) + np.sin(.7) * k + (1 - np.cos(.7)) * (k @ k)
    shift = np.array([2., -3., .7])
    for h in (.05, 1., 10.):
        p = np.array([h, 0, 0])
        expected = rot @ seg1(p, model, .5)[0]
        close(seg1(rot @ p + shift, model, .5, 1.0, rot @ A + shift, rot @ B + shift)[0], expected)


@pytest.mark.parametrize("model", range(6))
def test_k07_scale_covariance(model):
    for h in (.05, 1., 10.):
        p = np.array([h, 0, -.5])
        expected = seg1(p, model, .5)[0]
        for scale in (1e-3, 1., 1e3):
            got = scale * seg1(scale * p, model, scale * .5, 1.0, scale * A, scale * B)[0]
            close(got, expected, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("model", [0, 1, 2, 3, 5])
def test_k04_subdivision(model):
    for point in ([0, 0, 0], A, [.05, 0, -.5], [100, 0, -.5]):
        whole = seg1(point, model, .5, singularity="zero")[0]
        nodes = np.array([A, [0, 0, 0], B])
        split, _ = oracle_direct([point], nodes, [[0, 1], [1, 2]], 1.0, model, .5, singularity="zero")
        close(whole, split[0], rtol=0, atol=1e-4)
        if point[0] != 0:
            close(whole, split[0], rtol=1e-11, atol=1e-13)


def test_k08_singular_flags():
    for model in range(6):
        assert seg1([0, 0, 0], model, .5)[1] == O.FLAG_SINGULAR
        u, f = seg1(A, model, .5, singularity="zero")
        assert f == 0 and np.all(u == 0)
    u, f = seg1([0, 0, 2], 0, 0.0)   # collinear outside the segment: zero, no flag
    assert f == 0 and np.all(u == 0)
    # subnormal core: the core ratio overflows -> non-finite flag (test_vortex_numba.py:109-119)
    for model in (1, 2, 3):
        _, f = seg1([.5, 1., 0.], model, 1e-310, 1.0, [0, 0, 0], [1, 0, 0])
        assert f == O.FLAG_NONFINITE
    _, f = seg1([0, 1e160, 0], 0, 0.0, 1.0, [0, 0, 0], [1e-160, 0, 0])
    assert f == O.FLAG_NONFINITE


def a1_check(points, nodes, edges, gamma, model, radius, exclusions=None):
    """A1: |u_oracle - u_numpy| <= 1e-12 * sum_s |u_s| per point."""
    vortex_np = pytest.importorskip("oracle.vortex_np")   # CPU-only check: the module stays off the GPU box
    starts, directions, lengths = prepare(nodes, edges)
    m = len(starts)
    gamma = np.broadcast_to(np.asarray(gamma, float), (m,))
    radius = np.broadcast_to(np.asarray(radius, float), (m,))
    excluded = set(map(tuple, np.asarray(exclusions).tolist())) if exclusions is not None else None
    ref, scale = vortex_np.direct_sum(points, starts, directions, lengths, gamma, radius, model, excluded)
    out, flags = oracle_direct(points, nodes, edges, gamma, model, radius, exclusions, singularity="zero")
    err = np.linalg.norm(out - ref, axis=1)
    assert np.all(err <= 1e-12 * scale), float(np.max(err - 1e-12 * scale))
    return out


@pytest.mark.parametrize("seed", [3, 4])
@pytest.mark.parametrize("model", range(6))
def test_a1_against_numpy_restatement(seed, model):
    nodes, edges, gamma, radius, points = workloads.random_network(seed)
    for core in (0., .2, radius):
        a1_check(points, nodes, edges, gamma, model, core)


def test_a1_exclusions():
    nodes, edges, gamma, radius, _ = workloads.random_network(5)
    incident = [(n, s) for s, (a, b) in enumerate(edges) for n in (a, b)]
    rng = np.random.default_rng(9)
    extra = [(int(p), int(s)) for p, s in zip(rng.integers(0, len(nodes), 30), rng.integers(0, len(edges), 30))]
    pairs = np.asarray(incident + extra + incident[:5], dtype=np.int64)
    for model in (0, 2, 5):
        a1_check(nodes, nodes, edges, gamma, model, radius, pairs)


def test_thread_count_bitwise():
    nodes, edges, gamma, radius, points = workloads.random_network(6, nodes=200, segments=400, points=600)
    starts, directions, lengths = prepare(nodes, edges)
    outs = [O.direct_sum(points, starts, directions, lengths, gamma, radius, 2, None, "zero", threads=t)[0]
            for t in (1, 3, 8)]
    for o in outs[1:]:
        assert np.array_equal(outs[0].view(np.uint64), o.view(np.uint64))


def test_exclusion_csr_brute_force():
    rng = np.random.default_rng(1)
    for n, e in ((7, 0), (50, 300), (200, 1000)):
        pairs = np.stack([rng.integers(0, n, e), rng.integers(0, 40, e)], axis=1) if e else np.zeros((0, 2), int)
        ptr, seg = O.exclusion_csr(n, pairs)
        uniq = np.unique(pairs, axis=0) if e else np.zeros((0, 2), int)
        assert np.array_equal(seg, uniq[:, 1])
        assert np.array_equal(ptr, np.concatenate([[0], np.cumsum(np.bincount(uniq[:, 0], minlength=n))]))
