"""The synthetic wakes have the sizes the benchmarks are quoted on (CPU only)."""
import numpy as np
import pytest

from openfawe_amd import workloads


@pytest.mark.parametrize("rows,segments", [(468, 102382), (936, 204874), (1872, 409858)])
def test_sheet_sizes_match_the_reference_speed_table(rows, segments):
    case = workloads.sheet(rows)
    assert len(case["points"]) == 20570
    assert len(case["starts"]) == segments
    np.testing.assert_allclose(np.linalg.norm(case["directions"], axis=1), 1.0, rtol=1e-15)
    assert case["gamma"][0] == 0.5 and case["gamma"][-1] == 1.5


def test_loop_wake_shape_and_exclusions():
    case = workloads.loop(rows=468)
    assert len(case["points"]) == 20570
    assert 102000 <= len(case["starts"]) <= 105000
    # every target node excludes exactly its incident segments
    pairs = case["pairs"]
    edges = case["edges"]
    for p in (0, 109, 110, 5000):
        segs = set(pairs[pairs[:, 0] == p, 1].tolist())
        expect = set(np.flatnonzero((edges[:, 0] == p) | (edges[:, 1] == p)).tolist())
        assert segs == expect
    assert np.abs(case["gamma"]).max() <= 8.0 * 1.3
    eight = workloads.loop(rows=3744)
    assert 819000 <= len(eight["starts"]) <= 823000


def test_helix_config1_shape():
    case = workloads.helix()
    assert len(case["points"]) == 2048
    assert len(case["starts"]) == 10240
    assert case["model"] == 3 and np.all(case["radius"] == 0.05)
    assert case["gamma"].min() >= 0 and case["gamma"].max() <= 8
    # node targets never sit on a non-excluded segment: all their incident segments are excluded
    assert len(case["pairs"]) >= 2 * (2048 - 100) - 200
