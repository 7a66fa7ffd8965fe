"""Stored wake networks (openfawe_amd.stored_wake): the format of the reference's state_<step>.npz files
(paper1/scripts/convergence/run_history.py:128-133) written and read back, the frozen cases of
check_v1_frozen.py:35-38, and the loader's errors.  CPU only; the reference ships no state files, so the files are
written here from a seeded wing wake."""
import numpy as np
import pytest

from oracle import network_np as NP
from openfawe_amd import stored_wake as SW
from tests.test_oracle_convection import wing_wake


def write_state(path, rings=2):
    nodes, edges, gamma, first, _ = wing_wake(depth=8, width=9, rings=rings)
    wing = np.column_stack([np.full(16, 0.25), np.linspace(-3.9, 3.9, 16), np.zeros(16)])
    SW.save_state(path, nodes=nodes, edges=edges, gamma=gamma, radii=0.15, first=first, wing_points=wing,
                  speed=12.0, core_model=2, chord=0.5, exclusions=NP.incident_pairs(nodes, edges, first))
    return nodes, edges, gamma, first, wing


def test_round_trip_and_frozen_cases(tmp_path):
    path = tmp_path / "state_0007.npz"
    nodes, edges, gamma, first, wing = write_state(path)
    state = SW.load_state(path)
    assert np.array_equal(state["nodes"], nodes) and np.array_equal(state["edges"], edges)
    assert np.array_equal(state["gamma"], gamma) and state["radii"].shape == (len(edges),)
    assert state["first"] == first and state["speed"] == 12.0 and state["core_model"] == 2
    np.testing.assert_array_equal(state["exclusions"], NP.incident_pairs(nodes, edges, first))
    cases = SW.frozen_cases(state)
    points, mask, ex = cases["wake"]
    assert np.array_equal(points, nodes[first:]) and mask.all() and len(ex) > 0
    points, mask, ex = cases["wing"]
    assert np.array_equal(points, wing) and len(ex) == 0
    assert np.array_equal(mask, np.all(edges >= 2 * first, axis=1)) and 0 < mask.sum() < len(edges)


def test_loader_errors(tmp_path):
    path = tmp_path / "state_0001.npz"
    write_state(path)
    data = dict(np.load(path, allow_pickle=False))
    broken = dict(data)
    del broken["speed"]
    np.savez(tmp_path / "missing.npz", **broken)
    with pytest.raises(ValueError, match="lacks"):
        SW.load_state(tmp_path / "missing.npz")
    broken = dict(data, edges=data["edges"].reshape(-1))
    np.savez(tmp_path / "shape.npz", **broken)
    with pytest.raises(ValueError):
        SW.load_state(tmp_path / "shape.npz")
    broken = dict(data, first=np.int64(len(data["nodes"]) + 1))
    np.savez(tmp_path / "first.npz", **broken)
    with pytest.raises(ValueError):
        SW.load_state(tmp_path / "first.npz")
    np.savez(tmp_path / "pickled.npz", **dict(data, speed=np.array([{"a": 1}], dtype=object)))
    with pytest.raises(ValueError):
        SW.load_state(tmp_path / "pickled.npz")          # object arrays need pickle: refused
