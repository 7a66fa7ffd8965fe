"""Stored wakes on the GPU: the FP64 direct sum of both frozen cases against the oracle (A1 bar: 1e-12 of the
speed scale) and the tree code within its stated error; the writer's default exclusions are the incident pairs."""
import numpy as np
import pytest

from oracle import network_np as NP
from oracle import oracle as O
from openfawe_amd import stored_wake as SW
from openfawe_amd.vortex import TreeSettings
from tests.test_stored_wake import write_state

pytestmark = pytest.mark.gpu


def test_evaluate_a_stored_wake(tmp_path):
    path = tmp_path / "state_0042.npz"
    write_state(path, rings=3)
    state = SW.load_state(path)
    print(f"stored wake: {len(state['nodes'])} nodes, {len(state['edges'])} segments", flush=True)
    rows = SW.evaluate_state(state, trees=(TreeSettings(0.2, 2, 8),), repeats=1)
    assert [r["targets"] for r in rows] == ["wake", "wake", "wing", "wing"]
    for name, (points, mask, ex) in SW.frozen_cases(state).items():
        s, d, L = SW._geometry(state["nodes"], state["edges"][mask])
        ref, flags = O.direct_sum(points, s, d, L, state["gamma"][mask], state["radii"][mask], 2,
                                  O.exclusion_csr(len(points), ex))
        assert not flags.any()
        direct = [r for r in rows if r["targets"] == name and r["engine"] == "direct-fp64"][0]
        assert abs(direct["max_induced_over_U"] - np.linalg.norm(ref, axis=1).max() / 12.0) <= 1e-12
        tree = [r for r in rows if r["targets"] == name and r["engine"] == "tree"][0]
        assert tree["max_err_over_U"] < 1e-3 and tree["rms_err_over_U"] <= tree["max_err_over_U"]


def test_default_exclusions_are_the_incident_pairs(tmp_path):
    nodes, edges, gamma, first, wing = write_state(tmp_path / "a.npz")
    SW.save_state(tmp_path / "b.npz", nodes=nodes, edges=edges, gamma=gamma, radii=0.15, first=first,
                  wing_points=wing, speed=12.0, core_model=2)
    np.testing.assert_array_equal(SW.load_state(tmp_path / "b.npz")["exclusions"],
                                  NP.incident_pairs(nodes, edges, first))
