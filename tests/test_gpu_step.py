"""The geometry step recorded as one graph (convection.GeometryStepPlan, ofw_step_*) against the stage-by-stage
convect_geometry: bit for bit on every replay (with eager work and synchronisations between replays), after the
inputs change, without the graph, and through the fallbacks (a flagged ring target, a flagged row, a capacity
overflow that makes the plan again); the rings' sheet drift recorded into the graph, and the surface guard's calls
as the stage-by-stage step makes them."""
import numpy as np
import pytest

from openfawe_amd import _lib
from openfawe_amd import convection as CV
from openfawe_amd import vortex as V
from tests.test_oracle_convection import RecordingGuard, wing_wake

pytestmark = pytest.mark.gpu

TREE = CV.ConvectionSettings(model=2, backend="hip_tree", tree=V.TreeSettings(0.3, 2, 8, 3.0),
                             hip=V.HipSettings("mixed"))


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def case(depth=10, width=13, rings=3, seed=0):
    nodes, edges, gamma, first, ring_first = wing_wake(depth=depth, width=width, rings=rings, seed=seed,
                                                       coincide=False)
    disp = np.zeros((first, 3))
    disp[:, 2] = 0.02
    return nodes, edges, gamma, first, ring_first, disp


def eager(nodes, edges, gamma, first, settings, **kw):
    return CV.convect_geometry(dev(nodes[:first]), dev(nodes[first:]), dev(edges), dev(gamma), 0.15, settings,
                               **kw).cpu().numpy()


@pytest.mark.parametrize("integrator", ["euler", "heun", "rk4"])
@pytest.mark.parametrize("split", ["rows-rings", "limit-ring-free"])
def test_replays_equal_the_stage_by_stage_step(integrator, split):
    import torch
    print(f"step plan: {integrator} {split}", flush=True)
    nodes, edges, gamma, first, ring_first, disp = case()
    settings = TREE if split == "rows-rings" else CV.ConvectionSettings(
        model=2, backend="hip_tree", tree=TREE.tree, hip=TREE.hip, free_convection_rows=4)
    kw = dict(inflow=[-4.0, 0.1, 0.2], time=0.5, dt_s=0.05, substeps=2, integrator=integrator,
              displacement=dev(disp), ring_first=ring_first, ring_free=None if split == "rows-rings" else 11)
    want = eager(nodes, edges, gamma, first, settings, **kw)
    bound, pts, e, g = dev(nodes[:first]), dev(nodes[first:]), dev(edges), dev(gamma)
    plan = CV.GeometryStepPlan(bound, pts, e, g, 0.15, settings, **kw)
    for replay in range(3):
        bits, _, _ = plan.run()
        assert bits == 0
        got = plan.points_out.cpu().numpy()
        assert np.array_equal(got.view(np.int64), want.view(np.int64)), replay
        torch.cuda.synchronize()
        _ = eager(nodes, edges, gamma, first, settings, **kw)        # the program's own eager work between replays
    # new contents in the same tensors: the replay follows them
    moved = nodes.copy()
    moved[first:] += 0.01 * np.sin(np.arange(len(moved) - first))[:, None]
    pts.copy_(dev(moved[first:]))
    g.mul_(1.5)
    plan.run()
    want2 = eager(moved, edges, 1.5 * gamma, first, settings, **kw)
    assert np.array_equal(plan.points_out.cpu().numpy().view(np.int64), want2.view(np.int64))
    plan.close()
    # the same plan without the graph
    pts.copy_(dev(nodes[first:]))
    g.copy_(dev(gamma))
    plan = CV.GeometryStepPlan(bound, pts, e, g, 0.15, settings, use_graph=False, **kw)
    plan.run()
    assert np.array_equal(plan.points_out.cpu().numpy().view(np.int64), want.view(np.int64))
    plan.close()


def test_fallbacks_and_capacity_overflow():
    print("step plan: fallbacks", flush=True)
    nodes, edges, gamma, first, ring_first, _ = case()
    kw = dict(inflow=[-4.0, 0.0, 0.0], time=0.0, dt_s=0.05, integrator="rk4", ring_first=ring_first)
    bound, pts, e, g = dev(nodes[:first]), dev(nodes[first:]), dev(edges), dev(gamma)
    plan = CV.GeometryStepPlan(bound, pts, e, g, 0.15, TREE, **kw)
    base_pairs = plan.capacities["pairs"]
    # a ring node inside a wing filament: the ring flag sends the step to the stage-by-stage path ('zero' there)
    bad = nodes.copy()
    bad[ring_first + 2] = 0.5 * (nodes[edges[3, 0]] + nodes[edges[3, 1]])
    pts.copy_(dev(bad[first:]))
    bits, _, _ = plan.run()
    assert bits & _lib.STEP_RING_FLAG and not bits & _lib.STEP_ROW_FLAG
    got = plan.step().cpu().numpy()
    want = eager(bad, edges, gamma, first, TREE, **kw)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    # a wing-row node inside a filament: the row flag, and the step raises as the reference
    bad = nodes.copy()
    bad[first + 3 * 13 + 5] = 0.5 * (nodes[edges[0, 0]] + nodes[edges[0, 1]])
    pts.copy_(dev(bad[first:]))
    assert plan.run()[0] & _lib.STEP_ROW_FLAG
    with pytest.raises(V.FilamentSingularity):
        plan.step()
    # many free nodes put on one position (none joined by an edge): more incident pairs than the plan holds
    crowd = nodes.copy()
    width = 13
    chosen = [first + r * width + c for r in range(1, 8, 2) for c in range(0, width, 2)]
    crowd[chosen] = crowd[first + 2 * width + 7]
    pts.copy_(dev(crowd[first:]))
    bits, pairs, _ = plan.run()
    assert bits & _lib.STEP_PAIR_OVERFLOW and pairs > base_pairs
    got = plan.step().cpu().numpy()
    assert plan.capacities["pairs"] > base_pairs
    want = eager(crowd, edges, gamma, first, TREE, **kw)
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    plan.close()


@pytest.mark.parametrize("precision,model", [("fp64", 2), ("mixed", 2), ("fp64", 3)])
@pytest.mark.parametrize("integrator", ["euler", "rk4"])
def test_direct_engine_replays(precision, model, integrator):
    print(f"step plan, direct engine: {precision} model {model} {integrator}", flush=True)
    nodes, edges, gamma, first, ring_first, disp = case(depth=8, width=11, rings=2)
    settings = CV.ConvectionSettings(model=model, hip=V.HipSettings(precision, 128), free_convection_rows=5)
    kw = dict(inflow=[-3.0, 0.2, 0.0], time=0.25, dt_s=0.04, integrator=integrator, displacement=dev(disp),
              ring_first=ring_first)
    want = eager(nodes, edges, gamma, first, settings, **kw)
    plan = CV.GeometryStepPlan(dev(nodes[:first]), dev(nodes[first:]), dev(edges), dev(gamma), 0.15, settings, **kw)
    for _ in range(2):
        assert plan.run()[0] == 0
        assert np.array_equal(plan.points_out.cpu().numpy().view(np.int64), want.view(np.int64))
    plan.close()


@pytest.mark.parametrize("integrator,substeps", [("euler", 1), ("heun", 2)])
def test_sheet_drift_and_surface_guard(integrator, substeps):
    print(f"step plan: sheet drift and guard, {integrator} x{substeps}", flush=True)
    nodes, edges, gamma, first, ring_first, disp = case()
    drift = np.random.default_rng(8).normal(size=(len(nodes) - ring_first, 3))
    settings = CV.ConvectionSettings(model=2, backend="hip_tree", tree=TREE.tree, hip=TREE.hip,
                                     free_convection_rows=4)
    kw = dict(inflow=[-4.0, 0.1, 0.2], time=0.5, dt_s=0.05, substeps=substeps, integrator=integrator,
              displacement=dev(disp), ring_first=ring_first, ring_free=9, ring_drift=drift)
    want = eager(nodes, edges, gamma, first, settings, **kw)
    assert not np.array_equal(want, eager(nodes, edges, gamma, first, settings, **dict(kw, ring_drift=None)))
    bound, pts, e, g = dev(nodes[:first]), dev(nodes[first:]), dev(edges), dev(gamma)
    plan = CV.GeometryStepPlan(bound, pts, e, g, 0.15, settings, **kw)
    for _ in range(2):                                   # the drift is recorded into every stage's velocity
        assert plan.run()[0] == 0
        assert np.array_equal(plan.points_out.cpu().numpy().view(np.int64), want.view(np.int64))
    plan.close()
    # a guard: the plan's step calls it as the stage-by-stage step does, on the same rows
    eager_guard, plan_guard = RecordingGuard(), RecordingGuard()
    eager(nodes, edges, gamma, first, settings, guard=eager_guard, wing="wing", **kw)
    plan = CV.GeometryStepPlan(bound, pts, e, g, 0.15, settings, guard=plan_guard, wing="wing", **kw)
    got = plan.step().cpu().numpy()
    assert np.array_equal(got.view(np.int64), want.view(np.int64))
    assert [c[:2] for c in plan_guard.calls] == [c[:2] for c in eager_guard.calls] and plan_guard.calls
    for a, b in zip(plan_guard.calls, eager_guard.calls):
        assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3]) and a[4] == b[4] == "wing"
    plan.close()
    plan = CV.GeometryStepPlan(bound, pts, e, g, 0.15, settings, guard=RecordingGuard(fail_at=1), wing="wing", **kw)
    with pytest.raises(RuntimeError, match="surface crossed"):
        plan.step()
    plan.close()
