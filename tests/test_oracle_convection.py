"""The oracle of the free-wake geometry step (oracle/convection_np.py) against brute force and the properties the
reference's code fixes (wake.py:404-491, 586-616).  CPU only.

* Without rings or limits the velocity is the ambient plus the pair-by-pair sum over the non-incident segments.
* Per-point sums do not depend on the point set (the reference's own remark, wake.py:411-414, 441-444): the free
  rows under ``free_convection_rows`` and the rows beside rings equal the full evaluation bit for bit, the frozen
  rows keep the ambient exactly, ring nodes beyond ``ring_free`` too.
* 'ambient_rings': the frozen rows take the ring edges alone; ring part plus wing part equals the full sum.
* Without a free wake every integrator moves points by dt * U (closed form).
* Convergence orders 1, 2 and 4 of Euler, Heun and RK4 on a self-convecting ring lattice pin the stage weights.
* Euler never reads a moved bound row; Heun and RK4 do.
"""
import numpy as np
import pytest

from oracle import convection_np as C
from oracle import network_np as NP
from oracle import oracle as O


def wing_wake(depth=6, width=9, rings=2, seed=0, coincide=True):
    """A wing wake lattice (bound row, trailing-edge row, free rows) packed into rings, plus closed rotor rings of
    8 nodes after ``ring_first``; one ring node coincides with a wake node when ``coincide``."""
    rng = np.random.default_rng(seed)
    x = 0.8 * np.arange(depth)[:, None] * np.ones((1, width))
    y = np.ones((depth, 1)) * np.linspace(-4.0, 4.0, width)[None, :]
    z = 0.1 * np.sin(0.7 * x) + 0.02 * rng.normal(size=x.shape)
    rows = np.stack((x, y, z), axis=-1)
    strengths = 1.0 + 0.3 * rng.normal(size=(depth - 1, width - 1))
    panels = np.column_stack([np.arange(width - 1), np.arange(1, width)])
    edges, gamma = NP.pack_ring_edges(depth, width, panels, strengths)
    nodes = rows.reshape(-1, 3)
    ring_first = len(nodes)
    ring_nodes, ring_edges, ring_gamma = [], [], []
    for r in range(rings):
        angle = np.linspace(0, 2 * np.pi, 8, endpoint=False)
        centre = np.array([1.5 + 1.7 * r, -2.0 + 3.0 * r, 0.6])
        ring = centre + 0.5 * np.stack([np.zeros(8), np.cos(angle), np.sin(angle)], axis=1)
        base = ring_first + 8 * r
        ring_nodes.append(ring)
        ring_edges.append(np.stack([base + np.arange(8), base + (np.arange(8) + 1) % 8], axis=1))
        ring_gamma.append(np.full(8, 0.8 + 0.2 * r))
    if rings and coincide:
        ring_nodes[0][3] = nodes[2 * width + 4]          # a ring node on a wake node: incident both ways
    nodes = np.concatenate([nodes, *ring_nodes]) if rings else nodes
    if rings:
        edges = np.concatenate([edges, *ring_edges])
        gamma = np.concatenate([gamma, *ring_gamma])
    return nodes, edges, gamma, width, ring_first if rings else None


def test_plain_step_is_ambient_plus_pairwise_sum():
    nodes, edges, gamma, first, _ = wing_wake(rings=0)
    U = np.array([-5.0, 0.3, 0.1])
    got = C.network_convection(nodes, edges, gamma, first, 0.15, 2, np.broadcast_to(U, (len(nodes) - first, 3)))
    excluded = {(int(p), int(s)) for p, s in NP.incident_pairs(nodes, edges, first)}
    starts = nodes[edges[:, 0]]
    delta = nodes[edges[:, 1]] - starts
    length = np.linalg.norm(delta, axis=1)
    for i, p in enumerate(nodes[first:]):
        keep = [s for s in range(len(edges)) if (i, s) not in excluded]
        u, flags = O.pair_values(np.repeat(p[None], len(keep), 0), starts[keep], delta[keep] / length[keep, None],
                                 length[keep], gamma[keep], 2, np.full(len(keep), 0.15))
        assert not flags.any()
        np.testing.assert_allclose(got[i], U + u.sum(axis=0), rtol=1e-12, atol=1e-13)


def test_point_set_independence_limits_and_ring_free():
    nodes, edges, gamma, first, ring_first = wing_wake()
    n = len(nodes)
    amb = np.broadcast_to(np.array([-4.0, 0.0, 0.2]), (n - first, 3))
    full = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, ring_first=ring_first)
    rows_end = ring_first - first
    # no split at all: every point 'raise' (no point lies on a filament here)
    plain = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb)
    assert np.array_equal(plain, full)
    limited = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, ring_first=ring_first,
                                   free_convection_rows=2)
    free = 2 * first
    assert np.array_equal(limited[:free], full[:free])
    assert np.array_equal(limited[free:rows_end], amb[free:rows_end])
    assert np.array_equal(limited[rows_end:], full[rows_end:])
    partial = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, ring_first=ring_first, ring_free=5)
    assert np.array_equal(partial[:rows_end + 5], full[:rows_end + 5])
    assert np.array_equal(partial[rows_end + 5:], amb[rows_end + 5:])
    same = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, ring_first=ring_first,
                                ring_radii=np.full(len(edges), 0.15))
    assert np.array_equal(same, full)
    cored = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, ring_first=ring_first, ring_radii=0.4)
    assert np.array_equal(cored[:rows_end], full[:rows_end]) and not np.array_equal(cored[rows_end:], full[rows_end:])


def test_ambient_rings_takes_the_ring_edges_alone():
    nodes, edges, gamma, first, ring_first = wing_wake()
    n = len(nodes)
    amb = np.zeros((n - first, 3))
    full = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, ring_first=ring_first)
    drift = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, ring_first=ring_first,
                                 free_convection_rows=1, frozen_rows_drift="ambient_rings")
    free, rows_end = first, ring_first - first
    ring_edge = np.all(edges >= ring_first, axis=1)
    wing_only = C.network_convection(nodes, edges[~ring_edge], gamma[~ring_edge], first, 0.15, 2, amb,
                                     ring_first=ring_first)
    # linearity: the ring edges' part plus the wing edges' part is the full sum, on the frozen rows
    np.testing.assert_allclose(drift[free:rows_end] + wing_only[free:rows_end], full[free:rows_end],
                               rtol=1e-12, atol=1e-13 * np.abs(full).max())
    assert not np.allclose(drift[free:rows_end], 0.0)


@pytest.mark.parametrize("integrator", ["euler", "heun", "rk4"])
def test_integrators_without_a_free_wake(integrator):
    nodes, edges, gamma, first, ring_first = wing_wake()
    U = np.array([-7.0, 0.5, 0.25])
    kw = dict(free_wake=False, ring_first=ring_first)
    out = C.convect_geometry(nodes[:first], nodes[first:], edges, gamma, 0.15, 2, inflow=U, time=0.3, dt_s=0.05,
                             substeps=2, integrator=integrator, **kw)
    np.testing.assert_allclose(out, nodes[first:] + 0.05 * U, rtol=0, atol=1e-14)


def _step_error(integrator, substeps):
    nodes, edges, gamma, first, ring_first = wing_wake(depth=4, width=6, rings=1, coincide=False)
    kw = dict(ring_first=ring_first)
    args = (nodes[:first], nodes[first:], edges, 3.0 * gamma, 0.3, 2)
    fine = C.convect_geometry(*args, inflow=[-1.0, 0.0, 0.0], time=0.0, dt_s=0.4, substeps=64, integrator="rk4",
                              **kw)
    got = C.convect_geometry(*args, inflow=[-1.0, 0.0, 0.0], time=0.0, dt_s=0.4, substeps=substeps,
                             integrator=integrator, **kw)
    return np.abs(got - fine).max()


@pytest.mark.parametrize("integrator,order", [("euler", 1), ("heun", 2), ("rk4", 4)])
def test_convergence_orders_pin_the_stage_weights(integrator, order):
    e1, e2 = _step_error(integrator, 2), _step_error(integrator, 4)
    rate = np.log2(e1 / e2)
    assert order - 0.35 < rate < order + 0.6, rate


def test_displacement_reaches_only_the_later_stages():
    nodes, edges, gamma, first, ring_first = wing_wake(coincide=False)
    disp = np.zeros((first, 3))
    disp[:, 2] = 0.05
    base = dict(inflow=[-3.0, 0.0, 0.0], time=1.0, dt_s=0.05, ring_first=ring_first)
    args = (nodes[:first], nodes[first:], edges, gamma, 0.15, 2)
    euler = C.convect_geometry(*args, integrator="euler", displacement=disp, **base)
    assert np.array_equal(euler, C.convect_geometry(*args, integrator="euler", **base))
    zero = C.convect_geometry(*args, integrator="rk4", displacement=0 * disp, **base)
    assert np.array_equal(zero, C.convect_geometry(*args, integrator="rk4", **base))
    moved = C.convect_geometry(*args, integrator="rk4", displacement=disp, **base)
    assert not np.array_equal(moved, zero)


def test_convection_settings_validation():
    from openfawe_amd import convection as CV
    from openfawe_amd.vortex import TreeSettings
    for bad in (dict(backend="cuda"), dict(backend="hip_tree"), dict(tree=TreeSettings()), dict(model=9),
                dict(frozen_rows_drift="rings"), dict(free_convection_rows=0), dict(free_convection_rows=1.5),
                dict(free_convection_rows=True), dict(hip="fp64")):
        with pytest.raises(ValueError):
            CV.ConvectionSettings(**bad)
    assert CV.ConvectionSettings(backend="hip_tree", tree=TreeSettings()).model == 2


class RecordingGuard:
    """Stands in for the reference's StaticSurfaceGuard: records every check_rows call, raises on ``fail_at``."""

    def __init__(self, fail_at=None):
        self.calls, self.fail_at = [], fail_at

    def check_rows(self, old_rows, new_rows, wing, *, stage, time_s):
        self.calls.append((stage, time_s, np.array(old_rows), np.array(new_rows), wing))
        if self.fail_at is not None and len(self.calls) == self.fail_at:
            raise RuntimeError(f"surface crossed at {stage}")


@pytest.mark.parametrize("integrator", ["euler", "heun", "rk4"])
def test_sheet_drift_moves_the_frozen_rings_by_dt_times_drift(integrator):
    # without a free wake every stage's velocity is U (+ the drift of the rings beyond ring_free): closed form
    nodes, edges, gamma, first, ring_first = wing_wake(rings=3)
    rows, rings_nodes = ring_first - first, len(nodes) - ring_first
    drift = np.random.default_rng(4).normal(size=(rings_nodes, 3))
    U = np.array([-6.0, 0.25, -0.5])
    kw = dict(inflow=U, time=0.2, dt_s=0.08, substeps=2, integrator=integrator, free_wake=False,
              ring_first=ring_first, ring_free=8)
    out = C.convect_geometry(nodes[:first], nodes[first:], edges, gamma, 0.15, 2, ring_drift=drift, **kw)
    want = nodes[first:] + 0.08 * U
    want[rows + 8:] = nodes[ring_first + 8:] + 0.08 * (U + drift[8:])
    np.testing.assert_allclose(out, want, rtol=0, atol=1e-14)
    # no ring_free (every ring convects freely) or a drift array no longer than ring_free: no drift at all
    plain = C.convect_geometry(nodes[:first], nodes[first:], edges, gamma, 0.15, 2, **kw)
    for other in (dict(kw, ring_free=None), dict(kw, ring_free=rings_nodes)):
        assert np.array_equal(C.convect_geometry(nodes[:first], nodes[first:], edges, gamma, 0.15, 2,
                                                 ring_drift=drift, **other),
                              C.convect_geometry(nodes[:first], nodes[first:], edges, gamma, 0.15, 2, **other))
    assert not np.array_equal(out, plain)


def test_sheet_drift_with_a_free_wake_is_linear_for_euler():
    # Euler evaluates the velocity once, at the start points: the drift adds dt * drift to the frozen rings only
    nodes, edges, gamma, first, ring_first = wing_wake(rings=3, coincide=False)
    drift = np.random.default_rng(5).normal(size=(len(nodes) - ring_first, 3))
    kw = dict(inflow=[-3.0, 0.0, 0.1], time=0.0, dt_s=0.05, integrator="euler", ring_first=ring_first,
              ring_free=11)
    args = (nodes[:first], nodes[first:], edges, gamma, 0.15, 2)
    plain = C.convect_geometry(*args, **kw)
    drifted = C.convect_geometry(*args, ring_drift=drift, **kw)
    split = ring_first - first + 11
    assert np.array_equal(drifted[:split], plain[:split])
    np.testing.assert_allclose(drifted[split:] - plain[split:], 0.05 * drift[11:], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("integrator,labels", [
    ("euler", ["accepted_geometry"]),
    ("heun", ["heun_predictor", "accepted_geometry"]),
    ("rk4", ["rk4_stage2", "rk4_stage3", "rk4_stage4", "accepted_geometry"])])
def test_guard_sees_every_stage_point_of_the_wing_rows(integrator, labels):
    nodes, edges, gamma, first, ring_first = wing_wake(coincide=False)
    rows = ring_first - first
    guard, wing = RecordingGuard(), object()
    kw = dict(inflow=[-3.0, 0.0, 0.0], time=1.0, dt_s=0.1, substeps=2, integrator=integrator, ring_first=ring_first)
    args = (nodes[:first], nodes[first:], edges, gamma, 0.15, 2)
    out = C.convect_geometry(*args, guard=guard, wing=wing, **kw)
    assert np.array_equal(out, C.convect_geometry(*args, **kw))        # the guard changes nothing it accepts
    assert [c[0] for c in guard.calls] == labels * 2
    offsets = {"heun_predictor": 1.0, "rk4_stage2": 0.5, "rk4_stage3": 0.5, "rk4_stage4": 1.0,
               "accepted_geometry": 1.0}
    half = C.convect_geometry(*args, **dict(kw, dt_s=0.05, substeps=1))   # the state after the first substep
    for i, (label, t, old, new, seen) in enumerate(guard.calls):
        sub = i // len(labels)
        assert seen is wing
        assert t == pytest.approx(1.0 + 0.05 * (sub + offsets[label]), abs=1e-15)
        assert old.shape == new.shape == (rows // first, first, 3)
        start = nodes[first:ring_first] if sub == 0 else half[:rows]
        assert np.array_equal(old.reshape(-1, 3), start)
    # the accepted rows of each substep are the rows that substep hands on, the last ones the result
    accepted = [c for c in guard.calls if c[0] == "accepted_geometry"]
    assert np.array_equal(accepted[0][3].reshape(-1, 3), half[:rows])
    assert np.array_equal(accepted[1][3].reshape(-1, 3), out[:rows])
    # an exception from the guard stops the step
    with pytest.raises(RuntimeError, match="surface crossed"):
        C.convect_geometry(*args, guard=RecordingGuard(fail_at=len(labels)), wing=wing, **kw)
