"""The GPU free-wake geometry step (openfawe_amd.convection) against the oracle (oracle/convection_np.py): the device
pieces exactly (exclusion CSR subsets, stage arithmetic bit for bit, segment geometry), the target split of
_network_convection in every mode and the Euler / Heun / RK4 stage loop within the direct sum's A1 bar (1e-12 of
the velocity scale), the tree engine within its stated error, and the reference's errors."""
import numpy as np
import pytest

from oracle import convection_np as C
from openfawe_amd import _lib
from openfawe_amd import convection as CV
from openfawe_amd import vortex as V
from tests.test_oracle_convection import RecordingGuard, wing_wake

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    import torch
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a).to("cuda", dtype if dtype is not None else torch.from_numpy(a).dtype)


def test_exclusion_subset_matches_the_host_csr():
    rng = np.random.default_rng(0)
    pairs = np.stack([rng.integers(0, 300, 5000), rng.integers(0, 700, 5000)], axis=1)
    pairs = np.concatenate([pairs, pairs[:300]])                     # duplicates, any order
    for lo, hi in ((0, 300), (40, 41), (17, 250), (300, 300)):
        print(f"exclusion subset [{lo}, {hi})", flush=True)
        ptr, seg = CV.exclusion_subset(dev(pairs), lo, hi, 700)
        keep = (pairs[:, 0] >= lo) & (pairs[:, 0] < hi)
        want_ptr, want_seg = V.exclusion_csr(hi - lo, pairs[keep] - [lo, 0])
        np.testing.assert_array_equal(ptr.cpu().numpy(), want_ptr)
        np.testing.assert_array_equal(seg.cpu().numpy()[:want_ptr[-1]], want_seg)
    remap = np.where(rng.random(700) < 0.4, -1, 0)
    remap[remap == 0] = np.arange(int((remap == 0).sum()))
    ptr, seg = CV.exclusion_subset(dev(pairs), 10, 200, 700, dev(remap))
    keep = (pairs[:, 0] >= 10) & (pairs[:, 0] < 200) & (remap[pairs[:, 1]] >= 0)
    sub = np.column_stack([pairs[keep, 0] - 10, remap[pairs[keep, 1]]])
    want_ptr, want_seg = V.exclusion_csr(190, sub)
    np.testing.assert_array_equal(ptr.cpu().numpy(), want_ptr)
    np.testing.assert_array_equal(seg.cpu().numpy()[:want_ptr[-1]], want_seg)
    # two target ranges packed into one sum: [0, 100) -> [0, 100), [200, 250) -> [100, 150)
    point_map = np.full(250, -1)
    point_map[:100] = np.arange(100)
    point_map[200:] = np.arange(100, 150)
    ptr, seg = CV.exclusion_subset(dev(pairs), 0, 250, 700, point_map=dev(point_map), count=150)
    keep = (pairs[:, 0] < 250) & (point_map[np.minimum(pairs[:, 0], 249)] >= 0)
    sub = np.column_stack([point_map[pairs[keep, 0]], pairs[keep, 1]])
    want_ptr, want_seg = V.exclusion_csr(150, sub)
    np.testing.assert_array_equal(ptr.cpu().numpy(), want_ptr)
    np.testing.assert_array_equal(seg.cpu().numpy()[:want_ptr[-1]], want_seg)
    with pytest.raises(ValueError):
        CV.exclusion_subset(dev(pairs), 0, 300, 699)                  # a segment id beyond m
    with pytest.raises(ValueError):
        CV.exclusion_subset(dev(pairs), 0, 250, 700, point_map=dev(point_map), count=149)   # map beyond count


def test_stage_arithmetic_is_numpys_bit_for_bit():
    rng = np.random.default_rng(1)
    x, a, b, c, d = (rng.normal(size=(999, 3)) * 10.0 ** rng.integers(-3, 4, size=(999, 1)) for _ in range(5))
    h = 0.0173
    want = {_lib.STAGE_STEP: x + h * a, _lib.STAGE_HALF: x + h * a / 2, _lib.STAGE_HEUN: x + h * (a + b) / 2,
            _lib.STAGE_RK4: x + h * (a + 2 * b + 2 * c + d) / 6, _lib.STAGE_ADD: x + a}
    for mode, ref in want.items():
        got = CV.stage(mode, h, dev(x), dev(a), dev(b), dev(c), dev(d)).cpu().numpy()
        assert np.array_equal(got.view(np.int64), ref.view(np.int64)), mode


def test_network_segments_and_errors():
    nodes, edges, _, _, _ = wing_wake()
    s, d, L = CV.network_segments(dev(nodes), dev(edges))
    starts = nodes[edges[:, 0]]
    delta = nodes[edges[:, 1]] - starts
    length = np.hypot(np.hypot(delta[:, 0], delta[:, 1]), delta[:, 2])
    assert np.array_equal(s.cpu().numpy(), starts)
    np.testing.assert_allclose(L.cpu().numpy(), length, rtol=4e-16, atol=0)
    np.testing.assert_allclose(d.cpu().numpy(), delta / length[:, None], rtol=1e-15, atol=1e-16)
    bad = edges.copy()
    bad[3] = [5, 5]
    with pytest.raises(ValueError, match="zero-length"):
        CV.network_segments(dev(nodes), dev(bad))
    bad[3] = [5, len(nodes)]
    with pytest.raises(ValueError, match="out of range"):
        CV.network_segments(dev(nodes), dev(bad))


SPLITS = {
    "plain": dict(),
    "rings": dict(rings=True),
    "rings-free-5": dict(rings=True, ring_free=5),
    "rings-cores": dict(rings=True, ring_radii=0.4),
    "limit": dict(rings=True, free_convection_rows=2),
    "limit-ambient-rings": dict(rings=True, free_convection_rows=1, frozen_rows_drift="ambient_rings"),
    "no-free-wake": dict(rings=True, free_wake=False),
}


def _split_args(spec, ring_first):
    spec = dict(spec)
    rings = spec.pop("rings", False)
    settings = {k: spec.pop(k) for k in ("free_convection_rows", "frozen_rows_drift", "free_wake") if k in spec}
    return (dict(ring_first=ring_first if rings else None, **spec), settings)


def _scale(nodes, edges, gamma, first):
    """The velocity scale of the step: the largest induced speed at a free point."""
    amb = np.zeros((len(nodes) - first, 3))
    return np.abs(C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb)).max()


@pytest.mark.parametrize("name", list(SPLITS))
def test_network_convection_split_modes(name):
    print(f"network convection: {name}", flush=True)
    nodes, edges, gamma, first, ring_first = wing_wake()
    kw, opts = _split_args(SPLITS[name], ring_first)
    amb = np.tile([-4.0, 0.1, 0.2], (len(nodes) - first, 1))
    want = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, **opts, **kw)
    settings = CV.ConvectionSettings(model=2, **opts)
    got = CV.network_convection(dev(nodes), dev(edges), dev(gamma), first, 0.15, settings, dev(amb),
                                **{k: (v if not isinstance(v, np.ndarray) else dev(v)) for k, v in kw.items()})
    tol = 1e-12 * _scale(nodes, edges, gamma, first)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=tol)


@pytest.mark.parametrize("integrator", ["euler", "heun", "rk4"])
def test_convect_geometry_matches_the_oracle(integrator):
    print(f"convect geometry: {integrator}", flush=True)
    nodes, edges, gamma, first, ring_first = wing_wake(coincide=False)
    disp = np.zeros((first, 3))
    disp[:, 2] = 0.03
    drift = np.random.default_rng(3).normal(size=(len(nodes) - ring_first, 3))      # the rings' sheet drift
    kw = dict(inflow=[-4.0, 0.1, 0.2], time=0.5, dt_s=0.05, substeps=2, integrator=integrator, ring_drift=drift,
              wing="wing")
    host_guard, dev_guard = RecordingGuard(), RecordingGuard()
    want = C.convect_geometry(nodes[:first], nodes[first:], edges, gamma, 0.15, 2, displacement=disp,
                              ring_first=ring_first, ring_free=7, free_convection_rows=3, guard=host_guard, **kw)
    settings = CV.ConvectionSettings(model=2, free_convection_rows=3)
    got = CV.convect_geometry(dev(nodes[:first]), dev(nodes[first:]), dev(edges), dev(gamma), 0.15, settings,
                              displacement=dev(disp), ring_first=ring_first, ring_free=7, guard=dev_guard,
                              **kw).cpu().numpy()
    step = np.abs(want - nodes[first:]).max()
    tol = 1e-12 * step + 4 * np.spacing(np.abs(want).max())
    np.testing.assert_allclose(got, want, rtol=0, atol=tol)
    # the guard's calls: the same stages, times and wing, rows within the same bar
    assert [c[:2] for c in dev_guard.calls] == [c[:2] for c in host_guard.calls] and dev_guard.calls
    for (_, _, old_d, new_d, wing_d), (_, _, old_h, new_h, _) in zip(dev_guard.calls, host_guard.calls):
        assert wing_d == "wing" and old_d.shape == old_h.shape == new_d.shape
        np.testing.assert_allclose(old_d, old_h, rtol=0, atol=tol)
        np.testing.assert_allclose(new_d, new_h, rtol=0, atol=tol)
    # the guard's exception stops the device step too
    with pytest.raises(RuntimeError, match="surface crossed"):
        CV.convect_geometry(dev(nodes[:first]), dev(nodes[first:]), dev(edges), dev(gamma), 0.15, settings,
                            displacement=dev(disp), ring_first=ring_first, ring_free=7,
                            guard=RecordingGuard(fail_at=1), **kw)


def test_tree_engine_and_errors():
    nodes, edges, gamma, first, ring_first = wing_wake(depth=40, width=33, rings=3)
    amb = np.tile([-4.0, 0.0, 0.0], (len(nodes) - first, 1))
    want = C.network_convection(nodes, edges, gamma, first, 0.15, 2, amb, ring_first=ring_first)
    tree = CV.ConvectionSettings(model=2, backend="hip_tree", tree=V.TreeSettings(0.2, 2, 16, 3.0),
                                 hip=V.HipSettings("mixed"))
    got = CV.network_convection(dev(nodes), dev(edges), dev(gamma), first, 0.15, tree, dev(amb),
                                ring_first=ring_first).cpu().numpy()
    induced = want - amb
    rel = np.linalg.norm(got - want) / np.linalg.norm(induced)
    assert rel < 5e-4, rel
    # a wing-row point on a filament raises ('raise'); the same point as a ring node does not ('zero')
    bad = nodes.copy()
    bad[3 * first + 2] = 0.5 * (nodes[edges[0, 0]] + nodes[edges[0, 1]])
    fp64 = CV.ConvectionSettings(model=0)
    with pytest.raises(V.FilamentSingularity):
        CV.network_convection(dev(bad), dev(edges), dev(gamma), first, 0.15, fp64, dev(amb))
    ring_node = ring_first + 1
    bad = nodes.copy()
    bad[ring_node] = 0.5 * (nodes[edges[0, 0]] + nodes[edges[0, 1]])
    out = CV.network_convection(dev(bad), dev(edges), dev(gamma), first, 0.15, fp64, dev(amb),
                                ring_first=ring_first)
    assert bool(out.isfinite().all())


def stacked_rings(copies=4, segments=24, seed=0):
    """Rings released from one rim: ``copies`` identical rings (a starting-ring cluster, ring_wake.py:186-191),
    then a few displaced ones; every ring closed."""
    rng = np.random.default_rng(seed)
    phi = np.linspace(0.0, 2 * np.pi, segments, endpoint=False)
    rim = np.stack([np.zeros(segments), 1.2 * np.cos(phi), 1.2 * np.sin(phi)], axis=1)
    nodes, edges, gamma = [], [], []
    for r in range(copies + 3):
        offset = 0.0 if r < copies else 0.3 * (r - copies + 1)
        base = len(nodes) * segments
        nodes.append(rim + [-offset, 0.0, 0.0] + (0 if r < copies else 0.01 * rng.normal(size=rim.shape)))
        edges.append(np.stack([base + np.arange(segments), base + (np.arange(segments) + 1) % segments], axis=1))
        gamma.append(np.full(segments, 0.7 + 0.1 * r))
    return np.concatenate(nodes), np.concatenate(edges), np.concatenate(gamma)


def test_ring_self_velocity_and_ring_induction():
    from oracle import network_np as NP
    nodes, edges, gamma = stacked_rings()
    for model in (2, 1):
        print(f"ring self velocity: model {model}", flush=True)
        want = C.segments_velocity(nodes, nodes, edges, gamma, 0.3, model, NP.incident_pairs(nodes, edges, 0),
                                   "zero")
        got = CV.ring_self_velocity(dev(nodes), dev(edges), dev(gamma), 0.3, CV.ConvectionSettings(model=model))
        np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=1e-12 * np.abs(want).max())
    # stacked rings meet only at segment ends, which the incident pairs exclude: 'raise' passes there
    CV.ring_self_velocity(dev(nodes), dev(edges), dev(gamma), 0.3, CV.ConvectionSettings(model=2), "raise")
    # a node on the inside of another ring's filament: 'zero' gives the regularised limit, 'raise' reports it
    bad = nodes.copy()
    bad[-1] = 0.5 * (nodes[edges[0, 0]] + nodes[edges[0, 1]])
    out = CV.ring_self_velocity(dev(bad), dev(edges), dev(gamma), 0.3, CV.ConvectionSettings(model=2))
    assert bool(out.isfinite().all())
    with pytest.raises(V.FilamentSingularity):
        CV.ring_self_velocity(dev(bad), dev(edges), dev(gamma), 0.3, CV.ConvectionSettings(model=2), "raise")
    points = np.stack([np.full(30, 0.4), np.linspace(-2.0, 2.0, 30), np.zeros(30)], axis=1)
    want = C.segments_velocity(points, nodes, edges, gamma, 0.1, 2, np.zeros((0, 2), np.int64), "raise")
    got = CV.ring_induction(dev(points), dev(nodes), dev(edges), dev(gamma), 0.1, 2).cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
