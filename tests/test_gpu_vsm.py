"""VSM wake operators on the GPU (openfawe_amd.vsm) against the oracle (oracle/vsm_np.py): the material edge cores
exactly, the batched attached-ring operator and the history induction within the direct sum's A1 bar, and the
operator's None cases (a zero-length segment, a load point on a bound vortex that is not dropped)."""
import numpy as np
import pytest

from oracle import network_np as NP
from oracle import vsm_np as VO
from openfawe_amd import vsm

pytestmark = pytest.mark.gpu


def wing_rows(depth, width=11, seed=0):
    rng = np.random.default_rng(seed)
    y = np.linspace(-5.0, 5.0, width)
    rows = np.zeros((depth, width, 3))
    rows[0, :, 0] = 0.25                                    # bound vortex at quarter chord
    rows[1, :, 0] = 1.0                                     # trailing edge
    for r in range(2, depth):
        rows[r, :, 0] = 1.0 + 0.6 * (r - 1) + 0.02 * rng.normal(size=width)
        rows[r, :, 2] = -0.03 * (r - 1) + 0.01 * rng.normal(size=width)
    rows[..., 1] = y
    panels = np.column_stack([np.arange(width - 1), np.arange(1, width)])
    closure = np.stack([np.full(width - 1, 0.75), 0.5 * (y[:-1] + y[1:]), np.zeros(width - 1)], axis=1)
    load = np.stack([np.full(width - 1, 0.25), 0.5 * (y[:-1] + y[1:]), np.zeros(width - 1)], axis=1)
    return rows, panels, closure, load


def cores(depth, width, seed=1):
    rng = np.random.default_rng(seed)
    return 0.1 + 0.05 * rng.random((depth, width - 1)), 0.1 + 0.05 * rng.random((depth, width))


def test_edge_radii_exact_and_errors():
    import torch
    depth, width = 12, 11
    rows, panels, _, _ = wing_rows(depth, width)
    span, stream = cores(depth, width)
    strengths = np.random.default_rng(2).normal(size=(depth - 1, width - 1))
    edges, _ = NP.pack_ring_edges(depth, width, panels, strengths)
    want = VO.edge_radii(edges, width, panels, 0.15, span, stream)
    got = vsm.edge_radii(torch.from_numpy(edges).cuda(), width, panels, 0.15, span, stream).cpu().numpy()
    assert np.array_equal(got, want)
    hist = VO.edge_radii(edges + 2 * width, width, panels, 0.15, span, stream)
    got = vsm.edge_radii(torch.from_numpy(edges).cuda(), width, panels, 0.15, span, stream, offset=2 * width)
    assert np.array_equal(got.cpu().numpy(), hist)
    with pytest.raises(ValueError, match="spanwise or streamwise"):
        vsm.edge_radii(torch.tensor([[0, width + 1]]).cuda(), width, panels, 0.15, span, stream)   # a diagonal
    with pytest.raises(ValueError, match="core history"):
        vsm.edge_radii(torch.from_numpy(edges).cuda(), width, panels, 0.15, span[:3], stream[:3])


@pytest.mark.parametrize("depth", [3, 5])
@pytest.mark.parametrize("with_cores", [False, True])
def test_batched_ring_operator(depth, with_cores):
    rows, panels, closure, load = wing_rows(depth)
    span, stream = cores(depth, rows.shape[1]) if with_cores else (None, None)
    for points, at_load in ((closure, False), (load, True), (closure, True)):
        print(f"ring operator: depth={depth} cores={with_cores} at_load={at_load}", flush=True)
        got = vsm.batched_ring_operator(rows, points, panels, 2, 0.15, at_load, span, stream)
        want = VO.ring_operator(rows, points, panels, 2, 0.15, at_load, span, stream)
        assert got.shape == want.shape == (len(points), len(panels), 3)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * np.abs(want).max())
    # load points on the bound vortex with the bound edges kept: a pair raises -> None; a zero-length segment too
    assert vsm.batched_ring_operator(rows, load, panels, 0, 0.15, False) is None
    flat = rows.copy()
    flat[1, 3] = flat[0, 3]
    assert vsm.batched_ring_operator(flat, closure, panels, 2, 0.15, False) is None


@pytest.mark.parametrize("with_cores", [False, True])
def test_history_induction(with_cores):
    depth = 40
    rows, panels, closure, load = wing_rows(depth + 2)
    history = rows[2:]                                       # newest history row is row two of the network
    span, stream = cores(depth + 2, rows.shape[1]) if with_cores else (None, None)
    strengths = 1.0 + 0.2 * np.random.default_rng(3).normal(size=(depth - 1, len(panels)))
    got = vsm.history_induction(closure, load, history, strengths, panels, 2, 0.15, span, stream)
    want = VO.history_induction(closure, load, history, strengths, panels, 2, 0.15, span, stream)
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, w, rtol=0, atol=1e-11 * np.abs(w).max())
    zeros = vsm.history_induction(closure, load, history[:1], strengths[:0], panels, 2, 0.15)
    assert not np.any(zeros[0]) and not np.any(zeros[1])


def test_history_induction_with_surface_classes():
    depth = 24
    rows, _, closure_all, load_all = wing_rows(depth + 2, width=11)
    # class 0: nodes 0-6 (panels 0-5); class 1: nodes 7-10 (panels 7-9); no panel joins the two
    panels = np.array([[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [5, 6], [7, 8], [8, 9], [9, 10]])
    keep = [0, 1, 2, 3, 4, 5, 7, 8, 9]
    closure, load = closure_all[keep], load_all[keep]
    panel_class = np.array([0] * 6 + [1] * 3)
    node_class = np.array([0] * 7 + [1] * 4)
    panel_radius = 0.2 + 0.05 * np.arange(9)
    node_radius = np.zeros(11)
    for j, (a, b) in enumerate(panels):
        node_radius[a] = max(node_radius[a], panel_radius[j])
        node_radius[b] = max(node_radius[b], panel_radius[j])
    strengths = 1.0 + 0.2 * np.random.default_rng(4).normal(size=(depth - 1, len(panels)))
    history = rows[2:]
    span, stream = cores(depth + 2, 11)
    span = span[:, :len(panels)]
    for s_, st_ in ((None, None), (span, stream)):
        got = vsm.history_induction_classes(closure, load, history, strengths, panels, 2, 0.15, panel_class,
                                            node_class, panel_radius, node_radius, s_, st_)
        want = VO.history_induction_classes(closure, load, history, strengths, panels, 2, 0.15, panel_class,
                                            node_class, panel_radius, node_radius, s_, st_)
        for g, w in zip(got, want):
            np.testing.assert_allclose(g, w, rtol=0, atol=1e-11 * np.abs(w).max())
    # one class: the single-class history induction
    one = vsm.history_induction_classes(closure, load, history, strengths, panels, 2, 0.15, np.zeros(9, int),
                                        np.zeros(11, int), panel_radius, node_radius)
    plain = vsm.history_induction(closure, load, history, strengths, panels, 2, 0.15)
    for g, w in zip(one, plain):
        np.testing.assert_allclose(g, w, rtol=0, atol=1e-12 * np.abs(w).max())
