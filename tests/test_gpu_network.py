"""GPU incident-exclusion pairs (ofw_incident_pairs) against the oracle, element for element: the reference's
test networks (tests/test_network_numba.py), signed zeros, NaN and infinite coordinates, self-loops, a group of
20,000 coincident points, the 819,826-segment crosswind-loop wake, and the C ABI's overflow and range errors.
The ring-edge aggregation (ofw_pack_ring_edges) bitwise against the oracle on the reference's pack_rings cases and a
3,744 x 110 lattice, with its overflow and panel errors."""
import ctypes

import numpy as np
import pytest

from oracle import network_np as NP
from openfawe_amd import _lib, workloads
from openfawe_amd import network as NET
from tests.test_oracle_network import ring_network, wavy_rows

pytestmark = pytest.mark.gpu


def check(nodes, edges, first):
    print(f"incident pairs: n={len(nodes)} m={len(edges)} first={first}", flush=True)
    got = NET.incident_pairs(nodes, edges, first)
    want = NP.incident_pairs(nodes, edges, first)
    assert got.dtype == np.int64 and got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    return got


@pytest.mark.parametrize("depth,width", [(2, 3), (3, 5), (12, 9), (41, 17), (400, 37)])
def test_ring_rows(depth, width):
    rng = np.random.default_rng(7 + depth)
    nodes, edges = ring_network(wavy_rows(rng, depth, width))
    pairs = check(nodes, edges, width)
    assert len(pairs) == int(np.sum(edges >= width))
    check(nodes, edges, 0)


def test_coincident_column_signed_zeros_nan_inf_self_loops():
    rng = np.random.default_rng(3)
    rows = wavy_rows(rng, 9, 8)
    rows[:, 4, :] = rows[:, 3, :]
    rows[0, 1, 2], rows[1, 1, 2] = -0.0, 0.0
    nodes, edges = ring_network(rows)
    edges = edges[rng.permutation(len(edges))]
    edges[::3] = edges[::3, ::-1]
    for first in (0, 8, 30, len(nodes)):
        check(nodes, edges, first)
    palette = np.array([[0.0, 0.0, 0.0], [-0.0, 0.0, 0.0], [1.0, 2.0, 3.0], [np.nan, 1.0, 1.0],
                        [np.inf, 0.0, 1.0], [1.0, 2.0, 3.0 + 2 ** -50]])
    for seed in range(4):
        r = np.random.default_rng(seed)
        nodes = palette[r.integers(0, len(palette), 5000)]
        edges = r.integers(0, 5000, size=(700, 2))
        edges[::50, 1] = edges[::50, 0]        # self-loops
        for first in (0, 1, 2500):
            check(nodes, edges, first)


def test_large_coincident_group_and_empty_inputs():
    rng = np.random.default_rng(9)
    nodes = rng.normal(size=(40_000, 3))
    nodes[::2] = [1.5, -2.0, 0.25]             # 20,000 coincident points, ranked by index
    edges = rng.integers(0, len(nodes), size=(3000, 2))
    edges[:7, 0] = 0
    pairs = check(nodes, edges, 0)
    assert len(pairs) > 7 * 20_000
    check(nodes, edges, 17)
    check(nodes, edges, len(nodes))
    check(nodes, edges[:0], 0)


def test_crosswind_loop_wake_820k():
    case = workloads.loop(rows=3744)
    nodes, edges = case["nodes"], case["edges"]
    pairs = check(nodes, edges, 0)
    # distinct nodes: one pair per segment end
    assert len(pairs) == 2 * len(edges)
    check(nodes, edges, 110 * 187)


def test_c_abi_overflow_range_and_determinism():
    import torch
    rng = np.random.default_rng(1)
    nodes, edges = ring_network(wavy_rows(rng, 30, 11))
    dn = torch.from_numpy(nodes).cuda()
    de = torch.from_numpy(edges).cuda()
    lib = _lib.load()
    from openfawe_amd.vortex import _ctx
    count = ctypes.c_int64(-1)
    out = torch.zeros((4, 2), dtype=torch.int64, device="cuda")
    status = lib.ofw_incident_pairs(_ctx(), len(nodes), dn.data_ptr(), len(edges), de.data_ptr(), 0,
                                    out.data_ptr(), 4, ctypes.byref(count))
    assert status == _lib.OFW_ERR_OVERFLOW and count.value == 2 * len(edges)
    assert int(out.abs().sum()) == 0            # nothing written on overflow
    a = NET.incident_pairs_device(dn, de, 0, capacity=3)      # retried with the exact count
    b = NET.incident_pairs_device(dn, de, 0)
    assert torch.equal(a, b)
    np.testing.assert_array_equal(a.cpu().numpy(), NP.incident_pairs(nodes, edges, 0))
    bad = de.clone()
    bad[5, 1] = len(nodes)
    with pytest.raises(ValueError):
        NET.incident_pairs_device(dn, bad, 0)
    bad[5, 1] = -3
    with pytest.raises(ValueError):
        NET.incident_pairs_device(dn, bad, 0)
    # the context still works after the errors
    assert torch.equal(NET.incident_pairs_device(dn, de, 0), b)


# ------------------------------------------------------------------ ring-edge aggregation
from tests.test_oracle_network import ring_cases  # noqa: E402


def check_rings(depth, width, panels, strengths):
    print(f"pack rings: depth={depth} width={width} panels={len(panels)}", flush=True)
    rows = np.zeros((depth, width, 3))
    rows[..., 0] = np.arange(depth)[:, None]
    rows[..., 1] = np.arange(width)[None, :]
    nodes, edges, gamma = NET.pack_rings(rows, strengths, panels)
    want_e, want_g = NP.pack_ring_edges_loop(depth, width, panels, strengths)
    np.testing.assert_array_equal(nodes, rows.reshape(-1, 3))
    assert edges.dtype == np.int64 and gamma.dtype == np.float64
    np.testing.assert_array_equal(edges, want_e)
    np.testing.assert_array_equal(gamma.view(np.int64), want_g.view(np.int64))   # bitwise
    return edges, gamma


@pytest.mark.parametrize("case", list(ring_cases()), ids=lambda c: c[0])
def test_pack_rings_bitwise(case):
    check_rings(*case[1:])


def test_pack_rings_full_wake_and_errors():
    import torch
    rng = np.random.default_rng(2)
    strengths = rng.normal(size=(3743, 109))
    strengths[1:, ::3] = strengths[:-1, ::3]          # cancelling shed edges
    strengths[::11] = 0.0
    panels = np.column_stack([np.arange(109), np.arange(1, 110)])
    edges, _ = check_rings(3744, 110, panels, strengths)
    assert len(edges) < 4 * 3743 * 109
    dp = torch.from_numpy(panels).cuda()
    ds = torch.from_numpy(strengths).cuda()
    e1, g1 = NET.pack_ring_edges_device(3744, 110, dp, ds, capacity=5)    # retried with the exact count
    e2, g2 = NET.pack_ring_edges_device(3744, 110, dp, ds)
    assert torch.equal(e1, e2) and torch.equal(g1.view(torch.int64), g2.view(torch.int64))
    bad = dp.clone()
    bad[7, 1] = 110
    with pytest.raises(ValueError):
        NET.pack_ring_edges_device(3744, 110, bad, ds)
    bad[7, 1] = bad[7, 0]
    with pytest.raises(ValueError):
        NET.pack_ring_edges_device(3744, 110, bad, ds)
    e3, g3 = NET.pack_ring_edges_device(3744, 110, dp, ds)
    assert torch.equal(e3, e2) and torch.equal(g3.view(torch.int64), g2.view(torch.int64))
