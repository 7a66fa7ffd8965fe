"""Wake-network bookkeeping on the GPU: the incident-exclusion pairs of a free-wake step.

``incident_pairs(nodes, edges, first)`` is the counterpart of the reference's
``network_numba.incident_pairs`` (network_numba.py:242-253), which reproduces element for element the
exclusion loop of ``wake._network_convection`` (wake.py:427-438) and the rotor rings' self-convection call
(rotor/ring_wake.py:196): for the free points ``nodes[first:]``, every (point, segment) pair in which the
point coincides with an end of the segment, in segment order, the start end before the end end, and for one
end the coinciding points in increasing order.  Coordinates match by bit pattern (-0.0 and 0.0 differ), a
node with a NaN coordinate never matches, point indices count from ``first``.

``pack_rings(rows, strengths, panel_edges)`` is the counterpart of ``wake.pack_rings(..., compiled=True)``
(wake.py:278-307, network_numba.pack_ring_edges :112-117): the directed edges of a lattice of vortex rings, each
shared edge once with its circulations summed in ring order, in order of first appearance, exact zeros omitted.

Both are computed by the HIP library (``ofw_incident_pairs``, ``ofw_pack_ring_edges``); there is no CPU fallback.
Arguments the reference hands to its dictionary version of ``incident_pairs`` (wrong shapes, ``first`` outside
[0, n], edge indices outside [0, n)) raise ``ValueError`` here.
"""
import ctypes

import numpy as np

from . import _lib
from .vortex import _ctx, _real, _torch


def _check_host(nodes, edges, first):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    edges = np.ascontiguousarray(edges, dtype=np.int64)
    if nodes.ndim != 2 or nodes.shape[1] != 3:
        raise ValueError("nodes must have shape (n, 3)")
    if edges.ndim != 2 or edges.shape[1] != 2:
        raise ValueError("edges must have shape (e, 2)")
    first = int(first)
    if not 0 <= first <= len(nodes):
        raise ValueError("first must lie in [0, len(nodes)]")
    if edges.size and (edges.min() < 0 or edges.max() >= len(nodes)):
        raise ValueError("edge index out of range")
    return nodes, edges, first


def incident_pairs_device(nodes, edges, first, capacity=None):
    """Pairs as an (e, 2) int64 device tensor from device tensors nodes (n,3) float64 and edges (m,2) int64.

    ``capacity`` is the first guess of the pair count (default: one pair per segment end); a short guess costs
    a second call with the exact count.
    """
    torch = _torch()
    if nodes.dtype != torch.float64 or nodes.dim() != 2 or nodes.shape[1] != 3 or not nodes.is_cuda:
        raise ValueError("nodes must be a (n, 3) float64 GPU tensor")
    if edges.dtype != torch.int64 or edges.dim() != 2 or edges.shape[1] != 2 or not edges.is_cuda:
        raise ValueError("edges must be a (e, 2) int64 GPU tensor")
    nodes, edges = nodes.contiguous(), edges.contiguous()
    n, m, first = int(nodes.shape[0]), int(edges.shape[0]), int(first)
    if not 0 <= first <= n:
        raise ValueError("first must lie in [0, len(nodes)]")
    if capacity is None:
        capacity = 2 * m
    lib = _lib.load()
    count = ctypes.c_int64(0)
    for _ in range(2):
        pairs = torch.empty((max(int(capacity), 1), 2), dtype=torch.int64, device=nodes.device)
        status = lib.ofw_incident_pairs(_ctx(), n, nodes.data_ptr() if n else None, m,
                                        edges.data_ptr() if m else None, first, pairs.data_ptr(), int(capacity),
                                        ctypes.byref(count))
        if status != _lib.OFW_ERR_OVERFLOW:
            _lib.check(status, "ofw_incident_pairs")
            return pairs[:count.value]
        capacity = count.value
    _lib.check(status, "ofw_incident_pairs")
    return pairs[:count.value]


def incident_pairs(nodes, edges, first):
    """Exclusion pairs of wake._network_convection as an (e, 2) int64 NumPy array (network_numba.py:242-253)."""
    nodes, edges, first = _check_host(nodes, edges, first)
    torch = _torch()
    dev_nodes = torch.from_numpy(nodes).to("cuda")
    dev_edges = torch.from_numpy(edges).to("cuda")
    return incident_pairs_device(dev_nodes, dev_edges, first).cpu().numpy()


def _check_rings(rows, strengths, panel_edges):
    """The validation of wake.pack_rings (wake.py:286-295)."""
    rows, strengths = _real(rows, "rows"), _real(strengths, "ring strengths")
    if rows.ndim != 3 or rows.shape[2] != 3 or rows.shape[1] < 2 or len(rows) < 2:
        raise ValueError("rows must have shape (depth+1,span+1,3)")
    depth, width = rows.shape[:2]
    if panel_edges is None:
        panel_edges = np.column_stack([np.arange(width - 1), np.arange(1, width)])
    panel_edges = np.asarray(panel_edges)
    if (panel_edges.ndim != 2 or panel_edges.shape[1] != 2 or panel_edges.dtype.kind not in "iu"
            or np.any(panel_edges < 0) or np.any(panel_edges >= width)
            or np.any(panel_edges[:, 0] == panel_edges[:, 1])):
        raise ValueError("invalid spanwise panel connectivity")
    if strengths.shape != (depth - 1, len(panel_edges)):
        raise ValueError("ring strength shape mismatch")
    return rows, strengths, np.ascontiguousarray(panel_edges, dtype=np.int64)


def pack_ring_edges_device(depth, width, panel_edges, strengths, capacity=None):
    """(edges (e,2) int64, gamma (e,) float64) device tensors from device tensors panel_edges (p,2) int64 and
    strengths (depth-1, p) float64.  ``capacity`` is the first guess of the edge count (default: every
    contribution); a short guess costs a second call with the exact count."""
    torch = _torch()
    if panel_edges.dtype != torch.int64 or panel_edges.dim() != 2 or panel_edges.shape[1] != 2:
        raise ValueError("panel_edges must be a (p, 2) int64 GPU tensor")
    depth, width, panels = int(depth), int(width), int(panel_edges.shape[0])
    if strengths.dtype != torch.float64 or tuple(strengths.shape) != (depth - 1, panels):
        raise ValueError("strengths must be a (depth-1, p) float64 GPU tensor")
    panel_edges, strengths = panel_edges.contiguous(), strengths.contiguous()
    if capacity is None:
        capacity = 4 * max(depth - 1, 0) * panels
    lib = _lib.load()
    count = ctypes.c_int64(0)
    for _ in range(2):
        cap = max(int(capacity), 1)
        edges = torch.empty((cap, 2), dtype=torch.int64, device=strengths.device)
        gamma = torch.empty(cap, dtype=torch.float64, device=strengths.device)
        status = lib.ofw_pack_ring_edges(_ctx(), depth, width, panels, panel_edges.data_ptr() if panels else None,
                                         strengths.data_ptr() if strengths.numel() else None, edges.data_ptr(),
                                         gamma.data_ptr(), int(capacity), ctypes.byref(count))
        if status != _lib.OFW_ERR_OVERFLOW:
            _lib.check(status, "ofw_pack_ring_edges")
            return edges[:count.value], gamma[:count.value]
        capacity = count.value
    _lib.check(status, "ofw_pack_ring_edges")
    return edges[:count.value], gamma[:count.value]


def pack_rings(rows, strengths, panel_edges=None):
    """(nodes (n,3), edges (e,2) int64, gamma (e,)) of wake.pack_rings(..., compiled=True), as NumPy arrays."""
    rows, strengths, panel_edges = _check_rings(rows, strengths, panel_edges)
    torch = _torch()
    depth, width = rows.shape[:2]
    edges, gamma = pack_ring_edges_device(depth, width, torch.from_numpy(panel_edges).to("cuda"),
                                          torch.from_numpy(np.ascontiguousarray(strengths)).to("cuda"))
    return rows.reshape(-1, 3).copy(), edges.cpu().numpy(), gamma.cpu().numpy()
