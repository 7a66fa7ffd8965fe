"""Wake-network bookkeeping on the GPU: the incident-exclusion pairs of a free-wake step.

``incident_pairs(nodes, edges, first)`` is the counterpart of the reference's
``network_numba.incident_pairs`` (network_numba.py:242-253), which reproduces element for element the
exclusion loop of ``wake._network_convection`` (wake.py:427-438) and the rotor rings' self-convection call
(rotor/ring_wake.py:196): for the free points ``nodes[first:]``, every (point, segment) pair in which the
point coincides with an end of the segment, in segment order, the start end before the end end, and for one
end the coinciding points in increasing order.  Coordinates match by bit pattern (-0.0 and 0.0 differ), a
node with a NaN coordinate never matches, point indices count from ``first``.

The pairs are computed by ``ofw_incident_pairs`` of the HIP library; there is no CPU fallback.  Arguments the
reference hands to its dictionary version (wrong shapes, ``first`` outside [0, n], edge indices outside
[0, n)) raise ``ValueError`` here.
"""
import ctypes

import numpy as np

from . import _lib
from .vortex import _ctx, _torch


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
