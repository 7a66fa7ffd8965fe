# SPDX-License-Identifier: Apache-2.0
# Test oracle: a restatement of OpenFAWE's src/openfawe/aero/vsm_wake.py and cores.py operators (Copyright 2026 Yinong Tian, Apache License 2.0, itself
# adapted from OpenFAST/OLAF; see NOTICE).
"""NumPy restatement of the reference's VSM wake operators.  TEST INFRASTRUCTURE ONLY.

* ``ring_columns`` -- the definition of vsm_wake._ring_layout (vsm_wake.py:177-185) and of the per-panel path it
  replaces (wake.current_influence, wake.py:640-651): column j is pack_rings with a unit strength on panel j in
  every row interval and zero on the other panels (so edges keep the first-appearance order over all panels),
  exact zeros omitted, bound-row edges dropped at the load points.
* ``edge_radii`` -- cores.edge_radii (cores.py:129-149), the loop over edges.
* ``ring_operator`` -- matrix[:, j] = the direct sum of column j's edges with their summed unit strengths
  (the per-panel segments_velocity calls of wake.py:646-649), by the C oracle.
* ``history_induction`` -- history_vsm_induction (vsm_wake.py:255-278) for one class: pack_rings, the cores with the
  two-row offset, the direct sum at the closure and load points.
"""
import numpy as np

from . import network_np
from . import oracle as O


def ring_columns(depth, width, panel_edges, at_load_points):
    """[(edges (k,2), signs (k,))] per panel."""
    panel_edges = np.asarray(panel_edges, dtype=np.int64).reshape(-1, 2)
    columns = []
    for j in range(len(panel_edges)):
        strengths = np.zeros((depth - 1, len(panel_edges)))
        strengths[:, j] = 1.0
        edges, gamma = network_np.pack_ring_edges_loop(depth, width, panel_edges, strengths)
        if at_load_points:
            keep = ~((edges[:, 0] < width) & (edges[:, 1] < width))
            edges, gamma = edges[keep], gamma[keep]
        columns.append((edges, gamma))
    return columns


def edge_radii(edges, width, panel_edges, initial, span=None, stream=None):
    """cores.edge_radii: the material core of every globally numbered wake edge."""
    edges = np.asarray(edges)
    result = np.full(len(edges), float(initial))
    if span is None and stream is None:
        return result
    lookup = {}
    for j, (a, b) in enumerate(np.asarray(panel_edges)):
        lookup[(min(a, b), max(a, b))] = j
    for i, (first, last) in enumerate(edges):
        ra, a = divmod(int(first), width)
        rb, b = divmod(int(last), width)
        if ra == rb:
            if ra >= 2:
                result[i] = span[ra - 2, lookup[(min(a, b), max(a, b))]]
        elif a == b and abs(rb - ra) == 1:
            if min(ra, rb) >= 1:
                result[i] = stream[min(ra, rb) - 1, a]
        else:
            raise ValueError("edge is not a spanwise or streamwise wake segment")
    return result


def _geometry(nodes, edges):
    starts = nodes[edges[:, 0]]
    delta = nodes[edges[:, 1]] - starts
    lengths = np.hypot(np.hypot(delta[:, 0], delta[:, 1]), delta[:, 2])
    return starts, delta / lengths[:, None], lengths


def ring_operator(rows, points, panel_edges, model, core_radius, at_load_points, span=None, stream=None):
    rows = np.asarray(rows, dtype=np.float64)
    depth, width = rows.shape[:2]
    nodes = rows.reshape(-1, 3)
    columns = ring_columns(depth, width, panel_edges, at_load_points)
    matrix = np.zeros((len(points), len(columns), 3))
    for j, (edges, gamma) in enumerate(columns):
        radii = edge_radii(edges, width, panel_edges, core_radius, span, stream)
        out, flags = O.direct_sum(points, *_geometry(nodes, edges), gamma, radii, model)
        assert not flags.any()
        matrix[:, j] = out
    return matrix


def history_induction(closure, load, rows, strengths, panel_edges, model, core_radius, span=None, stream=None):
    rows = np.asarray(rows, dtype=np.float64)
    depth, width = rows.shape[:2]
    edges, gamma = network_np.pack_ring_edges_loop(depth, width, panel_edges, strengths)
    radii = edge_radii(edges + 2 * width, width, panel_edges, core_radius, span, stream)
    points = np.concatenate([closure, load])
    out, flags = O.direct_sum(points, *_geometry(rows.reshape(-1, 3), edges), gamma, radii, model)
    assert not flags.any()
    return out[:len(closure)], out[len(closure):]


def history_induction_classes(closure, load, rows, strengths, panel_edges, model, core_radius, panel_class, node_class,
                              panel_radius, node_radius, span=None, stream=None):
    """history_vsm_parts (vsm_wake.py:281-300) with the AVL radii of surface_cores.py:112-131."""
    rows = np.asarray(rows, dtype=np.float64)
    depth, width = rows.shape[:2]
    panel_edges = np.asarray(panel_edges, dtype=np.int64)
    edges, gamma = network_np.pack_ring_edges_loop(depth, width, panel_edges, strengths)
    glob = edges + 2 * width
    radii = edge_radii(glob, width, panel_edges, core_radius, span, stream)
    lookup = {}
    for j, (a, b) in enumerate(panel_edges):
        lookup[(a, b)] = lookup[(b, a)] = j
    avl = np.empty(len(edges))
    for i, (e0, e1) in enumerate(glob):
        ra, a = divmod(int(e0), width)
        rb, b = divmod(int(e1), width)
        avl[i] = panel_radius[lookup[(a, b)]] if ra == rb else node_radius[a]
    cross = np.maximum(radii, avl)
    source = np.asarray(node_class)[edges[:, 0] % width]
    points = np.concatenate([closure, load])
    point_class = np.concatenate([panel_class, panel_class])
    nodes = rows.reshape(-1, 3)
    same_part, cross_part = np.zeros_like(points), np.zeros_like(points)
    for c in np.unique(panel_class):
        target = np.flatnonzero(point_class == c)
        same = source == c
        for keep, part, kernel, rad in ((same, same_part, model, radii), (~same, cross_part, 5, cross)):
            if np.any(keep):
                out, flags = O.direct_sum(points[target], *_geometry(nodes, edges[keep]), gamma[keep], rad[keep],
                                          kernel)
                assert not flags.any()
                part[target] = out
    total = same_part + cross_part
    return total[:len(closure)], total[len(closure):]
