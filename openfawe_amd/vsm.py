"""VSM wake operators on the GPU: the attached-ring operator, the material edge cores and the history induction.

* ``ring_layout`` -- vsm_wake._ring_layout (vsm_wake.py:177-219): per panel, the (segment, sign) list of its rings
  over every row interval, equal to the edges of ``pack_rings`` with a unit strength on that panel and zero on the
  others (first-appearance order over all panels; edges shared by the panel's own rings cancel and are omitted;
  bound-row edges dropped at the load points).  Host integer bookkeeping on a few rows, cached per topology.
* ``batched_ring_operator`` -- vsm_wake._batched_ring_operator (vsm_wake.py:222-252): the (points, panels, 3)
  influence of every panel's unit rings at the points, by the pair matrix and the ordered column sums of this
  package's kernels; None when a segment has zero length or a pair would raise, as the reference.
* ``edge_radii`` -- cores.edge_radii (cores.py:129-149) on the GPU (``ofw_edge_radii``).
* ``history_induction`` -- history_vsm_induction (vsm_wake.py:255-278) for one surface class: the wake history's
  rings (``pack_rings`` on the GPU) with their material cores summed at the closure and the load points (FP64
  direct sum).
* ``history_induction_classes`` -- the same with surface classes (history_vsm_parts, vsm_wake.py:281-300): own-class
  edges with the material kernel, other-class edges with the Rosenhead-Moore kernel and the larger of the AVL and
  material cores.
"""
import numpy as np

from . import _lib
from . import adjoint as AD
from . import network as NET
from .convection import _check, network_segments
from .vortex import HipSettings, _ctx, _model, _ptr, _torch, hip_sum_device

_LAYOUTS = {}


def ring_layout(depth, width, panel_edges, at_load_points):
    """(segments (k,2), pointer (panels+1), column_segment, column_sign) of the unit rings of every panel."""
    panel_edges = np.ascontiguousarray(panel_edges, dtype=np.int64).reshape(-1, 2)
    key = (int(depth), int(width), panel_edges.tobytes(), bool(at_load_points))
    if key in _LAYOUTS:
        return _LAYOUTS[key]
    depth, width, panels = int(depth), int(width), len(panel_edges)
    # every ring edge in the pack_rings loop order: row interval outer, panel inner, the four edges of a ring
    i = np.repeat(np.arange(depth - 1), panels)
    j = np.tile(np.arange(panels), depth - 1)
    left, right = panel_edges[j, 0], panel_edges[j, 1]
    corners = np.stack([i * width + left, i * width + right, (i + 1) * width + right, (i + 1) * width + left], 1)
    a, b = corners.reshape(-1), np.roll(corners, -1, axis=1).reshape(-1)
    owner = np.repeat(j, 4)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    sign = np.where(a < b, 1.0, -1.0)
    # the first appearance of every undirected edge over all panels
    edge_key = lo * (depth * width) + hi
    _, first_at, edge_id = np.unique(edge_key, return_index=True, return_inverse=True)
    # per (panel, edge): the summed sign (small integers, exact)
    pk = owner * len(first_at) + edge_id
    cells, cell_first, cell_of = np.unique(pk, return_index=True, return_inverse=True)
    total = np.bincount(cell_of, weights=sign, minlength=len(cells))
    cell_panel, cell_edge = cells // len(first_at), cells % len(first_at)
    cell_lo, cell_hi = lo[cell_first], hi[cell_first]
    keep = total != 0.0
    if at_load_points:
        keep &= ~((cell_lo < width) & (cell_hi < width))
    cell_panel, cell_edge, total = cell_panel[keep], cell_edge[keep], total[keep]
    cell_lo, cell_hi = cell_lo[keep], cell_hi[keep]
    # columns in panel order, each in first-appearance order of its edges
    order = np.lexsort((first_at[cell_edge], cell_panel))
    cell_panel, cell_edge, total = cell_panel[order], cell_edge[order], total[order]
    cell_lo, cell_hi = cell_lo[order], cell_hi[order]
    # segments numbered by first use along the columns
    used, use_first, column_segment = np.unique(cell_edge, return_index=True, return_inverse=True)
    renumber = np.empty(len(used), dtype=np.int64)
    renumber[np.argsort(use_first, kind="stable")] = np.arange(len(used))
    column_segment = renumber[column_segment]
    segments = np.empty((len(used), 2), dtype=np.int64)
    segments[column_segment, 0] = cell_lo
    segments[column_segment, 1] = cell_hi
    pointer = np.zeros(panels + 1, dtype=np.int64)
    np.cumsum(np.bincount(cell_panel, minlength=panels), out=pointer[1:])
    layout = (segments, pointer, column_segment.astype(np.int64), total.astype(np.float64))
    _LAYOUTS[key] = layout
    return layout


def edge_radii(edges, width, panel_edges, initial, span=None, stream=None, offset=0):
    """Material core of every edge (device tensor (m,)), cores.edge_radii with WakeCores arrays span and stream
    (None: every edge takes ``initial``, as the reference without cores)."""
    torch = _torch()
    m = int(edges.shape[0])
    if span is None and stream is None:
        return torch.full((m,), float(initial), dtype=torch.float64, device=edges.device)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=edges.device)  # noqa: E731
    span = dev(np.zeros((0, len(panel_edges))) if span is None else np.asarray(span, dtype=np.float64))
    stream = dev(np.zeros((0, width)) if stream is None else np.asarray(stream, dtype=np.float64))
    panels = dev(np.asarray(panel_edges, dtype=np.int64).reshape(-1, 2))
    out = torch.empty(m, dtype=torch.float64, device=edges.device)
    status = _lib.load().ofw_edge_radii(_ctx(), m, _ptr(edges.contiguous()), int(offset), int(width), len(panels),
                                        _ptr(panels), int(span.shape[0]), _ptr(span), int(stream.shape[0]),
                                        _ptr(stream), float(initial), _ptr(out))
    _lib.check(status, "ofw_edge_radii")
    return out


def batched_ring_operator(rows, points, panel_edges, model, core_radius, at_load_points, span=None, stream=None):
    """Influence (points, panels, 3) of each panel's unit rings over ``rows`` (depth, width, 3) at ``points``, as
    vsm_wake._batched_ring_operator; None when a segment has zero length or a pair would raise."""
    torch = _torch()
    rows = np.asarray(rows, dtype=np.float64)
    depth, width = rows.shape[:2]
    segments, pointer, column_segment, column_sign = ring_layout(depth, width, panel_edges, at_load_points)
    dev = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)  # noqa: E731
    nodes = dev(rows.reshape(-1, 3))
    seg = dev(segments)
    try:
        s, d, L = network_segments(nodes, seg)
    except ValueError:
        return None                                   # a zero-length segment (vsm_wake.py:236-237)
    radii = edge_radii(seg, width, panel_edges, core_radius, span, stream)
    pts = dev(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    pairs, flags = AD.pair_matrix_device(pts, s, d, L, radii, int(_model(model)), "raise")
    if bool(torch.logical_or(flags.any(), torch.logical_not(torch.isfinite(pairs).all())).item()):
        return None
    return AD.assemble_columns(pairs, dev(pointer), dev(column_segment), dev(column_sign)).cpu().numpy()


def history_induction(closure_points, load_points, rows, strengths, panel_edges, model, core_radius, span=None,
                      stream=None, settings=HipSettings("fp64")):
    """(at closure points, at load points) of the wake history's rings, history_vsm_induction for one class.

    rows (k, width, 3) are the history rows with the newest first (row two of the complete wake network, so the
    cores are looked up with an offset of two rows), strengths (k-1, panels)."""
    torch = _torch()
    closure = np.asarray(closure_points, dtype=np.float64).reshape(-1, 3)
    load = np.asarray(load_points, dtype=np.float64).reshape(-1, 3)
    if len(rows) <= 1:
        return np.zeros_like(closure), np.zeros_like(load)
    rows = np.asarray(rows, dtype=np.float64)
    depth, width = rows.shape[:2]
    rows, strengths, panels = NET._check_rings(rows, strengths, panel_edges)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")  # noqa: E731
    edges, gamma = NET.pack_ring_edges_device(depth, width, dev(panels), dev(strengths))
    nodes = dev(rows.reshape(-1, 3))
    radii = edge_radii(edges, width, panels, core_radius, span, stream, offset=2 * width)
    s, d, L = network_segments(nodes, edges)
    points = dev(np.concatenate([closure, load]))
    ptr = torch.zeros(len(points) + 1, dtype=torch.int64, device="cuda")
    seg = torch.zeros(1, dtype=torch.int64, device="cuda")
    out, flags, flag_seg = hip_sum_device(points, s, d, L, gamma, radii, int(_model(model)), ptr, seg, "raise",
                                          settings)
    _check(out, flags, flag_seg)
    out = out.cpu().numpy()
    return out[:len(closure)], out[len(closure):]


CROSS_MODEL = 5  # Rosenhead-Moore: the finite core between surface classes (surface_cores.py:31)


def history_induction_classes(closure_points, load_points, rows, strengths, panel_edges, model, core_radius,
                              panel_class, node_class, panel_radius, node_radius, span=None, stream=None,
                              settings=HipSettings("fp64")):
    """history_vsm_induction with surface classes (vsm_wake.py:268-300, history_vsm_parts): the points of each class
    take the history edges of their own class with the material kernel and cores, and the edges of the other
    classes with the Rosenhead-Moore kernel and max(AVL radius, material core) (surface_cores.py:112-131: a spanwise
    edge its strip's radius, a streamwise edge its node's).  Closure and load points are one per panel, of the
    panel's class.  Returns (at closure points, at load points), the sum of the two parts."""
    torch = _torch()
    closure = np.asarray(closure_points, dtype=np.float64).reshape(-1, 3)
    load = np.asarray(load_points, dtype=np.float64).reshape(-1, 3)
    panel_class = np.asarray(panel_class, dtype=np.int64)
    node_class = np.asarray(node_class, dtype=np.int64)
    if len(closure) != len(panel_class) or len(load) != len(panel_class):
        raise ValueError("one closure and one load point per panel")
    if len(rows) <= 1:
        return np.zeros_like(closure), np.zeros_like(load)
    rows, strengths, panels = NET._check_rings(np.asarray(rows, dtype=np.float64), strengths, panel_edges)
    depth, width = rows.shape[:2]
    if len(node_class) != width or len(panel_radius) != len(panels) or len(node_radius) != width:
        raise ValueError("one class per node and panel, one radius per panel and node")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")  # noqa: E731
    edges, gamma = NET.pack_ring_edges_device(depth, width, dev(panels), dev(strengths))
    nodes = dev(rows.reshape(-1, 3))
    radii = edge_radii(edges, width, panels, core_radius, span, stream, offset=2 * width)
    # the AVL radius of every edge (rows from two on): the strips' radius per row for spanwise edges, the nodes'
    # for streamwise edges
    avl = edge_radii(edges, width, panels, 0.0, np.tile(np.asarray(panel_radius, np.float64), (depth + 2, 1)),
                     np.tile(np.asarray(node_radius, np.float64), (depth + 2, 1)), offset=2 * width)
    cross_radii = torch.maximum(radii, avl)
    source = dev(node_class)[edges[:, 0] % width]
    points = np.concatenate([closure, load])
    point_class = np.concatenate([panel_class, panel_class])
    out = torch.zeros((len(points), 3), dtype=torch.float64, device="cuda")
    geometry = network_segments(nodes, edges)

    def part(target, keep, radius, kernel):
        sub = torch.nonzero(keep).reshape(-1)
        if sub.numel() == 0:
            return None
        pts = dev(points[target])
        ptr = torch.zeros(len(target) + 1, dtype=torch.int64, device="cuda")
        seg = torch.zeros(1, dtype=torch.int64, device="cuda")
        s, d, L = (g[sub].contiguous() for g in geometry)
        res, flags, flag_seg = hip_sum_device(pts, s, d, L, gamma[sub].contiguous(), radius[sub].contiguous(),
                                              kernel, ptr, seg, "raise", settings)
        _check(res, flags, flag_seg)
        return res

    model = int(_model(model))
    for c in np.unique(panel_class):
        target = np.flatnonzero(point_class == c)
        same = source == int(c)
        index = torch.from_numpy(target).to("cuda")
        for keep, radius, kernel in ((same, radii, model), (~same, cross_radii, CROSS_MODEL)):
            res = part(target, keep, radius, kernel)
            if res is not None:
                out[index] += res
    out = out.cpu().numpy()
    return out[:len(closure)], out[len(closure):]
