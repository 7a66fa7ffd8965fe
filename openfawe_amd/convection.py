"""The free-wake geometry step on the GPU: ``_network_convection`` and the stage loop of ``convect_geometry``.

``network_convection`` is the counterpart of ``wake._network_convection`` (wake.py:404-491): the velocity of the
free points ``nodes[first:]`` of a frozen circulation graph is the ambient velocity plus the induced velocity of
every segment, with the segments incident to a point excluded (``incident_pairs``), split as the reference splits
the targets:

* wing rows ``[0, rows_end)`` with singularity 'raise' and the edge cores ``radii``;
* with ``free_convection_rows`` (WakeConfig), only the newest ``limit * first`` row points are targets; the
  frozen rows keep the ambient velocity or, with ``frozen_rows_drift='ambient_rings'``, take the induction of the
  ring edges alone (both nodes ring nodes; wake.py:453-462);
* ring nodes ``[rows_end, ring_end)`` (the leading ``ring_free`` of them, all by default) with singularity 'zero'
  and the cores ``ring_radii`` (default ``radii``).

``convect_geometry`` is the stage loop of ``wake.convect_geometry`` (wake.py:559-616) on that graph: Euler, Heun
or RK4, ``substeps`` substeps, the bound row moved linearly by ``displacement`` over the step.

Everything stays in HBM: the exclusion pairs (``ofw_incident_pairs``), the segment geometry of each stage
(``ofw_network_segments``), the exclusion CSR of each target range (``ofw_exclusion_subset``), the sums (the
direct sum or the tree code of this package) and the stage arithmetic (``ofw_stage_combine``, NumPy's order of
operations).  The host reads flags and error status only, except for the reference's surface guard, which is
given the rows of each stage on the host.  Not modelled: callable inflows that need host arrays (an inflow is a
3-vector or a callable on device tensors).
"""
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import network as NET
from .vortex import (HipSettings, Statistics, TreeSettings, _ctx, _model, _ptr, _raise_for_flags, _torch,
                     hip_sum_device)

INTEGRATORS = ("euler", "heun", "rk4")


@dataclass(frozen=True)
class ConvectionSettings:
    """The WakeConfig fields of the convection step (wake.py:67-133) and the engine of its sums.

    backend 'hip' is the direct sum, 'hip_tree' the tree code (``tree`` required), both with ``hip`` settings
    (the reference's kernel_options(config, large=True), wake.py:24-38)."""
    model: int = 2
    free_wake: bool = True
    free_convection_rows: int | None = None
    frozen_rows_drift: str = "ambient"
    backend: str = "hip"
    tree: TreeSettings | None = None
    hip: HipSettings = HipSettings("fp64")

    def __post_init__(self):
        object.__setattr__(self, "model", int(_model(self.model)))
        if self.backend not in ("hip", "hip_tree"):
            raise ValueError("backend must be 'hip' or 'hip_tree'")
        if (self.backend == "hip_tree") != isinstance(self.tree, TreeSettings):
            raise ValueError("tree settings are required exactly for backend='hip_tree'")
        if not isinstance(self.hip, HipSettings):
            raise ValueError("hip must be HipSettings")
        if self.frozen_rows_drift not in ("ambient", "ambient_rings"):
            raise ValueError("frozen_rows_drift must be 'ambient' or 'ambient_rings'")
        limit = self.free_convection_rows
        if limit is not None and (isinstance(limit, bool) or int(limit) != limit or limit < 1):
            raise ValueError("free_convection_rows must be a positive integer")


# ------------------------------------------------------------------ device pieces
def network_segments(nodes, edges):
    """(starts (m,3), directions (m,3), lengths (m)) of the network's segments (vortex.py:458-465)."""
    torch = _torch()
    m = int(edges.shape[0])
    starts = torch.empty((m, 3), dtype=torch.float64, device=nodes.device)
    directions = torch.empty_like(starts)
    lengths = torch.empty(m, dtype=torch.float64, device=nodes.device)
    status = _lib.load().ofw_network_segments(_ctx(), int(nodes.shape[0]), _ptr(nodes), m, _ptr(edges),
                                              _ptr(starts), _ptr(directions), _ptr(lengths))
    _lib.check(status, "ofw_network_segments")
    return starts, directions, lengths


def exclusion_subset(pairs, lo, hi, m, remap=None, point_map=None, count=None, segment_ordered=False):
    """Exclusion CSR (ptr (count+1), seg) of the pairs whose point lies in [lo, hi): points shifted by -lo, or
    mapped through ``point_map`` (hi - lo entries, negative = dropped) onto ``count`` output points.
    ``segment_ordered``: the pairs' segments are non-decreasing (incident pairs), which saves a sort."""
    torch = _torch()
    count = int(hi) - int(lo) if count is None else int(count)
    ptr = torch.empty(count + 1, dtype=torch.int64, device=pairs.device)
    seg = torch.empty(max(int(pairs.shape[0]), 1), dtype=torch.int64, device=pairs.device)
    status = _lib.load().ofw_exclusion_subset(_ctx(), int(pairs.shape[0]), _ptr(pairs), int(lo), int(hi), count,
                                              _ptr(point_map), int(m), _ptr(remap), int(bool(segment_ordered)),
                                              _ptr(ptr), _ptr(seg))
    _lib.check(status, "ofw_exclusion_subset")
    return ptr, seg


def stage(mode, h, x, a, b=None, c=None, d=None, out=None):
    """``out = `` the stage arithmetic ``mode`` (the _lib.STAGE_* values) over whole tensors."""
    torch = _torch()
    if out is None:
        out = torch.empty_like(x)
    status = _lib.load().ofw_stage_combine(_ctx(), int(mode), int(x.numel()), float(h), _ptr(x), _ptr(a), _ptr(b),
                                           _ptr(c), _ptr(d), _ptr(out))
    _lib.check(status, "ofw_stage_combine")
    return out


def _sum(points, geometry, gamma, radius, ptr, seg, singularity, settings):
    """The engine's sum on device tensors: (out, flags, flag_seg or None)."""
    starts, directions, lengths = geometry
    if settings.backend == "hip_tree":
        from .tree import hip_tree_sum_device
        out, flags, _, _ = hip_tree_sum_device(points, starts, directions, lengths, gamma, radius, settings.model,
                                               ptr, seg, singularity, settings.tree, settings.hip)
        return out, flags, None
    return hip_sum_device(points, starts, directions, lengths, gamma, radius, settings.model, ptr, seg, singularity,
                          settings.hip)


def _check(out, flags, flag_seg):
    """segments_velocity's error rules (vortex.py:488-497) for one sum; one host synchronisation."""
    torch = _torch()
    if not bool(torch.logical_or(flags.any(), torch.logical_not(torch.isfinite(out).all())).item()):
        return
    if bool(flags.any()):
        stats = dict(flags=flags.cpu().numpy())
        if flag_seg is not None:
            stats["flag_seg"] = flag_seg.cpu().numpy()
        Statistics.last = stats
        _raise_for_flags(stats)
    raise ValueError("velocity exceeds finite float64 range")


def _induced(points, geometry, gamma, radius, ptr, seg, singularity, settings):
    """segments_velocity's sum and error rules (vortex.py:466-497) on device tensors."""
    out, flags, flag_seg = _sum(points, geometry, gamma, radius, ptr, seg, singularity, settings)
    _check(out, flags, flag_seg)
    return out


def _per_edge(torch, value, m, device, name):
    t = torch.as_tensor(value, dtype=torch.float64, device=device)
    if t.dim() == 0:
        t = t.expand(m)
    if tuple(t.shape) != (m,):
        raise ValueError(f"{name} must be a scalar or have one value per segment")
    if not bool(torch.isfinite(t).all()) or bool((t < 0).any()):
        raise ValueError(f"{name} must be finite and nonnegative")
    return t.contiguous()


def ambient_velocity(inflow, points, time):
    """wake.ambient_velocity (wake.py:399-401) on device: a 3-vector, or a callable (points tensor, time)."""
    torch = _torch()
    if callable(inflow):
        values = torch.as_tensor(inflow(points.clone(), float(time)), dtype=torch.float64, device=points.device)
    else:
        values = torch.as_tensor(np.asarray(inflow, dtype=np.float64), device=points.device)
    if values.dim() == 1:
        values = values.reshape(1, 3).expand(points.shape[0], 3)
    if tuple(values.shape) != tuple(points.shape) or not bool(torch.isfinite(values).all()):
        raise ValueError("ambient velocity must be finite with shape (3,) or (n, 3)")
    return values.contiguous().clone()


# ------------------------------------------------------------------ the step
def network_convection(nodes, edges, gamma, first, radii, settings, ambient, ring_first=None, ring_free=None,
                       ring_radii=None):
    """wake._network_convection (wake.py:404-491) on device tensors: the velocity (n - first, 3) of nodes[first:].

    nodes (n,3) float64, edges (m,2) int64, gamma (m,) float64, radii and ring_radii scalars or (m,) per edge,
    ambient the (n - first, 3) ambient velocity of the points.
    """
    torch = _torch()
    n, m, first = int(nodes.shape[0]), int(edges.shape[0]), int(first)
    count = n - first
    velocity = ambient.clone()
    if not settings.free_wake or m == 0:
        return velocity
    nodes, edges, gamma = nodes.contiguous(), edges.contiguous(), gamma.contiguous()
    points = nodes[first:]
    pairs = NET.incident_pairs_device(nodes, edges, first)
    geometry = network_segments(nodes, edges)
    radii = _per_edge(torch, radii, m, nodes.device, "radii")
    target_radii = radii if ring_radii is None else _per_edge(torch, ring_radii, m, nodes.device, "ring_radii")
    rows_end = count if ring_first is None or ring_first >= n else int(ring_first) - first
    ring_end = count if ring_free is None else min(count, rows_end + int(ring_free))

    def add(lo, hi, radius, singularity):
        ptr, seg = exclusion_subset(pairs, lo, hi, m, segment_ordered=True)
        out = _induced(points[lo:hi], geometry, gamma, radius, ptr, seg, singularity, settings)
        part = velocity[lo:hi]
        stage(_lib.STAGE_ADD, 0.0, part, out, out=part)

    def rows_and_rings(rows, ring_lo, ring_hi):
        """Rows [0, rows) with 'raise' and ring nodes [ring_lo, ring_hi) with 'zero'.  With one set of cores both
        go through one sum (one tree): a singular pair adds nothing in either mode, only its flag differs, so every
        target's velocity is the one of its own call.  The rows' flags raise as in their own call; when any ring
        target is flagged, the ring targets are summed again in their own 'zero' call."""
        if ring_radii is not None or ring_lo >= ring_hi or rows == 0:
            add(0, rows, radii, "raise")
            if ring_lo < ring_hi:
                add(ring_lo, ring_hi, target_radii, "zero")
            return
        rings = ring_hi - ring_lo
        point_map = torch.full((ring_hi,), -1, dtype=torch.int64, device=nodes.device)
        point_map[:rows] = torch.arange(rows, dtype=torch.int64, device=nodes.device)
        point_map[ring_lo:] = torch.arange(rows, rows + rings, dtype=torch.int64, device=nodes.device)
        ptr, seg = exclusion_subset(pairs, 0, ring_hi, m, point_map=point_map, count=rows + rings,
                                    segment_ordered=True)
        targets = torch.cat((points[:rows], points[ring_lo:ring_hi]))
        out, flags, flag_seg = _sum(targets, geometry, gamma, radii, ptr, seg, "raise", settings)
        _check(out[:rows], flags[:rows], None if flag_seg is None else flag_seg[:rows])
        if bool(flags[rows:].any().item()):
            ring_out = None
            add(ring_lo, ring_hi, target_radii, "zero")
        else:
            ring_out = out[rows:]
            _check(ring_out, flags[rows:], None)
        part = velocity[:rows]
        stage(_lib.STAGE_ADD, 0.0, part, out[:rows], out=part)
        if ring_out is not None:
            part = velocity[ring_lo:ring_hi]
            stage(_lib.STAGE_ADD, 0.0, part, ring_out.contiguous(), out=part)

    limit = settings.free_convection_rows
    if limit is not None and limit * first < rows_end:
        free = limit * first
        frozen_rings = settings.frozen_rows_drift == "ambient_rings" and ring_first is not None and ring_first < n
        if frozen_rings:                          # the reference's order: rows, frozen rows, ring nodes
            add(0, free, radii, "raise")
        else:
            rows_and_rings(free, rows_end, ring_end)
        if frozen_rings:
            ring_edge = (edges >= int(ring_first)).all(dim=1)
            sub = torch.nonzero(ring_edge).reshape(-1)
            if sub.numel():
                remap = torch.full((m,), -1, dtype=torch.int64, device=nodes.device)
                remap[sub] = torch.arange(sub.numel(), dtype=torch.int64, device=nodes.device)
                ptr, seg = exclusion_subset(pairs, free, rows_end, m, remap, segment_ordered=True)
                sub_geometry = tuple(g[sub].contiguous() for g in geometry)
                out = _induced(points[free:rows_end], sub_geometry, gamma[sub].contiguous(), radii[sub].contiguous(),
                               ptr, seg, "raise", settings)
                part = velocity[free:rows_end]
                stage(_lib.STAGE_ADD, 0.0, part, out, out=part)
            if rows_end < ring_end:
                add(rows_end, ring_end, target_radii, "zero")
    elif ring_first is None or ring_first >= n:
        add(0, count, radii, "raise")
    else:
        split = int(ring_first) - first
        rows_and_rings(split, split, ring_end)
    return velocity


def convect_geometry(bound, points, edges, gamma, radii, settings, *, inflow, time, dt_s, substeps=1,
                     integrator="euler", displacement=None, ring_first=None, ring_free=None, ring_radii=None,
                     ring_drift=None, guard=None, wing=None):
    """The stage loop of wake.convect_geometry (wake.py:572-616) on device tensors: the new free points.

    bound (first,3) are the bound nodes at ``time`` (moved by ``displacement`` (first,3) over the step), points
    (k,3) the free nodes (wing rows, then ring nodes from ``ring_first``), edges (m,2) index the network
    concatenate(bound, points).

    * ``ring_drift`` (per ring node, as RingWake.drift_m_s): the rings' opt-in sheet drift (wake.py:584-590), a
      constant velocity of the ring nodes beyond the leading ``ring_free``, added to every stage's velocity.
    * ``guard``: the reference's surface guard object (``make_surface_guard(wing, config, next_wing)``,
      wake.py:619-631), or anything with its ``check_rows(old_rows, new_rows, wing, stage=, time_s=)``.  It is
      called where wake.convect_geometry calls it (wake.py:576-582, 601-615): on the wing rows of every Heun / RK4
      stage point and of the accepted geometry, before they are used, with host copies of the (depth, first, 3)
      rows.  Its exception stops the step.
    """
    torch = _torch()
    if integrator not in INTEGRATORS:
        raise ValueError(f"convection_integrator must be one of {INTEGRATORS}")
    if isinstance(substeps, bool) or int(substeps) != substeps or substeps < 1:
        raise ValueError("substeps must be a positive integer")
    first = int(bound.shape[0])
    bound, points = bound.contiguous(), points.contiguous()
    if displacement is not None:
        displacement = displacement.contiguous()
        if not bool(displacement.any()):
            displacement = None
    dt = dt_s / substeps
    rows = int(points.shape[0]) if ring_first is None else int(ring_first) - first  # row points precede rings

    def checked(candidate, previous, label, t):
        if guard is not None:
            old = previous[:rows].cpu().numpy().reshape(-1, first, 3)
            new = candidate[:rows].cpu().numpy().reshape(-1, first, 3)
            guard.check_rows(old, new, wing, stage=label, time_s=t)
        return candidate

    drift = None
    if ring_drift is not None and ring_first is not None and ring_free is not None:
        node_drift = torch.as_tensor(ring_drift, dtype=torch.float64, device=points.device)
        if int(node_drift.shape[0]) > int(ring_free):
            drift = torch.zeros_like(points)
            drift[rows + int(ring_free):] = node_drift[int(ring_free):]

    def rhs(current, t):
        stage_bound = bound
        if displacement is not None and t != time:
            stage_bound = stage(_lib.STAGE_STEP, (t - time) / dt_s, bound, displacement)
        nodes = torch.cat((stage_bound, current))
        ambient = ambient_velocity(inflow, current, t)
        velocity = network_convection(nodes, edges, gamma, first, radii, settings, ambient, ring_first, ring_free,
                                      ring_radii)
        return velocity if drift is None else velocity + drift

    for substep in range(int(substeps)):
        t = time + substep * dt
        k1 = rhs(points, t)
        if integrator == "euler":
            candidate = stage(_lib.STAGE_STEP, dt, points, k1)
        elif integrator == "heun":
            predicted = checked(stage(_lib.STAGE_STEP, dt, points, k1), points, "heun_predictor", t + dt)
            k2 = rhs(predicted, t + dt)
            candidate = stage(_lib.STAGE_HEUN, dt, points, k1, k2)
        else:
            k2 = rhs(checked(stage(_lib.STAGE_HALF, dt, points, k1), points, "rk4_stage2", t + dt / 2), t + dt / 2)
            k3 = rhs(checked(stage(_lib.STAGE_HALF, dt, points, k2), points, "rk4_stage3", t + dt / 2), t + dt / 2)
            k4 = rhs(checked(stage(_lib.STAGE_STEP, dt, points, k3), points, "rk4_stage4", t + dt), t + dt)
            candidate = stage(_lib.STAGE_RK4, dt, points, k1, k2, k3, k4)
        points = checked(candidate, points, "accepted_geometry", t + dt)
    return points


# ------------------------------------------------------------------ rotor rings
def ring_self_velocity(nodes, edges, gamma, radius, settings, singularity="zero"):
    """RingWake.self_velocity (rotor/ring_wake.py:183-199) on device tensors: the velocity of every ring node from
    all ring segments, the segments incident to a node excluded (incident_pairs with first = 0).  Identical rings
    released from one rim stack onto each other (coincident nodes on each other's filament ends); with a core,
    'zero' is the physical limit there, 'raise' reports such events."""
    torch = _torch()
    n, m = int(nodes.shape[0]), int(edges.shape[0])
    if m == 0:
        return torch.zeros((n, 3), dtype=torch.float64, device=nodes.device)
    nodes, edges, gamma = nodes.contiguous(), edges.contiguous(), gamma.contiguous()
    pairs = NET.incident_pairs_device(nodes, edges, 0)
    geometry = network_segments(nodes, edges)
    ptr, seg = exclusion_subset(pairs, 0, n, m, segment_ordered=True)
    radius = _per_edge(torch, radius, m, nodes.device, "core_radius")
    return _induced(nodes, geometry, gamma, radius, ptr, seg, singularity, settings)


def ring_induction(points, nodes, edges, gamma, radius, model):
    """wake.ring_induction (wake.py:353-375) for rings without a refined sheet: the rings' velocity at
    lifting-surface points, by the FP64 direct sum (float64 by contract near the surface core, wake.py:42-54)."""
    torch = _torch()
    k, m = int(points.shape[0]), int(edges.shape[0])
    if m == 0 or k == 0:
        return torch.zeros((k, 3), dtype=torch.float64, device=points.device)
    geometry = network_segments(nodes.contiguous(), edges.contiguous())
    radius = _per_edge(torch, radius, m, points.device, "core_radius")
    ptr = torch.zeros(k + 1, dtype=torch.int64, device=points.device)
    seg = torch.zeros(1, dtype=torch.int64, device=points.device)
    settings = ConvectionSettings(model=model, hip=HipSettings("fp64"))
    return _induced(points.contiguous(), geometry, gamma.contiguous(), radius, ptr, seg, "raise", settings)


# ------------------------------------------------------------------ the step as one graph
class GeometryStepPlan:
    """A geometry step recorded as one graph (ofw_step_*): the stage loop of wake.convect_geometry
    (wake.py:586-616) over the target split of wake._network_convection for either engine (the tree code or the
    direct sum), one set of cores (no ``ring_radii``), a uniform inflow and frozen rows that keep it
    (``frozen_rows_drift='ambient'``).

    Every stage's incident pairs, segment geometry, exclusion CSR, tree sum and RK arithmetic are replayed by one
    graph launch; the host reads one status block per step.  The plan keeps the tensors it was made with: a new
    step with the same topology writes new contents into them (positions, circulation, cores, inflow) and calls
    ``run`` again.  ``step`` returns the new free points, or takes the stage-by-stage ``convect_geometry`` when the
    status asks for the reference's exact behaviour (a flagged pair raises there, a flagged ring target repeats its
    'zero' call), and makes the plan again when a capacity was exceeded.

    ``ring_drift`` (per ring node, as RingWake.drift_m_s) is the rings' sheet drift (wake.py:584-590), recorded into
    every stage's velocity.  A ``guard`` (wake.py:576-582) sees the wing rows of every stage point: with one Euler
    substep its only call is on the accepted geometry, made by ``step`` on host copies after the replay; any other
    integrator calls it between stages, so ``step`` then takes the stage-by-stage path.
    """

    def __init__(self, bound, points, edges, gamma, radii, settings, *, inflow, time, dt_s, substeps=1,
                 integrator="euler", displacement=None, ring_first=None, ring_free=None, ring_drift=None,
                 guard=None, wing=None, use_graph=True):
        torch = _torch()
        if not settings.free_wake or settings.frozen_rows_drift != "ambient":
            raise ValueError("a recorded step needs a free wake and frozen_rows_drift='ambient'")
        if integrator not in INTEGRATORS:
            raise ValueError(f"convection_integrator must be one of {INTEGRATORS}")
        if isinstance(substeps, bool) or int(substeps) != substeps or substeps < 1:
            raise ValueError("substeps must be a positive integer")
        tensors = dict(bound=bound, points=points, edges=edges, gamma=gamma)
        for name, t in tensors.items():
            if not (t.is_cuda and t.is_contiguous()):
                raise ValueError(f"{name} must be a contiguous GPU tensor")
        first, k, m = int(bound.shape[0]), int(points.shape[0]), int(edges.shape[0])
        if first < 1 or k < 1 or m < 1:
            raise ValueError("a recorded step needs bound nodes, free points and segments")
        n = first + k
        rows_end = k if ring_first is None or ring_first >= n else int(ring_first) - first
        ring_end = k if ring_free is None else min(k, rows_end + int(ring_free))
        limit = settings.free_convection_rows
        rows = limit * first if limit is not None and limit * first < rows_end else rows_end
        ring_lo, ring_hi = (rows_end, ring_end) if rows_end < ring_end else (rows_end, rows_end)
        if rows + (ring_hi - ring_lo) < 1:
            raise ValueError("a recorded step needs targets")
        if displacement is not None:
            displacement = displacement.contiguous()
            if not bool(displacement.any()):
                displacement = None
        self.args = dict(bound=bound, points=points, edges=edges, gamma=gamma, radii=radii, settings=settings,
                         time=time, dt_s=dt_s, substeps=int(substeps), integrator=integrator,
                         displacement=displacement, ring_first=ring_first, ring_free=ring_free,
                         ring_drift=ring_drift, guard=guard, wing=wing)
        self.radii = _per_edge(torch, radii, m, points.device, "radii")
        self.drift = None
        if ring_drift is not None and ring_first is not None and ring_free is not None:
            node_drift = torch.as_tensor(ring_drift, dtype=torch.float64, device=points.device).reshape(-1, 3)
            if int(node_drift.shape[0]) > int(ring_free):
                self.drift = torch.zeros_like(points)
                self.drift[rows_end + int(ring_free):] = node_drift[int(ring_free):]
        self.inflow = torch.as_tensor(np.asarray(inflow, dtype=np.float64).reshape(3), device=points.device).clone()
        self.displacement = displacement
        self.points_out = torch.empty_like(points)
        self.shape = dict(first=first, k=k, m=m, rows=rows, ring_lo=ring_lo, ring_hi=ring_hi, split=rows_end)
        self.use_graph = bool(use_graph)
        self.handle = None
        self._create()

    def _create(self):
        import ctypes
        a, s = self.args, self.shape
        settings = a["settings"]
        tree = settings.tree if settings.backend == "hip_tree" else TreeSettings()
        handle = ctypes.c_void_p()
        stats = (ctypes.c_int64 * 3)()
        status = _lib.load().ofw_step_create(
            _ctx(), ctypes.byref(handle), _lib.PRECISIONS[settings.hip.precision], s["first"], _ptr(a["bound"]),
            _ptr(self.displacement), s["k"], _ptr(a["points"]), _ptr(self.points_out), s["m"], _ptr(a["edges"]),
            _ptr(a["gamma"]), _ptr(self.radii), settings.model, s["rows"], s["ring_lo"], s["ring_hi"],
            float(tree.theta), int(tree.order), int(tree.leaf_size), float(tree.core_clearance), _ptr(self.inflow),
            _ptr(self.drift), float(a["time"]), float(a["dt_s"]), a["substeps"], INTEGRATORS.index(a["integrator"]),
            0 if settings.backend == "hip_tree" else 1, int(settings.hip.block), int(self.use_graph), stats)
        _lib.check(status, "ofw_step_create")
        self.handle = handle
        self.capacities = dict(pairs=int(stats[1]), list_slots=int(stats[2]))

    def run(self):
        """Replay the step: (status bits, largest pair count, largest list count)."""
        import ctypes
        out = (ctypes.c_int64 * 3)()
        _lib.check(_lib.load().ofw_step_run(self.handle, out), "ofw_step_run")
        return int(out[0]), int(out[1]), int(out[2])

    def last_ms(self):
        import ctypes
        ms = ctypes.c_float()
        _lib.check(_lib.load().ofw_step_timings(self.handle, ctypes.byref(ms)), "ofw_step_timings")
        return float(ms.value)

    def step(self):
        """The new free points (``points_out``), with the reference's behaviour in every case."""
        a = self.args
        guard = a["guard"]
        staged = guard is not None and not (a["integrator"] == "euler" and a["substeps"] == 1)
        bits = 0
        if not staged:
            bits, _, _ = self.run()
            if bits & (_lib.STEP_PAIR_OVERFLOW | _lib.STEP_TREE_OVERFLOW):
                self.close()
                self._create()
                bits, _, _ = self.run()
        if bits or staged:
            out = convect_geometry(a["bound"], a["points"], a["edges"], a["gamma"], self.radii, a["settings"],
                                   inflow=self.inflow.cpu().numpy(), time=a["time"], dt_s=a["dt_s"],
                                   substeps=a["substeps"], integrator=a["integrator"],
                                   displacement=self.displacement, ring_first=a["ring_first"],
                                   ring_free=a["ring_free"], ring_drift=a["ring_drift"], guard=guard, wing=a["wing"])
            self.points_out.copy_(out)
        elif guard is not None:
            split, first = self.shape["split"], self.shape["first"]
            old = a["points"][:split].cpu().numpy().reshape(-1, first, 3)
            new = self.points_out[:split].cpu().numpy().reshape(-1, first, 3)
            guard.check_rows(old, new, a["wing"], stage="accepted_geometry", time_s=a["time"] + a["dt_s"])
        return self.points_out

    def close(self):
        if self.handle is not None:
            _lib.load().ofw_step_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
