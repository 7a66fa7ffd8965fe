# SPDX-License-Identifier: Apache-2.0
# Test oracle: a restatement of OpenFAWE's src/openfawe/aero/wake.py geometry step (Copyright 2026 Yinong Tian, Apache License 2.0, itself
# adapted from OpenFAST/OLAF; see NOTICE).
"""NumPy restatement of the reference's free-wake geometry step.  TEST INFRASTRUCTURE ONLY.

* ``segments_velocity`` -- vortex.py:458-497 for validated inputs: segment geometry from nodes and edges
  (hypot norms, direction = delta / length, zero-length error), the exclusion CSR, the sum of the C oracle (the
  direct sum, vortex_numba.direct_sum, or the tree code, vortex_tree.tree_sum), and the flag and finiteness
  errors.
* ``network_convection`` -- wake._network_convection (wake.py:404-491): ambient plus induced velocity of the free
  points nodes[first:], incident segments excluded, targets split into wing rows ('raise'), frozen rows beyond
  ``free_convection_rows`` (ambient, or with 'ambient_rings' the ring edges alone) and ring nodes ('zero',
  ``ring_radii``), the leading ``ring_free`` of them.
* ``convect_geometry`` -- the stage loop of wake.convect_geometry (wake.py:586-616): Euler, Heun, RK4 with
  ``substeps`` substeps, the bound row moved linearly by ``displacement``.
"""
import numpy as np

from . import network_np
from . import oracle as O


class Singular(ValueError):
    """A flagged pair (the reference raises FilamentSingularity or the float64 range error)."""


def direct_engine(points, starts, directions, lengths, gamma, radius, model, csr, singularity):
    """The reference's direct sum (vortex_numba.direct_sum) by the C oracle: (out, flags)."""
    return O.direct_sum(points, starts, directions, lengths, gamma, radius, model, csr, singularity)


def tree_engine(theta, order=2, leaf_size=32, clearance=3.0, threads=None):
    """The reference's tree code (vortex_tree.tree_sum) by the C oracle, as an engine: (out, flags)."""
    def engine(points, starts, directions, lengths, gamma, radius, model, csr, singularity):
        out, flags, _, _ = O.tree_sum(points, starts, directions, lengths, gamma, radius, model, csr, singularity,
                                      theta, order, leaf_size, clearance, threads)
        return out, flags
    return engine


def segments_velocity(points, nodes, edges, gamma, radius, model, exclusions, singularity, engine=direct_engine):
    """vortex.py:458-497: the induced velocity at points of the segments nodes[edges], pairs excluded."""
    nodes = np.asarray(nodes, dtype=np.float64)
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    starts = nodes[edges[:, 0]]
    delta = nodes[edges[:, 1]] - starts
    lengths = np.hypot(np.hypot(delta[:, 0], delta[:, 1]), delta[:, 2])
    if np.any(lengths == 0):
        raise ValueError("zero-length segment")
    directions = delta / lengths[:, None]
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (len(edges),))
    csr = O.exclusion_csr(len(points), np.asarray(exclusions, dtype=np.int64).reshape(-1, 2))
    out, flags = engine(points, starts, directions, lengths, gamma, radius, model, csr, singularity)
    if np.any(flags):
        raise Singular("flagged pair")
    if not np.isfinite(out).all():
        raise ValueError("velocity exceeds finite float64 range")
    return out


def network_convection(nodes, edges, gamma, first, radii, model, ambient, *, free_wake=True,
                       free_convection_rows=None, frozen_rows_drift="ambient", ring_first=None, ring_free=None,
                       ring_radii=None, engine=direct_engine):
    """wake._network_convection (wake.py:404-491): velocity (n - first, 3) of the points nodes[first:]."""
    nodes = np.asarray(nodes, dtype=np.float64)
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    gamma = np.asarray(gamma, dtype=np.float64)
    points = nodes[first:]
    target_radii = radii if ring_radii is None else ring_radii
    velocity = np.array(ambient, dtype=np.float64).reshape(-1, 3).copy()
    if not free_wake or not len(edges):
        return velocity
    exclusions = network_np.incident_pairs(nodes, edges, first)
    rows_end = len(points) if ring_first is None or ring_first >= len(nodes) else ring_first - first
    ring_end = len(points) if ring_free is None else min(len(points), rows_end + int(ring_free))

    def within(lo, hi):
        keep = (exclusions[:, 0] >= lo) & (exclusions[:, 0] < hi)
        return exclusions[keep] - np.array([lo, 0])

    limit = free_convection_rows
    if limit is not None and limit * first < rows_end:
        free = limit * first
        velocity[:free] += segments_velocity(points[:free], nodes, edges, gamma, radii, model, within(0, free),
                                             "raise", engine)
        if frozen_rows_drift == "ambient_rings" and ring_first is not None and ring_first < len(nodes):
            ring_edge = np.all(edges >= ring_first, axis=1)
            if np.any(ring_edge):
                sub = np.flatnonzero(ring_edge)
                remap = np.full(len(edges), -1, dtype=np.int64)
                remap[sub] = np.arange(len(sub))
                pairs = within(free, rows_end)
                pairs = pairs[ring_edge[pairs[:, 1]]]
                pairs = np.column_stack((pairs[:, 0], remap[pairs[:, 1]])).reshape(-1, 2)
                cores = np.broadcast_to(radii, (len(edges),))[ring_edge]
                velocity[free:rows_end] += segments_velocity(points[free:rows_end], nodes, edges[ring_edge],
                                                             gamma[ring_edge], cores, model, pairs, "raise", engine)
        if rows_end < ring_end:
            velocity[rows_end:ring_end] += segments_velocity(points[rows_end:ring_end], nodes, edges, gamma,
                                                             target_radii, model, within(rows_end, ring_end), "zero",
                                                             engine)
    elif ring_first is None or ring_first >= len(nodes):
        velocity += segments_velocity(points, nodes, edges, gamma, radii, model, exclusions, "raise", engine)
    else:
        split = ring_first - first
        velocity[:split] += segments_velocity(points[:split], nodes, edges, gamma, radii, model, within(0, split),
                                              "raise", engine)
        if split < ring_end:
            velocity[split:ring_end] += segments_velocity(points[split:ring_end], nodes, edges, gamma, target_radii,
                                                          model, within(split, ring_end), "zero", engine)
    return velocity


def convect_geometry(bound, points, edges, gamma, radii, model, *, inflow, time, dt_s, substeps=1,
                     integrator="euler", displacement=None, ring_drift=None, guard=None, wing=None, **split):
    """The stage loop of wake.convect_geometry (wake.py:572-616) with a uniform inflow: the new free points.

    ring_drift (per ring node, RingWake.drift_m_s): the rings' sheet drift (wake.py:584-590), a constant velocity
    of the ring nodes beyond the leading ring_free, added to every stage's velocity.  guard: an object with the
    reference's check_rows(old_rows, new_rows, wing, stage=, time_s=) (wake.py:576-582), called on the rows
    (the points before ring_first) of every stage point and of the accepted geometry, before they are used.
    """
    bound = np.asarray(bound, dtype=np.float64)
    points = np.asarray(points, dtype=np.float64).copy()
    first = len(bound)
    inflow = np.asarray(inflow, dtype=np.float64)
    if displacement is not None and not np.any(displacement):
        displacement = None
    dt = dt_s / substeps
    ring_first, ring_free = split.get("ring_first"), split.get("ring_free")
    rows = len(points) if ring_first is None else ring_first - first  # row points precede ring points

    def checked(candidate, previous, label, t):
        if guard is not None:
            guard.check_rows(previous[:rows].reshape(-1, first, 3), candidate[:rows].reshape(-1, first, 3), wing,
                             stage=label, time_s=t)
        return candidate

    drift = None
    if ring_drift is not None and ring_first is not None and ring_free is not None:
        node_drift = np.asarray(ring_drift, dtype=np.float64)
        if len(node_drift) > ring_free:
            drift = np.zeros_like(points)
            drift[rows + ring_free:] = node_drift[ring_free:]

    def rhs(current, t):
        stage_bound = bound
        if displacement is not None and t != time:
            stage_bound = bound + ((t - time) / dt_s) * displacement
        ambient = np.broadcast_to(inflow, current.shape)
        velocity = network_convection(np.concatenate((stage_bound, current)), edges, gamma, first, radii, model,
                                      ambient, **split)
        return velocity if drift is None else velocity + drift

    for substep in range(substeps):
        t = time + substep * dt
        k1 = rhs(points, t)
        if integrator == "euler":
            candidate = points + dt * k1
        elif integrator == "heun":
            k2 = rhs(checked(points + dt * k1, points, "heun_predictor", t + dt), t + dt)
            candidate = points + dt * (k1 + k2) / 2
        else:
            k2 = rhs(checked(points + dt * k1 / 2, points, "rk4_stage2", t + dt / 2), t + dt / 2)
            k3 = rhs(checked(points + dt * k2 / 2, points, "rk4_stage3", t + dt / 2), t + dt / 2)
            k4 = rhs(checked(points + dt * k3, points, "rk4_stage4", t + dt), t + dt)
            candidate = points + dt * (k1 + 2 * k2 + 2 * k3 + k4) / 6
        points = checked(candidate, points, "accepted_geometry", t + dt)
    return points
