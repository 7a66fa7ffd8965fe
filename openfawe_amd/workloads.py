# SPDX-License-Identifier: Apache-2.0
# ``sheet`` reproduces the synthetic sheet of OpenFAWE's apps/studio/openfawe_studio/selftest.py:243-272
# (Copyright 2026 Yinong Tian, Apache License 2.0; see NOTICE), whose random stream fixes the reference speed
# table's sizes; the other generators are this package's own.
"""Seeded synthetic wakes: the inputs of the parity tests and of bench.py.

* ``sheet`` -- the reference's synthetic sheet (apps/studio/openfawe_studio/selftest.py:243-272), the
  generator of the reference's GPU tree speed table: 110 nodes per row, spanwise and trailing segments,
  20,570 off-node targets; 468 / 936 / 1,872 / 3,744 rows give 102,382 / 204,874 / 409,858 / 819,826
  segments.
* ``loop`` -- a crosswind-loop wake: rows of wing shed and trailing segments along a circular loop of
  radius 145 m convected downwind, plus 8 rotor tip-vortex helices; targets are the first 187 rows of
  wake nodes with their incident segments excluded (the wake-on-wake convection call of wake.py:404-491).
* ``batch`` -- uneven snapshots of ``loop`` at random loop phases and wind speeds, for the batched call.
* ``helix`` -- 40 rows x 256 spanwise segments on a helix of radius 150 m and pitch 40 m, Vatistas cores.
* ``random_network`` -- the reference's backend-equivalence network (tests/test_vortex_numba.py:38-54).
* ``lattice`` and ``ring`` -- the reference's tree-test geometries (tests/test_vortex_tree.py:18-40).

Every generator returns a dict with the engine inputs: points, starts, directions, lengths, gamma,
radius (float64), pairs ((e,2) int64 exclusions), model.
"""
import numpy as np


def _segments(nodes, edges):
    starts = nodes[edges[:, 0]]
    vector = nodes[edges[:, 1]] - starts
    length = np.hypot(np.hypot(vector[:, 0], vector[:, 1]), vector[:, 2])
    return starts, vector / length[:, None], length


def incident_pairs(point_nodes, edges):
    """(point, segment) pairs that exclude every segment incident to a target node.

    ``point_nodes[i]`` is the node index of target i.  The rule of wake.py:427-438 (incident_pairs).
    """
    point_nodes = np.asarray(point_nodes, dtype=np.int64)
    edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    size = int(max(edges.max(initial=-1), point_nodes.max(initial=-1))) + 1
    lookup = np.full(size, -1, dtype=np.int64)
    lookup[point_nodes[::-1]] = np.arange(len(point_nodes))[::-1]   # the first target of a repeated node
    target = lookup[edges]                                            # (m, 2): start, end
    seg = np.broadcast_to(np.arange(len(edges))[:, None], target.shape)
    keep = target >= 0                                                # segment order, start before end
    return np.stack([target[keep], seg[keep]], axis=1).astype(np.int64).reshape(-1, 2)


def sheet(rows, nodes_per_row=110, span=26.0, dx=0.45, seed=None):
    """The reference's synthetic sheet; seed defaults to ``rows`` as in its GPU self-test."""
    rng = np.random.default_rng(rows if seed is None else seed)
    y = np.linspace(-span / 2, span / 2, nodes_per_row)
    grid = np.zeros((rows, nodes_per_row, 3))
    for r in range(rows):
        s = r * dx
        grid[r, :, 0] = s + 0.02 * rng.standard_normal(nodes_per_row)
        grid[r, :, 1] = y * (1 + 0.002 * r) + 0.5 * np.sin(0.01 * r)
        grid[r, :, 2] = -0.03 * s + 0.3 * np.sin(y / 3.0 + 0.02 * r)
    starts, ends = [], []
    for r in range(rows):
        starts.append(grid[r, :-1])
        ends.append(grid[r, 1:])
        if r + 1 < rows:
            starts.append(grid[r])
            ends.append(grid[r + 1])
    starts, ends = np.concatenate(starts), np.concatenate(ends)
    vector = ends - starts
    length = np.linalg.norm(vector, axis=1)
    targets = grid[:min(rows, 187)].reshape(-1, 3) + np.array([0.0, 0.0, 0.011])
    return dict(points=targets, starts=starts, directions=vector / length[:, None], lengths=length,
                gamma=np.linspace(0.5, 1.5, len(starts)), radius=np.full(len(starts), 0.075),
                pairs=np.zeros((0, 2), dtype=np.int64), model=2, name=f"sheet-{rows}")


def loop(rows=468, seed=0, nodes_per_row=110, span=26.0, radius_loop=145.0, wind=10.0, rows_per_loop=468,
         phase=0.0, kite_speed=50.0, rotor_segments=64, target_rows=187, core=0.1, rotor_core=0.05):
    """Crosswind-loop wake (see the module docstring); ``rows / rows_per_loop`` loops long."""
    rng = np.random.default_rng(seed)
    period = 2 * np.pi * radius_loop / kite_speed
    dt = period / rows_per_loop
    # row r was shed r steps ago at loop angle phase - omega * age and has drifted downwind since
    age = np.arange(rows) * dt
    angle = phase - 2 * np.pi * age / period
    centre = np.stack([wind * age, radius_loop * np.cos(angle), radius_loop * np.sin(angle)], axis=1)
    radial = np.stack([np.zeros(rows), np.cos(angle), np.sin(angle)], axis=1)
    eta = np.linspace(-0.5, 0.5, nodes_per_row)
    nodes = centre[:, None, :] + span * eta[None, :, None] * radial[:, None, :]
    nodes = nodes + 0.02 * rng.standard_normal(nodes.shape)
    nodes = nodes.reshape(-1, 3)
    # panel circulation: seeded elliptic spanwise profile, peak 8 m^2/s, small per-row modulation
    mid_eta = 0.5 * (eta[:-1] + eta[1:])
    profile = 8.0 * np.sqrt(np.clip(1 - (2 * mid_eta) ** 2, 0, None))
    panel = profile[None, :] * (1 + 0.05 * rng.standard_normal((rows, nodes_per_row - 1)))
    # per row: the shed (spanwise) segments, then the trailing segments towards the older row (none after
    # the oldest row)
    npr = nodes_per_row
    base = (np.arange(rows) * npr)[:, None]
    j1, j2 = np.arange(npr - 1)[None, :], np.arange(npr)[None, :]
    shed = np.stack([base + j1, base + j1 + 1], axis=-1)
    trail = np.stack([base + j2, base + npr + j2], axis=-1)
    older = np.vstack([panel[1:], np.zeros((1, npr - 1))])
    padded = np.pad(panel, ((0, 0), (1, 1)))
    kept = rows * (2 * npr - 1) - npr
    edges = np.concatenate([shed, trail], axis=1).reshape(-1, 2)[:kept].astype(np.int64)
    gamma = np.concatenate([panel - older, padded[:, 1:] - padded[:, :-1]], axis=1).reshape(-1)[:kept]
    radius = np.full(len(edges), core)
    # 8 rotors x 4 blades: tip-vortex helices of radius 1.15 m trailing from the rotors
    rotor_nodes, rotor_edges, rotor_gamma = [], [], []
    offset = len(nodes)
    flight = np.stack([np.zeros(rows), -np.sin(angle), np.cos(angle)], axis=1)
    hub_eta = np.linspace(-0.35, 0.35, 8)
    k = np.arange(rotor_segments + 1)
    for rotor in range(8):
        for blade in range(4):
            idx = np.minimum(k // 2, rows - 1)
            hub = centre[idx] + span * hub_eta[rotor] * radial[idx] + 1.5 * flight[idx]
            spin = 0.7 * k + 0.5 * np.pi * blade + rng.uniform(0, 2 * np.pi)
            e1, e2 = radial[idx], np.cross(flight[idx], radial[idx])
            tip = hub + 1.15 * (np.cos(spin)[:, None] * e1 + np.sin(spin)[:, None] * e2)
            rotor_nodes.append(tip)
            first = offset + sum(len(v) for v in rotor_nodes[:-1])
            for j in range(rotor_segments):
                rotor_edges.append((first + j, first + j + 1))
            rotor_gamma.append(np.full(rotor_segments, rng.uniform(0, 2)))
    nodes = np.concatenate([nodes, *rotor_nodes])
    edges = np.concatenate([edges, np.asarray(rotor_edges, dtype=np.int64)])
    gamma = np.concatenate([gamma, *rotor_gamma])
    radius = np.concatenate([radius, np.full(len(rotor_edges), rotor_core)])
    starts, directions, lengths = _segments(nodes, edges)
    target_nodes = np.arange(min(rows, target_rows) * nodes_per_row)
    return dict(points=nodes[target_nodes].copy(), starts=starts, directions=directions, lengths=lengths,
                gamma=gamma, radius=radius, pairs=incident_pairs(target_nodes, edges), model=2,
                nodes=nodes, edges=edges, name=f"loop-{rows}")


def batch(count=64, seed=4, m_lo=10_000, m_hi=410_000, n_lo=2_000, n_hi=20_000):
    """``count`` independent snapshots of the loop wake at random loop phases and wind speeds of 6-14 m/s.

    Segment counts are drawn log-uniform in [m_lo, m_hi] (the loop gives 219 segments a row plus 2,048
    rotor segments, so each snapshot has the row count nearest the draw); target counts follow the segment
    counts log-linearly from n_lo (at m_lo) to n_hi (at m_hi).  Targets are the first wake nodes with their
    incident segments excluded, as in ``loop``.
    """
    rng = np.random.default_rng(seed)
    snaps = []
    for k in range(count):
        m_want = float(np.exp(rng.uniform(np.log(m_lo), np.log(m_hi))))
        frac = (np.log(m_want) - np.log(m_lo)) / (np.log(m_hi) - np.log(m_lo))
        n = int(round(np.exp(np.log(n_lo) + frac * (np.log(n_hi) - np.log(n_lo)))))
        rows = max(2, int(round((m_want - 2048 + 110) / 219)))
        target_rows = int(np.ceil(n / 110))
        case = loop(rows=max(rows, target_rows + 1), seed=seed * 1000 + k, phase=float(rng.uniform(0, 2 * np.pi)),
                    wind=float(rng.uniform(6.0, 14.0)), target_rows=target_rows)
        pairs = case["pairs"][case["pairs"][:, 0] < n]
        snaps.append(dict(case, points=case["points"][:n].copy(), pairs=pairs, name=f"batch-{k}"))
    return snaps


def helix(seed=0, rows=40, spanwise=256, radius_helix=150.0, pitch=40.0, span=26.0, row_spacing=2.0,
          jitter=0.2, targets=2048, wing_points=100, core=0.05):
    """40 rows x 256 spanwise-connected segments on a helix, Vatistas n=2 cores, 2,048 targets."""
    rng = np.random.default_rng(seed)
    dphi = row_spacing / np.hypot(radius_helix, pitch / (2 * np.pi))
    phi = dphi * np.arange(rows)
    centre = np.stack([radius_helix * np.cos(phi), radius_helix * np.sin(phi), pitch * phi / (2 * np.pi)], axis=1)
    radial = np.stack([np.cos(phi), np.sin(phi), np.zeros(rows)], axis=1)
    eta = np.linspace(-0.5, 0.5, spanwise + 1)
    nodes = centre[:, None, :] + span * eta[None, :, None] * radial[:, None, :]
    nodes = (nodes + rng.normal(0.0, jitter, nodes.shape)).reshape(-1, 3)
    edges = np.asarray([(r * (spanwise + 1) + j, r * (spanwise + 1) + j + 1)
                        for r in range(rows) for j in range(spanwise)], dtype=np.int64)
    gamma = rng.uniform(0.0, 8.0, len(edges))
    node_targets = np.sort(rng.choice(len(nodes), targets - wing_points, replace=False))
    # the wing: a straight 26 m line one row ahead of the helix start
    wing_centre = np.array([radius_helix * np.cos(-dphi), radius_helix * np.sin(-dphi), pitch * -dphi / (2 * np.pi)])
    wing_dir = np.array([np.cos(-dphi), np.sin(-dphi), 0.0])
    wing = wing_centre + span * np.linspace(-0.5, 0.5, wing_points)[:, None] * wing_dir
    points = np.concatenate([nodes[node_targets], wing])
    starts, directions, lengths = _segments(nodes, edges)
    return dict(points=points, starts=starts, directions=directions, lengths=lengths, gamma=gamma,
                radius=np.full(len(edges), core), pairs=incident_pairs(node_targets, edges), model=3,
                nodes=nodes, edges=edges, name="helix")


def vsm_wing(panels=100, span=26.0, chord=1.3, rotors=8, ring_segments=32, rotor_radius=1.15, wake_rows=936,
             seed=5, core=0.05):
    """A VSM wing in a free wake: the differentiation case.

    Each of ``panels`` spanwise panels is a vortex ring over its bound vortex at 1/4 chord: quarter-chord
    segment, two chordwise legs to the trailing edge and the trailing-edge segment; neighbouring panels
    share their legs with opposite signs (the ring layout of vsm_wake._ring_layout, vsm_wake.py:196-220).
    Collocation points are at 3/4 chord, panel mid-span.  ``rotors`` rotor rings of ``ring_segments``
    segments and radius ``rotor_radius`` sit ahead of the wing, and a crosswind-loop wake of ``wake_rows``
    rows (about 205k segments at 936 rows) trails behind it.  Returns the collocation points, the wing ring
    segments and their column layout (ptr, segment, sign), and the wake segments (rotor rings included).
    """
    rng = np.random.default_rng(seed)
    y = np.linspace(-span / 2, span / 2, panels + 1)
    sweep = 0.02 * np.abs(y)
    qc = np.stack([sweep + 0.25 * chord, y, 0.01 * y ** 2 / span], axis=1)
    te = np.stack([sweep + chord, y, 0.01 * y ** 2 / span], axis=1)
    nodes = np.concatenate([qc, te])
    index, segments, ptr, col_seg, col_sign = {}, [], [0], [], []

    def edge(a, b):
        key = (min(a, b), max(a, b))
        if key not in index:
            index[key] = len(segments)
            segments.append(key)
        return index[key], (1.0 if (a, b) == key else -1.0)
    n1 = panels + 1
    for j in range(panels):
        for a, b in ((j, j + 1), (j + 1, n1 + j + 1), (n1 + j + 1, n1 + j), (n1 + j, j)):
            k, sign = edge(a, b)
            col_seg.append(k)
            col_sign.append(sign)
        ptr.append(len(col_seg))
    wing_edges = np.asarray(segments, dtype=np.int64)
    ws, wd, wl = _segments(nodes, wing_edges)
    ymid = 0.5 * (y[:-1] + y[1:])
    points = np.stack([0.02 * np.abs(ymid) + 0.75 * chord, ymid, 0.01 * ymid ** 2 / span], axis=1)
    # rotor rings ahead of the wing
    hubs = np.linspace(-0.35 * span, 0.35 * span, rotors)
    phi = np.linspace(0.0, 2 * np.pi, ring_segments + 1)
    rnodes, redges = [], []
    for k, yh in enumerate(hubs):
        ring = np.stack([np.full_like(phi, -1.5), yh + rotor_radius * np.cos(phi), rotor_radius * np.sin(phi)], axis=1)
        base = sum(len(v) for v in rnodes)
        rnodes.append(ring)
        redges += [(base + i, base + i + 1) for i in range(ring_segments)]
    rnodes = np.concatenate(rnodes)
    rs, rd, rl = _segments(rnodes, np.asarray(redges, dtype=np.int64))
    wake = loop(rows=wake_rows, seed=seed, target_rows=1)
    # place the wake's newest row (the kite's current position on the loop) one chord behind the wing
    shift = np.array([chord + 0.5, 0.0, 0.0]) - wake["points"].mean(axis=0)
    starts = np.concatenate([rs, wake["starts"] + shift])
    directions = np.concatenate([rd, wake["directions"]])
    lengths = np.concatenate([rl, wake["lengths"]])
    gamma = np.concatenate([rng.uniform(0.0, 2.0, len(rs)), wake["gamma"]])
    return dict(points=points, wing_starts=ws, wing_directions=wd, wing_lengths=wl,
                column_ptr=np.asarray(ptr, dtype=np.int64), column_segment=np.asarray(col_seg, dtype=np.int64),
                column_sign=np.asarray(col_sign), starts=starts, directions=directions, lengths=lengths,
                gamma=gamma, radius=np.full(len(starts), core), wing_radius=np.full(len(ws), core), model=2,
                name=f"vsm-{panels}-wake-{len(starts)}")


def random_network(seed, nodes=60, segments=90, points=70):
    """The reference's random network with near-axis targets (tests/test_vortex_numba.py:38-54)."""
    rng = np.random.default_rng(seed)
    xyz = rng.normal(size=(nodes, 3))
    edges = rng.integers(0, nodes, size=(segments, 2))
    edges = edges[edges[:, 0] != edges[:, 1]]
    gamma = rng.normal(size=len(edges))
    gamma[::7] = 0.
    radius = rng.uniform(.01, .6, size=len(edges))
    targets = rng.normal(size=(points, 3))
    a, b = xyz[edges[:5, 0]], xyz[edges[:5, 1]]
    normal = np.cross(b - a, [0.3, -0.2, 0.9])
    normal /= np.linalg.norm(normal, axis=1)[:, None]
    near = [a + .4 * (b - a) + 1e-9 * normal, a + .4 * (b - a) + 1e-13 * normal,
            a + 1.7 * (b - a), a - .5 * (b - a), a + .5 * (b - a) + .05 * normal]
    targets = np.vstack([targets, *near])
    return xyz, edges, gamma, radius, targets


def lattice(seed=11, rows=24, span=17):
    """Wavy wake-like ring lattice with shed circulation differences (tests/test_vortex_tree.py:25-40)."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0., 12., rows)[:, None] * np.ones((1, span))
    y = np.ones((rows, 1)) * np.linspace(-4., 4., span)[None, :]
    z = .3 * np.sin(.7 * x) * np.cos(.5 * y) + .05 * rng.normal(size=x.shape)
    nodes = np.column_stack((x.ravel(), y.ravel(), z.ravel()))
    edges, gamma = [], []
    strength = np.cos(np.linspace(-1.3, 1.3, span - 1))[None, :] * np.exp(-.05 * np.arange(rows - 1))[:, None]
    for i in range(rows - 1):
        for j in range(span - 1):
            a, b, c, d = i * span + j, i * span + j + 1, (i + 1) * span + j + 1, (i + 1) * span + j
            for p, q in ((a, b), (b, c), (c, d), (d, a)):
                edges.append((p, q))
                gamma.append(strength[i, j])
    return nodes, np.asarray(edges), np.asarray(gamma)


def ring(radius=1.0, count=64, centre=(0., 0., 0.)):
    """Polygonal ring of ``count`` segments (tests/test_vortex_tree.py:18-22)."""
    angle = np.linspace(0., 2 * np.pi, count, endpoint=False)
    nodes = np.column_stack((radius * np.cos(angle), radius * np.sin(angle), np.zeros(count))) + centre
    edges = np.column_stack((np.arange(count), (np.arange(count) + 1) % count))
    return nodes, edges


def engine_args(case):
    """(points, starts, directions, lengths, gamma, radius, model) of a generator's case."""
    return (case["points"], case["starts"], case["directions"], case["lengths"], case["gamma"], case["radius"],
            case["model"])
