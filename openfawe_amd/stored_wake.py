"""Stored wake networks: the ``state_<step>.npz`` files of the reference's convergence campaign, on the GPU.

The reference writes one file per kept step of a free-wake history (paper1/scripts/convergence/run_history.py:
121-133) with the keys ``nodes, edges, gamma, radii, first, wing_points, exclusions, chord, speed, core_model``,
and evaluates the frozen network two ways (check_v1_frozen.py:29-57, local_benchmark.py:57-85):

* ``wake`` -- the free nodes ``nodes[first:]`` as targets of every segment, the stored incident exclusions,
  singularity 'raise' (the geometry step's call);
* ``wing`` -- the lifting-surface points ``wing_points`` as targets of the history edges (both nodes at or after
  ``2 * first``, i.e. behind the bound and trailing-edge rows), no exclusions.

``load_state`` reads a file with NumPy's loader and no pickle; ``save_state`` writes the same format (the
exclusions default to the GPU ``incident_pairs`` of the network, the rule of save_states.py:20-29);
``evaluate_state`` times the FP64 direct sum and the tree code on both cases with the inputs resident in HBM and
states the tree's error as max and RMS of |v_tree - v_direct| / speed, as the reference's tables do.
"""
import time

import numpy as np

from . import network as NET
from .vortex import HipSettings, TreeSettings, _torch, exclusion_csr, hip_sum_device

KEYS = ("nodes", "edges", "gamma", "radii", "first", "wing_points", "exclusions", "speed", "core_model")


def load_state(path):
    """The arrays of a stored wake (no pickle); raises ValueError when a key is missing or malformed."""
    with np.load(path, allow_pickle=False) as data:
        missing = [k for k in KEYS if k not in data.files]
        if missing:
            raise ValueError(f"stored wake {path} lacks {missing}")
        state = {k: np.asarray(data[k]) for k in data.files}
    nodes, edges = state["nodes"].astype(np.float64), state["edges"].astype(np.int64)
    if nodes.ndim != 2 or nodes.shape[1] != 3 or edges.ndim != 2 or edges.shape[1] != 2:
        raise ValueError("stored wake: nodes must be (n,3) and edges (m,2)")
    m = len(edges)
    state.update(nodes=nodes, edges=edges, gamma=state["gamma"].astype(np.float64).reshape(m),
                 radii=np.broadcast_to(state["radii"].astype(np.float64), (m,)).copy(),
                 first=int(state["first"]), speed=float(state["speed"]), core_model=int(state["core_model"]),
                 exclusions=state["exclusions"].astype(np.int64).reshape(-1, 2),
                 wing_points=state["wing_points"].astype(np.float64).reshape(-1, 3))
    if not 0 <= state["first"] <= len(nodes) or (m and (edges.min() < 0 or edges.max() >= len(nodes))):
        raise ValueError("stored wake: first or an edge index is out of range")
    return state


def save_state(path, *, nodes, edges, gamma, radii, first, wing_points, speed, core_model, chord=1.0,
               exclusions=None):
    """Write a stored wake in the reference's format (run_history.py:128-133)."""
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    edges = np.ascontiguousarray(edges, dtype=np.int64)
    if exclusions is None:
        exclusions = NET.incident_pairs(nodes, edges, first)
    np.savez_compressed(path, nodes=nodes, edges=edges, gamma=np.asarray(gamma, dtype=np.float64),
                        radii=np.asarray(radii, dtype=np.float64), first=int(first),
                        wing_points=np.asarray(wing_points, dtype=np.float64), exclusions=exclusions,
                        chord=float(chord), speed=float(speed), core_model=int(core_model))


def frozen_cases(state):
    """{name: (points, edge mask, exclusions)} of check_v1_frozen.py:35-38."""
    first, edges = state["first"], state["edges"]
    history = np.all(edges >= 2 * first, axis=1)
    return dict(wake=(state["nodes"][first:], np.ones(len(edges), bool), state["exclusions"]),
                wing=(state["wing_points"], history, np.zeros((0, 2), np.int64)))


def _geometry(nodes, edges):
    starts = nodes[edges[:, 0]]
    delta = nodes[edges[:, 1]] - starts
    lengths = np.hypot(np.hypot(delta[:, 0], delta[:, 1]), delta[:, 2])
    if np.any(lengths == 0):
        raise ValueError("zero-length segment")
    return starts, delta / lengths[:, None], lengths


def evaluate_state(state, trees=(TreeSettings(0.3, 2, 32), TreeSettings(0.2, 2, 32)), precision="mixed",
                   repeats=3):
    """Rows of (case, engine, seconds, error over speed) for a loaded stored wake, as check_v1_frozen.evaluate."""
    from .tree import hip_tree_sum_device
    torch = _torch()
    dev = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)  # noqa: E731
    rows = []
    speed, model = state["speed"], state["core_model"]
    for name, (points, mask, exclusions) in frozen_cases(state).items():
        edges = state["edges"][mask]
        if len(points) == 0 or len(edges) == 0:
            continue
        s, d, L = _geometry(state["nodes"], edges)
        ptr, seg = exclusion_csr(len(points), exclusions)
        args = [dev(points), dev(s), dev(d), dev(L), dev(state["gamma"][mask]), dev(state["radii"][mask]), model,
                dev(ptr, torch.int64), dev(seg if len(seg) else np.zeros(1, np.int64), torch.int64), "raise"]

        def timed(fn):
            out = fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(repeats):
                out = fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / repeats, out

        direct_s, (direct, flags, _) = timed(lambda: hip_sum_device(*args, HipSettings("fp64", 256)))
        if bool(flags.any()):
            raise ValueError(f"stored wake case {name}: a flagged pair in the direct sum")
        ref = direct.cpu().numpy()
        rows.append(dict(targets=name, points=len(points), segments=len(edges), engine="direct-fp64",
                         seconds=direct_s, max_induced_over_U=float(np.linalg.norm(ref, axis=1).max() / speed)))
        for tree in trees:
            settings = HipSettings(precision, 256)
            tree_s, (out, tflags, far, _) = timed(lambda: hip_tree_sum_device(*args, tree, settings))
            if bool(tflags.any()):
                raise ValueError(f"stored wake case {name}: a flagged pair in the tree sum")
            error = np.linalg.norm(out.cpu().numpy() - ref, axis=1) / speed
            rows.append(dict(targets=name, points=len(points), segments=len(edges), engine="tree",
                             theta=tree.theta, order=tree.order, leaf=tree.leaf_size, precision=precision,
                             seconds=tree_s, max_err_over_U=float(error.max()),
                             rms_err_over_U=float(np.sqrt(np.mean(error ** 2))),
                             mean_far_nodes=float(far.double().mean().item())))
    return rows
