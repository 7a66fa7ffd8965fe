"""Tree code (Barnes-Hut type) on MI355X: the engine behind ``segments_velocity(backend="hip_tree")``.

``hip_tree_sum`` has the signature of the reference's ``vortex_tree.tree_sum`` / ``vortex_cuda_tree.
cuda_tree_sum`` (vortex_tree.py:588, vortex_cuda_tree.py:439) and returns (velocity, any_flag).  Build,
node quantities and traversal all run in libofw_hip.so on the GPU.  ``build_tree`` and ``tree_nodes``
expose the tree and its node quantities for the bit-exactness checks.
"""
import ctypes
import time

import numpy as np

from . import _lib
from .vortex import (HipSettings, Statistics, TREE_MODELS, TreeSettings, _ctx, _dev, _ptr, _torch, _upload,
                     exclusion_csr)


def capacity(m, leaf_size):
    """Node capacity of a tree over m segments (vortex_tree.py:59)."""
    return 4 * (m // max(int(leaf_size), 1) + 1) + 1


def build_tree(mid, leaf_size):
    """vortex_tree.build_tree on the GPU: (order, lo, hi, left, right, node_count) as host int64 arrays."""
    torch = _torch()
    mid = np.ascontiguousarray(mid, dtype=np.float64).reshape(-1, 3)
    m = len(mid)
    cap = capacity(m, leaf_size)
    d_mid = _dev(torch, mid if m else np.zeros((1, 3)), np.float64)
    order = torch.zeros(max(m, 1), dtype=torch.int64, device="cuda")
    arrays = [torch.full((cap,), -1, dtype=torch.int64, device="cuda") for _ in range(4)]
    count = ctypes.c_int64(0)
    status = _lib.load().ofw_build_tree(_ctx(), m, d_mid.data_ptr(), int(leaf_size), order.data_ptr(),
                                        *[a.data_ptr() for a in arrays], ctypes.byref(count))
    _lib.check(status, "ofw_build_tree")
    c = count.value
    return (order[:m].cpu().numpy(), *[a[:c].cpu().numpy() for a in arrays], c)


def tree_nodes(starts, directions, lengths, gamma, radius, leaf_size):
    """Node quantities (centre, bound, core, m0, m1, m2, w1, t2, w2) in the reference numbering."""
    torch = _torch()
    m = len(starts)
    cap = capacity(m, leaf_size)
    dev = [_dev(torch, a, np.float64) for a in (starts, directions, lengths, gamma, radius)]
    shapes = [(cap, 3), (cap,), (cap,), (cap, 3), (cap, 3, 3), (cap, 3, 3, 3), (cap, 3), (cap, 3), (cap, 3, 3)]
    outs = [torch.zeros(s, dtype=torch.float64, device="cuda") for s in shapes]
    count = ctypes.c_int64(0)
    status = _lib.load().ofw_tree_nodes(_ctx(), m, *[d.data_ptr() for d in dev], int(leaf_size),
                                        *[o.data_ptr() for o in outs], ctypes.byref(count))
    _lib.check(status, "ofw_tree_nodes")
    c = count.value
    return tuple(o[:c].cpu().numpy() for o in outs)


def hip_tree_sum_device(points, starts, directions, lengths, gamma, radius, model, ex_ptr, ex_seg, singularity,
                        tree, settings, out=None, flags=None, far_counts=None, near_counts=None, split_depth=-1):
    """Tree sum on device tensors; returns (out, flags, far_counts, near_counts) device tensors."""
    torch = _torch()
    n, m = int(points.shape[0]), int(starts.shape[0])
    if out is None:
        out = torch.zeros((n, 3), dtype=torch.float64, device=points.device)
    if flags is None:
        flags = torch.zeros(n, dtype=torch.int8, device=points.device)
    if far_counts is None:
        far_counts = torch.zeros(n, dtype=torch.int64, device=points.device)
    if near_counts is None:
        near_counts = torch.zeros(n, dtype=torch.int64, device=points.device)
    if n == 0:
        return out, flags, far_counts, near_counts
    status = _lib.load().ofw_tree_sum(
        _ctx(), _lib.PRECISIONS[settings.precision], n, _ptr(points), m, _ptr(starts), _ptr(directions),
        _ptr(lengths), _ptr(gamma), _ptr(radius), int(model), int(singularity == "raise"), _ptr(ex_ptr), _ptr(ex_seg),
        float(tree.theta), int(tree.order), int(tree.leaf_size), float(tree.core_clearance), int(split_depth),
        _ptr(out), _ptr(flags), _ptr(far_counts), _ptr(near_counts))
    _lib.check(status, "ofw_tree_sum")
    return out, flags, far_counts, near_counts


def hip_tree_sum(points, starts, directions, lengths, gamma, radius, model, csr, singularity, tree, settings, *,
                 split_depth=-1, targets_per_launch=None, stack_size=None, block=None):
    """Return (velocity, any_flag) like vortex_tree.tree_sum for validated inputs.

    ``targets_per_launch``, ``stack_size`` and ``block`` are accepted for the signature of
    vortex_cuda_tree.cuda_tree_sum (vortex_cuda_tree.py:439-440) and have no effect: all targets go in one launch,
    the walk is stackless (no traversal stack can overflow) and the workgroup shapes are fixed per kernel.
    ``split_depth`` is this engine's own knob (the depth of the traversal's subtree frontier, -1: chosen from the
    target count)."""
    if not isinstance(tree, TreeSettings):
        raise ValueError("tree must be TreeSettings")
    if settings is None:
        settings = HipSettings("fp64")
    radius = np.asarray(radius, dtype=np.float64)
    if int(model) not in TREE_MODELS:
        if np.any(radius > 0):
            raise ValueError("tree backend supports core models NONE, RANKINE and LAMB_OSEEN only")
        model = 0   # without a core every model is the ideal filament
    if tree.theta >= 1.0:
        raise ValueError("tree theta must be below one")
    n, m = len(points), len(starts)
    if n == 0 or m == 0:
        Statistics.last = dict(points=n, segments=m)
        return np.zeros((n, 3)), False
    # the upload / evaluation / download split of vortex_cuda_tree.py:581-594 (host clock, device synchronised)
    t0 = time.perf_counter()
    torch, *dev = _upload(points, starts, directions, lengths, gamma, radius, csr)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    out, flags, far, near = hip_tree_sum_device(*dev[:6], model, dev[6], dev[7], singularity, tree, settings,
                                                split_depth=split_depth)
    t2 = time.perf_counter()
    host_out = out.cpu().numpy()
    host_flags = flags.cpu().numpy()
    t3 = time.perf_counter()
    far = far.cpu().numpy()
    stats = (ctypes.c_int64 * 8)()
    _lib.load().ofw_ctx_stats(_ctx(), stats, 8)
    Statistics.last = dict(points=n, segments=m, precision=settings.precision, flags=host_flags, flag_seg=None,
                           far_counts=far, near_counts=near.cpu().numpy(), mean_far_nodes=float(far.mean()),
                           nodes=int(stats[4]), depth=int(stats[5]), subtrees=int(stats[2]),
                           split_depth=int(stats[3]), work_items=int(stats[6]), attempts=int(stats[7]),
                           phases_ms=_lib.timings(_ctx()), upload_ms=1e3 * (t1 - t0), evaluate_ms=1e3 * (t2 - t1),
                           download_ms=1e3 * (t3 - t2))
    return host_out, bool(np.any(host_flags))


class TreeBatch:
    """Many independent snapshots in one tree-code call, recorded once and replayed (ofw_tree_batch_*).

    ``snapshots`` is a list of dicts with the engine arrays of one snapshot each (``points``, ``starts``,
    ``directions``, ``lengths``, ``gamma``, ``radius``; host or device).  They are packed back to back on the
    device; ``arrays`` holds the packed device tensors, whose contents may be changed in place between runs
    (the recorded call reads them at every replay).  Each snapshot's result equals ``hip_tree_sum`` on the
    snapshot alone with the same ``split_depth``, bit for bit.
    """

    def __init__(self, snapshots, tree, *, model=0, precision="mixed", split_depth=-1, use_graph=True,
                 singularity="zero", counts=False):
        torch = _torch()
        if not isinstance(tree, TreeSettings):
            raise ValueError("tree must be TreeSettings")
        names = ("points", "starts", "directions", "lengths", "gamma", "radius")
        self.K = len(snapshots)
        if self.K == 0:
            raise ValueError("a batch needs at least one snapshot")
        n_k = [int(len(s["points"])) for s in snapshots]
        m_k = [int(len(s["starts"])) for s in snapshots]
        self.target_offsets = np.concatenate([[0], np.cumsum(n_k)]).astype(np.int64)
        self.segment_offsets = np.concatenate([[0], np.cumsum(m_k)]).astype(np.int64)
        packed = {}
        for name in names:
            parts = [torch.as_tensor(s[name], dtype=torch.float64).to("cuda").reshape(
                (-1, 3) if name in ("points", "starts", "directions") else (-1,)) for s in snapshots]
            packed[name] = torch.cat(parts).contiguous()
        self.arrays = packed
        n = int(self.target_offsets[-1])
        # exclusions: each snapshot's (target, segment) pairs, local to the snapshot, shifted to batch ids
        pairs = [np.asarray(s.get("pairs", np.zeros((0, 2))), dtype=np.int64).reshape(-1, 2)
                 + [self.target_offsets[k], self.segment_offsets[k]] for k, s in enumerate(snapshots)]
        pairs = np.concatenate(pairs) if pairs else np.zeros((0, 2), np.int64)
        self.ex_ptr = self.ex_seg = None
        if len(pairs):
            ptr, seg = exclusion_csr(n, pairs)
            self.ex_ptr = torch.as_tensor(ptr).to("cuda")
            self.ex_seg = torch.as_tensor(seg).to("cuda")
        self.out = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        self.flags = torch.zeros(n, dtype=torch.int8, device="cuda")
        self.far_counts = torch.zeros(n, dtype=torch.int64, device="cuda") if counts else None
        self.near_counts = torch.zeros(n, dtype=torch.int64, device="cuda") if counts else None
        self._lib = _lib.load()
        self._plan = ctypes.c_void_p()
        # the plan replays on the stream that is current now; run() orders it after the caller's current stream
        self._stream = torch.cuda.current_stream()
        toff = (ctypes.c_int64 * (self.K + 1))(*self.target_offsets.tolist())
        soff = (ctypes.c_int64 * (self.K + 1))(*self.segment_offsets.tolist())
        status = self._lib.ofw_tree_batch_create(
            _ctx(), ctypes.byref(self._plan), _lib.PRECISIONS[precision], self.K, toff, soff,
            *[_ptr(packed[name]) for name in names], int(model), int(singularity == "raise"), _ptr(self.ex_ptr),
            _ptr(self.ex_seg),
            float(tree.theta), int(tree.order), int(tree.leaf_size), float(tree.core_clearance), int(split_depth),
            int(bool(use_graph)), _ptr(self.out), _ptr(self.flags), _ptr(self.far_counts), _ptr(self.near_counts))
        _lib.check(status, "ofw_tree_batch_create")

    def run(self):
        """Replay the call; returns the packed device velocities (n, 3).

        The replay runs on the stream that was current when the batch was made.  When another stream is current
        now (where the caller may have changed ``arrays``), the replay waits for it, and it waits for the replay.
        """
        torch = _torch()
        current = torch.cuda.current_stream()
        other = current != self._stream
        if other:
            self._stream.wait_stream(current)
        _lib.check(self._lib.ofw_tree_batch_run(self._plan), "ofw_tree_batch_run")
        if other:
            current.wait_stream(self._stream)
        return self.out

    def split(self, array):
        """Per-snapshot views of a packed per-target array."""
        return [array[a:b] for a, b in zip(self.target_offsets[:-1], self.target_offsets[1:])]

    def timings(self):
        buf = (ctypes.c_float * 8)()
        phases = self._lib.ofw_tree_batch_timings(self._plan, buf, 8)
        return [float(buf[k]) for k in range(max(phases, 0))]

    def stats(self):
        buf = (ctypes.c_int64 * 8)()
        self._lib.ofw_tree_batch_stats(self._plan, buf, 8)
        return dict(subtrees=int(buf[2]), split_depth=int(buf[3]), nodes=int(buf[4]), depth=int(buf[5]))

    def close(self):
        if self._plan:
            self._lib.ofw_tree_batch_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
