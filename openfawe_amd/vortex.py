# SPDX-License-Identifier: Apache-2.0
# The settings classes, the input checks of segments_velocity (their order, exception types and messages) and the
# segment_velocity / ring_velocity helpers follow OpenFAWE's src/openfawe/aero/vortex.py (Copyright 2026 Yinong
# Tian, Apache License 2.0, itself adapted from OpenFAST/OLAF; see NOTICE): the drop-in boundary keeps them.
"""Directed straight vortex segments on MI355X: the reference's batch API and engine signatures.

Public API (mirrors src/openfawe/aero/vortex.py of the reference):

* ``CoreModel``, ``TreeSettings`` (vortex.py:21-55) and ``HipSettings`` (the counterpart of
  ``CudaSettings``, vortex.py:58-116);
* ``segments_velocity`` (vortex.py:368-497) with the backends ``"hip"`` (direct sum) and
  ``"hip_tree"`` (tree code), and ``segment_velocity`` / ``ring_velocity`` (vortex.py:500-524);
* the engines behind them, with the reference's engine signature
  ``engine(points, starts, directions, lengths, gamma, radius, model, csr, singularity, *settings)
  -> (out, any_flag)``: ``hip_sum`` (for ``vortex_cuda.cuda_sum``, vortex_cuda.py:381) and
  ``hip_tree_sum`` (for ``vortex_cuda_tree.cuda_tree_sum``, vortex_cuda_tree.py:439);
* ``exclusion_csr`` (vortex_numba.py:228-243).

All arithmetic runs in libofw_hip.so on the GPU.  The engines take host NumPy arrays (as the
reference's do) or device tensors (``*_device`` variants) and return owned results.
"""
from dataclasses import dataclass
from enum import IntEnum
import operator

import numpy as np

from . import _lib

FLAG_OK, FLAG_SINGULAR, FLAG_NONFINITE = 0, 1, 2
SINGULAR_MESSAGE = "Point on filament/end point; declare exclusions or singularity='zero'"
RANGE_MESSAGE = "geometry/core/velocity exceeds supported float64 arithmetic range"
TREE_MODELS = (0, 1, 2)  # vortex_tree.py:46: NONE, RANKINE, LAMB_OSEEN


class CoreModel(IntEnum):
    NONE = 0
    RANKINE = 1
    LAMB_OSEEN = 2
    VATISTAS = 3
    OFFSET = 4
    ROSENHEAD_MOORE = 5


class FilamentSingularity(ValueError):
    """Observation is on a segment/end point and needs an explicit policy (vortex.py:137-138)."""


@dataclass(frozen=True)
class TreeSettings:
    """Hierarchical (Barnes-Hut type) far-field settings, validated as vortex.py:30-55.

    theta: opening ratio in (0, 1); order: multipole order 0, 1 or 2; leaf_size: at most this many
    segments per leaf; core_clearance: a node is approximated only if the target lies at least this
    many largest core radii outside the node's bounding sphere.
    """
    theta: float = 0.5
    order: int = 2
    leaf_size: int = 32
    core_clearance: float = 3.0

    def __post_init__(self):
        if isinstance(self.theta, (bool, np.bool_)) or not (0 < float(self.theta) < 1):
            raise ValueError("tree theta must lie strictly between 0 and 1")
        if isinstance(self.order, (bool, np.bool_)) or self.order not in (0, 1, 2):
            raise ValueError("tree order must be 0, 1 or 2")
        if isinstance(self.leaf_size, (bool, np.bool_)) or operator.index(self.leaf_size) < 1:
            raise ValueError("tree leaf size must be a positive integer")
        if not np.isfinite(self.core_clearance) or self.core_clearance < 0:
            raise ValueError("tree core clearance must be finite and nonnegative")


@dataclass(frozen=True)
class HipSettings:
    """GPU settings, the counterpart of ``CudaSettings`` (vortex.py:58-116), validated the same way.

    precision: 'fp64' (FP64 throughout), 'mixed' (FP32 far pairs and far field, FP64 near pairs,
    moments and per-target accumulation) or 'fp32' (as mixed with an FP32 accumulator).
    block: threads per workgroup (32, 64, 128, 256).
    suspect / moments / build / control / rings are accepted with the reference's values for drop-in
    use.  This backend has one implementation for each: the tree is always built and its moments are
    always computed on the GPU (element for element the reference tree), and near pairs are always
    re-evaluated in FP64 on the GPU, so no target is recomputed on the CPU.
    """
    precision: str = "mixed"
    block: int = 128
    suspect: str = "cpu"
    control: str = "cpu"
    moments: str = "cpu"
    build: str = "cpu"
    rings: str = "cpu"

    def __post_init__(self):
        if self.precision not in ("fp64", "mixed", "fp32"):
            raise ValueError("hip precision must be 'fp64', 'mixed' or 'fp32'")
        if isinstance(self.block, (bool, np.bool_)) or self.block not in (32, 64, 128, 256):
            raise ValueError("hip block must be 32, 64, 128 or 256")
        if self.suspect not in ("cpu", "gpu64"):
            raise ValueError("hip suspect must be 'cpu' or 'gpu64'")
        if self.suspect == "gpu64" and self.precision == "fp64":
            raise ValueError("hip suspect='gpu64' is for the single precisions 'mixed' and 'fp32' only")
        for name in ("control", "moments", "build", "rings"):
            if getattr(self, name) not in ("cpu", "gpu"):
                raise ValueError(f"hip {name} must be 'cpu' or 'gpu'")


# ------------------------------------------------------------------ validation (vortex.py:141-180)
def _real(value, name):
    raw = np.asarray(value)
    if raw.dtype.kind not in "iuf":
        raise ValueError(f"{name} must contain real numbers")
    result = np.asarray(raw, dtype=np.float64)
    if not np.isfinite(result).all():
        raise ValueError(f"{name} must be finite")
    return result


def _matrix(value, width, name):
    result = _real(value, name)
    if result.ndim != 2 or result.shape[1] != width:
        raise ValueError(f"{name} must have shape (n, {width})")
    return result


def _indices(value, width, name):
    result = np.asarray(value)
    if result.ndim != 2 or result.shape[1] != width or result.dtype.kind not in "iu":
        raise ValueError(f"{name} must be an integer array of shape (n, {width})")
    return result


def _per_segment(value, count, name):
    result = _real(value, name)
    if result.ndim == 0:
        return np.full(count, result.item(), dtype=np.float64)
    if result.shape != (count,):
        raise ValueError(f"{name} must be scalar or have shape ({count},)")
    return result


def _model(value):
    if isinstance(value, (bool, np.bool_)):
        raise ValueError("model must be an integer CoreModel ID")
    try:
        return CoreModel(operator.index(value))
    except (TypeError, ValueError) as error:
        raise ValueError("model must be an integer CoreModel ID from 0 through 5") from error


def _norm(value):
    return np.hypot(np.hypot(value[..., 0], value[..., 1]), value[..., 2])


def exclusion_csr(point_count, pairs):
    """Per-point sorted unique excluded segment ids from an (e, 2) (point, segment) array.

    The arrays of vortex_numba.exclusion_csr (vortex_numba.py:180-243): pointer (n+1) int64 and
    segments (unique pairs sorted by point, then segment).
    """
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    count = operator.index(point_count)
    pointer = np.zeros(count + 1, dtype=np.int64)
    if len(pairs) == 0:
        return pointer, np.zeros(0, dtype=np.int64)
    if int(pairs[:, 0].min()) < 0 or int(pairs[:, 0].max()) >= count:
        raise ValueError("exclusion index out of range")
    key = np.unique(pairs[:, 0] * (int(pairs[:, 1].max()) + 1) + pairs[:, 1])
    width = int(pairs[:, 1].max()) + 1
    points, segments = key // width, key % width
    np.cumsum(np.bincount(points, minlength=count), out=pointer[1:])
    return pointer, np.ascontiguousarray(segments, dtype=np.int64)


# ------------------------------------------------------------------ device plumbing
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.LibraryError("the hip velocity backend needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def available():
    """True when the library is built and a GPU is visible (the counterpart of vortex_cuda.available)."""
    try:
        _lib.load()
        import torch
        return bool(torch.cuda.is_available())
    except Exception:  # noqa: BLE001 - any missing piece means "not available"
        return False


def _ctx():
    torch = _torch()
    stream = torch.cuda.current_stream()
    return _lib.context(stream.cuda_stream, torch.cuda.current_device())


def _dev(torch, array, dtype):
    t = torch.from_numpy(np.ascontiguousarray(array, dtype=dtype))
    return t.to("cuda")


def _ptr(t):
    return t.data_ptr() if t is not None and t.numel() else None


class Statistics:
    """Last call's counters (the counterpart of vortex_cuda.Statistics / vortex_cuda_tree.Statistics)."""
    last = {}


def hip_sum_device(points, starts, directions, lengths, gamma, radius, model, ex_ptr, ex_seg, singularity,
                   settings, out=None, flags=None, flag_seg=None, static_scratch=False):
    """Direct sum on device tensors (float64 / int64, contiguous, on the current GPU).

    Returns (out (n,3) float64, flags (n,) int8, flag_seg (n,) int32) as device tensors.  With
    ``static_scratch`` the call goes through ``ofw_direct_sum_static``, which allocates no device memory
    (its scratch is the library's fixed per-GPU array; with out / flags / flag_seg given, nothing is
    allocated at all).
    """
    torch = _torch()
    n, m = int(points.shape[0]), int(starts.shape[0])
    if out is None:
        out = torch.zeros((n, 3), dtype=torch.float64, device=points.device)
    if flags is None:
        flags = torch.zeros(n, dtype=torch.int8, device=points.device)
    if flag_seg is None:
        flag_seg = torch.full((n,), -1, dtype=torch.int32, device=points.device)
    if n == 0:
        return out, flags, flag_seg
    lib = _lib.load()
    name = "ofw_direct_sum_static" if static_scratch else "ofw_direct_sum"
    status = getattr(lib, name)(_ctx(), _lib.PRECISIONS[settings.precision], int(settings.block), n, _ptr(points), m,
                                _ptr(starts), _ptr(directions), _ptr(lengths), _ptr(gamma), _ptr(radius), int(model),
                                int(singularity == "raise"), _ptr(ex_ptr), _ptr(ex_seg), _ptr(out), _ptr(flags),
                                _ptr(flag_seg))
    _lib.check(status, name)
    return out, flags, flag_seg


def _upload(points, starts, directions, lengths, gamma, radius, csr):
    torch = _torch()
    pointer, segments = csr
    f64, i64 = np.float64, np.int64
    return (torch, _dev(torch, points, f64).reshape(-1, 3), _dev(torch, starts, f64).reshape(-1, 3),
            _dev(torch, directions, f64).reshape(-1, 3), _dev(torch, lengths, f64), _dev(torch, gamma, f64),
            _dev(torch, radius, f64), _dev(torch, pointer, i64),
            _dev(torch, segments if len(segments) else np.zeros(1, i64), i64))


def hip_sum(points, starts, directions, lengths, gamma, radius, model, csr, singularity, settings, *,
            pairs_per_launch=None):
    """Return (velocity, any_flag) for validated inputs, like vortex_cuda.cuda_sum (vortex_cuda.py:381).

    ``pairs_per_launch`` is accepted for the reference's signature (vortex_cuda.py:381-382) and has no effect:
    one call is one launch of the pair kernel here, whatever the pair count.
    ``csr`` is the (pointer, segments) exclusion structure from ``exclusion_csr``.  Per-target flags
    and the first flagged segment of each target are kept in ``Statistics.last`` for the caller's
    exception precedence.
    """
    n, m = len(points), len(starts)
    if n == 0 or m == 0:
        Statistics.last = dict(points=n, segments=m)
        return np.zeros((n, 3)), False
    torch, *dev = _upload(points, starts, directions, lengths, gamma, radius, csr)
    out, flags, flag_seg = hip_sum_device(*dev[:6], model, dev[6], dev[7], singularity, settings)
    host_flags = flags.cpu().numpy()
    Statistics.last = dict(points=n, segments=m, precision=settings.precision, flags=host_flags,
                           flag_seg=flag_seg.cpu().numpy())
    return out.cpu().numpy(), bool(np.any(host_flags))


def _raise_for_flags(stats):
    """Raise the reference's exception for flagged pairs (vortex.py:488-492 re-runs its NumPy sum to raise).

    The NumPy sum processes segments in blocks of 32 (vortex.py:326-359); the exception of the earliest
    block holding a flagged pair wins: FilamentSingularity when that block holds an on-segment pair,
    otherwise the float64 range error.
    """
    flags, flag_seg = stats.get("flags"), stats.get("flag_seg")
    if flags is None:
        raise ValueError(RANGE_MESSAGE)
    flagged = flags != 0
    if flag_seg is not None and np.any(flag_seg[flagged] >= 0):
        block = flag_seg[flagged] // 32
        first = block.min()
        kinds = flags[flagged][block == first]
    else:
        kinds = flags[flagged]
    if np.any(kinds == FLAG_SINGULAR):
        raise FilamentSingularity(SINGULAR_MESSAGE)
    raise ValueError(RANGE_MESSAGE)


def segments_velocity(points, nodes, connectivity, circulation, *, model=CoreModel.NONE, core_radius=0.0,
                      exclusions=None, singularity="raise", chunk_size=4096, backend="hip", tree=None, hip=None):
    """Sum directed segments into a new float64 (n,3) velocity array on the GPU.

    Same arguments, validation and errors as the reference's ``segments_velocity``
    (vortex.py:368-497).  ``backend='hip'`` is the direct sum, ``backend='hip_tree'`` the tree code
    (needs ``tree``).  ``hip`` is a ``HipSettings`` (default ``HipSettings('fp64')`` for the direct sum
    and the tree, so that the default result is the FP64 sum).
    """
    if backend not in ("hip", "hip_tree"):
        raise ValueError("backend must be 'hip' or 'hip_tree'")
    if backend == "hip_tree":
        if not isinstance(tree, TreeSettings):
            raise ValueError("backend='hip_tree' needs tree settings (TreeSettings)")
    elif tree is not None:
        raise ValueError("tree settings are required exactly for backend='hip_tree'")
    if hip is None:
        hip = HipSettings("fp64")
    if not isinstance(hip, HipSettings):
        raise ValueError("hip must be HipSettings")
    points, nodes = _matrix(points, 3, "points"), _matrix(nodes, 3, "nodes")
    edges = _indices(connectivity, 2, "connectivity")
    if np.any(edges < 0) or np.any(edges >= len(nodes)):
        raise ValueError("connectivity index out of range")
    gamma = _per_segment(circulation, len(edges), "circulation")
    radius = _per_segment(core_radius, len(edges), "core_radius")
    if np.any(radius < 0):
        raise ValueError("core_radius must be nonnegative")
    model = _model(model)
    if singularity not in ("raise", "zero"):
        raise ValueError("singularity must be 'raise' or 'zero'")
    if isinstance(chunk_size, (bool, np.bool_)):
        raise ValueError("chunk_size must be a positive integer")
    try:
        chunk_size = operator.index(chunk_size)
    except TypeError as error:
        raise ValueError("chunk_size must be a positive integer") from error
    if chunk_size <= 0:
        raise ValueError("chunk_size must be a positive integer")
    pairs = np.zeros((0, 2), dtype=np.int64)
    if exclusions is not None:
        pairs = _indices(exclusions, 2, "exclusions")
        if np.any(pairs < 0) or np.any(pairs[:, 0] >= len(points)) or np.any(pairs[:, 1] >= len(edges)):
            raise ValueError("exclusion index out of range")
    try:
        with np.errstate(over="raise", invalid="raise", divide="raise", under="ignore"):
            starts = nodes[edges[:, 0]]
            delta = nodes[edges[:, 1]] - starts
            lengths = _norm(delta)
            if np.any(lengths == 0):
                raise ValueError("zero-length segment")
            directions = delta / lengths[:, None]
    except FloatingPointError as error:
        raise ValueError(RANGE_MESSAGE) from error
    csr = exclusion_csr(len(points), pairs)
    if backend == "hip":
        out, flagged = hip_sum(points, starts, directions, lengths, gamma, radius, model, csr, singularity, hip)
    else:
        from .tree import hip_tree_sum
        out, flagged = hip_tree_sum(points, starts, directions, lengths, gamma, radius, model, csr, singularity,
                                    tree, hip)
    if flagged:
        _raise_for_flags(Statistics.last)
    if not np.isfinite(out).all():
        raise ValueError("velocity exceeds finite float64 range")
    return out


def segment_velocity(start, end, point, circulation, *, model=CoreModel.NONE, core_radius=0.0,
                     singularity="raise", hip=None):
    """Single-segment counterpart of segments_velocity; returns an owned (3,) array (vortex.py:500-507)."""
    vectors = [_real(value, name) for value, name in ((start, "start"), (end, "end"), (point, "point"))]
    if any(value.shape != (3,) for value in vectors):
        raise ValueError("start, end and point must each have shape (3,)")
    return segments_velocity(vectors[2][None, :], np.stack(vectors[:2]), [[0, 1]], circulation, model=model,
                             core_radius=core_radius, singularity=singularity, hip=hip)[0].copy()


def ring_velocity(corners, points, circulation, *, model=CoreModel.NONE, core_radius=0.0, singularity="raise",
                  exclusions=None, chunk_size=4096, hip=None):
    """Closed A->B->C->D->A ring (vortex.py:510-524); scalar circulation and core required."""
    corners = _real(corners, "corners")
    if corners.shape != (4, 3):
        raise ValueError("corners must have shape (4, 3)")
    if _real(circulation, "circulation").ndim or _real(core_radius, "core_radius").ndim:
        raise ValueError("ring circulation and core_radius must be scalars")
    return segments_velocity(points, corners, [[0, 1], [1, 2], [2, 3], [3, 0]], circulation, model=model,
                             core_radius=core_radius, singularity=singularity, exclusions=exclusions,
                             chunk_size=chunk_size, hip=hip)
