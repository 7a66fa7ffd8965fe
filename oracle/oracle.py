"""ctypes front end of the C oracle (libofw_oracle.so).  TEST INFRASTRUCTURE ONLY.

The oracle restates the reference algorithm of OpenFAWE's Biot-Savart segment
sum on the CPU (see ofw_oracle.c for the file:line each function follows).  It
checks the HIP product's results and is the CPU baseline of bench.py; the
product package ``openfawe_amd`` never imports it.

Signatures follow the reference's engine convention: float64 C-contiguous
arrays, ``csr = (pointer (n+1) int64, segments (e) int64)``.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libofw_oracle.so"

FLAG_OK, FLAG_SINGULAR, FLAG_NONFINITE = 0, 1, 2

_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p


def build():
    """Compile the oracle from ofw_oracle.c (make)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _lib.ofo_abi_version.restype = _i
        _lib.ofo_abi_version()
        _lib.ofo_pair_velocity.restype = _i
        _lib.ofo_pair_velocity.argtypes = [_d] * 11 + [_i, _d, _i, _p]
        _lib.ofo_exclusion_csr.restype = _i64
        _lib.ofo_tree_capacity.restype = _i64
        _lib.ofo_tree_capacity.argtypes = [_i64, _i64]
        _lib.ofo_build_tree_serial.restype = _i64
        _lib.ofo_build_tree.restype = _i64
        _lib.ofo_tree_sum.restype = _i64
        _lib.ofo_far_velocity.argtypes = [_d, _d, _d] + [_p] * 6 + [_i, _p]
        _lib.ofo_multipole_explicit.argtypes = [_d, _d, _d] + [_p] * 3 + [_i, _p]
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def default_threads():
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def pair_velocity(p, s, d, length, gamma, model, radius, raise_singular=True):
    """vortex_numba.pair_velocity for one pair: (u (3,), flag)."""
    out = np.zeros(3)
    f = lib().ofo_pair_velocity(float(p[0]), float(p[1]), float(p[2]), float(s[0]), float(s[1]), float(s[2]),
                                float(d[0]), float(d[1]), float(d[2]), float(length), float(gamma), int(model),
                                float(radius), int(bool(raise_singular)), _ptr(out))
    return out, f


def pair_values(points, starts, directions, lengths, gamma, model, radius, raise_singular=True):
    """One pair per row (k rows): (u (k,3), flags (k,))."""
    points, starts, directions = _f64(points, (-1, 3)), _f64(starts, (-1, 3)), _f64(directions, (-1, 3))
    k = len(points)
    lengths, gamma, radius = (_f64(np.broadcast_to(v, (k,))) for v in (lengths, gamma, radius))
    out = np.zeros((k, 3))
    flags = np.zeros(k, dtype=np.int8)
    lib().ofo_pair_values(_i64(k), _ptr(points), _ptr(starts), _ptr(directions), _ptr(lengths), _ptr(gamma),
                          _i(int(model)), _ptr(radius), _i(int(bool(raise_singular))), _ptr(out), _ptr(flags))
    return out, flags


def empty_csr(n):
    return np.zeros(n + 1, dtype=np.int64), np.zeros(0, dtype=np.int64)


def exclusion_csr(point_count, pairs):
    """vortex_numba.exclusion_csr: per-point sorted unique excluded segment ids."""
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if len(pairs) == 0:
        return empty_csr(point_count)
    pts, segs = _i64a(pairs[:, 0]), _i64a(pairs[:, 1])
    pointer = np.zeros(point_count + 1, dtype=np.int64)
    out = np.zeros(len(pairs), dtype=np.int64)
    k = lib().ofo_exclusion_csr(_i64(point_count), _i64(len(pairs)), _ptr(pts), _ptr(segs), _ptr(pointer), _ptr(out))
    return pointer, out[:k].copy()


def direct_sum(points, starts, directions, lengths, gamma, radius, model, csr=None, singularity="raise",
               threads=None, with_scale=False):
    """vortex_numba.direct_sum: (out (n,3), flags (n,) int8[, scale (n,) = sum_s |u_s|])."""
    points, starts, directions = _f64(points, (-1, 3)), _f64(starts, (-1, 3)), _f64(directions, (-1, 3))
    n, m = len(points), len(starts)
    lengths, gamma, radius = _f64(lengths), _f64(np.broadcast_to(gamma, (m,))), _f64(np.broadcast_to(radius, (m,)))
    ptr, seg = csr if csr is not None else empty_csr(n)
    ptr, seg = _i64a(ptr), _i64a(seg)
    out = np.zeros((n, 3))
    flags = np.zeros(n, dtype=np.int8)
    scale = np.zeros(n) if with_scale else None
    if n:
        lib().ofo_direct_sum(_i64(n), _ptr(points), _i64(m), _ptr(starts), _ptr(directions), _ptr(lengths),
                             _ptr(gamma), _ptr(radius), _i(int(model)), _i(int(singularity == "raise")), _ptr(ptr),
                             _ptr(seg if len(seg) else np.zeros(1, np.int64)), _ptr(out), _ptr(flags), _ptr(scale),
                             _i(default_threads() if threads is None else int(threads)))
    return (out, flags, scale) if with_scale else (out, flags)


def _tree_arrays(m, leaf):
    cap = lib().ofo_tree_capacity(m, leaf)
    return (np.zeros(m, np.int64), *(np.zeros(cap, np.int64) for _ in range(4)))


def build_tree(mid, leaf_size, threads=None):
    """vortex_tree.build_tree (level build + renumber): (order, lo, hi, left, right, count)."""
    mid = _f64(mid, (-1, 3))
    m = len(mid)
    order, lo, hi, left, right = _tree_arrays(m, int(leaf_size))
    count = lib().ofo_build_tree(_i64(m), _ptr(mid), _i64(int(leaf_size)), _ptr(order), _ptr(lo), _ptr(hi),
                                 _ptr(left), _ptr(right), _i(default_threads() if threads is None else int(threads)))
    return order, lo[:count].copy(), hi[:count].copy(), left[:count].copy(), right[:count].copy(), int(count)


def build_tree_serial(mid, leaf_size):
    """vortex_tree.build_tree_serial (depth-first build): (order, lo, hi, left, right, count)."""
    mid = _f64(mid, (-1, 3))
    m = len(mid)
    order, lo, hi, left, right = _tree_arrays(m, int(leaf_size))
    count = lib().ofo_build_tree_serial(_i64(m), _ptr(mid), _i64(int(leaf_size)), _ptr(order), _ptr(lo), _ptr(hi),
                                        _ptr(left), _ptr(right))
    return order, lo[:count].copy(), hi[:count].copy(), left[:count].copy(), right[:count].copy(), int(count)


def segment_arrays(starts, directions, lengths, gamma):
    """vortex_tree.py:608-611: (length_vectors, ends, alpha, mid), elementwise in float64."""
    starts, directions = _f64(starts, (-1, 3)), _f64(directions, (-1, 3))
    lv = directions * _f64(lengths)[:, None]
    ends = starts + lv
    alpha = _f64(gamma)[:, None] * lv
    mid = 0.5 * (starts + ends)
    return lv, ends, alpha, mid


def node_moments(order, lo, hi, starts, ends, alpha, length_vectors, radius, gamma, threads=None):
    """vortex_tree.node_moments: (centre, bound, core, m0, m1, m2)."""
    count = len(lo)
    order, lo, hi = _i64a(order), _i64a(lo), _i64a(hi)
    starts, ends, alpha, lv = (_f64(v, (-1, 3)) for v in (starts, ends, alpha, length_vectors))
    radius, gamma = _f64(radius), _f64(gamma)
    centre, bound, core = np.zeros((count, 3)), np.zeros(count), np.zeros(count)
    m0, m1, m2 = np.zeros((count, 3)), np.zeros((count, 3, 3)), np.zeros((count, 3, 3, 3))
    lib().ofo_node_moments(_i64(count), _ptr(order), _ptr(lo), _ptr(hi), _ptr(starts), _ptr(ends), _ptr(alpha),
                           _ptr(lv), _ptr(radius), _ptr(gamma), _ptr(centre), _ptr(bound), _ptr(core), _ptr(m0),
                           _ptr(m1), _ptr(m2), _i(default_threads() if threads is None else int(threads)))
    return centre, bound, core, m0, m1, m2


def node_contractions(m1, m2):
    """vortex_tree.node_contractions: (w1, t2, w2)."""
    m1, m2 = _f64(m1, (-1, 3, 3)), _f64(m2, (-1, 3, 3, 3))
    count = len(m1)
    w1, t2, w2 = np.zeros((count, 3)), np.zeros((count, 3)), np.zeros((count, 3, 3))
    lib().ofo_node_contractions(_i64(count), _ptr(m1), _ptr(m2), _ptr(w1), _ptr(t2), _ptr(w2))
    return w1, t2, w2


def far_velocity(r, m0, m1, m2, w1, t2, w2, order):
    """vortex_tree.far_velocity for one node's blocks."""
    out = np.zeros(3)
    blocks = [_f64(v) for v in (m0, m1, m2, w1, t2, w2)]
    lib().ofo_far_velocity(float(r[0]), float(r[1]), float(r[2]), *[_ptr(b) for b in blocks], int(order), _ptr(out))
    return out


def multipole_explicit(r, m0, m1, m2, order):
    """vortex_tree.multipole_velocity_explicit for one node's blocks."""
    out = np.zeros(3)
    blocks = [_f64(v) for v in (m0, m1, m2)]
    lib().ofo_multipole_explicit(float(r[0]), float(r[1]), float(r[2]), *[_ptr(b) for b in blocks], int(order),
                                 _ptr(out))
    return out


def tree_sum(points, starts, directions, lengths, gamma, radius, model, csr=None, singularity="raise", theta=0.5,
             order=2, leaf_size=32, clearance=3.0, threads=None):
    """vortex_tree.tree_sum: (out (n,3), flags (n,), far_counts (n,), near_counts (n,))."""
    points, starts, directions = _f64(points, (-1, 3)), _f64(starts, (-1, 3)), _f64(directions, (-1, 3))
    n, m = len(points), len(starts)
    lengths, gamma, radius = _f64(lengths), _f64(np.broadcast_to(gamma, (m,))), _f64(np.broadcast_to(radius, (m,)))
    ptr, seg = csr if csr is not None else empty_csr(n)
    ptr, seg = _i64a(ptr), _i64a(seg)
    out = np.zeros((n, 3))
    flags = np.zeros(n, dtype=np.int8)
    far = np.zeros(n, dtype=np.int64)
    near = np.zeros(n, dtype=np.int64)
    lib().ofo_tree_sum(_i64(n), _ptr(points), _i64(m), _ptr(starts), _ptr(directions), _ptr(lengths), _ptr(gamma),
                       _ptr(radius), _i(int(model)), _i(int(singularity == "raise")), _ptr(ptr),
                       _ptr(seg if len(seg) else np.zeros(1, np.int64)), _d(float(theta)), _i(int(order)),
                       _i64(int(leaf_size)), _d(float(clearance)), _ptr(out), _ptr(flags), _ptr(far), _ptr(near),
                       _i(default_threads() if threads is None else int(threads)))
    return out, flags, far, near


def pair_matrix(points, starts, directions, lengths, radius, model, singularity="raise", threads=None):
    """vortex_numba.pair_matrix: (out (n,m,3), flags (n,))."""
    points, starts, directions = _f64(points, (-1, 3)), _f64(starts, (-1, 3)), _f64(directions, (-1, 3))
    n, m = len(points), len(starts)
    lengths, radius = _f64(lengths), _f64(np.broadcast_to(radius, (m,)))
    out = np.zeros((n, m, 3))
    flags = np.zeros(n, dtype=np.int8)
    lib().ofo_pair_matrix(_i64(n), _ptr(points), _i64(m), _ptr(starts), _ptr(directions), _ptr(lengths),
                          _ptr(radius), _i(int(model)), _i(int(singularity == "raise")), _ptr(out), _ptr(flags),
                          _i(default_threads() if threads is None else int(threads)))
    return out, flags


def assemble_columns(pairs, column_ptr, column_segment, column_sign):
    """vortex_numba.assemble_columns: out[p, j] = sum_k sign_k * pairs[p, segment_k]."""
    pairs = _f64(pairs)
    n, m = pairs.shape[0], pairs.shape[1]
    column_ptr, column_segment, column_sign = _i64a(column_ptr), _i64a(column_segment), _f64(column_sign)
    columns = len(column_ptr) - 1
    out = np.zeros((n, columns, 3))
    lib().ofo_assemble_columns(_i64(n), _i64(m), _ptr(pairs), _i64(columns), _ptr(column_ptr),
                               _ptr(column_segment if len(column_segment) else np.zeros(1, np.int64)),
                               _ptr(column_sign if len(column_sign) else np.zeros(1)), _ptr(out))
    return out
