# SPDX-License-Identifier: Apache-2.0
# Test oracle: a restatement of OpenFAWE's src/openfawe/aero/vortex.py pair and summation code (Copyright 2026 Yinong Tian, Apache License 2.0, itself
# adapted from OpenFAST/OLAF; see NOTICE).
"""Vectorised NumPy restatement of the reference's pair kernel.  TEST INFRASTRUCTURE ONLY.

A second, independent CPU formulation of the segment velocity that follows the
reference's NumPy path (vortex.py:183-184 ``_norm``, vortex.py:187-199
``_rosenhead_geometry`` and vortex.py:258-307 ``_evaluate_pairs``: masks over
arrays instead of per-pair branches).  The reference checks its compiled kernel
against this path with bound A1 (|u_a - u_b| <= 1e-12 * sum_s |u_s| per point,
tests/test_vortex_numba.py:1-5 of the reference); the tests here apply the same
bound between this module and the C oracle.
"""
import numpy as np

LAMB = 1.25643


def _norm(v):
    return np.hypot(np.hypot(v[..., 0], v[..., 1]), v[..., 2])


def _rosenhead_geometry(na, nb, h, t, rc):
    na, nb = np.hypot(na, rc), np.hypot(nb, rc)
    product = na * nb
    transverse2 = h * h + rc * rc
    dot = transverse2 + t * (t - 1)
    plus = product + dot
    opposite = dot < 0
    plus[opposite] = transverse2[opposite] / (product[opposite] - dot[opposite])
    return na, nb, product * plus


def evaluate_pairs(points, starts, directions, lengths, gamma, model, radius, singularity="zero"):
    """Velocity of each (point_i, segment_i) pair; returns (out (k,3), on_segment (k,) bool)."""
    points, starts, directions = (np.asarray(v, dtype=np.float64) for v in (points, starts, directions))
    lengths, gamma, radius = (np.asarray(v, dtype=np.float64) for v in (lengths, gamma, radius))
    with np.errstate(all="ignore"):
        a = (points - starts) / lengths[:, None]
        t = (a[:, 0] * directions[:, 0] + a[:, 1] * directions[:, 1]) + a[:, 2] * directions[:, 2]
        cross = np.cross(directions, a)
        h = _norm(cross)
        na, nb = _norm(a), _norm(a - directions)
        axis = h <= 8 * np.finfo(np.float64).eps * np.maximum(1.0, np.maximum(na, nb))
        on_segment = axis & (t >= 0) & (t <= 1)
        out = np.zeros_like(points)
        active = (~axis) & (gamma != 0)
        if not np.any(active):
            return out, on_segment
        ha, taa = h[active], t[active]
        naa, nbb = na[active], nb[active]
        product = naa * nbb
        dot = ha * ha + taa * (taa - 1)
        plus = product + dot
        opposite = dot < 0
        plus[opposite] = ha[opposite] ** 2 / (product[opposite] - dot[opposite])
        denominator = product * plus
        length = lengths[active]
        radii = radius[active]
        rc = radii / length
        factor = np.ones_like(ha)
        cored = (radii > 0) & (model != 0)
        if np.any(cored):
            if model == 5:
                naa[cored], nbb[cored], denominator[cored] = _rosenhead_geometry(
                    naa[cored], nbb[cored], ha[cored], taa[cored], rc[cored])
            elif model == 4:
                denominator[cored] += rc[cored] * rc[cored]
            else:
                ratio = ha[cored] / rc[cored]
                correction = np.ones_like(ratio)
                if model == 1:
                    correction = np.minimum(ratio, 1.0) ** 2
                elif model == 2:
                    inside = ratio <= np.sqrt(10 / LAMB)
                    correction[inside] = -np.expm1(-LAMB * ratio[inside] ** 2)
                elif model == 3:
                    small = ratio < 1
                    q = ratio[small] ** 2
                    correction[small] = q / np.hypot(1, q)
                    correction[~small] = 1 / np.hypot(1, (1 / ratio[~small]) ** 2)
                factor[cored] = correction
        coefficient = (gamma[active] / length) / (4 * np.pi) * factor * (naa + nbb) / denominator
        out[active] = coefficient[:, None] * cross[active]
    return out, on_segment


def direct_sum(points, starts, directions, lengths, gamma, radius, model, excluded=None):
    """Per point, the ordered sum over segments (segment input order), skipping excluded pairs.

    ``excluded`` is a set of (point, segment) pairs.  Returns (out (n,3), scale (n,)) where
    scale = sum_s |u_s| is the A1 normaliser.
    """
    points = np.asarray(points, dtype=np.float64)
    starts, directions = np.asarray(starts, dtype=np.float64), np.asarray(directions, dtype=np.float64)
    n, m = len(points), len(starts)
    lengths = np.asarray(lengths, dtype=np.float64)
    gamma = np.broadcast_to(np.asarray(gamma, dtype=np.float64), (m,))
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), (m,))
    out = np.zeros((n, 3))
    scale = np.zeros(n)
    by_segment = {}
    for p, s in (excluded or ()):
        by_segment.setdefault(int(s), set()).add(int(p))
    for s in range(m):
        keep = np.ones(n, dtype=bool)
        if s in by_segment:
            keep[list(by_segment[s])] = False
        idx = np.flatnonzero(keep)
        k = len(idx)
        u, _ = evaluate_pairs(points[idx], np.broadcast_to(starts[s], (k, 3)), np.broadcast_to(directions[s], (k, 3)),
                              np.full(k, lengths[s]), np.full(k, gamma[s]), model, np.full(k, radius[s]))
        out[idx] += u
        scale[idx] += np.linalg.norm(u, axis=1)
    return out, scale
