"""The fast FP64 pair at the FP64 neighbours of its jump thresholds, against the oracle.

Targets from tests/threshold_cases.py sit on consecutive doubles across the axis guard
h <= 8 eps max(1, |a|, |b|) (vortex_numba.py:59-62) and across the Lamb-Oseen cutoff h / rc <= sqrt(10/1.25643)
(vortex_numba.py:95-97).  Each pair goes through ofw_pair_matrix, the FP64 and mixed direct sums and the FP64
tree's near field, and is compared with O.pair_matrix:
* where every faithfully rounded hypot gives the reference one decision, flags equal and values within
  1e-12 relative (the decision moves the value by 4.5e-5 at the cutoff and from 0 to ~1/h at the guard);
* in the one-ulp window where the reference's own decision depends on its libm's hypot, the value is one of
  the two the reference can give.
"""
import numpy as np
import pytest

from oracle import oracle as O
from openfawe_amd import adjoint as AD
from openfawe_amd import tree as T
from openfawe_amd import vortex as V
from tests import threshold_cases as TC

pytestmark = pytest.mark.gpu


def _close(u, v, rel=1e-12):
    return np.linalg.norm(u - v) <= rel * max(np.linalg.norm(v), 1e-300)


def _engines(points, s, d, L, r, model, k):
    """Per-target velocity of segment k alone through the direct sums and the tree's near field."""
    n = len(points)
    one = (s[k:k + 1], d[k:k + 1], L[k:k + 1], np.ones(1), r[k:k + 1])
    csr = V.exclusion_csr(n, np.zeros((0, 2), np.int64))
    out = {}
    for precision in ("fp64", "mixed"):
        out[precision] = V.hip_sum(points, *one, model, csr, "zero", V.HipSettings(precision))[0]
    out["tree"] = T.hip_tree_sum(points, *one, model, csr, "zero", V.TreeSettings(0.5, 2, 32, 3.0),
                                 V.HipSettings("fp64"))[0]
    return out


def test_axis_guard_neighbours():
    s, d, L, P, info = TC.axis_cases()
    r = np.zeros(len(s))
    ref, rflags = O.pair_matrix(P, s, d, L, r, 0, "raise")
    got, gflags = AD.pair_matrix(P, s, d, L, r, 0)
    assert np.array_equal(gflags, rflags)
    for i, (k, h, amb, axis) in enumerate(info):
        if amb:
            assert np.all(got[i, k] == 0.0) or np.linalg.norm(got[i, k]) > 0.0
        else:
            assert _close(got[i, k], ref[i, k]), (i, k, got[i, k], ref[i, k])
            assert np.all(got[i, k] == 0.0) == axis
    zref, _ = O.pair_matrix(P, s, d, L, r, 0, "zero")
    for k in range(len(s)):
        idx = [i for i, x in enumerate(info) if x[0] == k and not x[2]]
        engines = _engines(P[idx], s, d, L, r, 0, k)
        for name, out in engines.items():
            for j, i in enumerate(idx):
                assert _close(out[j], zref[i, k]), (name, i, out[j], zref[i, k])


def test_lamb_cutoff_neighbours():
    s, d, L, R, P, info = TC.lamb_cases()
    ref, rflags = O.pair_matrix(P, s, d, L, R, 2, "raise")
    plain, _ = O.pair_matrix(P, s, d, L, R, 0, "raise")
    got, gflags = AD.pair_matrix(P, s, d, L, R, 2)
    assert not rflags.any() and not gflags.any()
    for i, (k, ratio, amb, inside) in enumerate(info):
        if amb:
            inner = plain[i, k] * -np.expm1(-1.25643 * float(ratio) ** 2)
            assert _close(got[i, k], ref[i, k]) or _close(got[i, k], plain[i, k]) or _close(got[i, k], inner, 1e-9)
        else:
            assert _close(got[i, k], ref[i, k]), (i, k, inside)
    for k in range(len(s)):
        idx = [i for i, x in enumerate(info) if x[0] == k and not x[2]]
        engines = _engines(P[idx], s, d, L, R, 2, k)
        for name, out in engines.items():
            for j, i in enumerate(idx):
                assert _close(out[j], ref[i, k]), (name, i)
