"""The oracle at the jump thresholds of the pair kernel, and the walk's square-root-free acceptance test.

* Axis guard and Lamb-Oseen cutoff (vortex_numba.py:59-62, 95-97): at targets on the FP64 neighbours of each
  threshold (tests/threshold_cases.py), the oracle's decision equals the decision fixed by exact arithmetic on
  the reference's rounded inputs wherever every faithfully rounded hypot gives one answer.  Both sides of each
  threshold are populated, and a guard without its factor max(1, |a|, |b|) decides some cases differently.
* The walk's acceptance on squares (accept_sure in the tree evaluation kernel): restated here with NumPy's
  IEEE operations, which are the kernel's (contraction is off there), it either returns the reference's answer
  (bound < theta dist and dist - bound >= clearance core, vortex_tree.py:526-531) or defers (-1) on 1.2 M
  adversarial cases placed at the FP64 neighbours of both thresholds.
"""
from decimal import Decimal

import numpy as np

from oracle import oracle as O
from tests import threshold_cases as TC


def _oracle_pairs(points, starts, dirs, lengths, radius, model, singularity="raise"):
    return O.pair_matrix(points, starts, dirs, lengths, radius, model, singularity)


def test_axis_guard_neighbours_in_the_oracle():
    s, d, L, P, info = TC.axis_cases()
    pairs, flags = _oracle_pairs(P, s, d, L, np.zeros(len(s)), 0)
    checked = 0
    for i, (k, h, amb, axis) in enumerate(info):
        if amb:
            continue
        u = pairs[i, k]
        t = float(np.dot((P[i] - s[k]) / L[k], d[k]))
        beside = 0.0 <= t <= 1.0
        if axis:
            assert np.all(u == 0.0), (i, h)
            assert flags[i] == (1 if beside else 0), (i, h)
        else:
            assert np.linalg.norm(u) > 0.0 and flags[i] == 0, (i, h)
        checked += 1
    assert checked == len(info) and sum(1 for x in info if x[3]) >= len(info) // 3
    # a planted mistake: the guard without its factor max(1, |a|, |b|) disagrees beyond the segment's ends
    wrong = sum(1 for i, (k, h, amb, axis) in enumerate(info)
                if not amb and (h <= Decimal(TC.EPS8)) != axis)
    assert wrong > 0


def test_lamb_cutoff_neighbours_in_the_oracle():
    s, d, L, R, P, info = TC.lamb_cases()
    pairs, flags = _oracle_pairs(P, s, d, L, R, 2)
    plain, _ = _oracle_pairs(P, s, d, L, R, 0)
    assert not flags.any()
    inside_n = 0
    for i, (k, ratio, amb, inside) in enumerate(info):
        if amb:
            continue
        u, u0 = pairs[i, k], plain[i, k]
        f = float(np.dot(u, u0) / np.dot(u0, u0))
        if inside:
            want = -np.expm1(-1.25643 * float(ratio) ** 2)
            assert abs(f - want) < 1e-12 and f < 1.0 - 4e-5, (i, f)
            inside_n += 1
        else:
            assert f == 1.0 or abs(f - 1.0) < 1e-15, (i, f)
    # both sides are populated, so an oracle that applied the factor everywhere (outside: f < 1 - 4e-5) or
    # nowhere (inside: f = 1) fails above
    assert 0 < inside_n < len(info)


def _accept_reference(r, bound, theta, clear):
    dist = np.sqrt(r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1] + r[:, 2] * r[:, 2])
    return (bound < theta * dist) & (dist - bound >= clear)


def _accept_sure(r, bound, theta, clear):
    d2 = r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1] + r[:, 2] * r[:, 2]
    b2 = (bound * bound) * (1.0 / (theta * theta))
    s = bound + clear
    c2 = s * s
    sane = (d2 >= 1e-200) & (d2 <= 1e200) & (b2 <= 1e200) & (c2 <= 1e200)
    yes = (d2 > b2 * (1.0 + 1e-9)) & (d2 > c2 * (1.0 + 1e-9))
    no = (d2 < b2 * (1.0 - 1e-9)) | (d2 < c2 * (1.0 - 1e-9))
    return np.where(~sane, -1, np.where(yes, 1, np.where(no, 0, -1)))


def test_square_root_free_acceptance_never_contradicts_the_reference():
    rng = np.random.default_rng(12)
    n = 200_000
    total_sure = 0
    for which in ("theta", "clear"):
        for scale in (1e-3, 1.0, 1e3):
            bound = rng.uniform(0.01, 10.0, n) * scale
            theta = rng.choice([0.08, 0.11, 0.2, 0.5, 0.9], n) * (1 + rng.uniform(-0.01, 0.01, n))
            clear = rng.uniform(0.0, 3.0, n) * bound * rng.choice([0.0, 0.1, 1.0, 5.0], n)
            # the distance at which the chosen condition flips: bound / theta, or bound + clear
            edge = bound / theta if which == "theta" else bound + clear
            # k ulps around the edge (k in -40..40), along a random direction
            k = rng.integers(-40, 41, n).astype(float)
            dist = edge * (1 + k * 2.0 ** -52)
            v = rng.normal(size=(n, 3))
            v /= np.linalg.norm(v, axis=1)[:, None]
            r = v * dist[:, None]
            ref = _accept_reference(r, bound, theta, clear)
            sure = _accept_sure(r, bound, theta, clear)
            decided = sure >= 0
            assert np.array_equal(sure[decided] == 1, ref[decided])
            total_sure += int(decided.sum())
    # far from the edges the square test decides everything
    d = np.array([[3.0, 4.0, 0.0]])
    assert _accept_sure(d, np.array([1.0]), 0.5, np.array([0.5]))[0] == 1
    assert _accept_sure(d, np.array([4.0]), 0.5, np.array([0.5]))[0] == 0
    assert total_sure > 0
