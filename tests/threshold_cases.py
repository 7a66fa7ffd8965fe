"""Targets at the FP64 neighbours of the pair kernel's two jump thresholds (vortex_numba.py:28-29, 59-62, 95-97).

* The axis guard h <= 8 eps max(1, |a|, |a - d|): a target within h of a segment's line gets no velocity (and is
  flagged under 'raise' when it lies beside the segment), just outside it gets ~1/h.
* The Lamb-Oseen cutoff h / rc <= sqrt(10 / 1.25643): inside, the factor is -expm1(-1.25643 (h/rc)^2), which
  is 1 - 4.5e-5 at the cutoff; outside it is 1.

The reference norms are nested hypot calls.  glibc's and the device library's hypot are faithfully rounded
(either neighbour of the exact value), not correctly rounded, so a target whose exact norm lies strictly between
the threshold T and the next double above it can be decided either way by a conforming hypot.  Every case
records its exact norm (decimal arithmetic on the rounded inputs) and whether it is ambiguous in that sense; the
other cases have one decision for every faithful hypot.

Near the axis guard only axis-aligned segments reach the neighbours of T: for a direction with several nonzero
components the cross product d x a moves by about 10 % of h per ulp of the target, while along a coordinate axis
it is (0, -z, y) exactly, so h = hypot(y, z) can be placed ulp by ulp.  The Lamb-Oseen cases use generic
directions, stepping the target one ulp at a time across the cutoff.
"""
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np

EPS8 = 8 * float(np.finfo(np.float64).eps)
LAMB_CUT = float(np.sqrt(10 / 1.25643))
LAMB_Q = 10 / 1.25643
getcontext().prec = 90


def _exact_norm(*xs):
    return sum((Decimal(float(x)) ** 2 for x in xs), Decimal(0)).sqrt()


def _ambiguous(exact, threshold):
    """True when exact lies strictly between the threshold and the next double above it."""
    t = Decimal(float(threshold))
    return t < exact < Decimal(float(np.nextafter(threshold, np.inf)))


def reference_geometry(p, s, d, length):
    """a = (p - s) / L and c = d x a with the reference's separately rounded operations (vortex_numba.py:46-54)."""
    a = (np.asarray(p, float) - s) / length
    c = np.array([d[1] * a[2] - d[2] * a[1], d[2] * a[0] - d[0] * a[2], d[0] * a[1] - d[1] * a[0]])
    return a, c


def axis_cases(seed=0):
    """Segments along +-x, +-y, +-z; targets at hypot(y, z) = 8 eps (1 + j 2^-52) beside each segment and beyond
    its ends.  Returns (starts, directions, lengths, points, info); info[i] = (segment, exact h / threshold side,
    ambiguous, expected axis decision)."""
    rng = np.random.default_rng(seed)
    axes = [np.eye(3)[k] * sgn for k in range(3) for sgn in (1.0, -1.0)]
    starts, dirs, lengths, points, info = [], [], [], [], []
    for k, d in enumerate(axes):
        length = 2.0 ** int(rng.integers(-2, 3))
        s = np.round(rng.uniform(-4, 4, 3) * 64) / 64        # few bits: p - s is exact
        starts.append(s)
        dirs.append(d)
        lengths.append(length)
        ax = int(np.flatnonzero(d)[0])
        others = [q for q in range(3) if q != ax]
        for t in (0.25, 0.5, 0.75, 1.5, -0.5):
            scale = max(1.0, abs(t), abs(t - 1.0))
            threshold = EPS8 * scale                     # exact: scale has a few bits only
            for j in range(-6, 7):
                phi = rng.uniform(0, 2 * np.pi)
                rho = threshold * (1 + j * 2.0 ** -52)
                y, z = rho * np.cos(phi), rho * np.sin(phi)
                a = np.zeros(3)
                a[ax] = t * d[ax]
                a[others[0]], a[others[1]] = y, z
                p = s + length * a
                ra, c = reference_geometry(p, s, d, length)
                exact_h = _exact_norm(*c)
                na = _exact_norm(*ra)
                nb = _exact_norm(*(ra - d))
                # the threshold's own norms are t and |t - 1| to far below an ulp here
                assert abs(float(max(Decimal(1), na, nb)) - scale) == 0.0
                points.append(p)
                info.append((k, exact_h, _ambiguous(exact_h, threshold), exact_h <= Decimal(threshold)))
    return (np.array(starts), np.array(dirs), np.array(lengths), np.array(points), info)


def lamb_cases(seed=1, segments=6, steps=40):
    """Generic directions; targets stepped one ulp at a time (in one coordinate) across h / rc = sqrt(10/1.25643)
    beside each segment.  info[i] = (segment, exact ratio, ambiguous, expected inside)."""
    rng = np.random.default_rng(seed)
    starts, dirs, lengths, radii, points, info = [], [], [], [], [], []
    for k in range(segments):
        s = rng.uniform(-3, 3, 3)
        e = s + rng.normal(size=3) * rng.uniform(0.5, 3.0)
        delta = e - s
        length = np.hypot(np.hypot(delta[0], delta[1]), delta[2])
        d = delta / length
        radius = rng.uniform(0.02, 0.3) * length
        rc = radius / length
        n = np.cross(d, rng.normal(size=3))
        n /= np.linalg.norm(n)
        base = s + length * (rng.uniform(0.2, 0.8) * d + LAMB_CUT * rc * n)
        q = int(np.argmax(np.abs(n)))
        starts.append(s)
        dirs.append(d)
        lengths.append(length)
        radii.append(radius)
        x = base[q]
        lo = x
        for _ in range(steps // 2):
            lo = np.nextafter(lo, -np.inf * np.sign(n[q]))
        cur = lo
        for _ in range(steps):
            p = base.copy()
            p[q] = cur
            _, c = reference_geometry(p, s, d, length)
            h = _exact_norm(*c)
            ratio = h / Decimal(float(rc))
            # the reference decides fl(h_ref / rc) <= cut with h_ref from two nested faithful hypot calls, within
            # two ulps of h; the decision is monotone in h_ref, so the case is unambiguous when both ends of
            # h (1 -+ 2^-51) give the same answer
            hi = float(h * (1 + Decimal(2) ** -51)) / rc
            lo_r = float(h * (1 - Decimal(2) ** -51)) / rc
            amb = (lo_r <= LAMB_CUT) != (hi <= LAMB_CUT)
            points.append(p)
            info.append((k, ratio, amb, float(h) / rc <= LAMB_CUT))
            cur = np.nextafter(cur, np.inf * np.sign(n[q]))
    return (np.array(starts), np.array(dirs), np.array(lengths), np.array(radii), np.array(points), info)


# ---------------------------------------------------------------- a squares-based decision, emulated exactly
def _fl(x):
    return float(x)


def _fma(a, b, c):
    return _fl(Fraction(a) * Fraction(b) + Fraction(c))


def squares_axis(c, mx=1.0):
    """h2 <= (8 eps)^2 mx^2 with h2 = fma(cz, cz, fma(cy, cy, cx * cx)) (a decision on squares, no hypot)."""
    h2 = _fma(c[2], c[2], _fma(c[1], c[1], _fl(Fraction(c[0]) * Fraction(c[0]))))
    thr = _fl(Fraction(_fl(Fraction(EPS8) * Fraction(EPS8))) * Fraction(_fl(Fraction(mx) * Fraction(mx))))
    return h2 <= thr


def squares_lamb(c, rc):
    """q = h2 * (1 / rc^2) <= 10 / 1.25643 (a decision on squares)."""
    h2 = _fma(c[2], c[2], _fma(c[1], c[1], _fl(Fraction(c[0]) * Fraction(c[0]))))
    irc2 = _fl(1 / Fraction(_fl(Fraction(rc) * Fraction(rc))))
    return _fl(Fraction(h2) * Fraction(irc2)) <= LAMB_Q
