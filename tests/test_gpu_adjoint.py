"""Pair matrix, ordered column assembly and the adjoint kernel on the GPU.

Bars: pair matrix within 1e-13 relative per pair of the oracle's pair_matrix and the same flags; column
assembly bit for bit the oracle's on the same pairs and bit for bit the GPU direct sum of each column's
signed segments (the reference's bitwise check, tests/test_vsm_wake_batched.py:50-70); the adjoint's
dL/dGamma equal to pair_matrix^T dL/dV (linearity) to 1e-12; dL/d(end points) and dL/d(targets) against
central finite differences (h = 1e-6) of the forward, rtol 2e-6, for every core model (the Rosenhead-Moore core
included); the mixed-precision adjoint (far pairs in FP32) within 1e-5 relative L2 of the FP64 adjoint; the VSM
Newton Jacobian built from the GPU influence operator against central differences of the closure (rtol 2e-7,
atol 2e-9, the reference's tests/test_vsm_jacobian.py:67-88); the differentiation case at its full 207k-segment
wake against the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from openfawe_amd import adjoint as AD
from openfawe_amd import vortex as V
from openfawe_amd import workloads
from tests.helpers import prepare

pytestmark = pytest.mark.gpu


def network(seed=3):
    nodes, edges, gamma, radius, points = workloads.random_network(seed)
    s, d, L = prepare(nodes, edges)
    return nodes, edges, gamma, radius, points, s, d, L


@pytest.mark.parametrize("model", range(6))
def test_pair_matrix_matches_oracle(model):
    _, _, _, radius, points, s, d, L = network()
    ref, fref = O.pair_matrix(points, s, d, L, radius, model, "zero")
    got, flags = AD.pair_matrix(points, s, d, L, radius, model, "zero")
    assert np.array_equal(flags, fref)
    scale = np.linalg.norm(ref, axis=2, keepdims=True)
    assert np.all(np.abs(got - ref) <= 1e-13 * scale + 1e-300)
    _, fr = O.pair_matrix(points, s, d, L, radius, model, "raise")
    _, fg = AD.pair_matrix(points, s, d, L, radius, model, "raise")
    assert np.array_equal(fg, fr)


def ring_layout(rows, cols):
    """Panels of a lattice as rings of 4 signed edges over shared segments (as vsm_wake._ring_layout)."""
    nodes = np.array([[i * 0.4, j * 1.0, 0.05 * np.sin(j)] for i in range(rows + 1) for j in range(cols + 1)])
    index = {}
    segments = []

    def edge(a, b):
        key = (min(a, b), max(a, b))
        if key not in index:
            index[key] = len(segments)
            segments.append(key)
        return index[key], (1.0 if (a, b) == key else -1.0)
    ptr, seg, sign = [0], [], []
    for i in range(rows):
        for j in range(cols):
            a, b = i * (cols + 1) + j, i * (cols + 1) + j + 1
            c, dd = (i + 1) * (cols + 1) + j + 1, (i + 1) * (cols + 1) + j
            for p, q in ((a, b), (b, c), (c, dd), (dd, a)):
                k, sg = edge(p, q)
                seg.append(k)
                sign.append(sg)
            ptr.append(len(seg))
    return nodes, np.array(segments), np.array(ptr), np.array(seg), np.array(sign)


def test_columns_bitwise():
    nodes, edges, ptr, seg, sign = ring_layout(4, 25)
    s, d, L = prepare(nodes, edges)
    radius = np.full(len(s), 0.02)
    rng = np.random.default_rng(1)
    points = rng.random((100, 3)) * [1.6, 25, 0.5] + [0, 0, 0.3]
    pairs, _ = O.pair_matrix(points, s, d, L, radius, 2, "raise")
    ref = O.assemble_columns(pairs, ptr, seg, sign)
    got = AD.assemble_columns(pairs, ptr, seg, sign).cpu().numpy()
    assert np.array_equal(ref.view(np.uint64), got.view(np.uint64))
    gp, _ = AD.pair_matrix(points, s, d, L, radius, 2)
    matrix = AD.assemble_columns(gp, ptr, seg, sign).cpu().numpy()
    for j in range(0, len(ptr) - 1, 7):   # each column is the direct sum of its signed segments
        k = slice(ptr[j], ptr[j + 1])
        col, _ = V.hip_sum(points, s[seg[k]], d[seg[k]], L[seg[k]], sign[k], radius[seg[k]], 2,
                           V.exclusion_csr(len(points), np.zeros((0, 2), np.int64)), "raise", V.HipSettings("fp64"))
        assert np.array_equal(col.view(np.uint64), matrix[:, j].view(np.uint64))


def forward(points, a, b, gamma, radius, model):
    import torch
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).cuda()  # noqa: E731
    d, L = AD.segment_geometry(t(a), t(b))
    out, _ = V.hip_sum(points, a, d.cpu().numpy(), L.cpu().numpy(), gamma, radius, model,
                       V.exclusion_csr(len(points), np.zeros((0, 2), np.int64)), "zero", V.HipSettings("fp64"))
    return out


@pytest.mark.parametrize("model", range(6))
def test_vjp_against_linearity_and_finite_differences(model):
    nodes, edges, gamma, radius, points, s, d, L = network(4)
    points = points[:40]
    a = nodes[edges[:, 0]]
    b = nodes[edges[:, 1]]
    rng = np.random.default_rng(7)
    g = rng.normal(size=(len(points), 3))
    gg, ga, gb, gx, out = (v.cpu().numpy() if v is not None else None
                           for v in AD.direct_vjp(points, a, b, gamma, radius, model, g, fused_forward=True))
    # fused forward equals the direct sum
    ref, _, scale = O.direct_sum(points, s, d, L, gamma, radius, model, None, "zero", with_scale=True)
    assert np.all(np.linalg.norm(out - ref, axis=1) <= 1e-12 * scale)
    # dL/dGamma = pair_matrix^T g (the velocity is linear in Gamma)
    pm, _ = O.pair_matrix(points, s, d, L, radius, model, "zero")
    lin = np.einsum("psc,pc->s", pm, g)
    bound = 1e-12 * np.einsum("psc->s", np.abs(pm) * np.abs(g)[:, None, :]) + 1e-300
    assert np.all(np.abs(gg - lin) <= bound)
    # central differences of L = sum(g * V) in a sample of coordinates
    h = 1e-6
    for which, grad, base in (("a", ga, a), ("b", gb, b), ("x", gx, points)):
        for k in rng.choice(len(base), 6, replace=False):
            for c in range(3):
                plus, minus = base.copy(), base.copy()
                plus[k, c] += h
                minus[k, c] -= h
                args = dict(a=a, b=b, x=points)
                args[which] = plus
                vp = forward(args["x"], args["a"], args["b"], gamma, radius, model)
                args[which] = minus
                vm = forward(args["x"], args["a"], args["b"], gamma, radius, model)
                fd = np.sum(g * (vp - vm)) / (2 * h)
                assert abs(fd - grad[k, c]) <= 2e-6 * max(abs(fd), abs(grad[k, c])) + 1e-7, (which, k, c, fd,
                                                                                               grad[k, c])


def test_autograd_function():
    import torch
    nodes, edges, gamma, radius, points, s, d, L = network(5)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).cuda()  # noqa: E731
    x = t(points[:30]).requires_grad_(True)
    a = t(nodes[edges[:, 0]]).requires_grad_(True)
    b = t(nodes[edges[:, 1]]).requires_grad_(True)
    g = t(gamma).requires_grad_(True)
    v = AD.induced_velocity(x, a, b, g, t(radius), model=2, singularity="zero")
    ref, _ = O.direct_sum(points[:30], s, d, L, gamma, radius, 2, None, "zero")
    assert np.abs(v.detach().cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    w = t(np.random.default_rng(0).normal(size=(30, 3)))
    (v * w).sum().backward()
    gg, ga, gb, gx, _ = AD.direct_vjp(x.detach(), a.detach(), b.detach(), g.detach(), t(radius), 2, w)
    for auto, direct in ((g.grad, gg), (a.grad, ga), (b.grad, gb), (x.grad, gx)):
        assert torch.equal(auto, direct)


def test_vsm_closure_jacobian():
    """The reference's Newton Jacobian (vsm.py:333-349) from the GPU influence operator, against central
    differences of the closure (vsm.py:320-330) with h = 1e-6 (tests/test_vsm_jacobian.py:67-88)."""
    nodes, edges, ptr, seg, sign = ring_layout(1, 20)
    s, d, L = prepare(nodes, edges)
    radius = np.full(len(s), 0.05)
    n = len(ptr) - 1
    cp = np.array([[0.3, j + 0.5, 0.05 * np.sin(j + 0.5)] for j in range(n)])   # 3/4-chord points
    pairs, flags = AD.pair_matrix(cp, s, d, L, radius, 2)
    assert not flags.any()
    # the closure and its Jacobian are evaluated in extended precision from the FP64 operator, so that the
    # comparison measures the Jacobian's formula and not the rounding noise of the differences (|closure| / h)
    ld = np.longdouble
    matrix = AD.assemble_columns(pairs, ptr, seg, sign).cpu().numpy().astype(ld)
    chord = np.full(n, 0.4, dtype=ld)
    tangent = np.tile(np.array([1.0, 0.0, 0.0], dtype=ld), (n, 1))
    normal = np.tile(np.array([0.0, 0.0, 1.0], dtype=ld), (n, 1))
    span = np.tile(np.array([0.0, 1.0, 0.0], dtype=ld), (n, 1))
    ambient = np.tile(np.array([20.0, 0.0, 1.5], dtype=ld), (n, 1))
    known = np.zeros((n, 3), dtype=ld)
    cross = np.linalg.norm(np.cross(ambient, span), axis=1)
    slope = 2 * np.pi

    def closure(gamma):
        v = ambient + np.einsum("ijk,j->ik", matrix, gamma) + known
        vt, vn = np.sum(v * tangent, 1), np.sum(v * normal, 1)
        alpha = np.arctan2(vn, vt)
        speed2 = np.linalg.norm(np.cross(v, span), axis=1) ** 2
        return .5 * chord * (slope * alpha) * speed2 / cross - gamma

    def jacobian(gamma):
        v = ambient + np.einsum("ijk,j->ik", matrix, gamma) + known
        vt, vn = np.sum(v * tangent, 1), np.sum(v * normal, 1)
        alpha = np.arctan2(vn, vt)
        cl = slope * alpha
        speed2 = np.linalg.norm(np.cross(v, span), axis=1) ** 2
        an = np.einsum("ijk,ik->ij", matrix, normal)
        at = np.einsum("ijk,ik->ij", matrix, tangent)
        da = (vt[:, None] * an - vn[:, None] * at) / (vt ** 2 + vn ** 2)[:, None]
        dh = 2 * np.einsum("ik,ijk->ij", np.cross(v, span), np.cross(matrix, span[:, None, :]))
        return (.5 * chord / cross)[:, None] * (slope * speed2[:, None] * da + cl[:, None] * dh) - np.eye(n)

    gamma = np.linspace(1.0, 3.0, n).astype(ld)
    J = jacobian(gamma)
    h = ld(1e-6)
    fd = np.stack([(closure(gamma + h * e) - closure(gamma - h * e)) / (2 * h) for e in np.eye(n, dtype=ld)], axis=1)
    np.testing.assert_allclose(J.astype(np.float64), fd.astype(np.float64), rtol=2e-7, atol=2e-9)


def test_vsm_wing_in_wake_influence_and_adjoint():
    """The differentiation case at a reduced wake: the 100 x 100 ring influence Jacobian equals the direct sums
    of each panel's signed ring (bitwise), and the adjoint's dL/dGamma over the wake equals the unit-circulation
    pair velocities contracted with dL/dV (linearity)."""
    case = workloads.vsm_wing(wake_rows=12)
    pts = case["points"]
    pairs, flags = AD.pair_matrix(pts, case["wing_starts"], case["wing_directions"], case["wing_lengths"],
                                  case["wing_radius"], 2)
    assert not flags.any()
    jac = AD.assemble_columns(pairs, case["column_ptr"], case["column_segment"], case["column_sign"]).cpu().numpy()
    assert jac.shape == (100, 100, 3)
    ptr, seg, sign = case["column_ptr"], case["column_segment"], case["column_sign"]
    for j in (0, 37, 99):
        k = slice(ptr[j], ptr[j + 1])
        col, _ = V.hip_sum(pts, case["wing_starts"][seg[k]], case["wing_directions"][seg[k]],
                           case["wing_lengths"][seg[k]], sign[k], case["wing_radius"][seg[k]], 2,
                           V.exclusion_csr(len(pts), np.zeros((0, 2), np.int64)), "raise", V.HipSettings("fp64"))
        assert np.array_equal(col, jac[:, j])
    s, d, L, g, r = (case[k] for k in ("starts", "directions", "lengths", "gamma", "radius"))
    ends = s + d * L[:, None]
    grad = np.random.default_rng(3).normal(size=pts.shape)
    gg, ga, gb, gx, out = AD.direct_vjp(pts, s, ends, g, r, 2, grad, fused_forward=True)
    ref, _, scale = O.direct_sum(pts, s, d, L, g, r, 2, None, "raise", with_scale=True)
    assert np.all(np.linalg.norm(out.cpu().numpy() - ref, axis=1) <= 1e-12 * scale)
    pm, _ = O.pair_matrix(pts, s, d, L, r, 2, "raise")
    lin = np.einsum("psc,pc->s", pm, grad)
    bound = 1e-12 * np.einsum("psc->s", np.abs(pm) * np.abs(grad)[:, None, :]) + 1e-300
    assert np.all(np.abs(gg.cpu().numpy() - lin) <= bound)


@pytest.mark.parametrize("model", [0, 2, 5])
def test_mixed_adjoint_against_fp64(model):
    """Far pairs in FP32: every gradient within 1e-5 relative L2 of the FP64 adjoint (and the fused forward within
    the mixed direct sum's bound of the FP64 oracle), on a wake where most pairs are far."""
    case = workloads.loop(rows=60, target_rows=4, seed=2)
    pts = case["points"][::3]
    s, d, L = case["starts"], case["directions"], case["lengths"]
    ends = s + d * L[:, None]
    g, r = case["gamma"], case["radius"]
    grad = np.random.default_rng(5).normal(size=pts.shape)
    ptr, seg = V.exclusion_csr(len(case["points"]), case["pairs"])
    keep = np.arange(0, len(case["points"]), 3)
    pairs = case["pairs"][np.isin(case["pairs"][:, 0], keep)].copy()
    pairs[:, 0] //= 3
    import torch
    ptr, seg = V.exclusion_csr(len(pts), pairs)
    ptr_d = torch.from_numpy(ptr).cuda()
    seg_d = torch.from_numpy(seg if len(seg) else np.zeros(1, np.int64)).cuda()
    res = {}
    for precision in ("fp64", "mixed"):
        res[precision] = [v.cpu().numpy() for v in AD.direct_vjp(pts, s, ends, g, r, model, grad, ptr_d, seg_d,
                                                                  fused_forward=True, precision=precision)]
    for k, name in enumerate(("gamma", "starts", "ends", "points", "forward")):
        a, b = res["mixed"][k], res["fp64"][k]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel <= 1e-5, (name, rel)
        assert rel > 0.0 or name == "gamma"
    ref, _, _ = O.direct_sum(pts, s, d, L, g, r, model, (ptr, seg), "raise", with_scale=True)
    out = res["mixed"][4]
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) <= 1e-6


def test_rosenhead_moore_vjp_inside_the_core():
    """One segment with a core of 0.3 lengths, targets inside the core beside the segment and close to its line
    beyond an end, where the regularised norms dominate: gradients against central differences (h = 1e-6)."""
    a = np.array([[0.0, 0.0, 0.0]])
    b = np.array([[1.0, 0.0, 0.0]])
    pts = np.array([[0.5, 0.2, 0.1], [1.5, 0.05, 0.02], [-0.1, 0.01, -0.03]])
    g = np.random.default_rng(1).normal(size=pts.shape)
    gg, ga, gb, gx, out = (v.cpu().numpy() for v in AD.direct_vjp(pts, a, b, np.array([1.3]), np.array([0.3]), 5, g,
                                                                  fused_forward=True))
    h = 1e-6
    for base_name, grad in (("b", gb), ("x", gx)):
        for k in range(len(grad)):
            for c in range(3):
                args = dict(a=a.copy(), b=b.copy(), x=pts.copy())
                args[base_name][k, c] += h
                vp = forward(args["x"], args["a"], args["b"], np.array([1.3]), np.array([0.3]), 5)
                args[base_name][k, c] -= 2 * h
                vm = forward(args["x"], args["a"], args["b"], np.array([1.3]), np.array([0.3]), 5)
                fd = np.sum(g * (vp - vm)) / (2 * h)
                assert abs(fd - grad[k, c]) <= 2e-6 * max(abs(fd), abs(grad[k, c])) + 1e-9, (base_name, k, c)


def test_differentiation_case_full_wake():
    """Config 5 at its own size: a 100-panel VSM wing in a 207k-segment wake.  The fused forward within A1 of the
    oracle's direct sum, dL/dGamma equal to the oracle's unit-circulation pair velocities contracted with dL/dV
    (seeded subset of 4,000 segments, 1e-12), end-point gradients against central differences (h = 1e-6) of
    the GPU forward at sampled segments, and the mixed adjoint within 1e-5 of the FP64 one."""
    case = workloads.vsm_wing()
    pts, s, d, L, g, r = (case[k] for k in ("points", "starts", "directions", "lengths", "gamma", "radius"))
    assert len(pts) == 100 and len(s) > 200_000
    ends = s + d * L[:, None]
    grad = np.random.default_rng(11).normal(size=pts.shape)
    gg, ga, gb, gx, out = (v.cpu().numpy() for v in AD.direct_vjp(pts, s, ends, g, r, 2, grad, fused_forward=True))
    ref, flags, scale = O.direct_sum(pts, s, d, L, g, r, 2, None, "raise", with_scale=True)
    assert not flags.any()
    assert np.all(np.linalg.norm(out - ref, axis=1) <= 1e-12 * scale)
    rng = np.random.default_rng(12)
    sub = np.sort(rng.choice(len(s), 4000, replace=False))
    pm, _ = O.pair_matrix(pts, s[sub], d[sub], L[sub], r[sub], 2, "raise")
    lin = np.einsum("psc,pc->s", pm, grad)
    bound = 1e-12 * np.einsum("psc->s", np.abs(pm) * np.abs(grad)[:, None, :]) + 1e-300
    assert np.all(np.abs(gg[sub] - lin) <= bound)
    # the segments nearest the collocation points carry the largest end-point gradients: test those and a few
    near = np.argsort(-np.linalg.norm(ga, axis=1))[:4]
    h = 1e-6
    for k in np.concatenate([near, rng.choice(len(s), 2, replace=False)]):
        for which, grad_k in (("a", ga[k]), ("b", gb[k])):
            for c in range(3):
                sa, sb = s.copy(), ends.copy()
                arr = sa if which == "a" else sb
                arr[k, c] += h
                vp = forward(pts, sa, sb, g, r, 2)
                arr[k, c] -= 2 * h
                vm = forward(pts, sa, sb, g, r, 2)
                fd = np.sum(grad * (vp - vm)) / (2 * h)
                assert abs(fd - grad_k[c]) <= 2e-5 * max(abs(fd), abs(grad_k[c])) + 1e-9, (which, k, c, fd,
                                                                                            grad_k[c])
    mixed = [v.cpu().numpy() for v in AD.direct_vjp(pts, s, ends, g, r, 2, grad, fused_forward=True,
                                                     precision="mixed")]
    for a_, b_ in zip(mixed, (gg, ga, gb, gx, out)):
        assert np.linalg.norm(a_ - b_) / np.linalg.norm(b_) <= 1e-5
