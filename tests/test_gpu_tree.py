"""Tree code on the GPU against the oracle.

Bars: the tree (order, lo, hi, left, right, count) bit-exact; node centre, bound and core bit-exact, the
moments within 1e-12 node-relative (the reference's GPU-moments bound, tests/test_cuda_moments.py:90-105);
the accepted-node and near-pair counts of every target identical to the oracle's (same traversal); FP64
velocities within 1e-10 * max|u| of the oracle tree (the reference's GPU self-test bound,
apps/studio/openfawe_studio/selftest.py:413-414); mixed precision within 0.1 x the tree's own error against
the direct sum, in max and rms (selftest.py:415-417).
"""
import numpy as np
import pytest

from oracle import oracle as O
from openfawe_amd import tree as T
from openfawe_amd import vortex as V
from openfawe_amd import workloads
from tests.helpers import midpoint_cases, prepare

pytestmark = pytest.mark.gpu


def tree_cases():
    cases = [(name, mid, leaf) for name, mid, leaf in midpoint_cases()]
    zeros = np.zeros((9, 3))
    zeros[:, 0] = [0.0, -0.0, 1.0, -0.0, 0.0, 2.0, -1.0, 0.0, -0.0]
    cases.append(("signed-zeros", zeros, 2))
    for rows, leaf in ((60, 32), (468, 32), (200, 1), (468, 8)):
        c = workloads.sheet(rows)
        cases.append((f"sheet-{rows}-leaf-{leaf}", O.segment_arrays(c["starts"], c["directions"], c["lengths"],
                                                                    c["gamma"])[3], leaf))
    loop = workloads.loop(rows=60)
    cases.append(("loop-60", O.segment_arrays(loop["starts"], loop["directions"], loop["lengths"], loop["gamma"])[3],
                  16))
    rng = np.random.default_rng(3)
    ties = np.round(rng.random((9000, 3)) * 4) / 4   # many equal coordinates across big (merged) nodes
    cases.append(("ties-9000", ties, 4))
    # big nodes that cannot split (5,000 equal midpoints) and leaves larger than an LDS subtree, each ending at
    # an odd and at an even depth: the segments of such nodes stay in the order buffer of their depth
    same = np.zeros((5000, 3))
    spread = np.column_stack((1 + rng.random(15000), 0.5 * rng.random(15000), 0.5 * rng.random(15000)))
    cases.append(("unsplittable-depth-1", np.concatenate((spread[:5000], same)), 16))
    cases.append(("unsplittable-depth-2", np.concatenate((spread, same)), 16))
    cases.append(("leaf-3000-depth-1", rng.random((5000, 3)), 3000))
    cases.append(("leaf-3000-depth-2", rng.random((9000, 3)), 3000))
    # runs of equal coordinates on every axis (at most about 250 long) through the levels that split the
    # presorted lists: each partition puts the runs back in the order of their new positions
    cases.append(("runs-20000", np.round(rng.random((20000, 3)) * 128) / 128, 8))
    cases.append(("runs-200000", np.round(rng.random((200000, 3)) * 1024) / 1024, 32))
    # a run longer than the lists re-order (every z equal): the build is made with the per-level sorts
    planar = rng.random((20000, 3))
    planar[:, 2] = 0.0
    cases.append(("planar-20000", planar, 32))
    # the 821,874-segment loop wake: 32 rotor segments share the midpoint z = 1.5
    loop = workloads.loop(rows=3744)
    cases.append(("loop-3744", O.segment_arrays(loop["starts"], loop["directions"], loop["lengths"],
                                                loop["gamma"])[3], 32))
    return cases


@pytest.mark.parametrize("case", tree_cases(), ids=lambda c: c[0])
def test_build_element_for_element(case):
    name, mid, leaf = case
    ref = O.build_tree(mid, leaf)
    got = T.build_tree(mid, leaf)
    for a, b, what in zip(ref[:5], got[:5], ("order", "lo", "hi", "left", "right")):
        assert np.array_equal(a, b), what
    assert ref[5] == got[5]


def segments(seed, count=400, zero_every=7):
    rng = np.random.default_rng(seed)
    starts = rng.random((count, 3)) * [20.0, 5.0, 2.0]
    d = rng.normal(size=(count, 3))
    directions = d / np.linalg.norm(d, axis=1)[:, None]
    lengths = 0.2 + 0.3 * rng.random(count)
    gamma = rng.normal(size=count)
    if zero_every:
        gamma[::zero_every] = 0.0
    radius = rng.uniform(0.0, 0.08, count)
    return starts, directions, lengths, gamma, radius


def node_relative(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.abs(a).reshape(len(a), -1).max(axis=1)
    diff = np.abs(a - b).reshape(len(a), -1).max(axis=1)
    return np.where(scale > 0, diff / np.where(scale > 0, scale, 1.0), diff)


@pytest.mark.parametrize("seed,leaf,count", [(1, 8, 300), (2, 32, 300), (3, 1, 120), (4, 16, 20000), (5, 1, 150000)])
def test_node_quantities(seed, leaf, count):
    s, d, L, g, r = segments(seed, count)
    lv, ends, alpha, mid = O.segment_arrays(s, d, L, g)
    order, lo, hi, _, _, _ = O.build_tree(mid, leaf)
    centre, bound, core, m0, m1, m2 = O.node_moments(order, lo, hi, s, ends, alpha, lv, r, g)
    w1, t2, w2 = O.node_contractions(m1, m2)
    got = T.tree_nodes(s, d, L, g, r, leaf)
    assert np.array_equal(got[0], centre)
    assert np.array_equal(got[1], bound)
    assert np.array_equal(got[2], core)
    for name, a, b in (("m0", m0, got[3]), ("m1", m1, got[4]), ("m2", m2, got[5])):
        assert node_relative(a, b).max() <= 1e-12, name
    for name, a, b, made_of in (("w1", w1, got[6], m1), ("t2", t2, got[7], m2), ("w2", w2, got[8], m2)):
        scale = np.abs(made_of).reshape(len(made_of), -1).max(axis=1)
        diff = np.abs(a - b).reshape(len(a), -1).max(axis=1)
        assert np.all(diff <= 1e-12 * scale), name


def run_tree(case, theta, order=2, leaf=32, clearance=3.0, precision="fp64", split_depth=-1, singularity="raise"):
    args = workloads.engine_args(case)
    n = len(case["points"])
    csr = V.exclusion_csr(n, case["pairs"])
    out, flagged = T.hip_tree_sum(*args, csr, singularity, V.TreeSettings(theta, order, leaf, clearance),
                                  V.HipSettings(precision), split_depth=split_depth)
    return out, flagged, dict(V.Statistics.last)


def oracle_tree(case, theta, order=2, leaf=32, clearance=3.0, singularity="raise"):
    args = workloads.engine_args(case)
    csr = O.exclusion_csr(len(case["points"]), case["pairs"])
    return O.tree_sum(*args, csr, singularity, theta, order, leaf, clearance)


def small_cases():
    nodes, edges, gamma = workloads.lattice()
    s, d, L = prepare(nodes, edges)
    rng = np.random.default_rng(2)
    lattice = dict(points=nodes + .07 * rng.normal(size=nodes.shape), starts=s, directions=d, lengths=L, gamma=gamma,
                   radius=np.full(len(s), .1), pairs=np.zeros((0, 2), np.int64), model=2, name="lattice")
    incident = np.asarray([(n, k) for k, (a, b) in enumerate(edges) for n in (a, b)], dtype=np.int64)
    lattice_nodes = dict(lattice, points=nodes, pairs=incident, name="lattice-nodes")
    return [lattice, lattice_nodes, workloads.sheet(60), workloads.loop(rows=40)]


@pytest.mark.parametrize("case", small_cases(), ids=lambda c: c["name"])
@pytest.mark.parametrize("theta,order,leaf", [(0.5, 2, 32), (0.2, 2, 32), (0.4, 1, 8), (0.6, 0, 16)])
def test_fp64_tree_matches_oracle(case, theta, order, leaf):
    ref, fref, far_ref, near_ref = oracle_tree(case, theta, order, leaf)
    out, flagged, stats = run_tree(case, theta, order, leaf)
    assert not flagged and not fref.any()
    assert np.array_equal(stats["far_counts"], far_ref)
    assert np.array_equal(stats["near_counts"], near_ref)
    scale = np.linalg.norm(ref, axis=1).max()
    assert np.abs(out - ref).max() <= 1e-10 * scale


def test_split_depth_only_reassociates():
    case = workloads.sheet(60)
    base, _, st0 = run_tree(case, 0.3, split_depth=0)
    for depth in (1, 3, 6, -1):
        out, _, st = run_tree(case, 0.3, split_depth=depth)
        assert np.array_equal(st["far_counts"], st0["far_counts"])
        scale = np.linalg.norm(base, axis=1).max()
        assert np.abs(out - base).max() <= 1e-13 * scale


def test_tiny_theta_equals_direct_and_excluded_never_leak():
    nodes, edges, gamma = workloads.lattice(rows=12, span=9)
    s, d, L = prepare(nodes, edges)
    incident = np.asarray([(n, k) for k, (a, b) in enumerate(edges) for n in (a, b)], dtype=np.int64)
    case = dict(points=nodes, starts=s, directions=d, lengths=L, gamma=gamma, radius=np.full(len(s), .1),
                pairs=incident, model=2)
    direct, _ = O.direct_sum(nodes, s, d, L, gamma, case["radius"], 2, O.exclusion_csr(len(nodes), incident))
    exact, _, st = run_tree(case, 1e-9, leaf=4)
    assert st["far_counts"].sum() == 0
    scale = np.linalg.norm(direct, axis=1).max()
    np.testing.assert_allclose(exact, direct, rtol=0, atol=1e-13 * scale)
    loud, silent = gamma.copy(), gamma.copy()
    loud[-1], silent[-1] = 1e6, 0.0
    far_pair = np.array([[0, len(edges) - 1]], dtype=np.int64)
    one = dict(case, points=nodes[:1], pairs=far_pair)
    a, _, _ = run_tree(dict(one, gamma=loud), 0.6, leaf=4, singularity="zero")
    b, _, _ = run_tree(dict(one, gamma=silent), 0.6, leaf=4, singularity="zero")
    assert np.array_equal(a, b)
    leaked, _, _ = run_tree(dict(one, gamma=loud, pairs=np.zeros((0, 2), np.int64)), 0.6, leaf=4, singularity="zero")
    assert np.linalg.norm(leaked - b) > 1e3 * np.linalg.norm(b)


def test_mixed_within_a_tenth_of_the_tree_error():
    """selftest.py:415-417 on the reference's sheet: mixed vs FP64 tree within 0.1 x tree-vs-direct error."""
    case = workloads.sheet(120)
    args = workloads.engine_args(case)
    direct, _ = O.direct_sum(*args, None, "raise")
    ref, _, _, _ = oracle_tree(case, 0.2)
    err_tree = np.linalg.norm(ref - direct, axis=1)
    out, flagged, _ = run_tree(case, 0.2, precision="mixed")
    assert not flagged
    err = np.linalg.norm(out - ref, axis=1)
    assert err.max() <= 0.1 * err_tree.max()
    assert np.sqrt(np.mean(err ** 2)) <= 0.1 * np.sqrt(np.mean(err_tree ** 2))


def test_headline_820k_mixed_tree_accuracy():
    """The 20,570-target x 821,874-segment loop wake in mixed precision at theta 0.11: relative L2 and max-abs
    (relative to max |u|) error at most 1e-5 against the FP64 direct sum at 2,048 seeded sample targets.  The
    FP64 direct kernel is pinned to the oracle by the A1 tests of test_gpu_direct.py."""
    from openfawe_amd import autotune as AT
    case = workloads.loop(rows=3744)
    n = len(case["points"])
    assert n == 20570 and len(case["starts"]) == 821874
    out, flagged, stats = run_tree(case, 0.11, precision="mixed", singularity="raise")
    assert not flagged
    idx = np.sort(np.random.default_rng(5).choice(n, size=2048, replace=False))
    ref, ref_flagged = V.hip_sum(case["points"][idx], case["starts"], case["directions"], case["lengths"],
                                 case["gamma"], case["radius"], case["model"], AT._sample_csr(case, idx), "raise",
                                 V.HipSettings("fp64", 256))
    assert not ref_flagged
    err = out[idx] - ref
    assert np.linalg.norm(err) <= 1e-5 * np.linalg.norm(ref)
    assert np.abs(err).max() <= 1e-5 * np.abs(ref).max()


def test_api_errors_flags_and_determinism():
    nodes, edges, gamma = workloads.lattice(rows=6, span=5)
    point = .5 * (nodes[edges[3, 0]] + nodes[edges[3, 1]])[None]
    settings = V.TreeSettings(theta=.4, order=2, leaf_size=2)
    with pytest.raises(V.FilamentSingularity):
        V.segments_velocity(point, nodes, edges, gamma, backend="hip_tree", tree=settings)
    for model in (3, 4, 5):
        with pytest.raises(ValueError):
            V.segments_velocity(point + 1., nodes, edges, gamma, model=model, core_radius=.1, backend="hip_tree",
                                tree=settings)
    ok = V.segments_velocity(point + 1., nodes, edges, gamma, model=3, core_radius=0.0, backend="hip_tree",
                             tree=settings)
    ideal = V.segments_velocity(point + 1., nodes, edges, gamma, backend="hip_tree", tree=settings)
    assert np.array_equal(ok, ideal)
    case = workloads.sheet(40)
    a, _, _ = run_tree(case, 0.5, precision="mixed")
    b, _, _ = run_tree(case, 0.5, precision="mixed")
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    empty = V.segments_velocity(np.empty((0, 3)), nodes, edges, gamma, backend="hip_tree", tree=settings)
    assert empty.shape == (0, 3)


def batch_snapshots(sizes, seed=11):
    """Snapshots of (targets, segments) from the loop generator at several phases; empty ones included."""
    rng = np.random.default_rng(seed)
    snaps = []
    for k, (n, m) in enumerate(sizes):
        rows = max(2, int(np.ceil(m / 220)))
        case = workloads.loop(rows=rows, phase=float(rng.uniform(0, 2 * np.pi)), wind=float(rng.uniform(6, 14)))
        take = slice(0, m)
        pts = case["points"][rng.choice(len(case["points"]), size=n, replace=n > len(case["points"]))]
        snaps.append(dict(points=pts + 0.01 * rng.normal(size=pts.shape) if n else np.zeros((0, 3)),
                          starts=case["starts"][take], directions=case["directions"][take],
                          lengths=case["lengths"][take], gamma=case["gamma"][take],
                          radius=case["radius"][take]))
    return snaps


def single_call(snap, settings, precision, split_depth):
    torch = V._torch()
    n, m = len(snap["points"]), len(snap["starts"])
    if n == 0:
        return np.zeros((0, 3)), np.zeros(0, np.int64), np.zeros(0, np.int64)
    dev = [torch.as_tensor(np.ascontiguousarray(snap[k]), dtype=torch.float64).cuda()
           for k in ("points", "starts", "directions", "lengths", "gamma", "radius")]
    ptr, seg = V.exclusion_csr(n, snap.get("pairs", np.zeros((0, 2), np.int64)))
    ex_ptr = torch.as_tensor(ptr).cuda()
    ex_seg = torch.as_tensor(seg if len(seg) else np.zeros(1, np.int64)).cuda()
    out, _, far, near = T.hip_tree_sum_device(*dev, 2, ex_ptr, ex_seg, "zero", settings, V.HipSettings(precision),
                                              split_depth=split_depth)
    return out.cpu().numpy(), far.cpu().numpy(), near.cpu().numpy()


@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_single_calls_bitwise(use_graph, precision):
    """Every snapshot of a batch (empty ones included) equals its own single call, bit for bit, on every
    replay, also after the inputs change in place between replays."""
    sizes = [(300, 1500), (0, 700), (40, 0), (900, 6000), (64, 65), (1, 3000)]
    snaps = batch_snapshots(sizes)
    settings = V.TreeSettings(theta=0.5, order=2, leaf_size=16, core_clearance=3.0)
    batch = T.TreeBatch(snaps, settings, model=2, precision=precision, split_depth=3, use_graph=use_graph,
                        counts=True)
    try:
        for replay in range(3):
            if replay == 2:
                batch.arrays["gamma"].mul_(-0.75)
                for s in snaps:
                    s["gamma"] = s["gamma"] * -0.75
            out = batch.run().cpu().numpy()
            far = batch.far_counts.cpu().numpy()
            near = batch.near_counts.cpu().numpy()
            for k, snap in enumerate(snaps):
                a, b = batch.target_offsets[k], batch.target_offsets[k + 1]
                ref, far_ref, near_ref = single_call(snap, settings, precision, 3)
                assert np.array_equal(out[a:b].view(np.uint64), ref.view(np.uint64)), (replay, k)
                if len(snap["starts"]):
                    assert np.array_equal(far[a:b], far_ref) and np.array_equal(near[a:b], near_ref), (replay, k)
                else:
                    assert not out[a:b].any()
        assert batch.stats()["split_depth"] == 3
    finally:
        batch.close()


def test_batch_with_exclusions_and_default_depth():
    """The batch workload generator (uneven loop snapshots, incident segments excluded) at a reduced size:
    graph replay equals the single calls bit for bit at the batch's own split depth."""
    snaps = workloads.batch(count=5, m_lo=3_000, m_hi=30_000, n_lo=150, n_hi=1_200)
    settings = V.TreeSettings(theta=0.5, order=2, leaf_size=32, core_clearance=3.0)
    batch = T.TreeBatch(snaps, settings, model=2, precision="mixed", use_graph=True, counts=True)
    try:
        out = batch.run().cpu().numpy()
        depth = batch.stats()["split_depth"]
        for k, snap in enumerate(snaps):
            a, b = batch.target_offsets[k], batch.target_offsets[k + 1]
            ref, _, _ = single_call(snap, settings, "mixed", depth)
            assert np.array_equal(out[a:b].view(np.uint64), ref.view(np.uint64)), k
    finally:
        batch.close()


def coincident_snapshot(n_same, n_spread, seed):
    """Segments of which n_same share one midpoint (an unsplittable node of more than an LDS subtree) beside
    n_spread scattered further along x, so that the equal ones fill a whole child of the top splits."""
    rng = np.random.default_rng(seed)
    m = n_same + n_spread
    starts = np.zeros((m, 3))
    starts[n_same:] = np.column_stack((1 + rng.random(n_spread), 0.5 * rng.random(n_spread),
                                       0.5 * rng.random(n_spread)))
    directions = np.zeros((m, 3))
    directions[:, 1] = 1.0
    return dict(points=rng.random((200, 3)) * [2.0, 1.0, 1.0] + [0.0, 0.0, 0.3], starts=starts,
                directions=directions, lengths=np.full(m, 0.05), gamma=rng.uniform(0.5, 1.5, m),
                radius=np.full(m, 0.01))


@pytest.mark.parametrize("use_graph", [False, True], ids=["eager", "graph"])
def test_batch_with_unsplittable_nodes(use_graph):
    """A forest whose trees hold unsplittable nodes of 5,000 segments at depth 1 and at depth 2 (their segments
    stay in the order buffer of that depth until the final copy): every snapshot equals its single call."""
    snaps = [coincident_snapshot(5000, 5000, 1), batch_snapshots([(300, 6000)])[0], coincident_snapshot(5000, 15000, 2)]
    settings = V.TreeSettings(theta=0.5, order=2, leaf_size=32, core_clearance=3.0)
    batch = T.TreeBatch(snaps, settings, model=2, precision="mixed", split_depth=3, use_graph=use_graph)
    try:
        for replay in range(2):
            out = batch.run().cpu().numpy()
            for k, snap in enumerate(snaps):
                a, b = batch.target_offsets[k], batch.target_offsets[k + 1]
                ref, _, _ = single_call(snap, settings, "mixed", 3)
                assert np.array_equal(out[a:b].view(np.uint64), ref.view(np.uint64)), (replay, k)
                assert np.isfinite(out[a:b]).all()
    finally:
        batch.close()


def test_exclusion_capacity_grows_between_calls():
    """Calls after the first assume the exclusion capacity kept from earlier calls and check it on the device: a
    call with more exclusions than that reports the count, runs again with it, and matches the oracle; the same
    call again runs once and returns the same bits."""
    case = dict(workloads.sheet(60))
    n, m = len(case["points"]), len(case["starts"])
    rng = np.random.default_rng(11)
    per = 2_000_000 // n + 1  # more entries than any earlier call of the process kept room for
    seg = (rng.integers(0, m, size=(n, 1)) + np.arange(per)[None, :]) % m
    case["pairs"] = np.column_stack((np.repeat(np.arange(n), per), seg.ravel())).astype(np.int64)
    run_tree(workloads.sheet(8), 0.3)  # a small call first: the context's capacity is set
    out, flagged, st = run_tree(case, 0.3)
    assert st["attempts"] >= 2
    ref, fref, far_ref, near_ref = oracle_tree(case, 0.3)
    assert not flagged and not fref.any()
    assert np.array_equal(st["far_counts"], far_ref)
    assert np.array_equal(st["near_counts"], near_ref)
    scale = np.linalg.norm(ref, axis=1).max()
    assert np.abs(out - ref).max() <= 1e-10 * scale
    again, _, st_again = run_tree(case, 0.3)
    assert st_again["attempts"] == 1
    assert np.array_equal(out.view(np.uint64), again.view(np.uint64))


def test_list_slots_grow_between_calls():
    """A call whose near-field work items exceed the slots kept from the previous, smaller call reports the
    overflow on the device, runs again with its own count, and returns what its first call returned."""
    small, big = workloads.sheet(8), workloads.sheet(60)
    first, _, st_first = run_tree(big, 0.3)
    run_tree(small, 0.3)
    again, _, st_again = run_tree(big, 0.3)
    assert st_again["attempts"] == 2 and st_again["work_items"] == st_first["work_items"]
    assert np.array_equal(first.view(np.uint64), again.view(np.uint64))
    assert np.array_equal(st_first["near_counts"], st_again["near_counts"])
