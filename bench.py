"""Benchmark of the free-wake induced-velocity path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

A step is one induced-velocity call over one synthetic wake whose inputs are already resident in HBM:
targets (n,3), segment starts / directions / lengths / circulations / core radii and the exclusion CSR.
For the tree code a call is the whole pipeline (tree build, node moments, traversal).  The timed region is
K calls bracketed by a device synchronisation on both sides.  One JSON line is printed.

Workloads (--workload):
  tree-mixed-820k   (default) mixed-precision tree, 20,570 targets x 821,874 segments: crosswind-loop wake
                    of 3,744 rows over 8 loops + 8 rotor helices, theta 0.11 (relative L2 error against the
                    FP64 direct sum below 1e-5), order 2, leaf 32, Lamb-Oseen cores
  tree-mixed-820k-theta0.5  the same call at theta 0.5
  tree-mixed-410k   the same generator with 1,872 rows
  tree-fp64-102k    FP64 tree, 468 rows (one loop)
  sheet-mixed-820k  the reference's synthetic sheet (3,744 rows, 819,826 segments), theta 0.2
  direct-fp64-102k  FP64 direct sum over the reference's 468-row sheet (20,570 x 102,382)
  direct-mixed-102k mixed-precision direct sum, same wake
The headline metric is pair-interactions/s = n * m / (time per call), with ms per call.

Beside the headline workload the default line carries (skip them with --no-extra):
  legs          tree-mixed-410k (the metric's other size) and tree-fp64-102k (the FP64 tree of 20,570 x 104,430)
  sheet_table   the reference's speed table (README.md:83-91): the synthetic sheet at 468 / 936 / 1,872 / 3,744 rows,
                theta 0.2, FP64 and mixed, resident and as the drop-in call on host arrays, the oracle tree on the
                host's cores, and the accuracy against the CPU tree
  drop_in_host  the headline call on host NumPy arrays (upload / evaluation / download split)
  adjoint, batch, network, convection: the differentiation case, the 64-snapshot batch, the wake bookkeeping and the
                geometry step
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector (published, packed dual issue; the kernels issue unpacked FP32)
HBM_PEAK_GBS = 8000.0
# algorithmic ops per pair by core model (the operations of the reference's pair formula, vortex_numba.py:42-109,
# each add, multiply, divide or square root counted once, outside the Lamb-Oseen cutoff) and per accepted order-2
# node (far_velocity, vortex_tree.py:450-497, plus the acceptance test)
OPS_PER_PAIR = {0: 67, 1: 71, 2: 69, 3: 75, 4: 70, 5: 86}
OPS_PER_FAR = 178 + 4

WORKLOADS = {
    # theta 0.11 keeps the relative L2 error against the FP64 direct sum below 1e-5 on this wake
    "tree-mixed-820k": dict(kind="tree", gen="loop", rows=3744, precision="mixed", theta=0.11),
    "tree-mixed-820k-theta0.5": dict(kind="tree", gen="loop", rows=3744, precision="mixed", theta=0.5),
    "tree-mixed-410k": dict(kind="tree", gen="loop", rows=1872, precision="mixed", theta=0.11),
    "tree-fp64-102k": dict(kind="tree", gen="loop", rows=468, precision="fp64", theta=0.5),
    "tree-fp64-820k": dict(kind="tree", gen="loop", rows=3744, precision="fp64", theta=0.5),
    "sheet-mixed-820k": dict(kind="tree", gen="sheet", rows=3744, precision="mixed", theta=0.2),
    "sheet-fp64-102k": dict(kind="tree", gen="sheet", rows=468, precision="fp64", theta=0.2),
    "direct-fp64-102k": dict(kind="direct", gen="sheet", rows=468, precision="fp64"),
    "direct-mixed-102k": dict(kind="direct", gen="sheet", rows=468, precision="mixed"),
}


def usable_cpus():
    """min(affinity, floor(cgroup quota)), at least 1 (the rule of compute/hardware.py:38-47)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    quota = None
    try:
        q, p = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        quota = None if q == "max" else int(q) / int(p)
    except (OSError, ValueError):
        pass
    usable = affinity if quota is None else min(affinity, max(1, math.floor(quota + 1e-9)))
    return max(1, int(usable))


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_case(spec):
    from openfawe_amd import workloads
    return workloads.sheet(spec["rows"]) if spec["gen"] == "sheet" else workloads.loop(rows=spec["rows"])


def cpu_baseline(case, spec, threads):
    """The oracle (the reference algorithm restated in C, OpenMP over targets) on a bounded sample."""
    from oracle import oracle as O
    n, m = len(case["points"]), len(case["starts"])
    args = (case["starts"], case["directions"], case["lengths"], case["gamma"], case["radius"], case["model"])
    csr = O.exclusion_csr(n, case["pairs"])
    if spec["kind"] == "tree":
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.tree_sum(case["points"], *args, csr, "raise", spec["theta"], 2, 32, 3.0, threads=threads)
            times.append(time.perf_counter() - t0)
        dt = float(np.median(times))
        return dict(value=n * m / dt, unit="pair-interactions/s", cores=threads, kind="port", ms_per_call=1e3 * dt,
                    sample=f"oracle tree_sum (C, -O2, OpenMP over targets; build, moments, traversal) on the whole "
                           f"workload, median of 3 calls; {cpu_model()}")
    sample = 512
    while True:
        pts = case["points"][:sample]
        sub = O.exclusion_csr(sample, case["pairs"][case["pairs"][:, 0] < sample])
        t0 = time.perf_counter()
        O.direct_sum(pts, *args, sub, "raise", threads=threads)
        dt = time.perf_counter() - t0
        if dt > 4.0 or sample >= n:
            break
        sample = min(n, sample * 4)
    rate = sample * m / dt
    return dict(value=rate, unit="pair-interactions/s", cores=threads, kind="port", ms_per_call=1e3 * n * m / rate,
                sample=f"oracle direct sum (C, -O2, OpenMP) over the first {sample} of {n} targets x all {m} "
                       f"segments, {dt:.2f} s, extrapolated to the whole workload; {cpu_model()}")


def adjoint_bench(steps):
    """Differentiation case: 100-panel VSM wing in a ~205k-segment wake.  Times the forward direct sum at the
    3/4-chord points, the fused forward + adjoint (dV -> dGamma, d(end points), d(points)), each in FP64 and in
    mixed precision, and the 100 x 100 attached-ring influence Jacobian (pair matrix + ordered column sums)."""
    import torch
    from openfawe_amd import adjoint as AD
    from openfawe_amd import vortex as V
    from openfawe_amd import workloads

    case = workloads.vsm_wing()
    n, m = len(case["points"]), len(case["starts"])
    dev = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)  # noqa: E731
    pts, starts, dirs = dev(case["points"]), dev(case["starts"]), dev(case["directions"])
    lens, gam, rad = dev(case["lengths"]), dev(case["gamma"]), dev(case["radius"])
    ends = starts + dirs * lens[:, None]
    ptr, seg = dev(np.zeros(n + 1, np.int64), torch.int64), dev(np.zeros(1, np.int64), torch.int64)
    grad = dev(np.random.default_rng(0).normal(size=(n, 3)))
    ws, wd, wl, wr = (dev(case[k]) for k in ("wing_starts", "wing_directions", "wing_lengths", "wing_radius"))
    cptr, cseg, csign = dev(case["column_ptr"], torch.int64), dev(case["column_segment"], torch.int64), \
        dev(case["column_sign"])

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / steps

    res = {}
    for precision in ("fp64", "mixed"):
        hs = V.HipSettings(precision, 128)
        fwd = timed(lambda: V.hip_sum_device(pts, starts, dirs, lens, gam, rad, 2, ptr, seg, "zero", hs))
        fused = timed(lambda: AD.direct_vjp(pts, starts, ends, gam, rad, 2, grad, ptr, seg, fused_forward=True,
                                            precision=precision))
        res[precision] = dict(forward_ms=fwd, fused_forward_adjoint_ms=fused, adjoint_fwd_ratio=fused / fwd)

    def jac():
        pairs, _ = AD.pair_matrix_device(pts, ws, wd, wl, wr, 2)
        return AD.assemble_columns(pairs, cptr, cseg, csign)
    jms = timed(jac)
    return dict(workload=case["name"], targets=n, segments=m, **res["fp64"], mixed=res["mixed"],
                influence_jacobian_ms=jms, jacobian_shape=[n, n, 3],
                note="fused = forward velocities + dL/dGamma + dL/d(end points) + dL/d(targets) in one call; "
                     "mixed: far pairs in FP32, near pairs and accumulation in FP64, against the mixed forward")


def network_bench(case, steps, cpu=True):
    """Incident-exclusion pairs of the whole wake (the rotor-ring call, first = 0, network_numba.py:242-253) on
    the GPU with nodes and edges resident in HBM, and the oracle (NumPy, one core) on the same input."""
    import torch
    from openfawe_amd import network as NET
    nodes, edges = case["nodes"], case["edges"]
    dn = torch.from_numpy(np.ascontiguousarray(nodes)).cuda()
    de = torch.from_numpy(np.ascontiguousarray(edges)).cuda()
    pairs = NET.incident_pairs_device(dn, de, 0)
    cap = int(pairs.shape[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        NET.incident_pairs_device(dn, de, 0, capacity=cap)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    result = dict(call="incident_pairs(nodes, edges, first=0)", nodes=len(nodes), segments=len(edges),
                  pairs=cap, ms_per_call=ms, note="includes the one stream synchronisation that reads the count")
    # ring-edge aggregation of a lattice of the wake's size: 3,744 rows x 109 panels, seeded strengths
    rng = np.random.default_rng(2)
    strengths = rng.normal(size=(3743, 109))
    strengths[1:, ::3] = strengths[:-1, ::3]
    panels = np.column_stack([np.arange(109), np.arange(1, 110)])
    dp, ds = torch.from_numpy(panels).cuda(), torch.from_numpy(strengths).cuda()
    edges_n = int(NET.pack_ring_edges_device(3744, 110, dp, ds)[0].shape[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        NET.pack_ring_edges_device(3744, 110, dp, ds, capacity=edges_n)
    torch.cuda.synchronize()
    rings = dict(call="pack_rings (3,744 x 110 nodes, 109 panels)", contributions=4 * 3743 * 109, edges=edges_n,
                 ms_per_call=1e3 * (time.perf_counter() - t0) / steps)
    if cpu:
        from oracle import network_np as NP
        t0 = time.perf_counter()
        NP.incident_pairs(nodes, edges, 0)
        result["cpu_oracle_ms"] = 1e3 * (time.perf_counter() - t0)
        result["cpu_oracle"] = "oracle/network_np.py incident_pairs (NumPy sort and search, one core)"
        t0 = time.perf_counter()
        NP.pack_ring_edges(3744, 110, panels, strengths)
        rings["cpu_oracle_ms"] = 1e3 * (time.perf_counter() - t0)
        rings["cpu_oracle"] = "oracle/network_np.py pack_ring_edges (NumPy sort and ordered sums, one core)"
    result["pack_rings"] = rings
    return result


def convection_bench(case, rows, steps, tree, settings, cpu_threads=None):
    """One free-wake geometry step (wake.py:586-616) of the loop wake on the GPU: row 0 is the bound row, the newest
    187 rows are free targets (free_convection_rows=187; 20,570 points), the older rows frozen, the rotor helices ring
    nodes ('zero'); every stage runs the incident pairs, the segment geometry, the exclusion subsets and the tree sums
    of wake._network_convection."""
    import torch
    from openfawe_amd import convection as CV
    nodes, edges = case["nodes"], case["edges"]
    width = 110
    ring_first = rows * width          # the rotor helices follow the wake rows
    dn = torch.from_numpy(np.ascontiguousarray(nodes)).cuda()
    de = torch.from_numpy(np.ascontiguousarray(edges)).cuda()
    dg = torch.from_numpy(np.ascontiguousarray(case["gamma"])).cuda()
    dr = torch.from_numpy(np.ascontiguousarray(case["radius"])).cuda()
    cs = CV.ConvectionSettings(model=case["model"], free_convection_rows=187, backend="hip_tree", tree=tree,
                               hip=settings)
    result = dict(workload=f"{case['name']}: bound row + {rows - 1} rows x {width} nodes + rotor helices",
                  free_row_targets=187 * width, ring_targets=len(nodes) - ring_first, segments=len(edges),
                  settings=dict(theta=tree.theta, precision=settings.precision, free_convection_rows=187))
    for integrator in ("euler", "rk4"):
        def run():
            return CV.convect_geometry(dn[:width], dn[width:], de, dg, dr, cs, inflow=[10.0, 0.0, 0.0], time=0.0,
                                       dt_s=0.05, integrator=integrator, ring_first=ring_first)
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        torch.cuda.synchronize()
        result[f"{integrator}_step_ms"] = 1e3 * (time.perf_counter() - t0) / steps
        # the same step recorded as one graph (GeometryStepPlan): replay plus the status read-back
        plan = CV.GeometryStepPlan(dn[:width], dn[width:], de, dg, dr, cs, inflow=[10.0, 0.0, 0.0], time=0.0,
                                   dt_s=0.05, integrator=integrator, ring_first=ring_first)
        plan.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.run()
        torch.cuda.synchronize()
        result[f"{integrator}_graph_step_ms"] = 1e3 * (time.perf_counter() - t0) / steps
        result[f"{integrator}_graph_capacities"] = plan.capacities
        plan.close()
    if cpu_threads:
        # the same Euler step by the oracle (the reference's algorithm in C, FP64 tree code, OpenMP over targets)
        from oracle import convection_np as CO
        t0 = time.perf_counter()
        CO.convect_geometry(nodes[:width], nodes[width:], edges, case["gamma"], case["radius"], case["model"],
                            inflow=[10.0, 0.0, 0.0], time=0.0, dt_s=0.05, integrator="euler", ring_first=ring_first,
                            free_convection_rows=187, engine=CO.tree_engine(tree.theta, tree.order, tree.leaf_size,
                                                                             tree.core_clearance, cpu_threads))
        result["cpu_euler_step_ms"] = 1e3 * (time.perf_counter() - t0)
        result["cpu"] = (f"oracle/convection_np.py Euler step with the C oracle's tree code (FP64, theta "
                         f"{tree.theta}), {cpu_threads} threads, {cpu_model()}")
    return result


def traffic(workload, kernel):
    """HBM bytes per launch of the kernel from the committed PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy-kernel calibration), or None when none was measured for this workload."""
    path = ROOT / "profiles" / "traffic.json"
    try:
        table = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    entry = table.get(f"{workload}:{kernel}")
    return entry.get("bytes_per_launch") if entry else None


def copy_calibration(mib=1024, repeats=5):
    """A plain device copy of a known size, for scaling the PMC byte counters (FETCH_SIZE, WRITE_SIZE)."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps(dict(copy_bytes=n * 8, repeats=repeats)), flush=True)


def accuracy(case, out, spec, sample=1024, seed=0):
    """Relative L2 and max-abs error of the measured call against the FP64 direct sum (HIP kernel) at a seeded
    sample of targets, with the same exclusions."""
    from openfawe_amd import autotune as AT
    from openfawe_amd import vortex as V
    rng = np.random.default_rng(seed)
    n = len(case["points"])
    idx = np.sort(rng.choice(n, size=min(sample, n), replace=False))
    ref, _ = V.hip_sum(case["points"][idx], case["starts"], case["directions"], case["lengths"], case["gamma"],
                       case["radius"], case["model"], AT._sample_csr(case, idx), "raise", V.HipSettings("fp64", 256))
    got = out.cpu().numpy()[idx]
    err = got - ref
    return dict(reference="FP64 direct sum (HIP kernel) at sampled targets", sampled_targets=int(len(idx)),
                rel_l2=float(np.linalg.norm(err) / np.linalg.norm(ref)), max_abs=float(np.abs(err).max()),
                max_abs_rel=float(np.abs(err).max() / np.abs(ref).max()), theta=spec.get("theta"))


def batch_bench(steps, split_depth=-1, autotune=True):
    """Batched solve: 64 uneven loop snapshots (10k-410k segments, 2k-20k targets each) in one graph replay.
    Reports aggregate pair-interactions/s, the batch time, the amortised time per snapshot, and the time of
    the same 64 snapshots as separate calls."""
    import torch
    from openfawe_amd import autotune as AT
    from openfawe_amd import tree as T
    from openfawe_amd import vortex as V
    from openfawe_amd import workloads

    snaps = workloads.batch()
    n_k = np.array([len(s["points"]) for s in snaps])
    m_k = np.array([len(s["starts"]) for s in snaps])
    interactions = float((n_k * m_k).sum())
    tuned = None
    theta, leaf, depth = 0.5, 32, split_depth
    if autotune:
        tuned = AT.tune_batch(snaps, precision="mixed", tolerance=1e-5)
        theta, leaf, depth = tuned["theta"], tuned["leaf_size"], tuned["split_depth"]
    settings = V.TreeSettings(theta, 2, leaf, 3.0)
    batch = T.TreeBatch(snaps, settings, model=2, precision="mixed", split_depth=depth, use_graph=True)
    try:
        for _ in range(2):
            batch.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.run()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        graph_ms = batch.timings()[0] if batch.timings() else None
        stats = batch.stats()
    finally:
        batch.close()
    # the same snapshots as 64 separate calls (same settings), for the batching gain
    seq = AT.sequential_ms(snaps, settings, "mixed")
    return dict(workload="batch-64-loop", snapshots=len(snaps), targets=int(n_k.sum()), segments=int(m_k.sum()),
                targets_range=[int(n_k.min()), int(n_k.max())], segments_range=[int(m_k.min()), int(m_k.max())],
                interactions=interactions, ms_per_batch=ms, graph_ms=graph_ms,
                value=interactions / (ms * 1e-3), unit="pair-interactions/s (aggregate, n*m/t)",
                ms_per_snapshot_amortised=ms / len(snaps), sequential_calls_ms=seq,
                settings=dict(theta=theta, leaf_size=leaf, split_depth=stats["split_depth"], order=2,
                              precision="mixed"), autotune=tuned)


def upload_case(case):
    """The engine inputs of a case as device tensors (float64 / int64) and preallocated outputs."""
    import torch
    from openfawe_amd import vortex as V
    dev = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)  # noqa: E731
    n = len(case["points"])
    ptr, seg = V.exclusion_csr(n, case["pairs"])
    return dict(pts=dev(case["points"]), starts=dev(case["starts"]), dirs=dev(case["directions"]),
                lens=dev(case["lengths"]), gam=dev(case["gamma"]), rad=dev(case["radius"]), ptr=dev(ptr, torch.int64),
                seg=dev(seg if len(seg) else np.zeros(1, np.int64), torch.int64),
                out=torch.zeros((n, 3), dtype=torch.float64, device="cuda"),
                flags=torch.zeros(n, dtype=torch.int8, device="cuda"),
                far=torch.zeros(n, dtype=torch.int64, device="cuda"),
                near=torch.zeros(n, dtype=torch.int64, device="cuda"),
                fseg=torch.zeros(n, dtype=torch.int32, device="cuda"))


def timed_calls(step, steps, warmup):
    """Mean ms per call over `steps` calls after `warmup`, bracketed by device synchronisations, and the mean of
    the library's phase timings (HIP events on its stream) over the timed calls."""
    import torch
    from openfawe_amd import _lib
    from openfawe_amd import vortex as V
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctx = V._ctx()  # the library context of the current stream: the same for every call
    phases = []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        phases.append(_lib.timings(ctx))
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    return ms, np.mean(np.array(phases), axis=0).tolist()


def tree_step(case, d, spec, split_depth=-1):
    from openfawe_amd import tree as T
    from openfawe_amd import vortex as V
    settings = V.HipSettings(spec["precision"], 256)
    tree = V.TreeSettings(spec.get("theta", 0.5), 2, 32, 3.0)

    def step():
        T.hip_tree_sum_device(d["pts"], d["starts"], d["dirs"], d["lens"], d["gam"], d["rad"], case["model"], d["ptr"],
                              d["seg"], "raise", tree, settings, d["out"], d["flags"], d["far"], d["near"],
                              split_depth=split_depth)
    return step


def tree_leg(name, steps, warmup):
    """One more tree workload on the line: time per call (resident inputs), rate, phases and accuracy."""
    spec = WORKLOADS[name]
    case = make_case(spec)
    d = upload_case(case)
    n, m = len(case["points"]), len(case["starts"])
    ms, phase_ms = timed_calls(tree_step(case, d, spec), steps, warmup)
    near_n, far_n = int(d["near"].sum().item()), int(d["far"].sum().item())
    peak = FP64_PEAK_TFLOPS if spec["precision"] == "fp64" else FP32_PEAK_TFLOPS
    kernel_tflops = (near_n * OPS_PER_PAIR[case["model"]] + far_n * OPS_PER_FAR) / (phase_ms[5] * 1e-3) / 1e12
    return dict(workload=name, targets=n, segments=m, precision=spec["precision"], theta=spec["theta"],
                ms_per_call=ms, value=n * m / (ms * 1e-3), unit="pair-interactions/s",
                phases_ms=dict(zip(PHASES, phase_ms)), k_tree_near_tflops=kernel_tflops,
                k_tree_near_frac=kernel_tflops / peak, accuracy=accuracy(case, d["out"], spec))


def drop_in_host(case, spec, calls=5):
    """The drop-in call as a maintainer binds it (host NumPy arrays in and out, hip_tree_sum): median of `calls`
    calls, with the upload / evaluation / download split (vortex_cuda_tree.py:581-594)."""
    from openfawe_amd import tree as T
    from openfawe_amd import vortex as V
    args = (case["points"], case["starts"], case["directions"], case["lengths"], case["gamma"], case["radius"],
            case["model"])
    csr = V.exclusion_csr(len(case["points"]), case["pairs"])
    tree = V.TreeSettings(spec.get("theta", 0.5), 2, 32, 3.0)
    settings = V.HipSettings(spec["precision"], 256)
    T.hip_tree_sum(*args, csr, "raise", tree, settings)
    rows = []
    for _ in range(calls):
        t0 = time.perf_counter()
        T.hip_tree_sum(*args, csr, "raise", tree, settings)
        st = V.Statistics.last
        rows.append((1e3 * (time.perf_counter() - t0), st["upload_ms"], st["evaluate_ms"], st["download_ms"]))
    med = np.median(np.array(rows), axis=0)
    mb = sum(np.asarray(case[k]).nbytes for k in ("points", "starts", "directions", "lengths", "gamma", "radius"))
    return dict(call="hip_tree_sum on host NumPy arrays (exclusion CSR built on the host beforehand)",
                ms_per_call=float(med[0]), upload_ms=float(med[1]), evaluate_ms=float(med[2]),
                download_ms=float(med[3]), input_mb=mb / 1e6, calls=calls, statistic="median")


def sheet_table(steps, threads, cpu=True):
    """The reference's speed table (README.md:83-91): the synthetic sheet (selftest.py:243-272) at 468, 936,
    1,872 and 3,744 rows, 20,570 targets, theta 0.2, order 2, leaf 32, Lamb-Oseen cores; FP64 and mixed on the
    GPU (resident inputs, and the drop-in call on host arrays), the oracle tree on the host's cores, and the
    accuracy statements of README.md:93-94 (FP64 against the CPU tree, mixed within a tenth of the tree's own
    error against the direct sum)."""
    from oracle import oracle as O
    from openfawe_amd import autotune as AT
    from openfawe_amd import vortex as V
    from openfawe_amd import workloads
    table = []
    for rows in (468, 936, 1872, 3744):
        case = workloads.sheet(rows)
        n, m = len(case["points"]), len(case["starts"])
        args = workloads.engine_args(case)
        row = dict(rows=rows, targets=n, segments=m)
        csr_o = O.exclusion_csr(n, case["pairs"])
        if cpu:
            times = []
            for _ in range(3):
                t0 = time.perf_counter()
                ref_tree, _, _, _ = O.tree_sum(*args, csr_o, "raise", 0.2, 2, 32, 3.0, threads=threads)
                times.append(time.perf_counter() - t0)
            row["cpu_oracle_tree_ms"] = 1e3 * float(np.median(times))
            row["cpu_threads"] = threads
        else:
            ref_tree, _, _, _ = O.tree_sum(*args, csr_o, "raise", 0.2, 2, 32, 3.0, threads=threads)
        idx = np.sort(np.random.default_rng(0).choice(n, size=1024, replace=False))
        direct, _ = V.hip_sum(case["points"][idx], *args[1:], AT._sample_csr(case, idx), "raise",
                              V.HipSettings("fp64", 256))
        tree_err = np.linalg.norm(ref_tree[idx] - direct, axis=1).max()
        d = upload_case(case)
        for precision in ("fp64", "mixed"):
            spec = dict(kind="tree", precision=precision, theta=0.2)
            ms, _ = timed_calls(tree_step(case, d, spec), steps, 2)
            out = d["out"].cpu().numpy()
            host = drop_in_host(case, spec)
            entry = dict(ms_per_call=ms, value=n * m / (ms * 1e-3), drop_in_ms=host["ms_per_call"],
                         drop_in_split_ms=[host["upload_ms"], host["evaluate_ms"], host["download_ms"]])
            if precision == "fp64":
                entry["max_abs_vs_cpu_tree_m_s"] = float(np.abs(out - ref_tree).max())
            else:
                entry["max_err_vs_cpu_tree_over_tree_err"] = float(
                    np.linalg.norm(out[idx] - ref_tree[idx], axis=1).max() / tree_err)
            row[precision] = entry
        row["tree_err_vs_direct_max_m_s"] = float(tree_err)
        table.append(row)
    return dict(reference="README.md:83-94 (RTX 4090 D, 22 CPU threads: GPU double 0.14-0.53 s, mixed "
                          "0.07-0.52 s per call, host arrays in and out)", settings=dict(theta=0.2, order=2,
                                                                                        leaf_size=32),
                rows=table)


PHASES = ["build (midpoints + tree)", "node moments", "traversal plan (exclusions, tasks)", "walk (k_tree_eval)",
          "list plan", "interaction lists (k_tree_near)", "reduce"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="tree-mixed-820k", choices=sorted(WORKLOADS))
    ap.add_argument("--split-depth", type=int, default=-1)
    ap.add_argument("--theta", type=float, default=None, help="override the workload's opening angle")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (PMC byte-counter calibration)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the adjoint, batch and network legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              f"(targets shard across GPUs without communication; this measures one GPU's share)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return

    import torch
    from openfawe_amd import vortex as V

    spec = dict(WORKLOADS[args.workload])
    if args.theta is not None:
        spec["theta"] = args.theta
    torch.cuda.set_device(0)
    case = make_case(spec)
    n, m = len(case["points"]), len(case["starts"])
    settings = V.HipSettings(spec["precision"], 256)
    tree = V.TreeSettings(spec.get("theta", 0.5), 2, 32, 3.0)
    d = upload_case(case)
    if spec["kind"] == "tree":
        step = tree_step(case, d, spec, args.split_depth)
    else:
        def step():
            V.hip_sum_device(d["pts"], d["starts"], d["dirs"], d["lens"], d["gam"], d["rad"], case["model"], d["ptr"],
                             d["seg"], "raise", settings, d["out"], d["flags"], d["fseg"])
    ms, phase_ms = timed_calls(step, args.steps, args.warmup)
    value = n * m / (ms * 1e-3)
    fp64 = spec["precision"] == "fp64"
    peak = FP64_PEAK_TFLOPS if fp64 else FP32_PEAK_TFLOPS
    if spec["kind"] == "tree":
        far_n = int(d["far"].sum().item())
        near_n = int(d["near"].sum().item())
        # the roofline kernel: the interaction-list kernel, which evaluates every near pair and every accepted
        # node's far field of the tree code; the walk (acceptance tests only) is reported beside it
        flops = near_n * OPS_PER_PAIR[case["model"]] + far_n * OPS_PER_FAR
        kms = phase_ms[5]
        metric = "pair-interactions/s (tree code, direct-sum equivalent n*m/t)"
        kernel = "k_tree_near"
        extra = dict(far_nodes_per_target=far_n / n, near_pairs_per_target=near_n / n,
                     phases_ms=dict(zip(PHASES, phase_ms)),
                     walk=dict(kernel="k_tree_eval", kernel_ms=phase_ms[3],
                               note="latency-bound pre-order walk: acceptance tests, list appends"))
    else:
        pairs = n * m - len(case["pairs"])
        flops = pairs * OPS_PER_PAIR[case["model"]]
        kms = phase_ms[1]
        metric = "pair-interactions/s (direct sum, n*m/t)"
        kernel = "k_direct64" if fp64 else "k_direct_mixed"
        extra = dict(phases_ms=dict(zip(["pack", "pair kernel", "chunk reduce"], phase_ms)))
    achieved = flops / (kms * 1e-3) / 1e12
    result = dict(
        metric=metric, value=value, unit="pair-interactions/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64" if fp64 else "f32+f64", data=f"synthetic ({case['name']} wake, seeded)",
        config=dict(workload=args.workload, targets=n, segments=m, core_model="lamb-oseen",
                    precision=spec["precision"], **({"theta": spec["theta"], "order": 2, "leaf_size": 32,
                                                     "core_clearance": 3.0} if spec["kind"] == "tree" else {})),
        roofline=dict(bound="valu-fp64" if fp64 else "valu-fp32", achieved=achieved, peak=peak, unit="TFLOP/s",
                      frac=achieved / peak, traffic=traffic(args.workload, kernel), kernel=kernel, kernel_ms=kms,
                      flops_per_pair=OPS_PER_PAIR[case["model"]],
                      **({"flops_per_far_node": OPS_PER_FAR} if spec["kind"] == "tree" else {})),
        **extra,
    )
    if spec["kind"] == "tree":
        result["accuracy"] = accuracy(case, d["out"], spec)
        result["drop_in_host"] = drop_in_host(case, spec)
    if not args.no_extra and spec["kind"] == "tree":
        # the metric's other size and config 2, then the reference's own speed-table shape
        result["legs"] = [tree_leg(name, args.steps, args.warmup) for name in ("tree-mixed-410k", "tree-fp64-102k")]
        result["sheet_table"] = sheet_table(max(3, args.steps // 2), usable_cpus(), cpu=not args.no_cpu)
        result["adjoint"] = adjoint_bench(max(3, args.steps // 2))
        result["batch"] = batch_bench(max(3, args.steps // 2))
        if "nodes" in case:
            result["network"] = network_bench(case, max(3, args.steps // 2), cpu=not args.no_cpu)
            result["convection"] = convection_bench(case, spec["rows"], max(2, args.steps // 5), tree, settings,
                                                    None if args.no_cpu else usable_cpus())
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(case, spec, usable_cpus())
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
