"""Autotuner of the batched tree call (``TreeBatch``): opening angle, leaf size and traversal split depth.

The reference fixes its settings per call site (TreeSettings, vortex.py:30-55) and has no batched call.  For a
batch of snapshots the tuner picks the fastest settings whose velocities stay within a relative-L2 tolerance
of the FP64 direct sum:

1. a sample of targets of every snapshot is evaluated with the FP64 direct sum (the HIP direct kernel, with
   the snapshot's exclusions), once;
2. every (theta, leaf size) candidate runs the whole batch once at the default split depth; its error is the
   relative L2 norm of the velocity error over all sampled targets, and its time the best of a few runs;
3. among the candidates within the tolerance, the fastest is kept, and its split depth is swept around the
   default with the batch recorded in a HIP graph, as it runs in production.

Only settings change; every candidate computes the reference's tree sum for those settings.
"""
import time

import numpy as np

from . import tree as T
from . import vortex as V

_ENGINE = ("points", "starts", "directions", "lengths", "gamma", "radius")


def _sample(snaps, per_snapshot, seed):
    rng = np.random.default_rng(seed)
    return [np.sort(rng.choice(len(s["points"]), size=min(per_snapshot, len(s["points"])), replace=False))
            for s in snaps]


def _sample_csr(snap, idx):
    pairs = np.asarray(snap.get("pairs", np.zeros((0, 2))), dtype=np.int64).reshape(-1, 2)
    where = np.full(len(snap["points"]), -1, dtype=np.int64)
    where[idx] = np.arange(len(idx))
    keep = where[pairs[:, 0]] >= 0 if len(pairs) else np.zeros(0, bool)
    local = np.stack([where[pairs[keep, 0]], pairs[keep, 1]], axis=1) if keep.any() else np.zeros((0, 2), np.int64)
    return V.exclusion_csr(len(idx), local)


def direct_reference(snaps, samples):
    """FP64 direct-sum velocities at the sampled targets of every snapshot (one array per snapshot)."""
    settings = V.HipSettings("fp64", 256)
    refs = []
    for snap, idx in zip(snaps, samples):
        if len(idx) == 0 or len(snap["starts"]) == 0:
            refs.append(np.zeros((len(idx), 3)))
            continue
        out, _ = V.hip_sum(snap["points"][idx], snap["starts"], snap["directions"], snap["lengths"],
                           snap["gamma"], snap["radius"], 2, _sample_csr(snap, idx), "zero", settings)
        refs.append(out)
    return refs


def _error(batch, out, samples, refs):
    num = den = 0.0
    for k, idx in enumerate(samples):
        got = out[batch.target_offsets[k] + idx]
        num += float(np.sum((got - refs[k]) ** 2))
        den += float(np.sum(refs[k] ** 2))
    return float(np.sqrt(num / den)) if den > 0 else 0.0


def _best_ms(batch, repeats):
    import torch
    best = float("inf")
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.run()
        torch.cuda.synchronize()
        best = min(best, 1e3 * (time.perf_counter() - t0))
    return best


def tune_batch(snaps, *, precision="mixed", tolerance=1e-5, thetas=(0.08, 0.1, 0.12, 0.15), leaves=(16, 32, 64),
               depth_offsets=(-1, 0, 1, 2), per_snapshot=64, repeats=3, seed=0, order=2, clearance=3.0):
    """Fastest (theta, leaf size, split depth) within ``tolerance`` (relative L2 against the FP64 direct sum
    at sampled targets).  Returns the choice with the table of every candidate."""
    samples = _sample(snaps, per_snapshot, seed)
    refs = direct_reference(snaps, samples)
    table = []
    for leaf in leaves:
        for theta in thetas:
            batch = T.TreeBatch(snaps, V.TreeSettings(theta, order, leaf, clearance), model=2, precision=precision,
                                use_graph=False)
            try:
                out = batch.run().cpu().numpy()
                err = _error(batch, out, samples, refs)
                ms = _best_ms(batch, repeats)
                table.append(dict(theta=theta, leaf_size=leaf, split_depth=batch.stats()["split_depth"],
                                  rel_l2=err, ms=ms))
            finally:
                batch.close()
    feasible = [row for row in table if row["rel_l2"] <= tolerance]
    best = min(feasible, key=lambda r: r["ms"]) if feasible else min(table, key=lambda r: r["rel_l2"])
    sweep = []
    for off in depth_offsets:
        depth = best["split_depth"] + off
        if depth < 0 or depth > 14:
            continue
        batch = T.TreeBatch(snaps, V.TreeSettings(best["theta"], order, best["leaf_size"], clearance), model=2,
                            precision=precision, split_depth=depth, use_graph=True)
        try:
            batch.run()
            sweep.append(dict(split_depth=depth, ms=_best_ms(batch, repeats)))
        finally:
            batch.close()
    depth = min(sweep, key=lambda r: r["ms"])["split_depth"] if sweep else best["split_depth"]
    return dict(theta=best["theta"], leaf_size=best["leaf_size"], split_depth=depth, rel_l2=best["rel_l2"],
                tolerance=tolerance, within_tolerance=bool(feasible), sampled_targets=int(sum(map(len, samples))),
                candidates=table, depth_sweep=sweep)


def sequential_ms(snaps, settings, precision, depth=-1, repeats=2):
    """Time of the snapshots as separate tree calls (inputs resident on the device), best of ``repeats``."""
    import torch
    dev = []
    for s in snaps:
        arrays = [torch.as_tensor(np.ascontiguousarray(s[k]), dtype=torch.float64).cuda() for k in _ENGINE]
        ptr, seg = V.exclusion_csr(len(s["points"]), s.get("pairs", np.zeros((0, 2), np.int64)))
        dev.append((arrays, torch.as_tensor(ptr).cuda(),
                    torch.as_tensor(seg if len(seg) else np.zeros(1, np.int64)).cuda()))
    hip = V.HipSettings(precision)
    best = float("inf")
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for arrays, ptr, seg in dev:
            T.hip_tree_sum_device(*arrays, 2, ptr, seg, "zero", settings, hip, split_depth=depth)
        torch.cuda.synchronize()
        best = min(best, 1e3 * (time.perf_counter() - t0))
    return best
