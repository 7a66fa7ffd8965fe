"""Benchmark of the free-wake induced-velocity path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one induced-velocity call over one synthetic wake whose inputs are already resident in HBM:
targets (n,3), segment starts / directions / lengths / circulations / core radii.  The timed region is
K calls bracketed by a device synchronisation on both sides.  One JSON line is printed.

Workloads (--workload):
  direct-fp64-102k  FP64 direct sum, 20,570 targets x 102,382 segments (the reference's 468-row sheet)
  direct-mixed-102k mixed-precision direct sum, same wake
The headline metric is pair-interactions/s = n*m / (time per call).
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
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector (packed); non-packed issue rate is half of it
HBM_PEAK_GBS = 8000.0
OPS_PER_PAIR = {0: 67, 1: 71, 2: 69, 3: 75, 4: 70, 5: 86}   # algorithmic ops per far pair by core model


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


def make_case(workload):
    from openfawe_amd import workloads
    if workload.endswith("102k"):
        return workloads.sheet(468)
    if workload.endswith("410k"):
        return workloads.sheet(1872)
    if workload.endswith("820k"):
        return workloads.sheet(3744)
    raise SystemExit(f"unknown workload {workload}")


def cpu_baseline(case, workload, threads):
    """The oracle (the reference algorithm restated in C, OpenMP over targets) on a bounded sample."""
    from oracle import oracle as O
    n, m = len(case["points"]), len(case["starts"])
    args = (case["starts"], case["directions"], case["lengths"], case["gamma"], case["radius"], case["model"])
    sample = 512
    while True:
        pts = case["points"][:sample]
        t0 = time.perf_counter()
        O.direct_sum(pts, *args, None, "raise", threads=threads)
        dt = time.perf_counter() - t0
        if dt > 4.0 or sample >= n:
            break
        sample = min(n, sample * 4)
    rate = sample * m / dt
    return dict(value=rate, unit="pair-interactions/s", cores=threads, kind="port",
                sample=f"oracle direct sum (C, -O2, OpenMP) over the first {sample} of {n} targets x all {m} "
                       f"segments, {dt:.2f} s; {cpu_model()}",
                ms_per_call_extrapolated=1e3 * n * m / rate)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="direct-fp64-102k")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              f"(targets shard without communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    from openfawe_amd import _lib
    from openfawe_amd import vortex as V

    torch.cuda.set_device(0)
    case = make_case(args.workload)
    n, m = len(case["points"]), len(case["starts"])
    precision = "mixed" if "mixed" in args.workload else "fp64"
    settings = V.HipSettings(precision, 256)
    dev = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)  # noqa: E731
    pts, starts, dirs = dev(case["points"]), dev(case["starts"]), dev(case["directions"])
    lens, gam, rad = dev(case["lengths"]), dev(case["gamma"]), dev(case["radius"])
    ptr, seg = V.exclusion_csr(n, case["pairs"])
    ptr = dev(ptr, torch.int64)
    seg = dev(seg if len(seg) else np.zeros(1, np.int64), torch.int64)
    out = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    flags = torch.zeros(n, dtype=torch.int8, device="cuda")
    fseg = torch.zeros(n, dtype=torch.int32, device="cuda")

    def step():
        V.hip_sum_device(pts, starts, dirs, lens, gam, rad, case["model"], ptr, seg, "raise", settings,
                         out, flags, fseg)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        kernel_ms.append(_lib.timings(V._ctx())[1])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = 1e3 * elapsed / args.steps
    pairs = n * m - len(case["pairs"])
    value = n * m / (ms * 1e-3)
    kms = float(np.mean(kernel_ms))
    flops = pairs * OPS_PER_PAIR[case["model"]]
    achieved = flops / (kms * 1e-3) / 1e12
    peak = FP64_PEAK_TFLOPS if precision == "fp64" else FP32_PEAK_TFLOPS / 2
    result = dict(
        metric="pair-interactions/s (direct sum, 20,570 targets x 102k segments)",
        value=value, unit="pair-interactions/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64" if precision == "fp64" else "f32/f64", data="synthetic (seeded sheet wake of the reference)",
        config=dict(workload=args.workload, targets=n, segments=m, core_model="lamb-oseen", precision=precision),
        roofline=dict(bound="valu-fp64" if precision == "fp64" else "valu-fp32", achieved=achieved, peak=peak,
                      unit="TFLOP/s", frac=achieved / peak, traffic=None,
                      kernel="k_direct64" if precision == "fp64" else "k_direct_mixed", kernel_ms=kms,
                      ops_per_pair=OPS_PER_PAIR[case["model"]]),
    )
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(case, args.workload, usable_cpus())
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
