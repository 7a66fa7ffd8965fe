"""ofw_direct_sum_static, the direct sum that allocates no device memory, against ofw_direct_sum.

Both go through the C ABI.  Where the static scratch holds the allocating entry point's own chunk count the
two calls run the same kernels with the same launch plan, so the bar is bitwise equality (velocities, flags and
first flagged segments).  Where the chunk count has to be lowered to fit, the association of the per-target sum
changes and the bar is the direct sum's A1 bound, |u_static - u| <= 1e-12 * sum_s |u_s| per target.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from openfawe_amd import _lib
from openfawe_amd import vortex as V
from openfawe_amd import workloads

pytestmark = pytest.mark.gpu

SCRATCH = 128 << 20


def _device(case, torch):
    dev = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dt)  # noqa: E731
    n = len(case["points"])
    ptr, seg = V.exclusion_csr(n, case["pairs"])
    return (dev(case["points"]), dev(case["starts"]), dev(case["directions"]), dev(case["lengths"]),
            dev(case["gamma"]), dev(case["radius"]), case["model"], dev(ptr, torch.int64),
            dev(seg if len(seg) else np.zeros(1, np.int64), torch.int64))


def _run(args, precision, static, singularity="raise", block=128):
    out, flags, fseg = V.hip_sum_device(*args, singularity, V.HipSettings(precision, block), static_scratch=static)
    return out.cpu().numpy(), flags.cpu().numpy(), fseg.cpu().numpy()


def _chunks():
    buf = (ctypes.c_int64 * 2)()
    _lib.load().ofw_ctx_stats(V._ctx(), buf, 2)
    return int(buf[0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_two_calls_in_a_row_equal_the_allocating_entry_point(precision):
    import torch
    big = _device(workloads.loop(rows=400, target_rows=60, seed=3), torch)       # 6,600 x 89,698, exclusions
    small = _device(workloads.loop(rows=120, target_rows=30, seed=4), torch)     # 3,300 x 28,378
    first = _run(big, precision, True)
    second = _run(small, precision, True)          # the smaller call follows on the same scratch
    want_first = _run(big, precision, False)
    want_second = _run(small, precision, False)
    for got, want in ((first, want_first), (second, want_second)):
        assert np.array_equal(got[0].view(np.uint64), want[0].view(np.uint64))
        assert np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2])
        assert not got[1].any()


def test_flags_through_the_static_scratch():
    """A target on a filament with 'raise' is flagged with the same first segment by both entry points."""
    import torch
    case = workloads.loop(rows=60, target_rows=10, seed=5)
    case = dict(case, pairs=np.zeros((0, 2), np.int64))   # no exclusions: every wake node sits on its segments
    args = _device(case, torch)
    got = _run(args, "fp64", True)
    want = _run(args, "fp64", False)
    assert got[1].any() and np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2])
    assert np.array_equal(got[0].view(np.uint64), want[0].view(np.uint64))


def test_lowered_chunk_count_within_a1():
    """100 targets x 780k segments: the records take most of the scratch, so the partials get fewer chunks."""
    import torch
    case = workloads.vsm_wing(wake_rows=3560)
    n, m = len(case["points"]), len(case["starts"])
    assert 160 * m < SCRATCH < 160 * m + 8192 * 28 * n
    case = dict(case, pairs=np.zeros((0, 2), np.int64))
    args = _device(case, torch)
    want, wflags, _ = _run(args, "fp64", False)
    chunks_alloc = _chunks()
    got, gflags, _ = _run(args, "fp64", True)
    chunks_static = _chunks()
    assert chunks_static < chunks_alloc
    assert not wflags.any() and np.array_equal(gflags, wflags)
    ref, _, scale = O.direct_sum(*workloads.engine_args(case), None, "raise", with_scale=True)
    assert np.all(np.linalg.norm(got - ref, axis=1) <= 1e-12 * scale)
    assert np.all(np.linalg.norm(got - want, axis=1) <= 1e-12 * scale)


def test_rejects_a_call_beyond_the_scratch():
    import torch
    m, n = 900_000, 64                       # 144 MB of records alone
    starts = torch.zeros((m, 3), dtype=torch.float64, device="cuda")
    dirs = torch.zeros((m, 3), dtype=torch.float64, device="cuda")
    dirs[:, 0] = 1.0
    ones = torch.ones(m, dtype=torch.float64, device="cuda")
    pts = torch.ones((n, 3), dtype=torch.float64, device="cuda")
    ptr = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
    seg = torch.zeros(1, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError, match="static scratch"):
        V.hip_sum_device(pts, starts, dirs, ones, ones, ones, 2, ptr, seg, "raise", V.HipSettings("fp64"),
                         static_scratch=True)
    # the allocating entry point takes the same call
    out, flags, _ = V.hip_sum_device(pts, starts, dirs, ones, ones, ones, 2, ptr, seg, "raise", V.HipSettings("fp64"))
    assert out.shape == (n, 3)
