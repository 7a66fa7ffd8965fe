"""Derivatives of the direct sum and the VSM influence operator on MI355X.

* ``pair_matrix`` / ``assemble_columns``: the unit-circulation pair velocities and their signed, ordered
  column sums (vortex_numba.py:271-308), from which the reference builds the attached-ring influence operator
  dV/dGamma of the VSM (vsm_wake.py:222-252) and its Newton Jacobian (vsm.py:333-349).
* ``direct_vjp``: the hand-written adjoint kernel (no counterpart in the reference): given dL/dV it returns
  dL/dGamma, dL/d(segment start), dL/d(segment end) and dL/d(target), optionally with the forward velocities
  in the same pass (the fused forward + adjoint).
* ``InducedVelocity`` / ``induced_velocity``: a ``torch.autograd.Function`` over (targets, segment starts,
  segment ends, circulations) whose backward is that kernel.
"""
import ctypes

import numpy as np

from . import _lib
from .vortex import HipSettings, _ctx, _ptr, _torch, exclusion_csr


def _f64(torch, a):
    if isinstance(a, np.ndarray) or np.isscalar(a):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")
    return a.to(device="cuda", dtype=torch.float64).contiguous()


def _i64(torch, a):
    if isinstance(a, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.int64)).to("cuda")
    return a.to(device="cuda", dtype=torch.int64).contiguous()


def pair_matrix(points, starts, directions, lengths, radius, model, singularity="raise"):
    """Unit-circulation velocity of every segment at every target: (out (n,m,3), flags (n,)) host arrays."""
    torch = _torch()
    pts, s, d, L, r = (_f64(torch, a) for a in (points, starts, directions, lengths, radius))
    n, m = pts.shape[0], s.shape[0]
    out = torch.zeros((n, m, 3), dtype=torch.float64, device="cuda")
    flags = torch.zeros(max(n, 1), dtype=torch.int8, device="cuda")
    status = _lib.load().ofw_pair_matrix(_ctx(), n, _ptr(pts), m, _ptr(s), _ptr(d), _ptr(L), _ptr(r), int(model),
                                         int(singularity == "raise"), _ptr(out), _ptr(flags))
    _lib.check(status, "ofw_pair_matrix")
    return out.cpu().numpy(), flags[:n].cpu().numpy()


def pair_matrix_device(pts, s, d, L, r, model, singularity="raise", out=None, flags=None):
    torch = _torch()
    n, m = pts.shape[0], s.shape[0]
    if out is None:
        out = torch.zeros((n, m, 3), dtype=torch.float64, device="cuda")
    if flags is None:
        flags = torch.zeros(max(n, 1), dtype=torch.int8, device="cuda")
    status = _lib.load().ofw_pair_matrix(_ctx(), n, _ptr(pts), m, _ptr(s), _ptr(d), _ptr(L), _ptr(r), int(model),
                                         int(singularity == "raise"), _ptr(out), _ptr(flags))
    _lib.check(status, "ofw_pair_matrix")
    return out, flags


def assemble_columns(pairs, column_ptr, column_segment, column_sign):
    """out[p, j] = sum_k sign_k * pairs[p, segment_k] in k order from 0.0 (vortex_numba.py:290-308)."""
    torch = _torch()
    pr = _f64(torch, pairs)
    n, m = pr.shape[0], pr.shape[1]
    ptr, seg = _i64(torch, column_ptr), _i64(torch, column_segment if len(column_segment) else np.zeros(1, np.int64))
    sign = _f64(torch, column_sign if len(column_sign) else np.zeros(1))
    columns = int(ptr.shape[0]) - 1
    out = torch.zeros((n, columns, 3), dtype=torch.float64, device="cuda")
    status = _lib.load().ofw_assemble_columns(_ctx(), n, m, _ptr(pr), columns, _ptr(ptr), _ptr(seg), _ptr(sign),
                                              _ptr(out))
    _lib.check(status, "ofw_assemble_columns")
    return out


def segment_geometry(starts, ends):
    """(directions, lengths) device tensors from segment end points (vortex.py:460-465)."""
    torch = _torch()
    a, b = _f64(torch, starts), _f64(torch, ends)
    m = a.shape[0]
    d = torch.zeros((m, 3), dtype=torch.float64, device="cuda")
    L = torch.zeros(m, dtype=torch.float64, device="cuda")
    status = _lib.load().ofw_segment_geometry(_ctx(), m, _ptr(a), _ptr(b), _ptr(d), _ptr(L))
    _lib.check(status, "ofw_segment_geometry")
    return d, L


def direct_vjp(points, starts, ends, gamma, radius, model, grad_out, ex_ptr=None, ex_seg=None, fused_forward=False,
               want_points=True, precision="fp64"):
    """Adjoint of the direct sum on device tensors.

    ``precision`` 'fp64', or 'mixed' (far pairs in FP32, near pairs and accumulation in FP64).  Returns
    (grad_gamma (m,), grad_starts (m,3), grad_ends (m,3), grad_points (n,3) or None, out (n,3) or None).
    """
    if precision not in ("fp64", "mixed"):
        raise ValueError("the adjoint runs in precision 'fp64' or 'mixed'")
    torch = _torch()
    pts, a, b, g, r, go = (_f64(torch, v) for v in (points, starts, ends, gamma, radius, grad_out))
    n, m = pts.shape[0], a.shape[0]
    if ex_ptr is None:
        ex_ptr = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
        ex_seg = torch.zeros(1, dtype=torch.int64, device="cuda")
    gg = torch.zeros(m, dtype=torch.float64, device="cuda")
    ga = torch.zeros((m, 3), dtype=torch.float64, device="cuda")
    gb = torch.zeros((m, 3), dtype=torch.float64, device="cuda")
    gx = torch.zeros((n, 3), dtype=torch.float64, device="cuda") if want_points else None
    out = torch.zeros((n, 3), dtype=torch.float64, device="cuda") if fused_forward else None
    status = _lib.load().ofw_direct_vjp(_ctx(), _lib.PRECISIONS[precision], n, _ptr(pts), m, _ptr(a), _ptr(b), _ptr(g),
                                        _ptr(r), int(model), _ptr(ex_ptr), _ptr(ex_seg), _ptr(go), _ptr(gg), _ptr(ga),
                                        _ptr(gb), _ptr(gx) if gx is not None else None,
                                        _ptr(out) if out is not None else None)
    _lib.check(status, "ofw_direct_vjp")
    return gg, ga, gb, gx, out


def _forward(pts, a, b, g, r, model, ex_ptr, ex_seg, singularity, precision):
    torch = _torch()
    n, m = pts.shape[0], a.shape[0]
    d, L = segment_geometry(a, b)
    out = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    flags = torch.zeros(max(n, 1), dtype=torch.int8, device="cuda")
    fseg = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    status = _lib.load().ofw_direct_sum(_ctx(), _lib.PRECISIONS[precision], 128, n, _ptr(pts), m, _ptr(a), _ptr(d),
                                        _ptr(L), _ptr(g), _ptr(r), int(model), int(singularity == "raise"),
                                        _ptr(ex_ptr), _ptr(ex_seg), _ptr(out), _ptr(flags), _ptr(fseg))
    _lib.check(status, "ofw_direct_sum")
    if bool(flags[:n].any()):
        from .vortex import FilamentSingularity, RANGE_MESSAGE, SINGULAR_MESSAGE
        if bool((flags[:n] == 1).any()):
            raise FilamentSingularity(SINGULAR_MESSAGE)
        raise ValueError(RANGE_MESSAGE)
    return out


def _autograd_function():
    torch = _torch()

    class InducedVelocity(torch.autograd.Function):
        """V(points; starts, ends, gamma) by the direct sum; backward by the adjoint kernel."""

        @staticmethod
        def forward(ctx, points, starts, ends, gamma, radius, model, ex_ptr, ex_seg, singularity, precision):
            out = _forward(points.detach(), starts.detach(), ends.detach(), gamma.detach(), radius.detach(), model,
                           ex_ptr, ex_seg, singularity, precision)
            ctx.save_for_backward(points, starts, ends, gamma, radius, ex_ptr, ex_seg)
            ctx.model = model
            ctx.precision = precision
            return out

        @staticmethod
        def backward(ctx, grad_out):
            points, starts, ends, gamma, radius, ex_ptr, ex_seg = ctx.saved_tensors
            # the adjoint of the forward's own arithmetic: FP64, or the mixed split of far and near pairs
            gg, ga, gb, gx, _ = direct_vjp(points, starts, ends, gamma, radius, ctx.model, grad_out.contiguous(),
                                           ex_ptr, ex_seg, want_points=ctx.needs_input_grad[0],
                                           precision=ctx.precision)
            return gx, ga, gb, gg, None, None, None, None, None, None

    return InducedVelocity


_FN = []


def induced_velocity(points, starts, ends, gamma, radius, model=0, exclusions=None, singularity="raise",
                     precision="fp64"):
    """Differentiable induced velocity (n,3) of straight segments starts -> ends at points (torch tensors).

    ``precision`` 'fp64' or 'mixed': the forward is the direct sum in that precision and the backward the
    adjoint of the same split (far pairs in FP32, near pairs in FP64 for 'mixed').
    """
    torch = _torch()
    if not _FN:
        _FN.append(_autograd_function())
    n, m = points.shape[0], starts.shape[0]
    if torch.is_tensor(gamma) and gamma.dim() == 0 or np.isscalar(gamma):
        gamma = torch.full((m,), float(gamma), dtype=torch.float64, device="cuda")
    if np.isscalar(radius) or (torch.is_tensor(radius) and radius.dim() == 0):
        radius = torch.full((m,), float(radius), dtype=torch.float64, device="cuda")
    ptr, seg = exclusion_csr(n, exclusions if exclusions is not None else np.zeros((0, 2), np.int64))
    ex_ptr = torch.from_numpy(ptr).to("cuda")
    ex_seg = torch.from_numpy(seg if len(seg) else np.zeros(1, np.int64)).to("cuda")
    if HipSettings(precision).precision not in ("fp64", "mixed"):
        raise ValueError("induced_velocity is differentiable in precision 'fp64' or 'mixed'")
    return _FN[0].apply(points, starts, ends, gamma, radius, int(model), ex_ptr, ex_seg, singularity, precision)
