"""ctypes binding of libofw_hip.so (the C ABI declared in include/ofw_hip.h).

The library is loaded from this package directory only.  There is no fallback:
when the library is missing or no device is present, every entry point raises.
"""
import ctypes
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libofw_hip.so"

OFW_OK, OFW_ERR_ARG, OFW_ERR_HIP, OFW_ERR_UNSUPPORTED, OFW_ERR_OVERFLOW = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": 0, "mixed": 1, "fp32": 2}

_c_i = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_d = ctypes.c_double
_c_p = ctypes.c_void_p

# name -> (restype, argtypes); every symbol of include/ofw_hip.h
SIGNATURES = {
    "ofw_abi_version": (_c_i, []),
    "ofw_last_error": (ctypes.c_char_p, []),
    "ofw_ctx_create": (_c_i, [ctypes.POINTER(_c_p), _c_p]),
    "ofw_ctx_destroy": (_c_i, [_c_p]),
    "ofw_ctx_set_stream": (_c_i, [_c_p, _c_p]),
    "ofw_ctx_timings": (_c_i, [_c_p, ctypes.POINTER(ctypes.c_float), _c_i]),
    "ofw_device_info": (_c_i, [ctypes.c_char_p, _c_i, ctypes.POINTER(_c_i)]),
    "ofw_direct_sum": (_c_i, [_c_p, _c_i, _c_i, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i, _c_i,
                              _c_p, _c_p, _c_p, _c_p, _c_p]),
    "ofw_direct_sum_static": (_c_i, [_c_p, _c_i, _c_i, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i,
                                     _c_i, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "ofw_ctx_stats": (_c_i, [_c_p, ctypes.POINTER(_c_i64), _c_i]),
    "ofw_build_tree": (_c_i, [_c_p, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, ctypes.POINTER(_c_i64)]),
    "ofw_tree_nodes": (_c_i, [_c_p, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i64] + [_c_p] * 9
                       + [ctypes.POINTER(_c_i64)]),
    "ofw_tree_sum": (_c_i, [_c_p, _c_i, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i, _c_i, _c_p, _c_p,
                            _c_d, _c_i, _c_i64, _c_d, _c_i, _c_p, _c_p, _c_p, _c_p]),
    "ofw_tree_batch_create": (_c_i, [_c_p, ctypes.POINTER(_c_p), _c_i, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                     _c_p, _c_p, _c_i, _c_i, _c_p, _c_p, _c_d, _c_i, _c_i64, _c_d, _c_i, _c_i, _c_p,
                                     _c_p, _c_p, _c_p]),
    "ofw_tree_batch_run": (_c_i, [_c_p]),
    "ofw_tree_batch_timings": (_c_i, [_c_p, ctypes.POINTER(ctypes.c_float), _c_i]),
    "ofw_tree_batch_stats": (_c_i, [_c_p, ctypes.POINTER(_c_i64), _c_i]),
    "ofw_tree_batch_destroy": (_c_i, [_c_p]),
    "ofw_pair_matrix": (_c_i, [_c_p, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_i, _c_i, _c_p, _c_p]),
    "ofw_assemble_columns": (_c_i, [_c_p, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p]),
    "ofw_segment_geometry": (_c_i, [_c_p, _c_i64, _c_p, _c_p, _c_p, _c_p]),
    "ofw_direct_vjp": (_c_i, [_c_p, _c_i, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p,
                              _c_p, _c_p, _c_p]),
    "ofw_pack_ring_edges": (_c_i, [_c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_i64,
                                   ctypes.POINTER(_c_i64)]),
    "ofw_edge_radii": (_c_i, [_c_p, _c_i64, _c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _c_p, _c_d,
                              _c_p]),
    "ofw_incident_pairs": (_c_i, [_c_p, _c_i64, _c_p, _c_i64, _c_p, _c_i64, _c_p, _c_i64, ctypes.POINTER(_c_i64)]),
    "ofw_network_segments": (_c_i, [_c_p, _c_i64, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_p]),
    "ofw_exclusion_subset": (_c_i, [_c_p, _c_i64, _c_p, _c_i64, _c_i64, _c_i64, _c_p, _c_i64, _c_p, _c_i, _c_p,
                                    _c_p]),
    "ofw_stage_combine": (_c_i, [_c_p, _c_i, _c_i64, _c_d, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "ofw_step_create": (_c_i, [_c_p, ctypes.POINTER(_c_p), _c_i, _c_i64, _c_p, _c_p, _c_i64, _c_p, _c_p, _c_i64, _c_p,
                               _c_p, _c_p, _c_i, _c_i64, _c_i64, _c_i64, _c_d, _c_i, _c_i64, _c_d, _c_p, _c_p, _c_d,
                               _c_d, _c_i, _c_i, _c_i, _c_i, _c_i, ctypes.POINTER(_c_i64)]),
    "ofw_step_run": (_c_i, [_c_p, ctypes.POINTER(_c_i64)]),
    "ofw_step_timings": (_c_i, [_c_p, ctypes.POINTER(ctypes.c_float)]),
    "ofw_step_destroy": (_c_i, [_c_p]),
}
STAGE_STEP, STAGE_HALF, STAGE_HEUN, STAGE_RK4, STAGE_ADD = 0, 1, 2, 3, 4
(STEP_PAIR_OVERFLOW, STEP_EDGE_INDEX, STEP_ZERO_LENGTH, STEP_EXCLUSION, STEP_TREE_OVERFLOW, STEP_ROW_FLAG,
 STEP_RING_FLAG, STEP_NONFINITE) = (1, 2, 4, 8, 16, 32, 64, 128)

_lib = None


class LibraryError(RuntimeError):
    """The HIP library failed (missing, no device, or a HIP runtime error)."""


def load():
    """The loaded library (raises LibraryError when it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryError(f"{LIB_PATH} is missing: build it with `python -m openfawe_amd.build`")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
    return _lib


def last_error():
    msg = load().ofw_last_error()
    return msg.decode() if msg else ""


def check(status, what):
    if status == OFW_OK:
        return
    message = f"{what}: {last_error()}"
    if status == OFW_ERR_ARG or status == OFW_ERR_UNSUPPORTED:
        raise ValueError(message)
    if status == OFW_ERR_OVERFLOW:
        raise RuntimeError(message)
    raise LibraryError(message)


def timings(ctx, count=8):
    """Phase timings (ms) of the context's last call."""
    buf = (ctypes.c_float * count)()
    phases = load().ofw_ctx_timings(ctx, buf, count)
    if phases < 0:
        check(phases, "ofw_ctx_timings")
    return [float(buf[k]) for k in range(min(phases, count))]


_contexts = {}


def context(stream_handle, device_index):
    """The library context bound to (device, stream); created on first use."""
    key = (int(device_index), int(stream_handle))
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = _c_p()
        check(load().ofw_ctx_create(ctypes.byref(ctx), _c_p(int(stream_handle))), "ofw_ctx_create")
        _contexts[key] = ctx
    return ctx
