"""The C ABI: the library loads (no GPU needed) and exports every symbol include/ofw_hip.h declares; the
Python API validates its inputs as the reference does (CPU only, no compute calls)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from openfawe_amd import _lib
from openfawe_amd.vortex import CoreModel, HipSettings, TreeSettings, exclusion_csr, segments_velocity

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include/ofw_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ofw_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 6
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SIGNATURES), "the ctypes signature table must cover the header exactly"
    assert _lib.load().ofw_abi_version() == 3


def test_settings_validation():
    for bad in (dict(theta=1.), dict(theta=0.), dict(order=3), dict(leaf_size=0), dict(core_clearance=-1.)):
        with pytest.raises(ValueError):
            TreeSettings(**bad)
    for bad in (dict(precision="fp16"), dict(block=48), dict(suspect="gpu64", precision="fp64"), dict(build="x")):
        with pytest.raises(ValueError):
            HipSettings(**bad)


def test_invalid_inputs_raise_before_any_device_work():
    A, B = np.array([0., 0., -1.]), np.array([0., 0., 1.])
    default = dict(points=[[1, 0, 0]], nodes=[A, B], connectivity=[[0, 1]], circulation=1)
    bad = [dict(points=[1, 0, 0]), dict(points=[[np.nan, 0, 0]]), dict(nodes=[[0, 0, 0], [0, 0, 0]]),
           dict(connectivity=[[0., 1.]]), dict(connectivity=[[False, True]]), dict(connectivity=[[0, 2]]),
           dict(connectivity=[[0, -1]]), dict(circulation=np.inf), dict(circulation=[1, 2]),
           dict(core_radius=-1), dict(core_radius=[1, 2]), dict(model=6), dict(model=2.), dict(model=True),
           dict(exclusions=[[1, 0]]), dict(exclusions=[[0, 1]]), dict(exclusions=[[0., 0.]]),
           dict(chunk_size=0), dict(chunk_size=True), dict(chunk_size=1.1), dict(singularity="ignore"),
           dict(points=[[1j, 0, 0]]), dict(points=[["1", "0", "0"]]), dict(backend="cuda")]
    for change in bad:
        with pytest.raises(ValueError):
            segments_velocity(**(default | change))


def test_exclusion_csr_matches_unique_pairs():
    rng = np.random.default_rng(4)
    pairs = np.stack([rng.integers(0, 30, 400), rng.integers(0, 90, 400)], axis=1)
    ptr, seg = exclusion_csr(30, pairs)
    uniq = np.unique(pairs, axis=0)
    assert np.array_equal(seg, uniq[:, 1])
    assert np.array_equal(ptr, np.concatenate([[0], np.cumsum(np.bincount(uniq[:, 0], minlength=30))]))
    ptr, seg = exclusion_csr(5, np.zeros((0, 2), np.int64))
    assert np.array_equal(ptr, np.zeros(6)) and len(seg) == 0
