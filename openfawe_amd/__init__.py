"""openfawe_amd: OpenFAWE's free-wake induced-velocity path on AMD Instinct MI355X (gfx950).

The Biot-Savart sum over straight vortex segments (direct and tree code) in hand-written HIP kernels
behind a C ABI (include/ofw_hip.h), with the reference's batch-backend API on top
(``openfawe_amd.vortex``).
"""
from .vortex import (CoreModel, FilamentSingularity, HipSettings, TreeSettings, exclusion_csr, hip_sum,
                     ring_velocity, segment_velocity, segments_velocity)

__all__ = ["CoreModel", "FilamentSingularity", "HipSettings", "TreeSettings", "exclusion_csr", "hip_sum",
           "ring_velocity", "segment_velocity", "segments_velocity"]
