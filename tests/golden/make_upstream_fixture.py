"""Extract the recorded welib velocities of the reference's kernel fixture into tests/golden.

Source (data only): tests/fixtures/upstream-kernel-check.json of the reference, which records the
velocity of the segment (0,0,-1)->(0,0,1), circulation 1, core radius 0.5, at several points for core
models 0-4 (used by the reference's tests/test_vortex.py:50-57, tolerance rtol 1e-12, atol 1e-14).
Run in the development container only:  python tests/golden/make_upstream_fixture.py
"""
import json
from pathlib import Path

SOURCE = Path("/root/reference/tests/fixtures/upstream-kernel-check.json")
TARGET = Path(__file__).resolve().parent / "upstream_kernel_check.json"

data = json.loads(SOURCE.read_text(encoding="utf-8"))
rows = [dict(point=c["point"], core_model_id=c["core_model_id"], core_radius_m=c["core_radius_m"],
             velocity_m_per_s=c["velocity_m_per_s"]) for c in data["direct_reference_results"]]
out = dict(provenance="reference tests/fixtures/upstream-kernel-check.json (direct_reference_results)",
           segment_start=data["segment_start"], segment_end=data["segment_end"],
           circulation_m2_per_s=data["circulation_m2_per_s"], cases=rows)
TARGET.write_text(json.dumps(out, indent=1) + "\n", encoding="utf-8")
print(f"wrote {len(rows)} cases to {TARGET}")
