"""Ghost-cell IBM cylinder (examples/cylinder_drag_validation) end to end on the GPU.

The reference's own case (440 x 82, D/dx = 20, first-order HLLC + viscous + IBM +
sponge, 44 000 RK substeps, Cd averaged over the last 10 % of substeps) publishes
marker Cd = 5.6312 (examples/make_validation_plots.py:104-107,
examples/cylinder_drag_validation/config.toml:26) with a 2 % steadiness gate.
"""
import pytest

from dev_field_efvm_amd.cylinder import CylinderCase, run_substeps_reference

pytestmark = pytest.mark.gpu


def test_reference_cylinder_cd_matches_published():
    case = CylinderCase(nx=440, ny=82, precision="fp64")
    r = run_substeps_reference(case, 44000)
    print(f"Cd {r['cd']:.5f} Cl {r['cl']:.5f} ptp {100 * r['ptp']:.3f}% t {r['time']:.5f} wall {r['wall_s']:.1f}s")
    assert abs(r["cd"] - 5.6312) <= 2e-3 * 5.6312
    assert r["ptp"] <= 0.02
    assert abs(r["cd"] - 5.5795) / 5.5795 <= 0.05      # the DFG benchmark, 5 % gate (config.toml:76)


def test_reference_cylinder_fp32_tracks_fp64():
    r64 = run_substeps_reference(CylinderCase(nx=440, ny=82, precision="fp64"), 6000)
    r32 = run_substeps_reference(CylinderCase(nx=440, ny=82, precision="fp32"), 6000)
    print(f"6000 substeps: Cd fp64 {r64['cd']:.6f} fp32 {r32['cd']:.6f}")
    assert abs(r32["cd"] - r64["cd"]) <= 1e-3 * abs(r64["cd"])
