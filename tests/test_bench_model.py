"""The counting models behind bench.py's roofline figures, checked against
closed forms and against the committed rocprofv3 counter summaries (CPU only).

- Algorithmic bytes of the 3-D gather kernels on the 64^3 cube: the per-element
  and per-point terms written out, and the survey's 143 B per point-step; the
  M^-1 bytes of the update against a brute-force count of the rows it reads.
- FP32 operations of elemforce_3d: 220 per local point; the committed VALU
  instruction counts (FMA counted twice), scaled to the 125 active lanes of
  each 128-thread workgroup, agree within 1 %.
- The committed HBM byte counters: the element-force writes are exactly the
  1536 B per element of the force blocks, and the measured traffic bench.py
  reports is FETCH x 2 + WRITE x 1 in KiB.
"""
import json
import os

import pytest

import bench

N = 64
NSPEC = N ** 3
NGLOB = (4 * N + 1) ** 3
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_algorithmic_bytes_3d_closed_form():
    ab = bench.algorithmic_bytes_3d(NSPEC, NGLOB)
    assert NGLOB == 16974593
    assert ab["elemforce"] == NSPEC * 125 * (8 + 12) + NGLOB * 12
    assert ab["update"] == NSPEC * 125 * 12 + NGLOB * 60
    assert abs(ab["survey_per_pt_step"] - 143.16) < 0.01
    assert abs(ab["per_pt_step"] - (ab["elemforce"] + ab["update"]) / NGLOB) < 1e-9


def test_algorithmic_bytes_3d_element_properties():
    """kappa and mu read once per element: 8 B per element instead of 1000."""
    per_point = bench.algorithmic_bytes_3d(NSPEC, NGLOB)
    per_elem = bench.algorithmic_bytes_3d(NSPEC, NGLOB, element_properties=True)
    assert per_point["elemforce"] - per_elem["elemforce"] == NSPEC * (125 * 8 - 8)
    assert per_elem["update"] == per_point["update"]


def test_algorithmic_bytes_2d_layouts():
    """The 2-D element-force bytes for each geometry source and property layout,
    written out per element (25 local points) and per point."""
    nspec, nglob = 4000 ** 2, 16001 ** 2
    for geometry, gbytes in (("point", 500), ("element", 32), ("control_nodes_4", 32), ("control_nodes_9", 72)):
        for eprops, pbytes in ((False, 200), (True, 8)):
            ab = bench.algorithmic_bytes_2d(nspec, nglob, geometry, eprops, scatter="gather")
            assert ab["stiffness_2d"] == nspec * (25 * 12 + pbytes + gbytes) + nglob * 8
            assert ab["update_fused"] == nspec * 25 * 12 + nglob * 44
            col = bench.algorithmic_bytes_2d(nspec, nglob, geometry, eprops, scatter="colored")
            assert col["stiffness_2d"] == nspec * (25 * 4 + pbytes + gbytes) + nglob * 24
            assert col["update_fused"] == nglob * 56
    # the data model: 130 B per point-step with per-point geometry (25/16 local points per point)
    model = bench.algorithmic_bytes_2d(nspec, nglob)["model_per_pt_step"]
    assert abs(model - (25 * 32 * nspec + 80 * nglob) / nglob) < 1e-9


def _minv_bytes_brute(nx, ny, nz, absorbing):
    total = 0
    for tz in range(nz + 1):
        for ty in range(ny + 1):
            for tx in range(nx + 1):
                face = ((absorbing[0] and tx == 0) or (absorbing[1] and tx == nx) or (absorbing[2] and ty == 0)
                        or (absorbing[3] and ty == ny) or (absorbing[4] and tz == 0) or (absorbing[5] and tz == nz))
                for k in range(4):
                    for j in range(4):
                        if (ty == ny and j > 0) or (tz == nz and k > 0):
                            continue
                        total += 48 if face else 16
    return total


@pytest.mark.parametrize("dims", [(3, 2, 4), (5, 3, 7), (1, 1, 1)])
@pytest.mark.parametrize("absorbing", [(1, 1, 1, 1, 1, 0), (0,) * 6, (1,) * 6, (0, 1, 0, 1, 1, 0)])
def test_minv_bytes_match_the_rows_the_update_reads(dims, absorbing):
    assert bench.minv_bytes_3d(*dims, list(absorbing)) == _minv_bytes_brute(*dims, absorbing)


def test_elemforce_flops_match_the_counters():
    flops = bench.elemforce_3d_flops(NSPEC)
    assert flops == NSPEC * 125 * 220
    counted = bench.measured_valu_flops("3d_layered_cube_64^3_ngll5", "gather", "elemforce_3d_kernel")
    assert counted is not None
    active = counted * 125 / 128
    assert abs(active / flops - 1) < 0.01


def test_committed_byte_counters():
    with open(os.path.join(ROOT, "profiles", "gather3d_64cube_pmc.json")) as fh:
        pmc = json.load(fh)
    assert pmc["scale"] == {"FETCH_SIZE": 2.0, "WRITE_SIZE": 1.0, "basis": pmc["scale"]["basis"]}
    ef = pmc["kib_per_dispatch"]["elemforce_3d_kernel"]
    assert ef["WRITE_SIZE"] * 1024 == NSPEC * 1536
    t = bench.measured_traffic("3d_layered_cube_64^3_ngll5", "gather", "assemble_update_3d_kernel")
    au = pmc["kib_per_dispatch"]["assemble_update_3d_kernel"]
    assert t == 1024.0 * (2.0 * au["FETCH_SIZE"] + au["WRITE_SIZE"])
    # traffic above the algorithmic bytes is over-fetch, below it cache hits; both within 25 %
    alg = bench.algorithmic_bytes_3d(NSPEC, NGLOB, bench.minv_bytes_3d(N, N, N, [1, 1, 1, 1, 1, 0]))["update"]
    assert 0.75 < t / alg < 1.25
    assert bench.measured_traffic("other_workload", "gather", "assemble_update_3d_kernel") is None
