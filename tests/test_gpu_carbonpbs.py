"""The CarbonPBS Python contract on the GPU: cuFinalDose (and its aliases calcDose / cuCalDose)
on a CarbonPBS-style beam (WEQ volume, texel-coded spots, IDD table, per-spot directions) against
the FP64 oracle on the same context and placeholder CT, within 0.5 % of Dmax; flat and x-major
output buffers hold the same dose (flat ones accumulate); cuCalDose3 zeroes its CSC outputs and returns the dose."""
import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import carbonpbs as CP, fixtures as F

from .test_carbonpbs import carbonpbs_args

gpu = pytest.mark.gpu


@gpu
def test_cu_final_dose_against_oracle_and_layouts():
    case = F.make_weq_case(index_spots=True)
    args = carbonpbs_args(case)
    nx, ny, nz = case["dose_dims"]
    print("\n[gpu-test] cuFinalDose on the WEQ-index case", flush=True)
    flat = np.zeros(nx * ny * nz, np.float32)
    CP.cuFinalDose(flat, **args)
    zyx = flat.reshape(nz, ny, nx).copy()
    # the grid is cubic (64^3): a 3-D buffer is read as (Nx, Ny, Nz), as in the reference
    xyz = np.zeros((nx, ny, nz), np.float32)
    CP.calcDose(xyz, **args)
    assert np.array_equal(xyz.transpose(2, 1, 0), zyx)
    CP.cuCalDose(xyz, **args)  # (Nx, Ny, Nz) buffers are overwritten, not accumulated
    assert np.array_equal(xyz.transpose(2, 1, 0), zyx)
    ctx = CP.build_carbonpbs_context(**args)
    one = {"ct": CP.placeholder_ct(ctx["dims"], ctx["beam"]["roi_linear_indices"]).reshape(nz, ny, nx),
           "ct_dims": ctx["dims"], "ct_resolution": ctx["resolution"], "ct_corner": ctx["corner"],
           "dose_dims": ctx["dims"], "dose_resolution": ctx["resolution"], "dose_corner": ctx["corner"],
           "beams": [ctx["beam"]], "energy": ctx["energy"]}
    ref, _ = O.beam_dose(one)
    err = float(np.abs(zyx - ref).max() / ref.max())
    print(f"cuFinalDose vs FP64 oracle: max|dD|/Dmax = {err:.3e}")
    assert err <= 5e-3
    # accumulate semantics of the flat buffer
    CP.cuFinalDose(flat, **args)
    np.testing.assert_allclose(flat.reshape(nz, ny, nx), 2.0 * zyx, rtol=1e-6, atol=1e-12)


@gpu
def test_cu_cal_dose3_zeroes_csc():
    case = F.make_weq_case(index_spots=True)
    a = carbonpbs_args(case)
    csc = [np.ones(10, np.float32), np.ones(5, np.int32), np.ones(10, np.int32), np.ones(1, np.uint64)]
    res = CP.cuCalDose3(csc[0], csc[1], csc[2], a["rayweq"], a["roiIdx"], a["all_energies"], a["sourcePos"],
                        a["tmpBeamDir"], a["bmxdir"], a["bmydir"], a["corner"], a["resolution"], a["dims"],
                        a["longitudalCutoff"], a["enelist"], a["idddata"], a["iddsetting"], a["profiledata"],
                        a["profilesetting"], a["beamparadata"], a["subspotdata"], a["layerInfo"], a["layerEnergy"],
                        csc[3], a["idbeamxy"], a["nPar"], None, a["sad"], None, None)
    assert all(not c.any() for c in csc)
    nx, ny, nz = case["dose_dims"]
    assert res["dose_grid"].shape == (nx, ny, nz) and res["dose_grid"].max() > 0
