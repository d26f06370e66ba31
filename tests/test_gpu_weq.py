"""WEQ-volume input and WEQ-index spots on the GPU against the FP64 oracle (the CarbonPBS
contract: water-equivalent depths replace the CT trace; spot positions may be WEQ texel
coordinates). Integers (grids, steps, start plane, entry / exit per ray, cut-offs, passive steps,
transfer box) bit-exact; density and WEPL to float rounding; dose within 0.5 % of Dmax; the
Python contract takes the same dict keys; the adjoint refuses index-coded spots."""
import numpy as np
import pytest

from subsecondpbs_amd import fixtures as F, rtd

from .oracle_stages import run as oracle_run

gpu = pytest.mark.gpu
DOSE_TOL = 5e-3


def _engine(case):
    from subsecondpbs_amd.engine import Engine

    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    return eng


@gpu
@pytest.mark.parametrize("index_spots", [False, True])
def test_weq_case_against_oracle(index_spots):
    import torch

    case = F.make_weq_case(index_spots=index_spots)
    print(f"\n[gpu-test] {case['name']}", flush=True)
    ref = oracle_run(case)
    eng = _engine(case)
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(case["beams"][0], dose)
    eng.sync()
    dbg = eng.debug_fetch(("density", "wepl", "fi", "fo"))
    ints, p = dbg["ints"], ref["plan"]
    assert ints["rayDims"] == p["rayDims"] and ints["steps"] == p["steps"]
    assert ints["fanCorner"][2] == pytest.approx(67.0)
    assert ints["beamFirstInside"] == p["beamFirstInside"] and ints["beamFirstOutside"] == p["beamFirstOutside"]
    assert ints["firstPastCutoffAll"] == p["firstPastCutoffAll"]
    np.testing.assert_array_equal(dbg["fi"], ref["fi"])
    np.testing.assert_array_equal(dbg["fo"], ref["fo"])
    np.testing.assert_allclose(dbg["wepl"], ref["wepl"], rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(dbg["density"], ref["density"], rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(ints["afterLast"], p["afterLast"])
    np.testing.assert_array_equal(ints["layerPassive"], p["layerPassive"])
    assert ints["transferStart"] == p["transferStart"] and ints["transferMax"] == p["transferMax"]
    got = dose.cpu().numpy()
    err = float(np.abs(got - ref["dose"]).max() / ref["dose"].max())
    print(f"{case['name']}: max|dD|/Dmax = {err:.3e}")
    assert err <= DOSE_TOL
    if index_spots:
        g = torch.ones_like(dose)
        with pytest.raises(RuntimeError, match="WEQ indices"):
            eng.beam_adjoint(case["beams"][0], g)
    eng.close()


@gpu
def test_python_contract_takes_the_weq_keys():
    import oracle as O

    case = F.make_weq_case(index_spots=True)
    d = rtd.raytracedicom_wrapper(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"],
                                  case["dose_dims"], case["dose_resolution"], case["dose_corner"],
                                  case["beams"][0], case["energy"])
    ref, _ = O.beam_dose(case)
    assert float(np.abs(d - ref).max() / ref.max()) <= DOSE_TOL
