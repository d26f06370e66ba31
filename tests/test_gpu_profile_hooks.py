"""CarbonPBS profile hooks on the GPU against the FP64 oracle (config-1 beam with a synthetic
two-Gaussian profile table and beam parameters; raytracedicom_wrapper.cu:2080-2188, 4604-4832).

Hooks off (the reference's setting) with the arrays present: the dose equals the plain beam's
bitwise. Sigma override, overall weight and both: radius histograms equal the oracle's and the
dose agrees within 0.5 % of Dmax."""
import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import fixtures as F

from .test_profile_hooks import profile_beam

gpu = pytest.mark.gpu
DOSE_TOL = 5e-3


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _engine_dose(case, beam):
    torch = _torch()
    from subsecondpbs_amd.engine import Engine
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(beam, dose)
    eng.sync()
    dbg = eng.debug_fetch(("radHist",))
    out = dose.cpu().numpy()
    eng.close()
    return out, dbg


@gpu
@pytest.mark.parametrize("hooks", [0, 1, 2, 3])
def test_profile_hooks_against_oracle(hooks):
    case = F.make_config1()
    beam = case["beams"][0]
    L = len(beam["energies"])
    b = profile_beam(beam, L, hooks=hooks)
    c = dict(case)
    c["beams"] = [b]
    ref, p = O.beam_dose(c)
    assert p.profileHooks == hooks
    dose, dbg = _engine_dose(case, b)
    if hooks == 0:
        plain, _ = _engine_dose(case, beam)
        assert np.array_equal(dose, plain)
    np.testing.assert_array_equal(dbg["radHist"], np.array([list(p.radHist[l]) for l in range(L)], np.int32))
    err = float(np.abs(dose - ref).max() / ref.max())
    print(f"profile hooks {hooks}: max|dD|/Dmax = {err:.3e}")
    assert err <= DOSE_TOL
