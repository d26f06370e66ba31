"""Carbon-ion beams on the GPU against the FP64 oracle, on the project-defined 12C tables
(``subsecondpbs_amd.carbon``; not reference data, every result here uses that project LUT).

The four carbon beams of the batching case (config-1 CT, 20 layers within 100-400 MeV/u):
plan integers, transfer box, tile radii (the oracle's float rule applied to the GPU's sigma),
radius histograms and batched launch radii (the oracle's own plan) bit-exact; dose within 0.5 % of Dmax with the halo branch off and on (FLUKA constants; halo
radius histograms bit-exact under the same rule); and the whole 16-beam mixed batch (12 proton +
4 carbon beams in one ``Batch``, the carbon tables as a second table set) against the sum of the
oracle's per-beam doses.
"""
import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import fixtures as F

gpu = pytest.mark.gpu
DOSE_TOL = 5e-3


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def config3():
    case = F.make_config3()
    print(f"\n[gpu-test] config3: {len(case['beams'])} beams, kinds {case['kinds']}", flush=True)
    return case


def _engine_dose(one, model):
    torch = _torch()
    from subsecondpbs_amd.engine import Engine
    eng = Engine(one["energy"])
    eng.set_ct(one["ct"], one["ct_dims"], one["ct_resolution"], one["ct_corner"])
    eng.set_dose_grid(one["dose_dims"], one["dose_resolution"], one["dose_corner"])
    eng.set_nuclear(model)
    nx, ny, nz = one["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(one["beams"][0], dose)
    eng.sync()
    ints = eng.debug_fetch(("radHist", "iddSigma", "tileRad"))
    ints["effRad"] = eng.debug_bins()
    halo = eng.debug_halo() if model else None
    out = dose.cpu().numpy()
    eng.close()
    return out, ints, halo


@gpu
@pytest.mark.parametrize("i", [12, 13, 14, 15])
def test_carbon_beam_plain(config3, i):
    one = F.case_for_beam(config3, i)
    ref, p = O.beam_dose(one)
    dose, dbg, _ = _engine_dose(one, 0)
    ints = dbg["ints"]
    assert ints["rayDims"] == tuple(p.rayDims) and ints["steps"] == p.steps
    assert ints["beamFirstInside"] == p.beamFirstInside and ints["beamFirstOutside"] == p.beamFirstOutside
    L = ints["numLayers"]
    np.testing.assert_array_equal(ints["afterLast"], np.array(p.afterLast[:L]))
    np.testing.assert_array_equal(ints["layerPassive"], np.array(p.layerPassive[:L]))
    np.testing.assert_array_equal(ints["energyIdx"], np.array(p.energyIdx[:L], np.float32))
    # tile radii: the oracle's float rule applied to the GPU's own rSigmaEff gives the GPU's radii
    sx, sy = ints["superpDims"]
    first = ints["beamFirstInside"]
    for l in range(L):
        k1 = int(ints["layerPassive"][l])
        if k1 <= first:
            continue
        rs = np.ascontiguousarray(dbg["iddSigma"][l, :, :, :, 1])
        want = np.zeros((k1 - first, sy // 8, sx // 32), np.int32)
        O.lib().orc_tile_radii_f32(O._fp(rs), sx, sy, first, k1, O._ip(want))
        np.testing.assert_array_equal(dbg["tileRad"][l, first:k1] & 255, want)
    # radius histograms and batched launch radii: the oracle's own plan (its FP64 sigma)
    np.testing.assert_array_equal(dbg["radHist"], np.array([list(p.radHist[l]) for l in range(L)], np.int32))
    eff = dbg["effRad"]
    for l in range(L):
        e = np.zeros(34, np.int32)
        bt = np.zeros(33, np.int32)
        O.lib().orc_batch_bins(O._ip(np.ascontiguousarray(np.array(p.radHist[l]), np.int32)), O._ip(bt), O._ip(e))
        np.testing.assert_array_equal(eff[l], e, err_msg=f"layer {l}")
        np.testing.assert_array_equal(bt, np.array(p.batchRad[l]), err_msg=f"layer {l}")
    assert ints["transferStart"] == tuple(p.transferStart) and ints["transferMax"] == tuple(p.transferMax)
    err = _rel(dose, ref)
    print(f"carbon beam {i}: max|dD|/Dmax = {err:.3e} (project-defined 12C tables)")
    assert err <= DOSE_TOL


@gpu
@pytest.mark.parametrize("i", [12, 15])
def test_carbon_beam_with_halo(config3, i):
    one = F.case_for_beam(config3, i)
    model = config3["carbon_nuclear"]
    ref, p = O.beam_dose(one, nuclear=model)
    dose, dbg, halo = _engine_dose(one, model)
    L = dbg["ints"]["numLayers"]
    assert halo["nucDims"] == tuple(p.nucDims) and halo["overflow"] == 0
    nx, ny = halo["nucDims"]
    first = dbg["ints"]["beamFirstInside"]
    for l in range(L):
        k1 = int(dbg["ints"]["layerPassive"][l])
        if k1 <= first:
            continue
        rs = np.ascontiguousarray(halo["iddSigma"][l, :, :, :, 1])
        want = np.zeros((k1 - first, ny // 8, nx // 32), np.int32)
        O.lib().orc_tile_radii_f32(O._fp(rs), nx, ny, first, k1, O._ip(want))
        np.testing.assert_array_equal(halo["radHist"][l], np.bincount(want.ravel(), minlength=34))
    err = _rel(dose, ref)
    print(f"carbon beam {i} with halo (model {model}): max|dD|/Dmax = {err:.3e}")
    assert err <= DOSE_TOL


@gpu
def test_mixed_batch_config3(config3):
    torch = _torch()
    from subsecondpbs_amd.batch import Batch
    case = config3
    nx, ny, nz = case["dose_dims"]
    total = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    bt = Batch(case["energy"], lanes=4)
    carbon = bt.add_table(case["carbon_energy"], case["carbon_nuclear"])
    bt.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    bt.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    bt.dose(case["beams"], total, tables=[carbon if k == "carbon" else 0 for k in case["kinds"]])
    bt.sync()
    launches = bt.last_launches
    bt.close()
    print(f"config3 mixed batch: {launches} launches for {len(case['beams'])} beams")
    assert 0 < launches <= 40
    got = total.cpu().numpy()
    ref = np.zeros((nz, ny, nx))
    for i, kind in enumerate(case["kinds"]):
        O.beam_dose(F.case_for_beam(case, i), dose=ref, nuclear=case["carbon_nuclear"] if kind == "carbon" else 0)
    err = _rel(got, ref)
    print(f"config3 mixed batch (12 p + 4 C): max|dD|/Dmax = {err:.3e}")
    assert err <= DOSE_TOL
