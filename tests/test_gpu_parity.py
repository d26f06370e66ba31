"""HIP engine vs the FP64 oracle, stage by stage, through the C ABI.

Integers (grid sizes, entry/exit steps, cut-offs, passive steps, tile radii, radius
histograms, transfer box) must match exactly; floats within the tolerances written here.
End-to-end dose: max |dD| <= 0.5 % of Dmax.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import _abi, fixtures as F

from .oracle_stages import run as oracle_run

gpu = pytest.mark.gpu
DOSE_TOL = 5e-3  # max |dD| / Dmax


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _engine_run(case, beam_index=0, use_subspots=False):
    torch = _torch()
    from subsecondpbs_amd.engine import Engine
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(case["beams"][beam_index], dose, use_subspots=use_subspots)
    eng.sync()
    dbg = eng.debug_fetch()
    out = dose.cpu().numpy()
    eng.close()
    return out, dbg


def _rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.fixture(scope="module")
def config1():
    case = F.make_config1()
    ref = oracle_run(case)
    return case, ref


@pytest.fixture(scope="module")
def config2():
    case = F.make_config2()
    ref = oracle_run(case)
    return case, ref


def _stage_report(dbg, ref):
    """Per-stage relative errors (printed, for diagnosis)."""
    ints = dbg["ints"]
    first = ints["beamFirstInside"]
    rep = {
        "density": _rel_err(dbg["density"], ref["density"]),
        "wepl": _rel_err(dbg["wepl"], ref["wepl"]),
        "rayW": _rel_err(dbg["rayW"], ref["rayW"]),
        "fi": int((dbg["fi"] != ref["fi"]).sum()), "fo": int((dbg["fo"] != ref["fo"]).sum()),
        "firstPassive": int((dbg["firstPassive"] != ref["firstPassive"]).sum()),
    }
    worst = 0.0
    for l in range(ints["numLayers"]):
        k1 = ints["afterLast"][l]
        worst = max(worst, _rel_err(dbg["iddSigma"][l, first:k1, :, :, 0], ref["idd"][l, first:k1]))
    rep["idd"] = worst
    return rep


@gpu
def test_config1_integers_bit_exact(config1):
    case, ref = config1
    _, dbg = _engine_run(case)
    ints, p = dbg["ints"], ref["plan"]
    assert ints["rayDims"] == p["rayDims"] and ints["steps"] == p["steps"]
    assert ints["beamFirstInside"] == p["beamFirstInside"]
    assert ints["beamFirstOutside"] == p["beamFirstOutside"]
    assert ints["firstPastCutoffAll"] == p["firstPastCutoffAll"]
    assert ints["guaranteedPassive"] == p["guaranteedPassive"]
    np.testing.assert_array_equal(ints["afterLast"], p["afterLast"])
    np.testing.assert_array_equal(ints["layerPassive"], p["layerPassive"])
    np.testing.assert_array_equal(dbg["fi"], ref["fi"])
    np.testing.assert_array_equal(dbg["fo"], ref["fo"])
    np.testing.assert_array_equal(dbg["radHist"], p["radHist"])
    assert ints["transferStart"] == p["transferStart"]
    assert ints["transferMax"] == p["transferMax"]
    np.testing.assert_array_equal(ints["energyIdx"], p["energyIdx"])
    np.testing.assert_array_equal(ints["peakDepth"], p["peakDepth"])


@gpu
def test_config1_stage_floats(config1):
    case, ref = config1
    _, dbg = _engine_run(case)
    # trace: FP32 trilinear + compensated prefix vs FP64
    np.testing.assert_allclose(dbg["density"], ref["density"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dbg["wepl"], ref["wepl"], rtol=2e-6, atol=1e-4)
    np.testing.assert_allclose(dbg["weplMin"], ref["weplMin"], rtol=2e-6, atol=1e-4)
    # ray weights: relative to the largest weight
    assert _rel_err(dbg["rayW"], ref["rayW"]) < 1e-5
    # IDD per layer over the steps each layer defines
    ints = dbg["ints"]
    first = ints["beamFirstInside"]
    for l in range(ints["numLayers"]):
        k1 = ints["afterLast"][l]
        g_idd = dbg["iddSigma"][l, first:k1, :, :, 0]
        o_idd = ref["idd"][l, first:k1]
        assert _rel_err(g_idd, o_idd) < 2e-4, l
        g_rs = dbg["iddSigma"][l, first:k1, :, :, 1]
        o_rs = ref["rsigma"][l, first:k1]
        fin = np.isfinite(o_rs)
        np.testing.assert_array_equal(np.isfinite(g_rs), fin)
        np.testing.assert_allclose(g_rs[fin], o_rs[fin], rtol=1e-4)
    np.testing.assert_array_equal(dbg["firstPassive"], ref["firstPassive"])


@gpu
def test_config1_tile_radii_bit_exact_on_gpu_sigma(config1):
    """The oracle's float radius rule applied to the GPU's own rSigmaEff gives the GPU's radii."""
    case, _ = config1
    _, dbg = _engine_run(case)
    ints = dbg["ints"]
    sx, sy = ints["superpDims"]
    first = ints["beamFirstInside"]
    for l in range(ints["numLayers"]):
        k1 = ints["layerPassive"][l]
        if k1 <= first:
            continue
        rs = np.ascontiguousarray(dbg["iddSigma"][l, :, :, :, 1])
        want = np.zeros((k1 - first, sy // 8, sx // 32), np.int32)
        O.lib().orc_tile_radii_f32(O._fp(rs), sx, sy, first, k1, O._ip(want))
        got = dbg["tileRad"][l, first:k1] & 255
        np.testing.assert_array_equal(got, want)


@gpu
def test_config1_dose_within_half_percent(config1):
    case, ref = config1
    dose, dbg = _engine_run(case)
    d0 = ref["dose"]
    err = _rel_err(dose, d0)
    print(f"config1 max|dD|/Dmax = {err:.3e}  Dmax={d0.max():.4e}")
    assert err <= DOSE_TOL
    # integral dose agrees too
    assert abs(dose.sum() / d0.sum() - 1.0) < 1e-3


@gpu
def test_config1_subspot_mode(config1):
    case, _ = config1
    ref = oracle_run(case, use_subspots=True)
    dose, dbg = _engine_run(case, use_subspots=True)
    assert _rel_err(dbg["rayW"], ref["rayW"]) < 1e-5
    assert _rel_err(dose, ref["dose"]) <= DOSE_TOL


@gpu
def test_config1_repeatable_bitwise(config1):
    case, _ = config1
    d1, _ = _engine_run(case)
    d2, _ = _engine_run(case)
    assert np.array_equal(d1, d2)


@pytest.fixture(scope="module")
def config2_gpu(config2):
    case, ref = config2
    dose, dbg = _engine_run(case)
    return case, ref, dose, dbg


def _print_config2(dbg, ref, dose):
    ints, p = dbg["ints"], ref["plan"]
    print("config2 stages:", _stage_report(dbg, ref))
    print("config2 groups:", ints["numGroups"], "tail planes:", ints["tailPlanes"],
          "tail counts:", list(ints["groupTailCount"]))
    print(f"config2 max|dD|/Dmax = {_rel_err(dose, ref['dose']):.3e}")


@gpu
def test_config2_integers(config2_gpu):
    """Head-size phantom, beam along +x: every integer equals the FP64 oracle's, from the beam
    scalars to each ray's entry / exit step, each ray-layer's first passive step and each
    layer's radius histogram (the oracle's own plan, complete_superposition.cu:197-255)."""
    case, ref, dose, dbg = config2_gpu
    _print_config2(dbg, ref, dose)
    ints, p = dbg["ints"], ref["plan"]
    assert ints["rayDims"] == p["rayDims"] and ints["steps"] == p["steps"]
    for k in ("beamFirstInside", "beamFirstOutside", "firstPastCutoffAll", "guaranteedPassive"):
        assert ints[k] == p[k], k
    np.testing.assert_array_equal(ints["afterLast"], p["afterLast"])
    np.testing.assert_array_equal(ints["layerPassive"], p["layerPassive"])
    assert ints["transferStart"] == p["transferStart"] and ints["transferMax"] == p["transferMax"]
    assert ints["tailOverflow"] == 0
    np.testing.assert_array_equal(dbg["fi"], ref["fi"])
    np.testing.assert_array_equal(dbg["fo"], ref["fo"])
    np.testing.assert_array_equal(dbg["firstPassive"], ref["firstPassive"])
    np.testing.assert_array_equal(dbg["radHist"], p["radHist"])


@gpu
def test_config2_stage_floats(config2_gpu):
    case, ref, dose, dbg = config2_gpu
    # WEPL relative to the deepest WEPL; density over the samples away from CT knife edges
    assert _rel_err(dbg["wepl"], ref["wepl"]) < 1e-4
    dd = np.abs(dbg["density"] - ref["density"])
    assert np.quantile(dd, 0.999) < 1e-4 and dd.max() < 2e-2
    assert _rel_err(dbg["rayW"], ref["rayW"]) < 1e-5
    # IDD on rays whose entry step and passive steps agree
    ints = dbg["ints"]
    first = ints["beamFirstInside"]
    same_fi = dbg["fi"] == ref["fi"]
    for l in range(ints["numLayers"]):
        ok = same_fi & (dbg["firstPassive"][l] == ref["firstPassive"][l])
        k1 = ints["afterLast"][l]
        g = dbg["iddSigma"][l, first:k1, :ok.shape[0], :ok.shape[1], 0][:, ok]
        o = ref["idd"][l, first:k1][:, ok]
        assert _rel_err(g, o) < 5e-4, l


@gpu
def test_config2_tile_radii_bit_exact_on_gpu_sigma(config2_gpu):
    case, ref, dose, dbg = config2_gpu
    ints = dbg["ints"]
    sx, sy = ints["superpDims"]
    first = ints["beamFirstInside"]
    for l in range(ints["numLayers"]):
        k1 = ints["layerPassive"][l]
        if k1 <= first:
            continue
        rs = np.ascontiguousarray(dbg["iddSigma"][l, :, :, :, 1])
        want = np.zeros((k1 - first, sy // 8, sx // 32), np.int32)
        O.lib().orc_tile_radii_f32(O._fp(rs), sx, sy, first, k1, O._ip(want))
        np.testing.assert_array_equal(dbg["tileRad"][l, first:k1] & 255, want)


@gpu
def test_config2_dose_within_half_percent(config2_gpu):
    case, ref, dose, dbg = config2_gpu
    d0 = ref["dose"]
    err = _rel_err(dose, d0)
    print(f"config2 max|dD|/Dmax = {err:.3e}  Dmax={d0.max():.4e}")
    assert err <= DOSE_TOL
    assert abs(dose.sum() / d0.sum() - 1.0) < 1e-3


def _rule_bins(hist):
    """The oracle's batching rule (complete_superposition.cu:413-490) on a radius histogram."""
    eff = np.zeros(34, np.int32)
    bt = np.zeros(33, np.int32)
    h = np.ascontiguousarray(hist, np.int32)
    O.lib().orc_batch_bins(O._ip(h), O._ip(bt), O._ip(eff))
    return eff


@gpu
@pytest.mark.parametrize("which", ["config1", "config2"])
def test_batch_bins_bit_exact(which, request):
    """Batched launch radii from the device equal the oracle's rule on the oracle's own radius
    histograms, and on the GPU's histograms (which equal the oracle's)."""
    torch = _torch()
    from subsecondpbs_amd.engine import Engine
    case, ref = request.getfixturevalue(which)
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(case["beams"][0], dose)
    eng.sync()
    hist = eng.debug_fetch(("radHist",))["radHist"]
    eff = eng.debug_bins()
    eng.close()
    for l in range(hist.shape[0]):
        np.testing.assert_array_equal(eff[l], _rule_bins(hist[l]), err_msg=f"layer {l}")
        np.testing.assert_array_equal(eff[l], _rule_bins(np.array(ref["plan"]["radHist"][l], np.int32)),
                                      err_msg=f"layer {l}")
