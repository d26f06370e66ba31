"""Nuclear halo branch on the GPU against the FP64 oracle: halo grid, radius histograms (bit
exact when the oracle's float rule is applied to the GPU's own halo sigma), dose within
0.5 % of Dmax for the three halo models, and the Python contract with nuclear_correction.
"""
import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import fixtures as F, lut, rtd

gpu = pytest.mark.gpu
DOSE_TOL = 5e-3
EN = lut.load_energy_tables(nuclear=True)


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _run(case, model):
    torch = _torch()
    from subsecondpbs_amd.engine import Engine
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    eng.set_nuclear(model)
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(case["beams"][0], dose)
    eng.sync()
    halo = eng.debug_halo()
    ints = eng.debug_fetch(())["ints"]
    out = dose.cpu().numpy()
    eng.close()
    return out, halo, ints


def _case(make):
    case = make()
    case["energy"] = EN
    return case


@gpu
@pytest.mark.parametrize("model", [1, 2, 3])
def test_halo_config1_models(model):
    case = _case(F.make_config1)
    ref, p = O.beam_dose(case, nuclear=model)
    dose, halo, ints = _run(case, model)
    assert halo["nucDims"] == tuple(p.nucDims) and halo["spotDist"] == pytest.approx(p.spotDist)
    assert halo["overflow"] == 0
    np.testing.assert_array_equal(ints["layerPassive"], np.array(p.layerPassive[:len(ints["layerPassive"])]))
    L = ints["numLayers"]
    np.testing.assert_array_equal(halo["radHist"], np.array([list(p.nucRadHist[l]) for l in range(L)], np.int32))
    err = _rel(dose, ref)
    print(f"halo model {model} config1 max|dD|/Dmax = {err:.3e}")
    assert err <= DOSE_TOL


@gpu
def test_halo_config2_gauss_fit():
    case = _case(F.make_config2)
    ref, p = O.beam_dose(case, nuclear=1)
    dose, halo, ints = _run(case, 1)
    nx, ny = halo["nucDims"]
    first = ints["beamFirstInside"]
    # halo tile radii: the oracle's float rule on the GPU's own halo sigma
    for l in range(ints["numLayers"]):
        k1 = int(ints["layerPassive"][l])
        if k1 <= first:
            continue
        rs = np.ascontiguousarray(halo["iddSigma"][l, :, :, :, 1])
        want = np.zeros((k1 - first, ny // 8, nx // 32), np.int32)
        O.lib().orc_tile_radii_f32(O._fp(rs), nx, ny, first, k1, O._ip(want))
        hist = np.bincount(want.ravel(), minlength=34)
        np.testing.assert_array_equal(halo["radHist"][l], hist)
    err = _rel(dose, ref)
    print(f"halo GAUSS_FIT config2 max|dD|/Dmax = {err:.3e}")
    assert err <= DOSE_TOL


@gpu
def test_raytracedicom_wrapper_nuclear_fills_tables():
    _torch()
    case = F.make_config1()  # energy dict without nuclear tables: filled from the tables folder
    assert "nuclear_weight_matrix" not in case["energy"]
    d = rtd.raytracedicom_wrapper(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"],
                                  case["dose_dims"], case["dose_resolution"], case["dose_corner"],
                                  case["beams"][0], case["energy"], nuclear_correction=True)
    ref, _ = O.beam_dose(_case(F.make_config1), nuclear=1)
    assert _rel(d, ref) <= DOSE_TOL


def _oracle_halo_fields(case, model):
    """Oracle fill outputs with the halo branch, per layer: primary (IDD, rSigma), halo (IDD, rSigma)."""
    import ctypes as C
    keep = O._Keep()
    en = O.make_energy(case["energy"], keep)
    bm = O.make_beam(case["beams"][0], keep)
    bm.nuclear = model
    p = O.plan_beam(case)
    rx, ry, K = p.rayDims[0], p.rayDims[1], p.steps
    ct = np.ascontiguousarray(O.ct_as_hu_plus_1000(case["ct"]), np.float32)
    dims = np.array(case["ct_dims"], np.int32)
    dens = np.zeros((K, ry, rx)); wepl = np.zeros((K, ry, rx))
    fi = np.zeros((ry, rx), np.int32); fo = np.zeros((ry, rx), np.int32)
    L = O.lib()
    L.orc_trace(C.byref(p), O._fp(ct), O._ip(dims), C.byref(en), O._dp(dens), O._dp(wepl), O._ip(fi), O._ip(fo))
    wmin = np.zeros(K)
    L.orc_beam_cutoffs(C.byref(p), C.byref(bm), C.byref(en), O._dp(wepl), O._ip(fi), O._ip(fo), O._dp(wmin))
    nl = len(case["beams"][0]["energies"])
    rw = np.zeros((nl, ry, rx))
    L.orc_conv_lattice(C.byref(p), C.byref(bm), C.byref(en), O._dp(rw))
    L.orc_layer_params(C.byref(p), C.byref(bm), C.byref(en), O._dp(wmin))
    r2n = np.zeros(rx * ry, np.int32)
    nucw = np.zeros((nl, 64 * 64))
    assert L.orc_halo_plan(C.byref(p), C.byref(bm), O._ip(r2n), O._dp(nucw)) == 0
    nx, ny = p.nucDims
    out = []
    for l in range(nl):
        idd = np.zeros((K, ry, rx)); rs = np.zeros((K, ry, rx)); fp = np.zeros((ry, rx), np.int32)
        nidd = np.zeros((K, ny, nx)); nrs = np.full((K, ny, nx), np.inf)
        L.orc_fill_halo(C.byref(p), l, C.byref(en), O._dp(dens), O._dp(wepl), O._dp(rw[l]), O._ip(fi), O._ip(fo),
                        O._dp(idd), O._dp(rs), O._ip(fp), model, O._ip(r2n),
                        O._dp(np.ascontiguousarray(nucw.ravel()[l * nx * ny:(l + 1) * nx * ny])), O._dp(nidd),
                        O._dp(nrs))
        out.append((idd, rs, nidd, nrs))
    return p, rw, out


@gpu
@pytest.mark.parametrize("model", [1, 2, 3])
def test_halo_config1_fill_fields(model):
    """The fill with the halo branch, field by field: rays weights, primary IDD / rSigma, halo
    IDD / rSigma, over each layer's computed steps."""
    torch = _torch()
    case = _case(F.make_config1)
    p, rw, ref = _oracle_halo_fields(case, model)
    from subsecondpbs_amd.engine import Engine
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    eng.set_nuclear(model)
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(case["beams"][0], dose)
    eng.sync()
    dbg = eng.debug_fetch(("rayW", "iddSigma"))
    halo = eng.debug_halo()
    eng.close()
    ints = dbg["ints"]
    first = ints["beamFirstInside"]
    rep = {"rayW": _rel(dbg["rayW"], rw)}
    for l, (idd, rs, nidd, nrs) in enumerate(ref):
        k1 = int(ints["afterLast"][l])
        rx, ry = ints["rayDims"]
        g = dbg["iddSigma"][l, first:k1, :ry, :rx]
        rep[f"idd{l}"] = _rel(g[..., 0], idd[first:k1])
        fin = np.isfinite(rs[first:k1])
        rep[f"rs{l}"] = float(np.abs(g[..., 1][fin] / rs[first:k1][fin] - 1).max()) if fin.any() else 0.0
        h = halo["iddSigma"][l, first:k1]
        rep[f"nidd{l}"] = _rel(h[..., 0], nidd[first:k1]) if nidd[first:k1].any() else float(np.abs(h[..., 0]).max())
        fin = np.isfinite(nrs[first:k1])
        rep[f"nrs{l}"] = float(np.abs(h[..., 1][fin] / nrs[first:k1][fin] - 1).max()) if fin.any() else 0.0
    print(f"halo model {model} fields:", {k: f"{v:.2e}" for k, v in rep.items()})
    assert rep["rayW"] < 1e-5
    for l in range(len(ref)):
        assert rep[f"idd{l}"] < 2e-4 and rep[f"rs{l}"] < 1e-4, l
        assert rep[f"nidd{l}"] < 2e-4 and rep[f"nrs{l}"] < 1e-4, l


@gpu
@pytest.mark.parametrize("fallback", ["canonical", "densest-layer", "weq-header"])
def test_halo_index_spots(fallback):
    """The halo branch with WEQ-index spots: the halo cells come from the physical PB lattice
    view (canonical 11 x 11 at 5 mm for spots on texel centres; the densest layer's 5 mm step when
    odd layers are shifted by 2 mm; the WEQ header's 81 x 71 lattice when a quarter-texel shift
    breaks the inference and the beam states its 5 mm spacing), against the oracle's dose within
    0.5 % of Dmax."""
    from .test_weq import _texel_centre_index_case

    if fallback == "weq-header":
        case = F.make_weq_case(index_spots=True)
        case["beams"][0]["spot_delta"] = np.array([5.0, 5.0, 0.0], np.float32)
    else:
        case = _texel_centre_index_case()
        if fallback == "densest-layer":
            b = case["beams"][0]
            raw = np.asarray(b["spot_positions"], np.float32).reshape(-1, 2).copy()
            layer = np.repeat(np.arange(len(b["energies"])), np.asarray(b["layer_spot_counts"]))
            raw[layer % 2 == 1, 0] += 2.0
            b["spot_positions"] = raw.ravel()
    case["energy"] = EN
    print(f"\n[gpu-test] halo with index spots, fallback={fallback}", flush=True)
    ref, p = O.beam_dose(case, nuclear=1)
    dose, halo, ints = _run(case, 1)
    assert halo["nucDims"] == tuple(p.nucDims) and halo["spotDist"] == pytest.approx(p.spotDist)
    assert halo["overflow"] == 0
    err = _rel(dose, ref)
    print(f"halo index spots (fallback={fallback}) max|dD|/Dmax = {err:.3e}; nucDims {tuple(p.nucDims)}")
    assert err <= DOSE_TOL
