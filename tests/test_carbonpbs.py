"""The CarbonPBS Python contract (calcDose / cuCalDose / cuFinalDose / cuCalDose3 /
debug_build_carbonpbs_context) on the CPU: its host-side conversions, pinned by what the
arithmetic fixes, and the beam it builds through the FP64 oracle.

* IDD -> CIDD: an IDD table sampled from the proton CIDD tables on a common depth axis sums back
  to those tables (telescoping); its peaks sit at the tables' Bragg peaks; for layers on table
  energies the oracle's fill with the converted tables gives the same energy index, peak depths
  within 0.2 mm, last steps within one, and per-step IDD within 2.5 % of the layer maximum;
* depth units: a step below 1 over a short span is read as cm (x10), a compacted axis stretches
  the step by iddsetting[2] / columns; fewer IDD rows than energies take the tail (or head);
* the subspot LUT on another energy axis is interpolated linearly between bracketing layers;
* one longitudinal cut-off per layer (a varying layer raises); ROI index forms;
* the context's beam equals the WEQ fixture's beam: the oracle gives the same dose bitwise.
"""
import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import carbonpbs as CP, fixtures as F
from subsecondpbs_amd.lut import load_energy_tables


def _idd_from_tables(step=0.5, n=600, rows=None):
    t = load_energy_tables()
    E = t["energies_per_u"] if rows is None else t["energies_per_u"][rows]
    idx = np.arange(t["n_energies"]) if rows is None else np.arange(t["n_energies"])[rows]
    depth = np.arange(n + 1) * step
    cum = np.zeros((idx.size, n + 1))
    for r, e in enumerate(idx):
        cum[r] = np.interp(depth * t["scale_facts"][e], np.arange(t["n_energy_samples"]), t["cidd_matrix"][e])
    idd = (np.diff(cum, axis=1) / step).astype(np.float32)   # sample j covers (j, j+1] steps
    return t, E, idd, cum[:, 1:]


def test_idd_to_cidd_and_peaks():
    t, E, idd, cum = _idd_from_tables(step=0.5, n=700)  # 350 mm: past the deepest peak (319 mm)
    en = CP.build_energy_from_carbonpbs(E, idd, [0.5, 0.5, idd.shape[1]])
    np.testing.assert_allclose(en["cidd_matrix"], cum, rtol=2e-5, atol=1e-12)
    np.testing.assert_allclose(en["scale_facts"], 2.0)
    # the IDD maximum sits at the table's Bragg peak (a step of 0.5 mm, sample j spans (j, j+1])
    assert np.abs(en["peak_depths"] - t["peak_depths"]).max() <= 1.5


def test_converted_tables_drive_the_fill_like_the_tables():
    # layers on table energies: a common depth axis interpolates neighbouring rows at equal depth
    # (not at equal peak-relative depth as the per-energy scale of the tables does), which the
    # reference does too but which would blur the comparison
    from .oracle_stages import run

    case = F.make_config1()
    tab = load_energy_tables()["energies_per_u"]
    b = dict(case["beams"][0])
    b["energies"] = np.array([tab[np.argmin(np.abs(tab - e))] for e in b["energies"]], np.float32)
    case["beams"] = [b]
    r0 = run(case)
    t, E, idd, _ = _idd_from_tables(step=0.05, n=6000)
    conv = dict(case)
    conv["energy"] = CP.build_energy_from_carbonpbs(E, idd, [0.0, 0.05, idd.shape[1]])
    r1 = run(conv)
    p0, p1 = r0["plan"], r1["plan"]
    np.testing.assert_array_equal(p0["energyIdx"], p1["energyIdx"])
    # peak depths from the IDD maximum: within a step of the tables' (they set the 1.05 x peak
    # cut-off, so a layer's last step may move by one)
    assert np.abs(np.asarray(p0["peakDepth"]) - np.asarray(p1["peakDepth"])).max() <= 0.2
    assert np.abs(np.asarray(p0["afterLast"]) - np.asarray(p1["afterLast"])).max() <= 1
    first = p0["beamFirstInside"]
    for l in range(len(b["energies"])):
        k1 = min(p0["afterLast"][l], p1["afterLast"][l]) - 1
        a, c = r0["idd"][l, first:k1], r1["idd"][l, first:k1]
        err = np.abs(a - c).max() / a.max()
        print(f"layer {l}: IDD of converted tables vs tables {err:.2e}")
        assert err < 2.5e-2


def test_depth_units_and_compacted_axis(monkeypatch):
    E = np.array([100.0, 150.0, 200.0], np.float32)
    idd = np.ones((3, 100), np.float32)
    cm = CP.build_energy_from_carbonpbs(E, idd, [0.0, 0.05, 100])      # 5 cm span in cm units
    np.testing.assert_allclose(cm["scale_facts"], 1.0 / 0.5, rtol=1e-6)
    mm = CP.build_energy_from_carbonpbs(E, idd, [0.0, 2.0, 100])
    np.testing.assert_allclose(mm["scale_facts"], 0.5, rtol=1e-6)
    comp = CP.build_energy_from_carbonpbs(E, idd, [0.0, 1.0, 400])     # 100 columns of a 400-sample axis
    np.testing.assert_allclose(comp["scale_facts"], 0.25, rtol=1e-6)
    np.testing.assert_allclose(comp["cidd_matrix"][:, -1], 400.0, rtol=1e-6)
    tail = CP.build_energy_from_carbonpbs(np.array([80.0, 100.0, 150.0, 200.0], np.float32), idd, [0.0, 1.0, 100])
    np.testing.assert_array_equal(tail["energies_per_u"], E)
    monkeypatch.setenv("RTD_IDD_SLICE_MODE", "head")
    head = CP.build_energy_from_carbonpbs(np.array([80.0, 100.0, 150.0, 200.0], np.float32), idd, [0.0, 1.0, 100])
    np.testing.assert_array_equal(head["energies_per_u"], [80.0, 100.0, 150.0])
    with pytest.raises(RuntimeError, match="rows"):
        CP.build_energy_from_carbonpbs(E[:2], idd, [0.0, 1.0, 100])


def test_subspot_lut_remap_is_linear():
    lut_e = np.array([100.0, 120.0, 140.0], np.float32)
    lut = np.zeros((3, 2, 5), np.float32)
    for i in range(3):
        lut[i, :, :] = np.arange(10, dtype=np.float32).reshape(2, 5) + 10.0 * i
    sub, sig = CP.remap_subspot_lut_to_actual_layers(lut_e, lut, [110.0, 120.0, 150.0, 90.0])
    np.testing.assert_allclose(sub[0], 0.5 * (lut[0] + lut[1]))
    np.testing.assert_array_equal(sub[1], lut[1])
    np.testing.assert_array_equal(sub[2], lut[2])   # clamped past the axis
    np.testing.assert_array_equal(sub[3], lut[0])
    np.testing.assert_allclose(sig[0], 0.5 * (lut[0, 0, 3:5] + lut[1, 0, 3:5]))


def test_layer_cutoffs_and_roi_forms():
    np.testing.assert_array_equal(CP.derive_layer_longitudinal_cutoffs([5, 5, 7, 7, 7], [2, 3]), [5.0, 7.0])
    with pytest.raises(RuntimeError, match="varies within layer 1"):
        CP.derive_layer_longitudinal_cutoffs([5, 5, 7, 8, 7], [2, 3])
    dims = (4, 5, 6)
    xyz = np.array([[1, 2, 3], [3, 4, 5]])
    want = [1 * 30 + 2 * 6 + 3, 3 * 30 + 4 * 6 + 5]
    np.testing.assert_array_equal(CP.to_roi_linear_indices(xyz, dims), want)
    np.testing.assert_array_equal(CP.to_roi_linear_indices(xyz.T, dims), want)
    np.testing.assert_array_equal(CP.to_roi_linear_indices(xyz.ravel(), dims), want)
    np.testing.assert_array_equal(CP.to_roi_linear_indices(np.array([7, 99, 100, 101]), dims), [7, 99, 100, 101])


def carbonpbs_args(case):
    """The CarbonPBS arrays of the WEQ-index fixture's beam (the proton tables as its IDD table)."""
    b = case["beams"][0]
    n = int(np.asarray(b["spot_weights"]).size)
    t = load_energy_tables()
    E = t["energies_per_u"]
    _, _, idd, _ = _idd_from_tables(step=0.25, n=1200)
    L = len(b["energies"])
    sig = np.asarray(b["spot_sigmas"], np.float32).reshape(L, 2)
    return dict(
        rayweq=b["water_equivalence"], roiIdx=np.zeros((0, 3), np.int32),
        all_energies=np.repeat(b["energies"], b["layer_spot_counts"]),
        sourcePos=np.tile(b["source_position"], (n, 1)), tmpBeamDir=np.tile(b["beam_direction"], (n, 1)),
        bmxdir=b["beam_xdir"], bmydir=b["beam_ydir"], corner=np.array(case["dose_corner"], np.float32),
        resolution=np.array(case["dose_resolution"], np.float32), dims=np.array(case["dose_dims"], np.int32),
        longitudalCutoff=None, enelist=E, idddata=idd, iddsetting=np.array([0.0, 0.25, 1200.0], np.float32),
        profiledata=None, profilesetting=None, beamparadata=None, subspotdata=F.single_subspot_lut(sig),
        layerInfo=b["layer_spot_counts"], layerEnergy=b["energies"], idbeamxy=np.asarray(b["spot_positions"]),
        nPar=b["spot_weights"], sad=np.array([b["sad"]], np.float32))


def test_context_beam_equals_the_fixture_beam():
    case = F.make_weq_case(index_spots=True)
    ctx = CP.build_carbonpbs_context(**carbonpbs_args(case))
    info = CP.debug_build_carbonpbs_context(**carbonpbs_args(case))
    assert info["totalSpots"] == int(case["beams"][0]["spot_weights"].size) and info["weq_len"] > 9
    one = dict(case)
    one["beams"] = [ctx["beam"]]
    d_ctx, _ = O.beam_dose(one)
    d_fix, _ = O.beam_dose(case)
    assert np.array_equal(d_ctx, d_fix)


def test_context_profile_model(monkeypatch):
    """The profile arrays travel with the beam as the reference stores them (pybind:1269-1302):
    a table of 20 stored depths for 40 logical ones doubles the depth step and stores 20; the LUT
    energy axis is cut to the table's rows (tail by default, head with RTD_PROFILE_SLICE_MODE);
    the hooks stay off unless asked for."""
    case = F.make_weq_case(index_spots=True)
    args = carbonpbs_args(case)
    E = np.asarray(args["enelist"], np.float32)
    table = np.ones((6, 20, 5), np.float32)
    args.update(profiledata=table, profilesetting=np.array([0.0, 2.0, 40.0], np.float32),
                beamparadata=np.ones((6, 3), np.float32), beamParaPos=12.5)
    b = CP.build_carbonpbs_context(**args)["beam"]
    np.testing.assert_array_equal(b["profile_setting"], np.array([0.0, 4.0, 20.0], np.float32))
    np.testing.assert_array_equal(b["profile_energies"], E[-6:])
    assert b["profile_data"].size == table.size and b["beam_para_pos"] == 12.5
    assert b["carbon_profile_hooks"] == 0
    monkeypatch.setenv("RTD_PROFILE_SLICE_MODE", "head")
    b = CP.build_carbonpbs_context(**args, carbon_profile_hooks=3)["beam"]
    np.testing.assert_array_equal(b["profile_energies"], E[:6])
    assert b["carbon_profile_hooks"] == 3
    # with the hooks off the profile arrays change nothing
    one = dict(case)
    one["beams"] = [CP.build_carbonpbs_context(**args)["beam"]]
    d_prof, _ = O.beam_dose(one)
    d_fix, _ = O.beam_dose(case)
    assert np.array_equal(d_prof, d_fix)
