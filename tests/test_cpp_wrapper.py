"""The C++ form of the drop-in: tests/cpp/rtd_wrapper_main.cpp fills the reference's
RTDBeamSettings / RTDEnergyStruct aggregates (std::vector members,
include/core/raytracedicom_integration.h:31-144 of the reference) and calls subsecondWrapper.
This file writes its inputs as records, runs it, and compares its dose with the FP64 oracle
(max |dD| <= 0.5 % of Dmax); a failing entry check prints [RTD_ENTRY_ASSERT] and leaves the
dose untouched (raytracedicom_wrapper.cu:1099-1162, 2467-2479)."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import fixtures as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_bin", "rtd_wrapper_main")
DOSE_TOL = 5e-3


def _need_exe():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/_bin/rtd_wrapper_main is not built: run __graft_entry__.build()")


def _rec(out, name, arr, kind):
    a = np.ascontiguousarray(np.asarray(arr, np.float32 if kind == 0 else np.int32).ravel())
    nb = name.encode()
    out += struct.pack("<I", len(nb)) + nb + struct.pack("<IQ", kind, a.size) + a.tobytes()


def write_records(path, case, beams, dose_in=None, nuclear=0):
    """The record file rtd_wrapper_main reads: grid, energy struct and beams under the
    RTDBeamSettings / RTDEnergyStruct member names."""
    out = bytearray(b"PBSR")
    f, i = 0, 1
    _rec(out, "numBeams", [len(beams)], i)
    _rec(out, "nuclear", [nuclear], i)
    _rec(out, "ct", case["ct"], f)
    _rec(out, "ctDims", case["ct_dims"], i)
    _rec(out, "ctRes", case["ct_resolution"], f)
    _rec(out, "ctCorner", case["ct_corner"], f)
    _rec(out, "doseDims", case["dose_dims"], i)
    _rec(out, "doseRes", case["dose_resolution"], f)
    _rec(out, "doseCorner", case["dose_corner"], f)
    if dose_in is not None:
        _rec(out, "doseIn", dose_in, f)
    e = case["energy"]
    for k, key in (("nEnergySamples", "n_energy_samples"), ("nEnergies", "n_energies"),
                   ("nDensitySamples", "n_density_samples"), ("nSpSamples", "n_sp_samples"),
                   ("nRRlSamples", "n_rrl_samples")):
        _rec(out, "e." + k, [int(e[key])], i)
    for k, key in (("energiesPerU", "energies_per_u"), ("peakDepths", "peak_depths"),
                   ("scaleFacts", "scale_facts"), ("ciddMatrix", "cidd_matrix"),
                   ("densityScaleFact", "density_scale_fact"), ("densityVector", "density_vector"),
                   ("spScaleFact", "sp_scale_fact"), ("spVector", "sp_vector"),
                   ("rRlScaleFact", "rrl_scale_fact"), ("rRlVector", "rrl_vector"),
                   ("nucWeightMatrix", "nuclear_weight_matrix"), ("nucSqSigmaMatrix", "nuclear_sq_sigma_matrix")):
        if e.get(key) is not None:
            _rec(out, "e." + k, np.atleast_1d(e[key]), f)
    for n, b in enumerate(beams):
        p = f"b{n}."
        ss = np.asarray(b["subspot_data"], np.float32)
        _rec(out, p + "energies", b["energies"], f)
        _rec(out, p + "spotSigmas", b["spot_sigmas"], f)
        _rec(out, p + "maxSubspotsPerLayer", [ss.shape[1]], i)
        _rec(out, p + "subspotData", ss, f)
        for k, key in (("raySpacing", "ray_spacing"), ("sourceDist", "source_dist"), ("spotOffset", "spot_offset"),
                       ("spotDelta", "spot_delta"), ("spotPositions", "spot_positions"),
                       ("spotWeights", "spot_weights"), ("spotBeamDirections", "spot_beam_directions"),
                       ("layerLongitudinalCutoffs", "layer_longitudinal_cutoffs"),
                       ("beamDirection", "beam_direction"), ("bmXDirection", "beam_xdir"),
                       ("bmYDirection", "beam_ydir"), ("sourcePosition", "source_position"),
                       ("sad", "sad"), ("refPlaneZ", "ref_plane_z")):
            if b.get(key) is not None:
                _rec(out, p + k, np.atleast_1d(b[key]), f)
        for k, key in (("steps", "steps"), ("layerSpotCounts", "layer_spot_counts"),
                       ("roiLinearIndices", "roi_linear_indices")):
            if b.get(key) is not None:
                _rec(out, p + k, np.atleast_1d(b[key]), i)
        if b.get("spot_positions_are_indices"):
            _rec(out, p + "spotPositionsAreIndices", [1], i)
    with open(path, "wb") as fp:
        fp.write(out)


def test_records_parse_in_the_cpp_caller(tmp_path):
    """No device: the C++ caller reads the records into the std::vector aggregates (it links
    libpbs_hip.so, so this also checks that the library resolves from the executable)."""
    _need_exe()
    case = F.make_config1()
    path = tmp_path / "in.bin"
    write_records(path, case, case["beams"])
    r = subprocess.run([EXE, "--parse", str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    b = case["beams"][0]
    ss = np.asarray(b["subspot_data"])
    nd = int(np.prod(case["dose_dims"]))
    assert r.stdout.split() == ["beams", "1", "layers", str(len(b["energies"])), "subspots", str(ss.size),
                                "ct", str(np.asarray(case["ct"]).size), "dose", str(nd),
                                "energies", str(int(case["energy"]["n_energies"]))]


@pytest.mark.gpu
def test_cpp_subsecond_wrapper_config1_vs_oracle(tmp_path):
    _need_exe()
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    case = F.make_config1()
    ref, _ = O.beam_dose(case)
    nd = tuple(reversed(case["dose_dims"]))
    path, out = tmp_path / "in.bin", tmp_path / "dose.f32"
    # the wrapper accumulates into the caller's dose: start from 1 % of the oracle's maximum
    base = np.full(nd, 0.01 * float(ref.max()), np.float32)
    write_records(path, case, case["beams"], dose_in=base)
    print("[gpu-test] rtd_wrapper_main config1", flush=True)
    r = subprocess.run([EXE, str(path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "RTD_ENTRY_ASSERT" not in r.stderr
    d = np.fromfile(out, np.float32).reshape(nd) - base
    err = float(np.abs(d - ref).max() / np.abs(ref).max())
    print(f"C++ subsecondWrapper config1 max|dD|/Dmax = {err:.3e}")
    assert err <= DOSE_TOL
    # an entry check that fails: the dose comes back untouched, with the reference's message
    bad = dict(case["beams"][0])
    bad["spot_sigmas"] = np.asarray(bad["spot_sigmas"], np.float32).ravel()[:-2]
    write_records(path, case, [bad], dose_in=base)
    r = subprocess.run([EXE, str(path), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    assert "[RTD_ENTRY_ASSERT] beamIdx=0" in r.stderr
    assert np.array_equal(np.fromfile(out, np.float32).reshape(nd), base)


@pytest.mark.parametrize("field,value,needle", [
    ("layer_spot_counts", "drop_one", "one count per energy layer"),
    ("spot_positions", "odd", "length must be even"),
    ("spot_beam_directions", "short", "one direction per spot position"),
    ("layer_longitudinal_cutoffs", "short", "one value per energy layer"),
])
def test_cpp_entry_shape_checks(tmp_path, field, value, needle):
    """The entry validation's vector-shape checks (raytracedicom_wrapper.cu:1113-1150) run
    before any device work: the dose comes back untouched with an [RTD_ENTRY_ASSERT] line."""
    _need_exe()
    case = F.make_config1()
    b = dict(case["beams"][0])
    L = len(b["energies"])
    npos = int(np.sum(b["layer_spot_counts"]))
    if value == "drop_one":
        b[field] = np.asarray(b[field])[:-1]
    elif value == "odd":
        b[field] = np.asarray(b[field], np.float32).ravel()[:-1]
    elif field == "spot_beam_directions":
        b[field] = np.tile(np.float32([0, 0, 1]), npos - 1)
    else:
        b[field] = np.full(L - 1, 100.0, np.float32)
    nd = tuple(reversed(case["dose_dims"]))
    base = np.full(nd, 3.0, np.float32)
    path, out = tmp_path / "in.bin", tmp_path / "dose.f32"
    write_records(path, case, [b], dose_in=base)
    r = subprocess.run([EXE, str(path), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "[RTD_ENTRY_ASSERT] beamIdx=0" in r.stderr and needle in r.stderr, r.stderr
    assert np.array_equal(np.fromfile(out, np.float32).reshape(nd), base)
