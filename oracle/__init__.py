"""ctypes front end of the FP64 CPU oracle (``oracle/pbs_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``, always as the checker. The engine package never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpbs_oracle.so")
MAX_LAYERS = 256
MAX_RAD = 32

_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)


class Energy(C.Structure):
    _fields_ = [
        ("nEnergies", C.c_int), ("nSamples", C.c_int),
        ("energies", _f), ("peaks", _f), ("scales", _f), ("cidd", _f),
        ("nDensity", C.c_int), ("densityScale", C.c_float), ("density", _f),
        ("nSp", C.c_int), ("spScale", C.c_float), ("sp", _f),
        ("nRrl", C.c_int), ("rrlScale", C.c_float), ("rrl", _f),
        ("nucWeight", _f), ("nucSqSigma", _f),
        ("ionRangeCoef", C.c_float), ("ionRangeExp", C.c_float), ("ionRestMass", C.c_float),
        ("ionScatterScale", C.c_float),
    ]


class Beam(C.Structure):
    _fields_ = [
        ("numLayers", C.c_int), ("energies", _f), ("spotSigmas", _f),
        ("raySpacing", C.c_float * 2), ("sourceDist", C.c_float * 2), ("spotDelta", C.c_float * 3),
        ("maxSubspots", C.c_int), ("subspotData", _f),
        ("layerSpotCounts", _i), ("numSpots", C.c_int), ("spotPositions", _f), ("spotWeights", _f),
        ("spotBeamDirections", _f),
        ("beamDirection", C.c_float * 3), ("bmXDirection", C.c_float * 3),
        ("bmYDirection", C.c_float * 3), ("sourcePosition", C.c_float * 3),
        ("sad", C.c_float), ("useSubspotConvolution", C.c_int), ("nuclear", C.c_int),
        ("weq", _f), ("weqCount", C.c_size_t), ("spotIndices", C.c_int),
        ("cpbOverride", C.c_int), ("cpbOvCorner", C.c_float * 2), ("cpbOvRes", C.c_float * 2),
        ("cpbOvDims", C.c_int * 2),
        ("energyDepthMode", C.c_int), ("energyDepthToMm", C.c_float), ("cpbCoverage", C.c_float),
        ("profileEnergies", _f), ("numProfileEnergies", C.c_int), ("profileData", _f),
        ("profileDataCount", C.c_size_t), ("profileSetting", _f), ("numProfileSetting", C.c_int),
        ("beamParaData", _f), ("numBeamParaData", C.c_int), ("beamParaPos", C.c_float),
        ("carbonProfileHooks", C.c_int),
    ]


class Plan(C.Structure):
    _fields_ = [
        ("ok", C.c_int), ("legacyPath", C.c_int), ("lenToMm", C.c_float), ("energyDepthToMm", C.c_float),
        ("bmX", C.c_float * 3), ("bmY", C.c_float * 3), ("bmZ", C.c_float * 3), ("iso", C.c_float * 3),
        ("g2w", C.c_float * 9), ("g2wOff", C.c_float * 3),
        ("g2im", C.c_float * 9), ("g2imOff", C.c_float * 3),
        ("g2dose", C.c_float * 9), ("g2doseOff", C.c_float * 3),
        ("sourceDist", C.c_float * 2),
        ("spotDims", C.c_int * 2), ("spotOffset", C.c_float * 2), ("spotDelta", C.c_float * 2),
        ("cpbCorner", C.c_float * 3), ("cpbRes", C.c_float * 3),
        ("rayDims", C.c_int * 2), ("superpDims", C.c_int * 2), ("steps", C.c_int),
        ("fanCorner", C.c_float * 3), ("fanDelta", C.c_float * 3),
        ("beamFirstInside", C.c_int), ("beamFirstOutside", C.c_int),
        ("firstPastCutoffAll", C.c_int), ("guaranteedPassive", C.c_int),
        ("maxPeakDepth", C.c_float),
        ("energyIdx", C.c_float * MAX_LAYERS), ("scaleFact", C.c_float * MAX_LAYERS),
        ("peakDepth", C.c_float * MAX_LAYERS), ("convPeakDepth", C.c_float * MAX_LAYERS),
        ("entrySigma", (C.c_float * 2) * MAX_LAYERS),
        ("sigmaSqAirLin", C.c_float * MAX_LAYERS), ("sigmaSqAirQuad", C.c_float * MAX_LAYERS),
        ("volConst", C.c_float), ("volLin", C.c_float), ("volSq", C.c_float),
        ("spotDistInRays", C.c_float),
        ("afterLast", C.c_int * MAX_LAYERS), ("layerPassive", C.c_int * MAX_LAYERS),
        ("radHist", (C.c_int * (MAX_RAD + 2)) * MAX_LAYERS),
        ("batchRad", (C.c_int * (MAX_RAD + 1)) * MAX_LAYERS),
        ("transferStart", C.c_int * 3), ("transferMax", C.c_int * 3),
        ("numLayers", C.c_int), ("error", C.c_char * 256),
        ("nucDims", C.c_int * 2), ("spotDist", C.c_float), ("nucRadHist", (C.c_int * 34) * 256),
        ("entrySpotSigmaX", C.c_float * 256),
        ("weqVolume", C.c_int), ("spotIndices", C.c_int), ("weqDims", C.c_int * 3),
        ("weqStepMm", C.c_float), ("weqX0", C.c_float), ("weqY0", C.c_float), ("weqDx", C.c_float),
        ("weqDy", C.c_float), ("weqBody", _f),
        ("nucRaw", C.c_int * 2), ("nucFallback", C.c_int), ("nucOffset", C.c_float * 2), ("nucDelta", C.c_float * 2),
        ("profileHooks", C.c_int), ("profileRows", C.c_int), ("profileDepthN", C.c_int),
        ("profileChannels", C.c_int), ("profileDepth0", C.c_float), ("profileDepthStep", C.c_float),
        ("beamParaPos", C.c_float), ("profileLayer", (C.c_float * 4) * MAX_LAYERS), ("profileTable", _f),
    ]


_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_profile_hooks.restype = None
        L.orc_profile_hooks.argtypes = [C.POINTER(Plan), C.c_int, _d, _d, _d]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_ct_adapter.restype = C.c_int
        L.orc_ct_adapter.argtypes = [_f, C.c_size_t, C.c_int, C.POINTER(Energy), _f]
        L.orc_sizeof.restype = C.c_long
        L.orc_sizeof.argtypes = [C.c_int]
        L.orc_find_decimal_ordered.restype = C.c_float
        L.orc_find_decimal_ordered.argtypes = [_f, C.c_int, C.c_float]
        L.orc_vector_interpolate.restype = C.c_float
        L.orc_vector_interpolate.argtypes = [_f, C.c_int, C.c_float]
        L.orc_find_first_larger_ordered.restype = C.c_int
        L.orc_find_first_larger_ordered.argtypes = [_d, C.c_int, C.c_double]
        L.orc_align_grid.restype = C.c_int
        L.orc_align_grid.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                     _f, _i, _i, _i]
        L.orc_tile_radius_f32.restype = C.c_int
        L.orc_tile_radius_f32.argtypes = [C.c_float]
        L.orc_batch_bins.argtypes = [_i, _i, _i]
        L.orc_tile_radii_f32.argtypes = [_f, C.c_int, C.c_int, C.c_int, C.c_int, _i]
        L.orc_sample3d_border.restype = C.c_double
        L.orc_sample3d_border.argtypes = [_f, _i, C.c_double, C.c_double, C.c_double]
        L.orc_plan_beam.argtypes = [_i, _f, _f, _i, _f, _f, C.POINTER(Beam), C.POINTER(Energy), C.POINTER(Plan)]
        L.orc_trace.argtypes = [C.POINTER(Plan), _f, _i, C.POINTER(Energy), _d, _d, _i, _i]
        L.orc_beam_cutoffs.argtypes = [C.POINTER(Plan), C.POINTER(Beam), C.POINTER(Energy), _d, _i, _i, _d]
        L.orc_conv_lattice.argtypes = [C.POINTER(Plan), C.POINTER(Beam), C.POINTER(Energy), _d]
        L.orc_subspot_to_cpb.argtypes = [_f, C.c_int, C.c_int, _f, _f, _i, _f, _f, _f, _f, C.c_float, _d]
        L.orc_cpb_to_rays.argtypes = [_d, _i, _f, _f, _d, _i, _f, _f, C.c_int]
        L.orc_layer_params.argtypes = [C.POINTER(Plan), C.POINTER(Beam), C.POINTER(Energy), _d]
        L.orc_fill.argtypes = [C.POINTER(Plan), C.c_int, C.POINTER(Energy), _d, _d, _d, _i, _i, _d, _d, _i]
        L.orc_fill_halo.argtypes = [C.POINTER(Plan), C.c_int, C.POINTER(Energy), _d, _d, _d, _i, _i, _d, _d, _i,
                                    C.c_int, _i, _d, _d, _d]
        L.orc_halo_plan.restype = C.c_int
        L.orc_halo_plan.argtypes = [C.POINTER(Plan), C.POINTER(Beam), _i, _d]
        L.orc_superpose.argtypes = [C.POINTER(Plan), C.c_int, _d, _d, _d]
        L.orc_transfer.argtypes = [C.POINTER(Plan), C.c_int, _d, _d, _i]
        L.orc_beam_dose.restype = C.c_int
        L.orc_beam_dose.argtypes = [_f, _i, _f, _f, _d, _i, _f, _f, C.POINTER(Beam), C.POINTER(Energy), C.POINTER(Plan)]
        L.orc_beam_dose_entry.restype = C.c_int
        L.orc_beam_dose_entry.argtypes = [_f, _i, _f, _f, _d, _i, _f, _f, C.POINTER(Beam), C.POINTER(Energy), _i,
                                          C.POINTER(Plan)]
        L.orc_trace_entry_margin.argtypes = [C.POINTER(Plan), _f, _i, _d]
        L.orc_beam_adjoint.restype = C.c_int
        L.orc_beam_adjoint.argtypes = [_f, _i, _f, _f, _d, _i, _f, _f, C.POINTER(Beam), C.POINTER(Energy), _d,
                                       C.POINTER(Plan)]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(_f)


def _ip(a):
    return a.ctypes.data_as(_i)


def _dp(a):
    return a.ctypes.data_as(_d)


class _Keep:
    """Holds numpy buffers alive while a ctypes struct points at them."""

    def __init__(self):
        self.bufs = []

    def f32(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self.bufs.append(a)
        return _fp(a)

    def i32(self, a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        self.bufs.append(a)
        return _ip(a)


def make_energy(energy: dict, keep: _Keep) -> Energy:
    e = Energy()
    e.nEnergies = int(energy["n_energies"])
    e.nSamples = int(energy["n_energy_samples"])
    e.energies = keep.f32(energy["energies_per_u"])
    e.peaks = keep.f32(energy["peak_depths"])
    e.scales = keep.f32(energy["scale_facts"])
    e.cidd = keep.f32(energy["cidd_matrix"])
    e.nDensity = int(energy["n_density_samples"])
    e.densityScale = float(energy["density_scale_fact"])
    e.density = keep.f32(energy["density_vector"])
    e.nSp = int(energy["n_sp_samples"])
    e.spScale = float(energy["sp_scale_fact"])
    e.sp = keep.f32(energy["sp_vector"])
    e.nRrl = int(energy["n_rrl_samples"])
    e.rrlScale = float(energy["rrl_scale_fact"])
    e.rrl = keep.f32(energy["rrl_vector"])
    if "nuclear_weight_matrix" in energy:
        e.nucWeight = keep.f32(energy["nuclear_weight_matrix"])
        e.nucSqSigma = keep.f32(energy["nuclear_sq_sigma_matrix"])
    ion = energy.get("ion")
    if ion:
        e.ionRangeCoef = float(ion["range_coef"])
        e.ionRangeExp = float(ion["range_exp"])
        e.ionRestMass = float(ion["rest_mass"])
        e.ionScatterScale = float(ion["scatter_scale"])
    return e


def make_beam(beam: dict, keep: _Keep, use_subspots: bool = False) -> Beam:
    b = Beam()
    L = len(beam["energies"])
    b.numLayers = L
    b.energies = keep.f32(beam["energies"])
    b.spotSigmas = keep.f32(np.asarray(beam["spot_sigmas"]).reshape(L, 2))
    b.raySpacing[:] = [float(v) for v in beam.get("ray_spacing", (0.0, 0.0))]
    b.sourceDist[:] = [float(v) for v in beam.get("source_dist", (0.0, 0.0))]
    b.spotDelta[:] = [float(v) for v in beam.get("spot_delta", (0.0, 0.0, 0.0))]
    ss = np.ascontiguousarray(beam["subspot_data"], dtype=np.float32)
    b.maxSubspots = int(ss.shape[1])
    b.subspotData = keep.f32(ss)
    if "layer_spot_counts" in beam and beam["layer_spot_counts"] is not None:
        b.layerSpotCounts = keep.i32(beam["layer_spot_counts"])
        b.numSpots = int(np.asarray(beam["spot_weights"]).size)
        b.spotPositions = keep.f32(beam["spot_positions"])
        b.spotWeights = keep.f32(beam["spot_weights"])
    if beam.get("spot_beam_directions") is not None:
        b.spotBeamDirections = keep.f32(beam["spot_beam_directions"])
    b.beamDirection[:] = [float(v) for v in beam["beam_direction"]]
    b.bmXDirection[:] = [float(v) for v in beam["beam_xdir"]]
    b.bmYDirection[:] = [float(v) for v in beam["beam_ydir"]]
    b.sourcePosition[:] = [float(v) for v in beam["source_position"]]
    b.sad = float(beam["sad"])
    b.useSubspotConvolution = 1 if use_subspots else 0
    for key in ("water_equivalence", "ray_weq_header"):  # the WEQ header in use, then its samples
        v = beam.get(key)
        if v is not None and np.asarray(v).size >= 9:
            a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).ravel())
            b.weq = keep.f32(a)
            b.weqCount = int(a.size)
            break
    b.spotIndices = 1 if beam.get("spot_positions_are_indices", False) else 0
    # unit overrides as values (what the RTD_ENERGY_DEPTH_* / RTD_CPB_COVERAGE_CM variables select)
    b.energyDepthMode = int(beam.get("energy_depth_mode", 0))
    b.energyDepthToMm = float(beam.get("energy_depth_to_mm", 0.0))
    b.cpbCoverage = float(beam.get("cpb_coverage", 0.0))
    for key, ptr, cnt in (("profile_energies", "profileEnergies", "numProfileEnergies"),
                          ("profile_data", "profileData", "profileDataCount"),
                          ("profile_setting", "profileSetting", "numProfileSetting"),
                          ("beam_para_data", "beamParaData", "numBeamParaData")):
        v = beam.get(key)
        if v is not None:
            a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).ravel())
            setattr(b, ptr, keep.f32(a))
            setattr(b, cnt, int(a.size))
    if beam.get("beam_para_pos") is not None:
        b.beamParaPos = float(beam["beam_para_pos"])
    b.carbonProfileHooks = int(beam.get("carbon_profile_hooks", 0))
    ov = beam.get("cpb_override")  # (corner[2], res[2], dims[2]) in the gantry frame
    if ov is not None:
        b.cpbOverride = 1
        b.cpbOvCorner[:] = [float(v) for v in ov[0]]
        b.cpbOvRes[:] = [float(v) for v in ov[1]]
        b.cpbOvDims[:] = [int(v) for v in ov[2]]
    return b


def plan_to_dict(p: Plan) -> dict:
    L = p.numLayers
    return {
        "ok": p.ok, "legacy": p.legacyPath, "error": p.error.decode(),
        "rayDims": tuple(p.rayDims), "superpDims": tuple(p.superpDims), "steps": p.steps,
        "cpbCorner": tuple(p.cpbCorner), "cpbRes": tuple(p.cpbRes),
        "spotDims": tuple(p.spotDims), "spotOffset": tuple(p.spotOffset), "spotDelta": tuple(p.spotDelta),
        "fanCorner": tuple(p.fanCorner), "fanDelta": tuple(p.fanDelta), "sourceDist": tuple(p.sourceDist),
        "beamFirstInside": p.beamFirstInside, "beamFirstOutside": p.beamFirstOutside,
        "firstPastCutoffAll": p.firstPastCutoffAll, "guaranteedPassive": p.guaranteedPassive,
        "maxPeakDepth": p.maxPeakDepth,
        "energyIdx": np.array(p.energyIdx[:L], np.float32), "scaleFact": np.array(p.scaleFact[:L], np.float32),
        "peakDepth": np.array(p.peakDepth[:L], np.float32),
        "afterLast": np.array(p.afterLast[:L], np.int32), "layerPassive": np.array(p.layerPassive[:L], np.int32),
        "radHist": np.array([list(p.radHist[l]) for l in range(L)], np.int32),
        "batchRad": np.array([list(p.batchRad[l]) for l in range(L)], np.int32),
        "transferStart": tuple(p.transferStart), "transferMax": tuple(p.transferMax),
        "entrySigma": np.array([list(p.entrySigma[l]) for l in range(L)], np.float32),
        "volConst": p.volConst, "volLin": p.volLin, "volSq": p.volSq,
    }


def case_args(case: dict):
    ct_dims = np.array(case["ct_dims"], np.int32)
    ct_res = np.array(case["ct_resolution"], np.float32)
    ct_corner = np.array(case["ct_corner"], np.float32)
    dose_dims = np.array(case["dose_dims"], np.int32)
    dose_res = np.array(case["dose_resolution"], np.float32)
    dose_corner = np.array(case["dose_corner"], np.float32)
    return ct_dims, ct_res, ct_corner, dose_dims, dose_res, dose_corner


def plan_beam(case: dict, beam_index: int = 0, use_subspots: bool = False) -> Plan:
    keep = _Keep()
    en = make_energy(case["energy"], keep)
    bm = make_beam(case["beams"][beam_index], keep, use_subspots)
    ct_dims, ct_res, ct_corner, dose_dims, dose_res, dose_corner = case_args(case)
    p = Plan()
    lib().orc_plan_beam(_ip(ct_dims), _fp(ct_res), _fp(ct_corner), _ip(dose_dims), _fp(dose_res),
                        _fp(dose_corner), C.byref(bm), C.byref(en), C.byref(p))
    p._keep = keep  # the plan points at the beam's WEQ samples
    return p


def beam_dose(case: dict, beam_index: int = 0, dose: np.ndarray | None = None,
              use_subspots: bool = False, ct_hu_plus_1000: np.ndarray | None = None, nuclear: int = 0,
              entry_override: np.ndarray | None = None):
    """FP64 dose of one beam accumulated into ``dose`` ([nz][ny][nx] float64); returns (dose, plan).
    ``nuclear``: halo branch off (0) or its model: 1 GAUSS_FIT, 2 SOUKUP, 3 FLUKA.
    ``entry_override`` ([ry][rx] int32): a ray's entry step where >= 0 in place of the trace's."""
    keep = _Keep()
    en = make_energy(case["energy"], keep)
    bm = make_beam(case["beams"][beam_index], keep, use_subspots)
    bm.nuclear = int(nuclear)
    ct_dims, ct_res, ct_corner, dose_dims, dose_res, dose_corner = case_args(case)
    ct = ct_hu_plus_1000 if ct_hu_plus_1000 is not None else ct_as_hu_plus_1000(case["ct"], case["energy"])
    ct = np.ascontiguousarray(ct, dtype=np.float32)
    if dose is None:
        dose = np.zeros((dose_dims[2], dose_dims[1], dose_dims[0]), np.float64)
    p = Plan()
    ov = None if entry_override is None else np.ascontiguousarray(entry_override, np.int32)
    rc = lib().orc_beam_dose_entry(_fp(ct), _ip(ct_dims), _fp(ct_res), _fp(ct_corner), _dp(dose), _ip(dose_dims),
                                   _fp(dose_res), _fp(dose_corner), C.byref(bm), C.byref(en),
                                   None if ov is None else _ip(ov), C.byref(p))
    if rc != 0:
        raise RuntimeError("oracle: " + p.error.decode())
    return dose, p


def entry_margin(case: dict, beam_index: int = 0, ct_hu_plus_1000: np.ndarray | None = None) -> np.ndarray:
    """Per ray ([ry][rx]), how far the cumulative HU+1000 around the entry step lies from the
    150 threshold (orc_trace_entry_margin)."""
    p = plan_beam(case, beam_index)
    ct_dims = case_args(case)[0]
    ct = ct_hu_plus_1000 if ct_hu_plus_1000 is not None else ct_as_hu_plus_1000(case["ct"], case["energy"])
    ct = np.ascontiguousarray(ct, dtype=np.float32)
    m = np.zeros((p.rayDims[1], p.rayDims[0]))
    lib().orc_trace_entry_margin(C.byref(p), _fp(ct), _ip(ct_dims), _dp(m))
    return m


def beam_adjoint(case: dict, g_dose: np.ndarray, beam_index: int = 0, use_subspots: bool = False,
                 ct_hu_plus_1000: np.ndarray | None = None) -> np.ndarray:
    """FP64 dJ/d(spot weights) of one beam for dJ/dDose = ``g_dose`` ([nz][ny][nx]), at the
    beam's weights (spot-lattice path, halo off). Returns float64 [numSpots]."""
    keep = _Keep()
    en = make_energy(case["energy"], keep)
    bm = make_beam(case["beams"][beam_index], keep, use_subspots)
    ct_dims, ct_res, ct_corner, dose_dims, dose_res, dose_corner = case_args(case)
    ct = ct_hu_plus_1000 if ct_hu_plus_1000 is not None else ct_as_hu_plus_1000(case["ct"], case["energy"])
    ct = np.ascontiguousarray(ct, dtype=np.float32)
    g = np.ascontiguousarray(g_dose, dtype=np.float64).reshape(dose_dims[2], dose_dims[1], dose_dims[0])
    out = np.zeros(int(bm.numSpots), np.float64)
    p = Plan()
    rc = lib().orc_beam_adjoint(_fp(ct), _ip(ct_dims), _fp(ct_res), _fp(ct_corner), _dp(g), _ip(dose_dims),
                                _fp(dose_res), _fp(dose_corner), C.byref(bm), C.byref(en), _dp(out), C.byref(p))
    if rc != 0:
        raise RuntimeError("oracle: " + p.error.decode())
    return out


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's following calls."""
    lib().orc_set_threads(int(n))


def ct_as_hu_plus_1000(ct: np.ndarray, energy: dict | None = None, forced_type: int = -1) -> np.ndarray:
    """CT adapter of the reference (orc_ct_adapter; wrapper:2307-2441): the sampled min/max decide
    the input type unless ``forced_type`` (0 HU, 1 HU+1000, 2 density, 3 SPR) is given; HU is
    shifted by +1000, density / SPR are inverted through the tables of ``energy``."""
    flat = np.ascontiguousarray(ct, dtype=np.float32).ravel()
    out = np.empty_like(flat)
    keep = _Keep()
    en = make_energy(energy, keep) if energy is not None else Energy()
    t = lib().orc_ct_adapter(_fp(flat), flat.size, int(forced_type), C.byref(en), _fp(out))
    if t < 0 or (t >= 2 and energy is None):
        raise ValueError("density / SPR CT: the adapter needs the energy dict's tables")
    return out.reshape(np.shape(ct))
