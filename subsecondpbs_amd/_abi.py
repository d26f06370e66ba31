"""ctypes bindings of the engine's C ABI (include/pbs_abi.h).

The library is the in-tree ``subsecondpbs_amd/_lib/libpbs_hip.so`` built by
``subsecondpbs_amd.build``. Importing this module does not touch the GPU; loading fails
loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpbs_hip.so")

_f = C.POINTER(C.c_float)
_i = C.POINTER(C.c_int)

PBS_OK = 0
ERRORS = {-1: "invalid argument", -2: "geometry", -3: "device", -4: "out of memory", -5: "unsupported"}


class BeamSettings(C.Structure):
    _fields_ = [
        ("numLayers", C.c_int), ("energies", _f), ("spotSigmas", _f),
        ("raySpacing", C.c_float * 2), ("steps", C.c_uint), ("sourceDist", C.c_float * 2),
        ("spotOffset", C.c_float * 3), ("spotDelta", C.c_float * 3),
        ("maxSubspotsPerLayer", C.c_int), ("subspotData", _f),
        ("layerSpotCounts", _i), ("numSpots", C.c_int), ("spotPositions", _f), ("spotWeights", _f),
        ("spotBeamDirections", _f), ("layerLongitudinalCutoffs", _f),
        ("beamDirection", C.c_float * 3), ("bmXDirection", C.c_float * 3),
        ("bmYDirection", C.c_float * 3), ("sourcePosition", C.c_float * 3),
        ("sad", C.c_float), ("refPlaneZ", C.c_float),
        ("roiLinearIndices", _i), ("numRoiLinearIndices", C.c_int),
        ("useSubspotConvolution", C.c_int),
        ("waterEquivalence", _f), ("waterEquivalenceCount", C.c_size_t), ("spotPositionsAreIndices", C.c_int),
        ("profileEnergies", _f), ("numProfileEnergies", C.c_int), ("profileData", _f),
        ("profileDataCount", C.c_size_t), ("profileSetting", _f), ("numProfileSetting", C.c_int),
        ("beamParaData", _f), ("numBeamParaData", C.c_int), ("beamParaPos", C.c_float),
        ("carbonProfileHooks", C.c_int),
    ]


class EnergyStruct(C.Structure):
    _fields_ = [
        ("nEnergySamples", C.c_int), ("nEnergies", C.c_int),
        ("energiesPerU", _f), ("peakDepths", _f), ("scaleFacts", _f), ("ciddMatrix", _f),
        ("nDensitySamples", C.c_int), ("densityScaleFact", C.c_float), ("densityVector", _f),
        ("nSpSamples", C.c_int), ("spScaleFact", C.c_float), ("spVector", _f),
        ("nRRlSamples", C.c_int), ("rRlScaleFact", C.c_float), ("rRlVector", _f),
        ("nucWeightMatrix", _f), ("nucSqSigmaMatrix", _f),
        ("ionRangeCoef", C.c_float), ("ionRangeExp", C.c_float), ("ionRestMass", C.c_float),
        ("ionScatterScale", C.c_float),
    ]


class DebugInts(C.Structure):
    _fields_ = [
        ("rayDims", C.c_int * 2), ("superpDims", C.c_int * 2), ("steps", C.c_int), ("numLayers", C.c_int),
        ("legacyPath", C.c_int),
        ("beamFirstInside", C.c_int), ("beamFirstOutside", C.c_int), ("firstPastCutoffAll", C.c_int),
        ("guaranteedPassive", C.c_int),
        ("afterLast", C.c_int * 256), ("layerPassive", C.c_int * 256),
        ("transferStart", C.c_int * 3), ("transferMax", C.c_int * 3),
        ("cpbCorner", C.c_float * 3), ("fanCorner", C.c_float * 3), ("fanDelta", C.c_float * 3),
        ("sourceDist", C.c_float * 2),
        ("energyIdx", C.c_float * 256), ("scaleFact", C.c_float * 256), ("peakDepth", C.c_float * 256),
        ("numGroups", C.c_int), ("tailPlanes", C.c_int), ("tailOverflow", C.c_int), ("mainEndMax", C.c_int),
        ("mainEnd", C.c_int * 256), ("layerGroup", C.c_int * 256),
        ("groupPassive", C.c_int * 256), ("groupTailStart", C.c_int * 256), ("groupTailCount", C.c_int * 256),
    ]


class PlanInfo(C.Structure):
    """pbs_plan_info: the host planner's grids (pbs_plan_debug)."""
    _fields_ = [
        ("legacyPath", C.c_int), ("spotIndices", C.c_int), ("weqVolume", C.c_int), ("numLayers", C.c_int),
        ("rayDims", C.c_int * 2), ("superpDims", C.c_int * 2), ("spotDims", C.c_int * 2), ("steps", C.c_int),
        ("spotOffset", C.c_float * 2), ("spotDelta", C.c_float * 2), ("cpbCorner", C.c_float * 2),
        ("cpbRes", C.c_float * 2), ("fanCorner", C.c_float * 3), ("fanDelta", C.c_float * 3),
        ("sourceDist", C.c_float * 2), ("volConst", C.c_float), ("volLin", C.c_float), ("volSq", C.c_float),
        ("lenToMm", C.c_float), ("energyDepthToMm", C.c_float),
        ("profileHooks", C.c_int), ("profileRows", C.c_int), ("profileDepthN", C.c_int),
        ("profileChannels", C.c_int),
    ]


class Vec3f(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Vec3i(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class CPBGrid(C.Structure):
    _fields_ = [("corner", Vec3f), ("resolution", Vec3f), ("dims", Vec3i), ("nLayers", C.c_int)]


# every symbol include/pbs_abi.h declares
EXPORTED = [
    "pbs_subsecond_wrapper", "pbs_last_error", "pbs_device_count",
    "pbs_engine_create", "pbs_engine_destroy", "pbs_engine_set_ct", "pbs_engine_set_dose_grid",
    "pbs_abi_sizeof", "pbs_engine_beam_dose", "pbs_engine_beam_dose_noalloc", "pbs_engine_beam_adjoint", "pbs_engine_keep_beam", "pbs_engine_update_beam",
    "pbs_engine_update_adjoint", "pbs_resident_beam_use_graph", "pbs_resident_beam_destroy", "pbs_engine_sync", "pbs_engine_stream", "pbs_engine_debug_fetch",
    "pbs_engine_set_profiling", "pbs_engine_stage_times", "pbs_engine_set_tuning", "pbs_engine_get_tuning", "pbs_engine_set_nuclear", "pbs_engine_debug_halo", "pbs_engine_debug_bins",
    "pbs_plan_debug", "pbs_ct_to_hu_plus_1000",
    "pbs_final_dose_bind", "cudaFinalDose", "pbs_final_dose_roi",
    "pbs_batch_create", "pbs_batch_destroy", "pbs_batch_set_ct", "pbs_batch_set_dose_grid", "pbs_batch_set_nuclear",
    "pbs_batch_dose", "pbs_batch_submit", "pbs_batch_query", "pbs_batch_wait", "pbs_batch_sync", "pbs_batch_lanes",
    "pbs_batch_add_table", "pbs_batch_dose_tables", "pbs_batch_last_launches", "pbs_batch_graph_captures",
    "pbs_server_create", "pbs_server_destroy", "pbs_server_add_ct", "pbs_server_remove_ct", "pbs_server_submit",
    "pbs_server_submit_weights", "pbs_server_query", "pbs_server_wait", "pbs_server_result",
    "pbs_server_result_device", "pbs_server_timing",
    "pbs_server_release_state", "pbs_server_memory",
]

_lib = None


def lib():
    """Load libpbs_hip.so (raises when it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP engine library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.pbs_last_error.restype = C.c_char_p
        L.pbs_abi_sizeof.restype = C.c_long
        L.pbs_abi_sizeof.argtypes = [C.c_int]
        L.pbs_device_count.restype = C.c_int
        L.pbs_subsecond_wrapper.restype = C.c_int
        L.pbs_subsecond_wrapper.argtypes = [_f, _i, _f, _f, _f, _i, _f, _f, C.POINTER(BeamSettings), C.c_size_t,
                                            C.POINTER(EnergyStruct), C.c_int, C.c_int, C.c_int]
        L.pbs_engine_create.restype = C.c_void_p
        L.pbs_engine_create.argtypes = [C.c_int, C.POINTER(EnergyStruct)]
        L.pbs_engine_destroy.argtypes = [C.c_void_p]
        L.pbs_engine_set_ct.restype = C.c_int
        L.pbs_engine_set_ct.argtypes = [C.c_void_p, C.c_void_p, _i, _f, _f, C.c_int]
        L.pbs_engine_set_dose_grid.restype = C.c_int
        L.pbs_engine_set_dose_grid.argtypes = [C.c_void_p, _i, _f, _f]
        L.pbs_engine_beam_dose.restype = C.c_int
        L.pbs_engine_beam_dose.argtypes = [C.c_void_p, C.POINTER(BeamSettings), C.c_void_p]
        L.pbs_engine_beam_dose_noalloc.restype = C.c_int
        L.pbs_engine_beam_dose_noalloc.argtypes = [C.c_void_p, C.POINTER(BeamSettings), C.c_void_p]
        L.pbs_engine_beam_adjoint.restype = C.c_int
        L.pbs_engine_beam_adjoint.argtypes = [C.c_void_p, C.POINTER(BeamSettings), C.c_void_p, C.c_void_p]
        L.pbs_engine_keep_beam.restype = C.c_int
        L.pbs_engine_keep_beam.argtypes = [C.c_void_p, C.POINTER(BeamSettings), C.c_void_p, C.POINTER(C.c_void_p)]
        L.pbs_engine_update_beam.restype = C.c_int
        L.pbs_engine_update_beam.argtypes = [C.c_void_p, C.c_void_p, _f, C.c_int, C.c_void_p]
        L.pbs_engine_update_adjoint.restype = C.c_int
        L.pbs_engine_update_adjoint.argtypes = [C.c_void_p, C.c_void_p, _f, C.c_int, C.c_void_p, C.c_void_p]
        L.pbs_resident_beam_use_graph.restype = C.c_int
        L.pbs_resident_beam_use_graph.argtypes = [C.c_void_p, C.c_int]
        L.pbs_resident_beam_destroy.restype = None
        L.pbs_resident_beam_destroy.argtypes = [C.c_void_p]
        L.pbs_engine_sync.restype = C.c_int
        L.pbs_engine_sync.argtypes = [C.c_void_p]
        L.pbs_engine_stream.restype = C.c_void_p
        L.pbs_engine_stream.argtypes = [C.c_void_p]
        L.pbs_engine_debug_fetch.restype = C.c_int
        L.pbs_engine_debug_fetch.argtypes = [C.c_void_p, C.POINTER(DebugInts)] + [C.c_void_p] * 11
        L.pbs_engine_set_profiling.restype = C.c_int
        L.pbs_engine_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.pbs_engine_stage_times.restype = C.c_int
        L.pbs_engine_stage_times.argtypes = [C.c_void_p, _f, _i]
        L.pbs_engine_set_tuning.restype = C.c_int
        L.pbs_engine_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.pbs_engine_get_tuning.restype = C.c_int
        L.pbs_engine_get_tuning.argtypes = [C.c_void_p, C.c_int, _i]
        L.pbs_engine_set_nuclear.restype = C.c_int
        L.pbs_engine_set_nuclear.argtypes = [C.c_void_p, C.c_int]
        L.pbs_engine_debug_bins.restype = C.c_int
        L.pbs_engine_debug_bins.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.pbs_ct_to_hu_plus_1000.restype = C.c_int
        L.pbs_ct_to_hu_plus_1000.argtypes = [_f, C.c_size_t, C.POINTER(EnergyStruct), _f, _i]
        L.pbs_plan_debug.restype = C.c_int
        L.pbs_plan_debug.argtypes = [_i, _f, _f, _i, _f, _f, C.POINTER(BeamSettings), C.POINTER(EnergyStruct),
                                     C.POINTER(PlanInfo), C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                     C.c_void_p, C.c_void_p]
        L.pbs_engine_debug_halo.restype = C.c_int
        L.pbs_engine_debug_halo.argtypes = [C.c_void_p, _i, _f, C.c_void_p, C.c_void_p, _i]
        L.pbs_batch_create.restype = C.c_void_p
        L.pbs_batch_create.argtypes = [C.c_int, C.POINTER(EnergyStruct), C.c_int]
        L.pbs_batch_destroy.argtypes = [C.c_void_p]
        L.pbs_batch_set_ct.restype = C.c_int
        L.pbs_batch_set_ct.argtypes = [C.c_void_p, C.c_void_p, _i, _f, _f, C.c_int]
        L.pbs_batch_set_dose_grid.restype = C.c_int
        L.pbs_batch_set_dose_grid.argtypes = [C.c_void_p, _i, _f, _f]
        L.pbs_batch_set_nuclear.restype = C.c_int
        L.pbs_batch_set_nuclear.argtypes = [C.c_void_p, C.c_int]
        L.pbs_batch_dose.restype = C.c_int
        L.pbs_batch_dose.argtypes = [C.c_void_p, C.POINTER(BeamSettings), C.c_size_t, C.c_void_p]
        L.pbs_batch_add_table.restype = C.c_int
        L.pbs_batch_add_table.argtypes = [C.c_void_p, C.POINTER(EnergyStruct), C.c_int]
        L.pbs_batch_dose_tables.restype = C.c_int
        L.pbs_batch_dose_tables.argtypes = [C.c_void_p, C.POINTER(BeamSettings), C.POINTER(C.c_int), C.c_size_t,
                                            C.c_void_p]
        L.pbs_batch_last_launches.restype = C.c_int
        L.pbs_batch_last_launches.argtypes = [C.c_void_p]
        L.pbs_batch_graph_captures.restype = C.c_int
        L.pbs_batch_graph_captures.argtypes = [C.c_void_p]
        L.pbs_batch_submit.restype = C.c_int
        L.pbs_batch_submit.argtypes = [C.c_void_p, C.POINTER(BeamSettings), C.c_void_p, C.POINTER(C.c_longlong)]
        L.pbs_batch_query.restype = C.c_int
        L.pbs_batch_query.argtypes = [C.c_void_p, C.c_longlong]
        L.pbs_batch_wait.restype = C.c_int
        L.pbs_batch_wait.argtypes = [C.c_void_p, C.c_longlong]
        L.pbs_batch_sync.restype = C.c_int
        L.pbs_batch_sync.argtypes = [C.c_void_p]
        L.pbs_batch_lanes.restype = C.c_int
        L.pbs_batch_lanes.argtypes = [C.c_void_p]
        LL = C.POINTER(C.c_longlong)
        L.pbs_server_create.restype = C.c_void_p
        L.pbs_server_create.argtypes = [C.c_int, C.POINTER(EnergyStruct), C.c_int, C.c_double]
        L.pbs_server_destroy.restype = None
        L.pbs_server_destroy.argtypes = [C.c_void_p]
        L.pbs_server_add_ct.restype = C.c_int
        L.pbs_server_add_ct.argtypes = [C.c_void_p, C.c_void_p, _i, _f, _f, C.c_int, C.POINTER(C.c_int)]
        L.pbs_server_remove_ct.restype = C.c_int
        L.pbs_server_remove_ct.argtypes = [C.c_void_p, C.c_int]
        L.pbs_server_submit.restype = C.c_int
        L.pbs_server_submit.argtypes = [C.c_void_p, C.c_int, _i, _f, _f, C.POINTER(BeamSettings), C.c_int, LL]
        L.pbs_server_submit_weights.restype = C.c_int
        L.pbs_server_submit_weights.argtypes = [C.c_void_p, C.c_longlong, _f, C.c_int, LL]
        for fn in ("pbs_server_query", "pbs_server_wait", "pbs_server_release_state"):
            getattr(L, fn).restype = C.c_int
            getattr(L, fn).argtypes = [C.c_void_p, C.c_longlong]
        L.pbs_server_result.restype = C.c_int
        L.pbs_server_result.argtypes = [C.c_void_p, C.c_longlong, C.c_void_p]
        L.pbs_server_result_device.restype = C.c_int
        L.pbs_server_result_device.argtypes = [C.c_void_p, C.c_longlong, C.c_void_p]
        L.pbs_server_timing.restype = C.c_int
        L.pbs_server_timing.argtypes = [C.c_void_p, C.c_longlong, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.pbs_server_memory.restype = C.c_int
        L.pbs_server_memory.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.pbs_final_dose_bind.restype = C.c_int
        L.pbs_final_dose_bind.argtypes = [C.c_void_p, _f, _f, C.c_int, C.c_float]
        L.cudaFinalDose.restype = C.c_int
        L.cudaFinalDose.argtypes = [C.c_void_p, C.c_void_p, C.c_int, Vec3f, Vec3f, Vec3f, Vec3f, C.c_float,
                                    C.c_float, C.c_int, C.c_int, C.c_void_p, Vec3i]
        L.pbs_final_dose_roi.restype = C.c_int
        L.pbs_final_dose_roi.argtypes = [C.POINTER(Vec3i), C.c_int, Vec3f, Vec3f, Vec3f, Vec3f, C.c_int,
                                         C.POINTER(Vec3f), C.POINTER(Vec3f), C.POINTER(Vec3i), C.POINTER(CPBGrid)]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().pbs_last_error().decode(errors="replace")


def check(rc: int, what: str):
    if rc != PBS_OK:
        raise RuntimeError(f"{what} failed ({ERRORS.get(rc, rc)}): {last_error()}")


class Keep:
    """Owns the numpy buffers a ctypes struct points at."""

    def __init__(self):
        self.bufs = []

    def _ptr(self, a, ptype):
        # a pointer object viewing a c_void_p that holds the buffer's address (several times
        # cheaper than ndarray.ctypes.data_as); the c_void_p is kept with the buffer
        self.bufs.append(a)
        holder = C.c_void_p(a.__array_interface__["data"][0])
        self.bufs.append(holder)
        return ptype.from_address(C.addressof(holder))

    def f32(self, a):
        if a is None:
            return None
        return self._ptr(np.ascontiguousarray(a, dtype=np.float32), _f)

    def i32(self, a):
        if a is None:
            return None
        return self._ptr(np.ascontiguousarray(a, dtype=np.int32), _i)


def energy_struct(energy: dict, keep: Keep) -> EnergyStruct:
    e = EnergyStruct()
    e.nEnergies = int(energy["n_energies"])
    e.nEnergySamples = int(energy["n_energy_samples"])
    e.energiesPerU = keep.f32(energy["energies_per_u"])
    e.peakDepths = keep.f32(energy["peak_depths"])
    e.scaleFacts = keep.f32(energy["scale_facts"])
    cidd = np.asarray(energy["cidd_matrix"], dtype=np.float32)
    if cidd.size != e.nEnergies * e.nEnergySamples:
        raise ValueError("cidd_matrix shape mismatch with n_energies/n_energy_samples")
    e.ciddMatrix = keep.f32(cidd)
    e.nDensitySamples = int(energy["n_density_samples"])
    e.densityScaleFact = float(energy["density_scale_fact"])
    e.densityVector = keep.f32(energy["density_vector"])
    e.nSpSamples = int(energy["n_sp_samples"])
    e.spScaleFact = float(energy["sp_scale_fact"])
    e.spVector = keep.f32(energy["sp_vector"])
    e.nRRlSamples = int(energy["n_rrl_samples"])
    e.rRlScaleFact = float(energy["rrl_scale_fact"])
    e.rRlVector = keep.f32(energy["rrl_vector"])
    if energy.get("nuclear_weight_matrix") is not None:
        e.nucWeightMatrix = keep.f32(energy["nuclear_weight_matrix"])
        e.nucSqSigmaMatrix = keep.f32(energy["nuclear_sq_sigma_matrix"])
    ion = energy.get("ion")
    if ion:
        e.ionRangeCoef = float(ion["range_coef"])
        e.ionRangeExp = float(ion["range_exp"])
        e.ionRestMass = float(ion["rest_mass"])
        e.ionScatterScale = float(ion["scatter_scale"])
    return e


def _vec(beam, key, n, default):
    v = beam.get(key, default)
    v = np.asarray(v, dtype=np.float32).ravel()
    if v.size != n:
        raise ValueError(f"beam['{key}'] must have {n} values")
    return v.tolist()


def beam_settings(beam: dict, keep: Keep, use_subspots: bool | None = None) -> BeamSettings:
    """Beam dict (reference binding key names, raytracedicom_pybind.cpp:1712-1821) -> POD struct."""
    b = BeamSettings()
    energies = np.asarray(beam["energies"], dtype=np.float32).ravel()
    L = energies.size
    b.numLayers = L
    b.energies = keep.f32(energies)
    sig = np.asarray(beam["spot_sigmas"], dtype=np.float32)
    if sig.size != 2 * L:
        raise ValueError("beam.spot_sigmas length must equal num_layers")
    b.spotSigmas = keep.f32(sig.reshape(L, 2))
    b.raySpacing[:] = _vec(beam, "ray_spacing", 2, (0.0, 0.0))
    b.steps = int(beam.get("steps", 0))
    b.sourceDist[:] = _vec(beam, "source_dist", 2, (0.0, 0.0))
    b.spotOffset[:] = _vec(beam, "spot_offset", 3, (0.0, 0.0, 0.0))
    b.spotDelta[:] = _vec(beam, "spot_delta", 3, (0.0, 0.0, 0.0))
    ss = np.asarray(beam["subspot_data"], dtype=np.float32)
    if ss.ndim != 3 or ss.shape[2] != 5:
        raise ValueError("subspot_data must have shape (num_layers, max_subspots_per_layer, 5)")
    if ss.shape[0] != L:
        raise ValueError("beam.energies length must equal subspot_data.shape[0] (num_layers)")
    b.maxSubspotsPerLayer = int(ss.shape[1])
    b.subspotData = keep.f32(ss)
    if beam.get("layer_spot_counts") is not None:
        b.layerSpotCounts = keep.i32(beam["layer_spot_counts"])
        w = np.asarray(beam["spot_weights"], dtype=np.float32).ravel()
        b.numSpots = int(w.size)
        b.spotWeights = keep.f32(w)
        b.spotPositions = keep.f32(beam["spot_positions"])
    if beam.get("spot_beam_directions") is not None:
        b.spotBeamDirections = keep.f32(beam["spot_beam_directions"])
    if beam.get("layer_longitudinal_cutoffs") is not None:
        lc = np.asarray(beam["layer_longitudinal_cutoffs"], dtype=np.float32).ravel()
        if lc.size != L:
            raise ValueError("beam.layer_longitudinal_cutoffs length must equal num_layers")
        b.layerLongitudinalCutoffs = keep.f32(lc)
    b.beamDirection[:] = _vec(beam, "beam_direction", 3, None)
    b.bmXDirection[:] = _vec(beam, "beam_xdir", 3, None)
    b.bmYDirection[:] = _vec(beam, "beam_ydir", 3, None)
    b.sourcePosition[:] = _vec(beam, "source_position", 3, None)
    b.sad = float(beam["sad"])
    b.refPlaneZ = float(beam.get("ref_plane_z", 0.0))
    if beam.get("roi_linear_indices") is not None:
        roi = np.asarray(beam["roi_linear_indices"], dtype=np.int32).ravel()
        b.roiLinearIndices = keep.i32(roi)
        b.numRoiLinearIndices = int(roi.size)
    if use_subspots is None:
        use_subspots = bool(beam.get("use_subspot_convolution", False))
    b.useSubspotConvolution = 1 if use_subspots else 0
    weq = weq_vector(beam)
    if weq is not None:
        b.waterEquivalence = keep.f32(weq)
        b.waterEquivalenceCount = int(weq.size)
    b.spotPositionsAreIndices = 1 if beam.get("spot_positions_are_indices", False) else 0
    # CarbonPBS profile model (reference keys, raytracedicom_pybind.cpp:1782-1796) and the hooks
    # switch of this build (0: off, as in the reference)
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
    return b


def plan_debug(case: dict, beam: dict, use_subspots: bool | None = None) -> dict:
    """The host planner's result for ``beam`` on the case's CT and dose grid (pbs_plan_debug; no
    device needed): the PlanInfo fields, per-layer energy index and peak depth, the dense spot
    lattice [L][ny][nx] and each spot's cell."""
    keep = Keep()
    es = energy_struct(case["energy"], keep)
    bs = beam_settings(beam, keep, use_subspots)
    d, dp = i3(case["ct_dims"])
    r, rp = f3(case["ct_resolution"])
    c, cp = f3(case["ct_corner"])
    dd, ddp = i3(case["dose_dims"])
    dr, drp = f3(case["dose_resolution"])
    dc, dcp = f3(case["dose_corner"])
    info = PlanInfo()
    L = int(bs.numLayers)
    rc = lib().pbs_plan_debug(dp, rp, cp, ddp, drp, dcp, C.byref(bs), C.byref(es), C.byref(info), None, None,
                              None, 0, None, None)
    check(rc, "pbs_plan_debug")
    eidx = np.zeros(L, np.float32)
    peak = np.zeros(L, np.float32)
    nx, ny = info.spotDims
    dense = np.zeros(max(L * nx * ny, 1), np.float32)
    cells = np.zeros(max(int(bs.numSpots), 1), np.int32)
    prof = np.zeros((L, 4), np.float32)
    rc = lib().pbs_plan_debug(dp, rp, cp, ddp, drp, dcp, C.byref(bs), C.byref(es), C.byref(info), eidx.ctypes.data,
                              peak.ctypes.data, dense.ctypes.data, dense.size, cells.ctypes.data, prof.ctypes.data)
    check(rc, "pbs_plan_debug")
    out = {k: (list(getattr(info, k)) if not isinstance(getattr(info, k), (int, float)) else getattr(info, k))
           for k, _ in PlanInfo._fields_}
    out["energyIdx"], out["peakDepth"] = eidx, peak
    out["dense"] = dense[:L * nx * ny].reshape(L, ny, nx) if not info.legacyPath else None
    out["spotCell"] = cells[:int(bs.numSpots)]
    out["profileLayer"] = prof
    return out


def ct_to_hu_plus_1000(ct, energy: dict):
    """The engine's CT adapter on the host (pbs_ct_to_hu_plus_1000): (HU+1000 volume, input type
    used: 0 HU, 1 HU+1000, 2 density, 3 SPR). RTD_CT_INPUT_TYPE forces the type."""
    keep = Keep()
    es = energy_struct(energy, keep)
    a = np.ascontiguousarray(ct, dtype=np.float32)
    out = np.empty_like(a)
    t = C.c_int(-1)
    check(lib().pbs_ct_to_hu_plus_1000(a.ctypes.data_as(_f), a.size, C.byref(es), out.ctypes.data_as(_f),
                                       C.byref(t)), "pbs_ct_to_hu_plus_1000")
    return out, t.value


def weq_vector(beam: dict):
    """The beam's WEQ header (+ samples): ``water_equivalence``, else ``ray_weq_header``
    (raytracedicom_pybind.cpp:1767-1781); None when neither holds the 9-value header."""
    for key in ("water_equivalence", "ray_weq_header"):
        v = beam.get(key)
        if v is not None:
            a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).ravel())
            if a.size >= 9:
                return a
    return None


def i3(v):
    a = np.ascontiguousarray(v, dtype=np.int32).ravel()
    return a, a.ctypes.data_as(_i)


def f3(v):
    a = np.ascontiguousarray(v, dtype=np.float32).ravel()
    return a, a.ctypes.data_as(_f)
