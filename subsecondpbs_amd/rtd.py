"""Python entry points with the reference binding's names and argument meaning.

``raytracedicom_wrapper`` mirrors ``cudaCalDoseRTD.raytracedicom_wrapper``
(src/bindings/raytracedicom_pybind.cpp:1679-1871, 2057-2086): the same positional
arguments, the same ``beam`` / ``energy`` dict keys, a float32 dose of shape (Z, Y, X).
Differences that a caller can observe:

* the returned dose starts from zero (the reference accumulates into an uninitialised
  buffer, pybind:1705);
* errors raise ``RuntimeError`` with the engine's message instead of printing and returning.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi
from .lut import TABLES_DIR, load_energy_tables


def _complete_energy(energy: dict, tables_dir: str | None) -> dict:
    """Fill density / SP / 1/X0 LUTs from the tables when the dict lacks them
    (maybe_fill_density_sp_rrl, pybind:758-794)."""
    need = ("density_vector", "sp_vector", "rrl_vector")
    if all(k in energy for k in need):
        return energy
    ref = load_energy_tables(tables_dir or TABLES_DIR)
    out = dict(energy)
    for k in ("n_density_samples", "density_scale_fact", "density_vector", "n_sp_samples", "sp_scale_fact",
              "sp_vector", "n_rrl_samples", "rrl_scale_fact", "rrl_vector"):
        out[k] = ref[k]
    return out


def _complete_nuclear(energy: dict, tables_dir: str | None) -> dict:
    """Nuclear halo tables on the energy dict's primary axis when the dict lacks them
    (maybe_fill_nuclear_tables, pybind:796-837; alignNuclearTablesToPrimaryAxis,
    include/utils/nuclear_table_utils.h:126-189): copied when the axes agree, otherwise each
    primary energy row is resampled at the same depths from the bracketing reference rows."""
    if "nuclear_weight_matrix" in energy and "nuclear_sq_sigma_matrix" in energy:
        return energy
    ref = load_energy_tables(tables_dir or TABLES_DIR, nuclear=True)
    out = dict(energy)
    e_out = np.asarray(energy["energies_per_u"], np.float32)
    same = (int(energy["n_energies"]) == ref["n_energies"] and int(energy["n_energy_samples"]) == ref["n_energy_samples"]
            and np.allclose(e_out, ref["energies_per_u"]) and np.allclose(energy["peak_depths"], ref["peak_depths"])
            and np.allclose(energy["scale_facts"], ref["scale_facts"]))
    if same:
        out["nuclear_weight_matrix"] = ref["nuclear_weight_matrix"]
        out["nuclear_sq_sigma_matrix"] = ref["nuclear_sq_sigma_matrix"]
        return out
    nE, nS = int(energy["n_energies"]), int(energy["n_energy_samples"])
    e_ref = np.asarray(ref["energies_per_u"], np.float64)
    sc_ref = np.asarray(ref["scale_facts"], np.float64)
    sc_out = np.asarray(energy["scale_facts"], np.float64)
    w = np.zeros((nE, nS), np.float32)
    s2 = np.zeros((nE, nS), np.float32)
    xs = np.arange(ref["n_energy_samples"])
    for e in range(nE):
        idx = float(np.interp(e_out[e], e_ref, np.arange(e_ref.size)))
        e0 = int(math.floor(idx)); fr = idx - e0
        e1 = min(e0 + (1 if fr > 0 else 0), e_ref.size - 1)
        depth = np.arange(nS) / sc_out[e] if sc_out[e] > 0 else np.zeros(nS)
        for m_ref, m_out in ((ref["nuclear_weight_matrix"], w), (ref["nuclear_sq_sigma_matrix"], s2)):
            a = np.interp(depth * sc_ref[e0], xs, m_ref[e0])
            b = np.interp(depth * sc_ref[e1], xs, m_ref[e1])
            m_out[e] = a + (b - a) * fr
    out["nuclear_weight_matrix"] = w
    out["nuclear_sq_sigma_matrix"] = s2
    return out


def _check_dose(dose: np.ndarray, beam: dict, where: str):
    # require_nonempty_final_dose (pybind:1069-1081)
    if not np.isfinite(dose).all():
        raise RuntimeError(f"{where}: dose contains NaN/Inf")
    if not (dose != 0).any():
        raise RuntimeError(f"{where}: dose is all zero")


def subsecond_wrapper(ct, ct_dims, ct_resolution, ct_corner, dose, dose_dims, dose_resolution, dose_corner,
                      beams, energy, gpu_id: int = 0, nuclear_correction: bool = False, verbose: int = 0):
    """Accumulate the dose of ``beams`` (list of beam dicts) into ``dose`` (float32, x fastest).
    ``nuclear_correction``: False/0 off, True/1 halo branch with GAUSS_FIT constants, 2 SOUKUP, 3 FLUKA."""
    keep = _abi.Keep()
    es = _abi.energy_struct(energy, keep)
    arr = (_abi.BeamSettings * len(beams))()
    for i, b in enumerate(beams):
        arr[i] = _abi.beam_settings(b, keep)
    ct = np.ascontiguousarray(ct, dtype=np.float32)
    if ct.size != int(np.prod(ct_dims)):
        raise RuntimeError("ct buffer size does not match ct_dims")
    if dose.dtype != np.float32 or not dose.flags.c_contiguous or dose.size != int(np.prod(dose_dims)):
        raise RuntimeError("dose must be a C-contiguous float32 buffer of dose_dims")
    cd, cdp = _abi.i3(ct_dims)
    cr, crp = _abi.f3(ct_resolution)
    cc, ccp = _abi.f3(ct_corner)
    dd, ddp = _abi.i3(dose_dims)
    dr, drp = _abi.f3(dose_resolution)
    dc, dcp = _abi.f3(dose_corner)
    rc = _abi.lib().pbs_subsecond_wrapper(ct.ctypes.data_as(_abi._f), cdp, crp, ccp, dose.ctypes.data_as(_abi._f),
                                          ddp, drp, dcp, arr, len(beams), es, int(gpu_id), int(nuclear_correction),
                                          int(verbose))
    _abi.check(rc, "subsecondWrapper")
    return dose


def raytracedicom_wrapper(ct, ct_dims, ct_resolution, ct_corner, dose_dims, dose_resolution, dose_corner,
                          beam: dict, energy: dict, gpu_id: int = 0, nuclear_correction: bool = False,
                          verbose: int = 0, tables_dir: str | None = None) -> np.ndarray:
    energy = _complete_energy(energy, tables_dir)
    if nuclear_correction:
        energy = _complete_nuclear(energy, tables_dir)
    for key in ("energies", "spot_sigmas", "steps", "subspot_data", "beam_direction", "beam_xdir", "beam_ydir",
                "sad", "source_position"):
        if key not in beam:
            raise KeyError(f"beam dict is missing required key '{key}'")
    dose = np.zeros((int(dose_dims[2]), int(dose_dims[1]), int(dose_dims[0])), np.float32)
    subsecond_wrapper(ct, ct_dims, ct_resolution, ct_corner, dose, dose_dims, dose_resolution, dose_corner, [beam],
                      energy, gpu_id, nuclear_correction, verbose)
    _check_dose(dose, beam, "raytracedicom_wrapper")
    return dose
