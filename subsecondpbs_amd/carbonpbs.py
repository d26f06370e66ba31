"""The CarbonPBS Python contract of the reference binding: ``calcDose`` / ``cuCalDose`` /
``cuFinalDose``, ``cuCalDose3`` and ``debug_build_carbonpbs_context``
(src/bindings/raytracedicom_pybind.cpp:1121-1373, 1469-1635, 1873-2050, 2088-2313).

A CarbonPBS caller hands over the beam as CarbonPBS arrays: a WEQ volume (``rayweq``: 9-value
header + water-equivalent depths), spots as WEQ texel coordinates (``idbeamxy``) with particle
counts (``nPar``) and per-spot beam directions, an IDD table (``enelist`` / ``idddata`` /
``iddsetting``) instead of the CIDD tables, and a subspot LUT on the IDD table's energy axis.
``build_carbonpbs_context`` turns them into the beam / energy dicts of ``raytracedicom_wrapper``
the way the reference does:

* CIDD = running sum of IDD x step per energy row; peak depth = depth of the IDD maximum
  (derive_peak_depths_from_idd, pybind:492-519; build_energy_from_carbonpbs, pybind:521-638),
  with the reference's depth-unit heuristics (cm tables x10) and a compacted depth axis
  (iddsetting[2] > columns: the step stretched); fewer IDD rows than ``enelist`` entries take the
  tail (or, with RTD_IDD_SLICE_MODE=head, the head) of ``enelist``;
* the density / SP / 1/X0 tables from ``tables_dir``;
* per-layer spot sigmas = the LUT's first subspot sigmas (pybind:640-657); a LUT on another
  energy axis is interpolated to the beam's layer energies (pybind:717-756);
* ROI indices as (N, 3) / (3, N) / flat xyz triplets or flat x-major linear indices
  (pybind:856-929);
* one longitudinal cut-off per layer (pybind:659-693; audit only, as in the reference).

``cuFinalDose`` builds the placeholder CT of the reference (1000 at the ROI's x-major linear
indices read as a z-major buffer, 0 elsewhere; all 1000 without a ROI; pybind:1601-1624), which
only bounds the geometry, since the WEQ volume replaces the trace; then runs the beam and
writes the dose into ``finalDose``: flat or C-contiguous (Nz, Ny, Nx) buffers accumulate,
(Nx, Ny, Nz) buffers receive the dose in that order (pybind:980-1028; the (Nx, Ny, Nz) shape is
tested first, so a cubic grid's 3-D buffer is filled x-major). The profile / beam-
parameter arrays travel with the beam; their profile-sigma and overall-weight hooks stay off
unless ``carbon_profile_hooks`` turns them on (the reference gates them off,
wrapper:2812-2814). ``cuCalDose3`` zeroes the CSC outputs (no spot matrix is formed, as in the
reference) and returns the dose.
"""
from __future__ import annotations

import os

import numpy as np

from .lut import TABLES_DIR, load_energy_tables
from .rtd import _check_dose, _complete_nuclear, subsecond_wrapper


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class ContractError(RuntimeError):
    """A CarbonPBS argument of the wrong shape or size. A RuntimeError, like the exception the
    reference's bound C++ raises for its argument checks."""


def to_float_vector(a, name: str = "array") -> np.ndarray:
    """``a`` as a flat float32 vector; a 2-D array counts as a vector when one extent is 1."""
    v = _f32(a)
    if v.ndim <= 1 or (v.ndim == 2 and min(v.shape) == 1):
        return v.ravel()
    raise ContractError(f"{name}: got shape {v.shape}, a vector was expected")


def to_vec_columns(a, k: int, expected: int = 0, name: str = "points") -> np.ndarray:
    """k-component points as a flat [N][k] float32 array. Accepted forms, tried in this order
    (pybind:141-223): a flat run of N*k values, rows (N, k), columns (k, N)."""
    v = _f32(a)
    if v.ndim == 1 and v.size % k == 0:
        flat = v
    elif v.ndim == 2 and v.shape[1] == k:
        flat = v.ravel()
    elif v.ndim == 2 and v.shape[0] == k:
        flat = np.ascontiguousarray(v.T).ravel()
    else:
        raise ContractError(f"{name}: {k}-component points take shape (N, {k}), ({k}, N) or a flat length "
                            f"divisible by {k}; got {v.shape}")
    if expected and flat.size != expected * k:
        raise ContractError(f"{name}: {flat.size // k} points given where {expected} are needed")
    return flat


def normalized_mean_direction(dirs: np.ndarray) -> np.ndarray:
    """Unit mean of the per-spot directions: components summed in double, divided by the count,
    then normalised in float (pybind:233-259)."""
    d = np.asarray(dirs, np.float32).reshape(-1, 3)
    if not d.shape[0]:
        raise ContractError("spot directions: none given")
    m = (d.astype(np.float64).sum(axis=0) / d.shape[0]).astype(np.float32)
    n = np.float32(np.sqrt(np.float32(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])))
    return m / n if n > 0 else m


def _slice_mode(var: str = "RTD_IDD_SLICE_MODE") -> str:
    v = os.environ.get(var, "").strip().lower()
    return "head" if v in ("head", "prefix") else "tail"


def _lround(v: float) -> int:
    """std::lround: halves away from zero."""
    return int(np.sign(v) * np.floor(abs(v) + 0.5))


def build_energy_from_carbonpbs(enelist, idddata, iddsetting, tables_dir: str | None = None) -> dict:
    """Energy dict from a CarbonPBS IDD table (pybind:521-638)."""
    energies = to_float_vector(enelist, "enelist")
    if not energies.size:
        raise ContractError("enelist: no energies given")
    setting = to_float_vector(iddsetting, "iddsetting")
    if setting.size < 3:
        raise ContractError(f"iddsetting: {setting.size} values given; (start, step, samples) are needed")
    idd = _f32(idddata)
    if idd.ndim != 2:
        raise ContractError(f"idddata: got {idd.ndim} dimensions; a table (energies, depth samples) is needed")
    rows, cols = idd.shape
    if rows <= 0 or rows > energies.size:
        raise ContractError(f"idddata: {rows} rows for {energies.size} energies (1 .. {energies.size} rows allowed)")
    if rows != energies.size:
        energies = energies[:rows] if _slice_mode() == "head" else energies[energies.size - rows:]
    start_raw, step_raw = np.float32(setting[0]), np.float32(setting[1])
    logical = int(np.round(float(setting[2])))
    compacted = cols > 0 and logical > cols
    slice_scale = np.float32(logical) / np.float32(cols) if compacted else np.float32(1.0)
    span = np.float32((cols - 1) * step_raw * slice_scale) if cols > 1 else np.float32(step_raw * slice_scale)
    unit = np.float32(1.0)
    if not compacted and 0.0 < step_raw < 1.0 and span < 20.0:
        unit = np.float32(10.0)
    elif not compacted and float(energies.max()) >= 250.0 and 20.0 < span <= 100.0:
        unit = np.float32(10.0)
    start = np.float32(start_raw * unit)
    step = np.float32(step_raw * slice_scale * unit)
    # running float sum per row, as the reference accumulates
    cidd = np.zeros_like(idd)
    acc = np.zeros(rows, np.float32)
    for j in range(cols):
        acc = (acc + idd[:, j] * step).astype(np.float32)
        cidd[:, j] = acc
    ref = load_energy_tables(tables_dir or TABLES_DIR)
    out = {
        "n_energies": rows, "n_energy_samples": cols, "energies_per_u": energies.astype(np.float32),
        "scale_facts": np.full(rows, (np.float32(1.0) / step) if step > 0 else np.float32(1.0), np.float32),
        "cidd_matrix": cidd,
    }
    for k in ("n_density_samples", "density_scale_fact", "density_vector", "n_sp_samples", "sp_scale_fact",
              "sp_vector", "n_rrl_samples", "rrl_scale_fact", "rrl_vector"):
        out[k] = ref[k]
    peaks = np.zeros(rows, np.float32)
    if step > 0:
        fin = np.where(np.isfinite(idd), idd, -np.inf)
        peaks = (start + step * np.argmax(fin, axis=1).astype(np.float32)).astype(np.float32)
    if not (peaks > 0).any():  # no usable IDD maximum: the reference tables' peak at each energy
        peaks = np.interp(energies, ref["energies_per_u"], ref["peak_depths"]).astype(np.float32)
    out["peak_depths"] = peaks
    return out


def derive_layer_sigmas_from_subspots(subspots: np.ndarray) -> np.ndarray:
    """[L][2]: the first subspot's (sigma_x, sigma_y) of every LUT layer (pybind:640-657)."""
    return np.ascontiguousarray(subspots[:, 0, 3:5], dtype=np.float32)


def derive_layer_longitudinal_cutoffs(per_spot, layer_counts) -> np.ndarray:
    """One cut-off per layer; spots of a layer must agree within 1e-3 (pybind:659-693)."""
    out, off = [], 0
    per_spot = np.asarray(per_spot, np.float32)
    for layer, c in enumerate(layer_counts):
        c = max(0, int(c))
        vals = per_spot[off:off + c]
        cut = float(vals[0]) if vals.size else 0.0
        bad = np.nonzero(np.abs(vals - np.float32(cut)) > 1e-3)[0]
        if bad.size:
            raise ContractError(f"longitudal_cutoff varies within layer {layer} by more than 1e-3; "
                                "one cut-off per layer is supported")
        out.append(cut)
        off += c
    return np.asarray(out, np.float32)


def _ordered_index(xs: np.ndarray, x: float) -> float:
    """Fractional index of x in a monotone axis, clamped (find_ordered_float_index, pybind:695-715)."""
    n = xs.size
    if n <= 1:
        return 0.0
    asc = xs[0] <= xs[-1]
    lo, hi = (xs[0], xs[-1]) if asc else (xs[-1], xs[0])
    if x <= lo:
        return 0.0 if asc else float(n - 1)
    if x >= hi:
        return float(n - 1) if asc else 0.0
    for i in range(n - 1):
        a, b = float(xs[i]), float(xs[i + 1])
        if x < min(a, b) or x > max(a, b):
            continue
        t = np.float32((x - a) / (b - a)) if abs(b - a) > 1e-6 else np.float32(0.0)
        return float(np.float32(i) + min(np.float32(1.0), max(np.float32(0.0), t)))
    return float(n - 1) if asc else 0.0


def remap_subspot_lut_to_actual_layers(lut_energies, lut_subspots: np.ndarray, layer_energies):
    """Subspot LUT and spot sigmas at the beam's layer energies: linear between the bracketing
    LUT layers (pybind:717-756). Returns (subspots [L][S][5], sigmas [L][2])."""
    lut_e = np.asarray(lut_energies, np.float32)
    n_lut, S = lut_subspots.shape[0], lut_subspots.shape[1]
    sig = derive_layer_sigmas_from_subspots(lut_subspots)
    L = len(layer_energies)
    out = np.zeros((L, S, 5), np.float32)
    out_sig = np.zeros((L, 2), np.float32)
    if n_lut <= 0 or S <= 0 or lut_e.size == 0:
        return out, out_sig
    for layer, e in enumerate(np.asarray(layer_energies, np.float32)):
        idx = _ordered_index(lut_e, float(e))
        i_int = float(np.floor(idx))
        frac = np.float32(idx - i_int)
        i0 = max(0, min(n_lut - 1, int(i_int)))
        i1 = max(0, min(n_lut - 1, i0 + (1 if frac > 0 else 0)))
        out[layer] = lut_subspots[i0] + (lut_subspots[i1] - lut_subspots[i0]) * frac
        out_sig[layer] = sig[min(i0, n_lut - 1)] + (sig[min(i1, n_lut - 1)] - sig[min(i0, n_lut - 1)]) * frac
    return out, out_sig


def to_roi_linear_indices(roi, dims) -> np.ndarray:
    """ROI voxels as x-major linear indices x ny nz + y nz + z (pybind:856-929)."""
    nx, ny, nz = (int(v) for v in dims)
    a = np.asarray(roi)
    if a.ndim == 1:
        v = a.astype(np.int64)
        if v.size and v.size % 3 == 0:
            t = v.reshape(-1, 3)
            if ((t[:, 0] >= 0) & (t[:, 0] < nx) & (t[:, 1] >= 0) & (t[:, 1] < ny) & (t[:, 2] >= 0)
                    & (t[:, 2] < nz)).all():
                return (t[:, 0] * ny * nz + t[:, 1] * nz + t[:, 2]).astype(np.int32)
        return v.astype(np.int32)
    if a.ndim != 2:
        raise ContractError(f"roi_index: got {a.ndim} dimensions; flat triplets, (N, 3) or (3, N) are accepted")
    t = a.astype(np.int64)
    if t.shape[1] == 3:
        return (t[:, 0] * ny * nz + t[:, 1] * nz + t[:, 2]).astype(np.int32)
    if t.shape[0] == 3:
        return (t[0] * ny * nz + t[1] * nz + t[2]).astype(np.int32)
    raise ContractError(f"roi_index: shape {a.shape} is neither (N, 3) nor (3, N)")


def build_carbonpbs_context(rayweq, roiIdx, all_energies, sourcePos, tmpBeamDir, bmxdir, bmydir, corner,
                            resolution, dims, longitudalCutoff, enelist, idddata, iddsetting, profiledata,
                            profilesetting, beamparadata, subspotdata, layerInfo, layerEnergy, idbeamxy, nPar,
                            sad, beamParaPos=None, tables_dir: str | None = None, carbon_profile_hooks: int = 0) -> dict:
    """Beam / energy dicts and grid of a CarbonPBS call (pybind:1121-1373). The profile model
    goes into the beam as the reference stores it (pybind:1269-1302: the depth step rescaled for a
    compacted table, the LUT energy axis sliced to the table's rows by RTD_PROFILE_SLICE_MODE,
    tail by default); ``carbon_profile_hooks`` (0, the reference's setting) switches its hooks."""
    d = [int(v) for v in np.asarray(dims).ravel()]
    if len(d) != 3:
        raise ContractError(f"dims: {len(d)} values given, 3 are needed")
    corner_v, res_v = to_float_vector(corner, "corner"), to_float_vector(resolution, "resolution")
    if corner_v.size != 3 or res_v.size != 3:
        raise ContractError(f"corner / resolution: {corner_v.size} / {res_v.size} values given, 3 each are needed")
    weq = _f32(rayweq).ravel()
    if weq.size < 9:
        raise ContractError(f"rayweq: {weq.size} values, shorter than its 9-value header")
    h = weq[:9]
    weights = to_float_vector(nPar, "nPar")
    n = weights.size
    spot_e = to_float_vector(all_energies, "all_energies")
    if spot_e.size and spot_e.size != n:
        raise ContractError(f"all_energies: {spot_e.size} values for {n} spot weights")
    counts = np.asarray(layerInfo).astype(np.int32).ravel()
    layer_e = to_float_vector(layerEnergy, "layerEnergy")
    if counts.size != layer_e.size:
        raise ContractError(f"layerEnergy has {layer_e.size} layers, layerInfo {counts.size}")
    dirs = to_vec_columns(tmpBeamDir, 3, n, "tmpBeamDir")
    sad_v = to_float_vector(sad, "sad") if np.ndim(sad) else np.asarray([sad], np.float32)
    if not sad_v.size:
        raise ContractError("sad: no value given")
    beam = {
        "water_equivalence": weq, "ray_spacing": np.array([h[7], h[4]], np.float32),
        "spot_delta": np.zeros(3, np.float32), "spot_offset": np.zeros(3, np.float32),
        "steps": max(1, int(np.round(float(h[2])))), "roi_linear_indices": to_roi_linear_indices(roiIdx, d),
        "energies": layer_e, "layer_spot_counts": counts, "spot_weights": weights,
        "spot_positions": to_vec_columns(idbeamxy, 2, n, "idbeamxy"), "spot_positions_are_indices": True,
        "spot_beam_directions": dirs, "beam_direction": normalized_mean_direction(dirs),
        "beam_xdir": to_vec_columns(bmxdir, 3, name="bmxdir")[:3], "beam_ydir": to_vec_columns(bmydir, 3, name="bmydir")[:3],
        "source_position": to_vec_columns(sourcePos, 3, name="sourcePos")[:3], "ref_plane_z": 0.0, "sad": float(sad_v[0]),
        "source_dist": np.zeros(2, np.float32),
    }
    if longitudalCutoff is not None:
        per_spot = to_float_vector(longitudalCutoff, "longitudal_cutoff")
        if per_spot.size:
            if per_spot.size != int(counts.sum()):
                raise ContractError(f"longitudal_cutoff: {per_spot.size} values for {int(counts.sum())} spots")
            beam["layer_longitudinal_cutoffs"] = derive_layer_longitudinal_cutoffs(per_spot, counts)
    sub = _f32(subspotdata)
    if sub.ndim != 3 or sub.shape[2] != 5:
        raise ContractError(f"subspotdata: shape {sub.shape}; (layers, subspots per layer, 5) is needed")
    lut_e = to_float_vector(enelist, "enelist")
    if sub.shape[0] == layer_e.size:
        beam["subspot_data"], beam["spot_sigmas"] = sub, derive_layer_sigmas_from_subspots(sub)
    else:
        beam["subspot_data"], beam["spot_sigmas"] = remap_subspot_lut_to_actual_layers(lut_e, sub, layer_e)
    prof_rows = prof_depth = 0
    if profiledata is not None:
        pa = _f32(profiledata)
        prof_rows = int(pa.shape[0]) if pa.ndim >= 1 else 0
        prof_depth = int(pa.shape[1]) if pa.ndim >= 2 else 0
        beam["profile_data"] = np.ascontiguousarray(pa).ravel()
    if profilesetting is not None:
        ps = to_float_vector(profilesetting, "profilesetting").copy()
        if ps.size >= 3 and prof_depth > 0:
            logical = max(0, _lround(float(ps[2])))
            if logical > prof_depth:
                ps[1] = np.float32(ps[1] * (np.float32(logical) / np.float32(prof_depth)))
            ps[2] = np.float32(prof_depth)
        beam["profile_setting"] = ps
    prof_e = lut_e
    if prof_rows > 0 and lut_e.size and prof_rows != lut_e.size:
        keep = min(lut_e.size, prof_rows)
        prof_e = lut_e[:keep] if _slice_mode("RTD_PROFILE_SLICE_MODE") == "head" else lut_e[lut_e.size - keep:]
    beam["profile_energies"] = prof_e
    if beamparadata is not None:
        beam["beam_para_data"] = np.ascontiguousarray(_f32(beamparadata)).ravel()
    if beamParaPos is not None:
        beam["beam_para_pos"] = float(beamParaPos)
    beam["carbon_profile_hooks"] = int(carbon_profile_hooks)
    energy = build_energy_from_carbonpbs(enelist, idddata, iddsetting, tables_dir)
    return {"dims": tuple(d), "corner": tuple(float(v) for v in corner_v),
            "resolution": tuple(float(v) for v in res_v), "beam": beam, "energy": energy,
            "profile": {"data": profiledata, "setting": profilesetting, "beam_para": beamparadata,
                        "beam_para_pos": beamParaPos}}


def debug_build_carbonpbs_context(*args, **kw) -> dict:
    """Summary of the context a CarbonPBS call builds (pybind:1375-1467)."""
    ctx = build_carbonpbs_context(*args, **kw)
    b, e = ctx["beam"], ctx["energy"]
    roi = b["roi_linear_indices"]
    out = {
        "dims": list(ctx["dims"]), "rayWeqHeader": [float(v) for v in b["water_equivalence"][:9]],
        "weq_len": int(b["water_equivalence"].size), "layer_energies": [float(v) for v in b["energies"]],
        "layerSpotCounts": [int(v) for v in b["layer_spot_counts"]], "totalSpots": int(b["layer_spot_counts"].sum()),
        "spotPositions": int(b["spot_positions"].size // 2), "spotWeights": int(b["spot_weights"].size),
        "spotDirections": int(b["spot_beam_directions"].size // 3), "spotPositionsAreIndices": True,
        "subspotDataSize": int(b["subspot_data"].size), "maxSubspotsPerLayer": int(b["subspot_data"].shape[1]),
        "roiLinearIndices": int(roi.size), "sourcePosition": [float(v) for v in b["source_position"]],
        "sad": float(b["sad"]), "energyData_nEnergies": int(e["n_energies"]),
        "energyData_nEnergySamples": int(e["n_energy_samples"]),
        "energyData_ciddMatrixSize": int(e["cidd_matrix"].size),
    }
    if roi.size:
        out["roiLinearMin"], out["roiLinearMax"] = int(roi.min()), int(roi.max())
    return out


def placeholder_ct(dims, roi_linear) -> np.ndarray:
    """The reference's stand-in CT: all 1000 without a ROI, else 1000 at the ROI's x-major linear
    indices of a buffer read z-major (so only the grid matters) and 0 elsewhere (pybind:1601-1624)."""
    total = int(np.prod(dims))
    if roi_linear.size == 0:
        return np.full(total, 1000.0, np.float32)
    ct = np.zeros(total, np.float32)
    r = roi_linear[(roi_linear >= 0) & (roi_linear < total)]
    ct[r] = 1000.0
    return ct


def cu_final_dose(finalDose, rayweq, roiIdx, all_energies, sourcePos, tmpBeamDir, bmxdir, bmydir, corner,
                  resolution, dims, longitudalCutoff, enelist, idddata, iddsetting, profiledata, profilesetting,
                  beamparadata, subspotdata, layerInfo, layerEnergy, idbeamxy, nPar, sad, cutoff=None,
                  beamParaPos=None, gpuId: int = 0, nuclear_correction: bool = False, verbose: int = 0,
                  debug: bool = False, tables_dir: str | None = None, carbon_profile_hooks: int = 0):
    """Dose of a CarbonPBS beam written into ``finalDose`` (pybind:1873-1953). ``cutoff`` is
    accepted for the contract and not used, as in the reference. ``carbon_profile_hooks`` (not in
    the reference's signature; 0 keeps its behaviour): bit 0 sigma from the profile model, bit 1
    the overall-weight channel."""
    ctx = build_carbonpbs_context(rayweq, roiIdx, all_energies, sourcePos, tmpBeamDir, bmxdir, bmydir, corner,
                                  resolution, dims, longitudalCutoff, enelist, idddata, iddsetting, profiledata,
                                  profilesetting, beamparadata, subspotdata, layerInfo, layerEnergy, idbeamxy, nPar,
                                  sad, beamParaPos, tables_dir, carbon_profile_hooks)
    nx, ny, nz = ctx["dims"]
    out = finalDose
    if not isinstance(out, np.ndarray) or out.dtype != np.float32 or out.size != nx * ny * nz:
        raise ContractError(f"finalDose: a float32 array of {nx * ny * nz} voxels is needed for dims {ctx['dims']}")
    xyz = False
    if out.ndim == 1:
        native = out
    elif out.ndim == 3 and out.flags.c_contiguous and out.shape == (nx, ny, nz):
        # checked first, as the reference does: a cubic grid's 3-D buffer is read as (Nx, Ny, Nz)
        native = np.zeros(nx * ny * nz, np.float32)
        xyz = True
    elif out.ndim == 3 and out.flags.c_contiguous and out.shape == (nz, ny, nx):
        native = out.reshape(-1)
    else:
        raise ContractError(f"finalDose: shape {out.shape}; flat, or C-contiguous ({nz}, {ny}, {nx}) or "
                            f"({nx}, {ny}, {nz}) is accepted")
    energy = _complete_nuclear(ctx["energy"], tables_dir) if nuclear_correction else ctx["energy"]
    ct = placeholder_ct(ctx["dims"], ctx["beam"]["roi_linear_indices"])
    subsecond_wrapper(ct, ctx["dims"], ctx["resolution"], ctx["corner"], native, ctx["dims"], ctx["resolution"],
                      ctx["corner"], [ctx["beam"]], energy, gpuId, nuclear_correction, verbose)
    if xyz:
        out[...] = native.reshape(nz, ny, nx).transpose(2, 1, 0)
    if (np.asarray(ctx["beam"]["spot_weights"]) > 0).any():
        _check_dose(native, ctx["beam"], "cuFinalDose/calcDose/cuCalDose")
    return finalDose


calcDose = cu_final_dose
cuCalDose = cu_final_dose
cuFinalDose = cu_final_dose


def cuCalDose3(out_cscdata, out_cscptr, out_cscrowind, weq_data, roi_index, source_energies, source, beam_dir,
               beam_xdir, beam_ydir, corner, resolution, dims, longitudal_cutoff, enelist, idddata, iddsetting,
               profiledata, profilesetting, beamparadata, subspot_data, layer_info, layer_energy, outNNZ, idbeamxy,
               num_particles_per_beam, cutoff, sad, beam_para_pos, python_nnz_size, gpu_id: int = 0,
               nuclear_correction: bool = False, verbose: int = 0, tables_dir: str | None = None,
               operator_mode: bool = False) -> dict:
    """The final dose in ``dose_grid`` (Nx, Ny, Nz) and the CSC outputs (pybind:1955-2050).

    By default the CSC outputs are zeroed, as the reference does (pybind:1995-1998). With
    ``operator_mode`` they hold the spot matrix (spot_matrix.spot_dose_matrix), in the CarbonPBS
    engine's layout (carbonPBS/cudaCalDose.cpp:930-1019): column j = dose per unit weight of spot j
    at the ROI voxels, row indices = positions in the ROI list (every voxel, x-major, when the ROI
    is empty), entries <= ``cutoff`` (0 when None) left out. Columns are written while they fit
    ``python_nnz_size`` (and the arrays); ``outNNZ[0]`` = entries written, ``status`` = -1 when a
    column did not fit, else 0."""
    for a in (out_cscdata, out_cscptr, out_cscrowind, outNNZ):
        a[...] = 0
    nx, ny, nz = (int(v) for v in np.asarray(dims).ravel())
    dose = np.zeros((nx, ny, nz), np.float32)
    cu_final_dose(dose, weq_data, roi_index, source_energies, source, beam_dir, beam_xdir, beam_ydir, corner,
                  resolution, dims, longitudal_cutoff, enelist, idddata, iddsetting, profiledata, profilesetting,
                  beamparadata, subspot_data, layer_info, layer_energy, idbeamxy, num_particles_per_beam, sad,
                  cutoff, beam_para_pos, gpu_id, nuclear_correction, verbose, False, tables_dir)
    out = {"dose_grid": dose, "out_cscdata": out_cscdata, "out_cscptr": out_cscptr,
           "out_cscrowind": out_cscrowind, "outNNZ": outNNZ, "status": 0}
    if not operator_mode:
        return out
    if nuclear_correction:
        raise RuntimeError("cuCalDose3 operator_mode: the halo branch has no resident (weight-only) form")
    from .engine import Engine
    from .spot_matrix import roi_from_xyz, spot_dose_matrix

    ctx = build_carbonpbs_context(weq_data, roi_index, source_energies, source, beam_dir, beam_xdir, beam_ydir,
                                  corner, resolution, dims, longitudal_cutoff, enelist, idddata, iddsetting,
                                  profiledata, profilesetting, beamparadata, subspot_data, layer_info, layer_energy,
                                  idbeamxy, num_particles_per_beam, sad, beam_para_pos, tables_dir)
    # ROI rows in the caller's order: CarbonPBS linear indices (x ny nz + y nz + z) -> engine order
    lin = to_roi_linear_indices(roi_index, (nx, ny, nz)) if np.asarray(roi_index).size else \
        np.arange(nx * ny * nz, dtype=np.int64)
    lin = np.asarray(lin, np.int64)
    roi = roi_from_xyz(np.stack([lin // (ny * nz), (lin // nz) % ny, lin % nz], axis=1), (nx, ny, nz))
    eng = Engine(ctx["energy"], device=gpu_id)
    try:
        eng.set_ct(placeholder_ct(ctx["dims"], ctx["beam"]["roi_linear_indices"]), ctx["dims"], ctx["resolution"],
                   ctx["corner"])
        eng.set_dose_grid(ctx["dims"], ctx["resolution"], ctx["corner"])
        data, rows, indptr, _ = spot_dose_matrix(eng, ctx["beam"], roi, 0.0 if cutoff is None else float(cutoff))
    finally:
        eng.close()
    cap = min(int(python_nnz_size) if python_nnz_size is not None else data.size, out_cscdata.size,
              out_cscrowind.size)
    S = indptr.size - 1
    if out_cscptr.size < S + 1:
        raise ContractError(f"out_cscptr: {out_cscptr.size} entries for {S} spots ({S + 1} needed)")
    done = S
    while done > 0 and indptr[done] > cap:
        done -= 1
    n = int(indptr[done])
    out_cscdata.reshape(-1)[:n] = data[:n]
    out_cscrowind.reshape(-1)[:n] = rows[:n]
    out_cscptr.reshape(-1)[:done + 1] = indptr[:done + 1]
    outNNZ.reshape(-1)[0] = n
    out["status"] = 0 if done == S else -1
    return out
