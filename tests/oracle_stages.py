"""Runs the FP64 oracle stage by stage and keeps every intermediate array (test helper)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle as O


def run(case: dict, beam_index: int = 0, use_subspots: bool = False, ct_hu_plus_1000=None):
    keep = O._Keep()
    en = O.make_energy(case["energy"], keep)
    bm = O.make_beam(case["beams"][beam_index], keep, use_subspots)
    ct_dims, ct_res, ct_corner, dose_dims, dose_res, dose_corner = O.case_args(case)
    L = O.lib()
    p = O.Plan()
    rc = L.orc_plan_beam(O._ip(ct_dims), O._fp(ct_res), O._fp(ct_corner), O._ip(dose_dims), O._fp(dose_res),
                         O._fp(dose_corner), C.byref(bm), C.byref(en), C.byref(p))
    assert rc == 0, p.error
    rx, ry, K = p.rayDims[0], p.rayDims[1], p.steps
    nl = bm.numLayers
    ct = ct_hu_plus_1000 if ct_hu_plus_1000 is not None else O.ct_as_hu_plus_1000(case["ct"])
    ct = np.ascontiguousarray(ct, np.float32)
    dens = np.zeros((K, ry, rx))
    wepl = np.zeros((K, ry, rx))
    fi = np.zeros((ry, rx), np.int32)
    fo = np.zeros((ry, rx), np.int32)
    L.orc_trace(C.byref(p), O._fp(ct), O._ip(ct_dims), C.byref(en), O._dp(dens), O._dp(wepl), O._ip(fi), O._ip(fo))
    wmin = np.zeros(K)
    L.orc_beam_cutoffs(C.byref(p), C.byref(bm), C.byref(en), O._dp(wepl), O._ip(fi), O._ip(fo), O._dp(wmin))
    rw = np.zeros((nl, ry, rx))
    if not p.legacyPath:
        L.orc_conv_lattice(C.byref(p), C.byref(bm), C.byref(en), O._dp(rw))
    L.orc_layer_params(C.byref(p), C.byref(bm), C.byref(en), O._dp(wmin))
    idd = np.zeros((nl, K, ry, rx))
    rs = np.full((nl, K, ry, rx), np.inf)
    fp = np.zeros((nl, ry, rx), np.int32)
    for l in range(nl):
        L.orc_fill(C.byref(p), l, C.byref(en), O._dp(dens), O._dp(wepl), O._dp(rw[l]), O._ip(fi), O._ip(fo),
                   O._dp(idd[l]), O._dp(rs[l]), O._ip(fp[l]))
    dose, pfull = O.beam_dose(case, beam_index, use_subspots=use_subspots, ct_hu_plus_1000=ct)
    return {
        "plan": O.plan_to_dict(pfull), "density": dens, "wepl": wepl, "fi": fi, "fo": fo, "weplMin": wmin,
        "rayW": rw, "idd": idd, "rsigma": rs, "firstPassive": fp, "dose": dose,
    }


# A float trace samples the CT at positions rounded to ~1e-4 voxel; HU+1000 changes by up to
# ~3000 across a voxel, so a sample, and the cumulative sum at the entry step, can move by a few
# tenths. A ray whose cumulative HU+1000 passes within this distance of the 150 threshold may
# enter one step earlier or later than in the FP64 trace.
ENTRY_MARGIN = 1.0


def dose_following_entry(case: dict, fi_gpu: np.ndarray, beam_index: int = 0):
    """FP64 dose with the float trace's entry step on the rays where the two traces disagree.

    Asserts that every such ray sits on the threshold (oracle margin < ENTRY_MARGIN) and moves by
    one step at most; returns (dose, number of rays followed)."""
    p = O.plan_beam(case, beam_index)
    ct_dims = O.case_args(case)[0]
    ct = np.ascontiguousarray(O.ct_as_hu_plus_1000(case["ct"]), np.float32)
    keep = O._Keep()
    en = O.make_energy(case["energy"], keep)
    K, ry, rx = p.steps, p.rayDims[1], p.rayDims[0]
    dens = np.zeros((K, ry, rx))
    wepl = np.zeros((K, ry, rx))
    fi = np.zeros((ry, rx), np.int32)
    fo = np.zeros((ry, rx), np.int32)
    O.lib().orc_trace(C.byref(p), O._fp(ct), O._ip(ct_dims), C.byref(en), O._dp(dens), O._dp(wepl), O._ip(fi),
                      O._ip(fo))
    fi_gpu = np.asarray(fi_gpu, np.int32)[:ry, :rx]
    diff = fi_gpu != fi
    if diff.any():
        m = O.entry_margin(case, beam_index)
        assert (m[diff] < ENTRY_MARGIN).all(), f"entry steps differ on rays off the threshold: {m[diff]}"
        assert (np.abs(fi_gpu[diff] - fi[diff]) <= 1).all()
    ov = np.where(diff, fi_gpu, -1).astype(np.int32)
    dose, _ = O.beam_dose(case, beam_index, entry_override=ov)
    return dose, int(diff.sum())
