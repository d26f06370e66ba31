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
