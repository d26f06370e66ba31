"""Spot-preserving operator mode: the dose of every spot at unit weight, as a CSC matrix.

The reference plans this export (include/core/rtd_operator_mode.h:43-112, spot batches over the
beam's layers) and its ``cuCalDose3`` returns the CSC arrays zeroed
(raytracedicom_pybind.cpp:1995-1998). The layout follows the CarbonPBS engine that the contract
comes from (carbonPBS/cudaCalDose.cpp:930-1019, deviceCalDose.cu:76-88): one column per spot in
spot order, row indices are positions in the ROI list, values are float32, and the nnz count is
returned beside the arrays.

Column j is the beam's dose with spot j at weight 1 and every other spot at 0, gathered at the
ROI voxels. It is computed on a resident beam (``Engine.keep_beam``): the trace and the IDD /
sigma fill run once, and each column is a weight-only update (convolution, superposition,
merge, transfer) into one device buffer, gathered on the engine's stream. Index-coded spots
(WEQ lattice coordinates) spread over four lattice cells, which the resident re-binning does not
hold; for them every column runs the whole chain from one prepared struct whose weights are
rewritten in place.
"""
from __future__ import annotations

import numpy as np

from .engine import Engine


def roi_from_xyz(xyz, dose_dims) -> np.ndarray:
    """Engine-order linear indices (z ny nx + y nx + x) of ROI voxels given as (x, y, z) rows."""
    a = np.asarray(xyz, dtype=np.int64).reshape(-1, 3)
    nx, ny, nz = (int(v) for v in dose_dims)
    return (a[:, 2] * ny + a[:, 1]) * nx + a[:, 0]


def spot_dose_matrix(engine: Engine, beam: dict, roi=None, cutoff: float = 0.0, chunk: int = 256):
    """(data, indices, indptr, roi) of the CSC matrix D [len(roi) x numSpots]: D[i, j] = dose at
    ROI voxel i per unit weight of spot j, entries <= ``cutoff`` left out. ``roi``: engine-order
    linear voxel indices (default: every voxel of the engine's dose grid)."""
    import torch

    if engine.dose_dims is None:
        raise RuntimeError("spot_dose_matrix: the engine has no dose grid")
    nx, ny, nz = engine.dose_dims
    nvox = nx * ny * nz
    roi = np.arange(nvox, dtype=np.int64) if roi is None else np.asarray(roi, dtype=np.int64).ravel()
    if roi.size and (roi.min() < 0 or roi.max() >= nvox):
        raise ValueError("spot_dose_matrix: ROI index outside the dose grid")
    w = np.asarray(beam["spot_weights"], np.float32).ravel()
    S = int(w.size)
    buf = torch.zeros(nvox, dtype=torch.float32, device="cuda")
    dose3 = buf.view(nz, ny, nx)
    stream = torch.cuda.ExternalStream(engine.stream_ptr)
    torch.cuda.current_stream().synchronize()
    unit = np.zeros(S, np.float32)
    rb = None
    if beam.get("spot_positions_are_indices", False):
        b = dict(beam)
        b["spot_weights"] = unit  # the prepared struct points at this array
        prepared = engine.prepare(b)

        def column():
            engine.beam_dose(prepared, dose3)
    else:
        # the resident state (trace, fill) at the beam's own weights: the dose it adds is discarded
        rb = engine.keep_beam(beam, dose3)

        def column():
            rb.update(unit, dose3)
    roi_t = torch.as_tensor(roi, device="cuda")
    data, rows, counts = [], [], []
    try:
        for c0 in range(0, S, chunk):
            c1 = min(S, c0 + chunk)
            block = torch.empty((c1 - c0, roi.size), dtype=torch.float32, device="cuda")
            for j in range(c0, c1):
                with torch.cuda.stream(stream):
                    buf.zero_()
                # the previous column's uploads have left the struct's arrays (the engine waits
                # for them before staging the next beam)
                unit[:] = 0.0
                unit[j] = 1.0
                column()
                with torch.cuda.stream(stream):
                    block[j - c0] = buf[roi_t]
            stream.synchronize()
            keep = block > cutoff
            counts.append(keep.sum(dim=1).cpu().numpy())
            nzr = torch.nonzero(keep)  # (spot, roi position), spot-major and rows ascending
            rows.append(nzr[:, 1].to(torch.int32).cpu().numpy())
            data.append(block[keep].cpu().numpy())
    finally:
        if rb is not None:
            rb.close()
    counts = np.concatenate(counts) if counts else np.zeros(0, np.int64)
    indptr = np.zeros(S + 1, np.int64)
    np.cumsum(counts, out=indptr[1:])
    data = np.concatenate(data) if data else np.zeros(0, np.float32)
    rows = np.concatenate(rows) if rows else np.zeros(0, np.int32)
    return data.astype(np.float32), rows, indptr, roi
