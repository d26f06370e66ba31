"""Small host-side helpers shared by tests and the benchmark."""
from __future__ import annotations

import numpy as np

MAX_RAD = 32
MIN_TILES_IN_BATCH = 16


def batch_bins(hist34) -> np.ndarray:
    """Launch radius of each radius bin (complete_superposition.cu:413-490): bin 33 is
    clamped into 32, then bins merge downwards until a batch holds >= 16 tiles and every
    tile of the batch is superposed with the batch's largest radius."""
    h = np.array(hist34, dtype=np.int64).copy()
    h[MAX_RAD] += h[MAX_RAD + 1]
    h[MAX_RAD + 1] = 0
    nz = np.nonzero(h)[0]
    layer_max = int(nz.max()) if nz.size else 0
    eff = np.arange(MAX_RAD + 2)
    eff[MAX_RAD + 1] = MAX_RAD
    batched, rec = 0, layer_max
    for r in range(layer_max, 0, -1):
        batched += int(h[r])
        eff[r] = rec
        if batched >= MIN_TILES_IN_BATCH:
            rec = r - 1
            batched = 0
    return eff
