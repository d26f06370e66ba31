"""The host planner (the part of pbs_engine_beam_dose that runs before any kernel) against the
FP64 oracle's plan, on the CPU through pbs_plan_debug.

For every beam of the fixtures (the small water beam, the head beam, the 16-beam proton/carbon
mix, the serving mix on three CT sizes, the WEQ beam with physical and index-coded spots):
* grids and step counts equal the oracle's (rays, 32x8 padding, spot lattice, steps, legacy
  flag), and the fan corner, steps and source distances agree to float rounding;
* per-layer energy index and peak depth agree with the oracle's;
* each spot's lattice cell is (position - origin) / step rounded half away from zero (std::lround,
  restated here in float64; a wrong origin or step, or a swapped axis, moves cells), and the
  dense lattice is the spots' weights added into their cells in spot order. The origin is the
  smallest position and the step the smallest gap, so accepted spots sit at or just above their
  lattice points: the rounding direction itself is not exercised.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import _abi
from subsecondpbs_amd import fixtures as F


def _cases():
    out = [("config1", F.make_config1()), ("config2", F.make_config2())]
    c3 = F.make_config3()
    for i, b in enumerate(c3["beams"]):
        out.append((f"config3-{i}", F.case_for_beam(c3, i)))
    c5 = F.make_config5()
    for ri in (0, 3, 4, 7, 11):
        r = c5["requests"][ri]
        c = dict(c5["cts"][r["ct"]])
        c["beams"] = [r["beam"]]
        c["energy"] = c5["energy"]
        out.append((f"serving-{ri}", c))
    out.append(("weq", F.make_weq_case()))
    out.append(("weq-index", F.make_weq_case(index_spots=True)))
    return out


CASES = _cases()


def _lround(v):
    """std::lround in float64 (exact for float32 inputs)."""
    v = np.asarray(v, np.float64)
    return np.trunc(v + np.copysign(0.5, v)).astype(np.int64)


@pytest.mark.parametrize("name,case", CASES, ids=[n for n, _ in CASES])
def test_host_plan_matches_the_oracle(name, case):
    beam = case["beams"][0]
    got = _abi.plan_debug(case, beam)
    p = O.plan_beam(case)
    assert p.ok == 1, p.error
    # per-layer parameters (energy index, peak depth) come from the energies alone; the minimum
    # WEPL per step only sets the layers' last steps, which this test does not read
    keep = O._Keep()
    O.lib().orc_layer_params(C.byref(p), C.byref(O.make_beam(beam, keep)), C.byref(O.make_energy(case["energy"], keep)),
                             O._dp(np.zeros(max(p.steps, 1))))
    ref = lambda k: np.atleast_1d(np.asarray(getattr(p, k)))
    for k in ("rayDims", "superpDims", "steps", "spotDims", "legacyPath", "weqVolume"):
        assert tuple(np.atleast_1d(got[k])) == tuple(ref(k)), k
    for k, n in (("spotOffset", 2), ("spotDelta", 2), ("cpbCorner", 2), ("fanCorner", 3), ("fanDelta", 3),
                 ("sourceDist", 2)):
        np.testing.assert_allclose(np.asarray(got[k])[:n], ref(k)[:n], rtol=1e-6, atol=1e-5, err_msg=k)
    L = got["numLayers"]
    np.testing.assert_allclose(got["energyIdx"], ref("energyIdx")[:L], rtol=0, atol=1e-5)
    np.testing.assert_allclose(got["peakDepth"], ref("peakDepth")[:L], rtol=1e-6)
    if got["legacyPath"] or got["spotIndices"]:
        return
    # lattice cells and the dense lattice restated from the spot list
    nx, ny = got["spotDims"]
    ox, oy = np.float32(got["spotOffset"][0]), np.float32(got["spotOffset"][1])
    dx, dy = np.float32(got["spotDelta"][0]), np.float32(got["spotDelta"][1])
    pos = np.asarray(beam["spot_positions"], np.float32).reshape(-1, 2)
    w = np.asarray(beam["spot_weights"], np.float32)
    layer = np.repeat(np.arange(L), np.asarray(beam["layer_spot_counts"]))
    ix = _lround((pos[:, 0] - ox) / dx)
    iy = _lround((pos[:, 1] - oy) / dy)
    inside = (ix >= 0) & (ix < nx) & (iy >= 0) & (iy < ny)
    want = np.where(inside, (layer * ny + iy) * nx + ix, -1)
    np.testing.assert_array_equal(got["spotCell"], want)
    dense = np.zeros(L * ny * nx, np.float32)
    live = (w > 0) & inside
    np.add.at(dense, want[live], w[live])
    np.testing.assert_array_equal(got["dense"].ravel(), dense)
