"""WEQ-volume input and WEQ-index spots in the FP64 oracle (the CarbonPBS contract:
RTDBeamSettings::waterEquivalence / spotPositionsAreIndices, wrapper:931-968, 1295-1364,
1923-1988, 3305-3370).

Pins:
* the struct mirrors have the compiled sizes;
* the trace stage of a WEQ volume: on the lateral wedge of the fixture (density 1 + 0.01 ix,
  so a transposed [x][y] index or a wrong column rounding shows) the oracle's density per ray is
  the fixture's closed form at the column nearest the ray, its WEPL telescopes back from the
  densities, and rays off the lattice never enter (entry K, exit 0);
* the start plane: moved to (mean source distance - depth_start) within max(depth span, 50 mm)
  of the CT's, left at the CT's otherwise or for depth_start <= 0;
* end to end: a water WEQ volume on a water CT gives the CT path's dose within 1 % of Dmax (the
  WEQ path feeds the same chain; only the entry rule and the off-axis step length differ);
* index-coded spots: texel-centre spots give the physical-position dose (to the tiling
  difference of the two ray grids), and a spot a quarter texel off is the bilinear mix of its
  four neighbours' ray weights (the lattice convolution is linear in the spot plane, which
  fixes the split exactly).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import fixtures as F


def test_struct_mirrors_match_the_oracle():
    L = O.lib()
    assert L.orc_sizeof(0) == C.sizeof(O.Beam)
    assert L.orc_sizeof(1) == C.sizeof(O.Plan)
    assert L.orc_sizeof(2) == C.sizeof(O.Energy)


def _trace(case):
    keep = O._Keep()
    en = O.make_energy(case["energy"], keep)
    p = O.plan_beam(case)
    rx, ry, K = p.rayDims[0], p.rayDims[1], p.steps
    dens = np.zeros((K, ry, rx)); wepl = np.zeros((K, ry, rx))
    fi = np.zeros((ry, rx), np.int32); fo = np.zeros((ry, rx), np.int32)
    ct = np.ascontiguousarray(O.ct_as_hu_plus_1000(case["ct"]), np.float32)
    dims = np.array(case["ct_dims"], np.int32)
    O.lib().orc_trace(C.byref(p), O._fp(ct), O._ip(dims), C.byref(en), O._dp(dens), O._dp(wepl), O._ip(fi),
                      O._ip(fo))
    return p, dens, wepl, fi, fo


def test_weq_trace_follows_the_volume():
    case = F.make_weq_case()
    h = case["weq_header"]
    p, dens, wepl, fi, fo = _trace(case)
    K, ry, rx = dens.shape
    assert p.weqVolume == 1 and K == 64
    xs = p.cpbCorner[0] + np.arange(rx) * p.cpbRes[0]
    ys = p.cpbCorner[1] + np.arange(ry) * p.cpbRes[1]
    qx = np.floor((xs - h[6]) / h[7] + 0.5).astype(int)   # half away from zero for the positive half
    qy = np.floor((ys - h[3]) / h[4] + 0.5).astype(int)
    inx = (qx >= 0) & (qx < 81)
    iny = (qy >= 0) & (qy < 71)
    k = np.arange(K)
    for j in range(ry):
        for i in range(rx):
            if inx[i] and iny[j]:
                want = 1.0 + 0.01 * qx[i] + 0.5 * ((k >= 20) & (k < 30) & (qy[j] >= 35))
                np.testing.assert_allclose(dens[:, j, i], want, rtol=1e-5)
                np.testing.assert_allclose(np.cumsum(dens[:, j, i]) * 2.0, wepl[:, j, i], rtol=1e-5)
                assert fi[j, i] == 1 and fo[j, i] == K
            else:
                assert not dens[:, j, i].any() and fi[j, i] == K and fo[j, i] == 0


@pytest.mark.parametrize("shift,moved", [(4.0, True), (80.0, True), (200.0, False)])
def test_start_plane_alignment(shift, moved):
    p = O.plan_beam(F.make_weq_case(start_shift_mm=shift))
    assert p.fanCorner[2] == pytest.approx(63.0 + shift if moved else 63.0)
    c = F.make_weq_case()
    c["beams"][0]["water_equivalence"] = c["beams"][0]["water_equivalence"].copy()
    c["beams"][0]["water_equivalence"][0] = -1.0
    assert O.plan_beam(c).fanCorner[2] == pytest.approx(63.0)


def test_water_weq_matches_the_ct_path():
    case = F.make_config1()
    case["ct"] = np.full_like(case["ct"], 1000.0)
    ct_dose, _ = O.beam_dose(case)
    weq = dict(case)
    b = dict(case["beams"][0])
    header = [2000.0 - 63.0, 2.0, 64.0, -60.0, 1.0, 121.0, -60.0, 1.0, 121.0]
    b["water_equivalence"] = F.weq_volume(header, lambda ix, iy, k: np.ones(ix.shape))
    weq["beams"] = [b]
    weq_dose, p = O.beam_dose(weq)
    assert p.weqVolume == 1 and p.fanCorner[2] == pytest.approx(63.0)
    err = np.abs(weq_dose - ct_dose).max() / ct_dose.max()
    print(f"water WEQ vs CT path: max|dD|/Dmax = {err:.2e}")
    assert err < 1e-2


def test_index_spots_at_texel_centres_equal_physical_spots():
    phys = F.make_weq_case()
    idx = F.make_weq_case(index_spots=True)
    b = idx["beams"][0]
    raw = np.asarray(b["spot_positions"], np.float32).reshape(-1, 2).copy()
    per = len(raw) // len(b["energies"])
    raw[::per] -= 0.25  # back onto texel centres
    b["spot_positions"] = raw.ravel()
    d_phys, _ = O.beam_dose(phys)
    d_idx, p = O.beam_dose(idx)
    assert tuple(p.spotDims) == (81, 71)
    # same spot weights on the same ray positions; the ray grids differ in extent (the index
    # lattice spans the whole WEQ grid), which moves the 32x8 tiles and so the per-tile radii:
    # 1e-3 of Dmax (a half-texel decoding slip shifts the field by 0.5 mm, ~5e-2 of Dmax)
    err = np.abs(d_idx - d_phys).max() / d_phys.max()
    print(f"texel-centre index spots vs physical: {err:.2e}")
    assert err <= 2e-3


def test_index_spot_split_is_bilinear():
    base = F.make_weq_case(index_spots=True)
    b0 = base["beams"][0]

    def ray_weights(x, y):
        # the lattice -> ray convolution is linear in the dense spot plane (no thresholds); the
        # index lattice is the whole WEQ grid, so every spot position gives the same ray grid
        c = dict(base)
        b = dict(b0)
        L = len(b["energies"])
        b["layer_spot_counts"] = np.ones(L, np.int32)
        b["spot_positions"] = np.tile(np.array([x, y], np.float32), L)
        b["spot_weights"] = np.ones(L, np.float32)
        c["beams"] = [b]
        keep = O._Keep()
        en = O.make_energy(c["energy"], keep)
        bm = O.make_beam(b, keep)
        p = O.plan_beam(c)
        rw = np.zeros((L, p.rayDims[1], p.rayDims[0]))
        O.lib().orc_conv_lattice(C.byref(p), C.byref(bm), C.byref(en), O._dp(rw))
        return rw

    x0, y0 = 40.5, 35.5  # texel centre of lattice cell (40, 35)
    mix = ray_weights(x0 + 0.25, y0 + 0.75)
    corners = [ray_weights(x0, y0), ray_weights(x0 + 1, y0), ray_weights(x0, y0 + 1), ray_weights(x0 + 1, y0 + 1)]
    wts = [0.75 * 0.25, 0.25 * 0.25, 0.75 * 0.75, 0.25 * 0.75]
    want = sum(w * d for w, d in zip(wts, corners))
    assert mix.max() > 0.0
    assert np.abs(mix - want).max() <= 1e-6 * want.max()
    # a plausible slip (no -1/2 texel offset) puts the mass one texel over
    wrong = ray_weights(x0 + 0.75, y0 + 1.25)
    assert np.abs(wrong - want).max() > 1e-2 * want.max()


def _texel_centre_index_case(**kw):
    case = F.make_weq_case(index_spots=True, **kw)
    b = case["beams"][0]
    raw = np.asarray(b["spot_positions"], np.float32).reshape(-1, 2).copy()
    raw[::len(raw) // len(b["energies"])] -= 0.25  # every spot on a texel centre
    b["spot_positions"] = raw.ravel()
    return case


def _halo_plan(case):
    keep = O._Keep()
    bm = O.make_beam(case["beams"][0], keep)
    p = O.plan_beam(case)
    rx, ry = p.rayDims[0], p.rayDims[1]
    L = bm.numLayers
    nx = (max(p.spotDims[0], 1) + 31) // 32 * 32
    ny = (max(p.spotDims[1], 1) + 7) // 8 * 8
    ray_to_nuc = np.zeros(rx * ry, np.int32)
    w = np.zeros(L * nx * ny)
    rc = O.lib().orc_halo_plan(C.byref(p), C.byref(bm), O._ip(ray_to_nuc), O._dp(w))
    return rc, p, ray_to_nuc, w


def test_halo_lattice_view_of_index_spots():
    """Index-coded spots at texel centres decode to the physical beam's 5 mm lattice
    (buildPhysicalPBLatticeView, wrapper:1478-1820): the halo cells, their weights, the spot
    spacing and the number of rays that centre a spot equal the physical-position beam's."""
    rc_p, pp, r2n_p, w_p = _halo_plan(F.make_weq_case())
    rc_i, pi_, r2n_i, w_i = _halo_plan(_texel_centre_index_case())
    assert rc_p == 0 and rc_i == 0, pi_.error
    assert tuple(pi_.nucRaw) == tuple(pp.spotDims) == (11, 11) and pi_.nucFallback == 0
    np.testing.assert_allclose(pi_.nucDelta, pp.spotDelta[:2], rtol=1e-6)
    np.testing.assert_allclose(pi_.nucOffset, pp.spotOffset[:2], atol=1e-4)
    assert tuple(pi_.nucDims) == tuple(pp.nucDims) and pi_.spotDist == pytest.approx(pp.spotDist)
    n = pp.nucDims[0] * pp.nucDims[1] * 5
    np.testing.assert_array_equal(w_i[:n], w_p[:n])
    assert (r2n_i >= 0).sum() == (r2n_p >= 0).sum() == 121


def test_halo_lattice_view_fallback_and_gate():
    """A spot a quarter texel off its centre breaks the canonical axes (a 0.25 mm gap is not a
    whole number of 5 mm steps) on all spots and on the densest layer, so the WEQ header's 1 mm
    lattice stands in: its spacing fails the spotDistInRays >= 1.5 gate unless the beam states its
    spot spacing, and then the halo grid is the 81 x 71 texel lattice."""
    case = F.make_weq_case(index_spots=True)
    rc, p, _, _ = _halo_plan(case)
    assert rc != 0 and b"spotDistInRays" in p.error
    case["beams"][0]["spot_delta"] = np.array([5.0, 5.0, 0.0], np.float32)
    rc, p, r2n, w = _halo_plan(case)
    assert rc == 0 and p.nucFallback == 1 and tuple(p.nucRaw) == (81, 71)
    assert abs(w.sum() - float(np.asarray(case["beams"][0]["spot_weights"]).sum())) < 1e-3


def test_halo_lattice_view_densest_layer():
    """Odd layers shifted by 2 mm (two texels): over all spots the gaps alternate 2 and 3 mm (median
    2, a 3 mm gap is 1.5 steps), so no single step fits, but each layer alone is the 5 mm lattice.
    (A 1 mm shift would not do: gaps of 1 and 4 mm are whole multiples of their 1 mm median.) The view then takes the step of
    the first layer with the most spots and the origin and extent of all spots
    (wrapper:1647-1707): 5 mm cells from the smallest decoded x, one column more to reach the
    shifted layers, and every spot binned to its nearest cell."""
    case = _texel_centre_index_case()
    b = case["beams"][0]
    raw = np.asarray(b["spot_positions"], np.float32).reshape(-1, 2).copy()
    counts = np.asarray(b["layer_spot_counts"])
    layer = np.repeat(np.arange(counts.size), counts)
    raw[layer % 2 == 1, 0] += 2.0
    b["spot_positions"] = raw.ravel()
    rc, p, r2n, w = _halo_plan(case)
    assert rc == 0, p.error
    assert p.nucFallback == 0 and p.nucDelta[0] == pytest.approx(5.0) and p.nucDelta[1] == pytest.approx(5.0)
    h = case["weq_header"]
    xs = h[6] + (raw[:, 0] - 0.5) * h[7]
    assert p.nucOffset[0] == pytest.approx(float(xs.min()), abs=1e-5)
    assert p.nucRaw[0] == int(round((xs.max() - xs.min()) / 5.0)) + 1 == 11 and p.nucRaw[1] == 11
    assert abs(w.sum() - float(np.asarray(b["spot_weights"]).sum())) < 1e-3
