"""Pins of the oracle's superposition and BEV -> dose transfer beyond the one-pixel and
identity-geometry cases (CPU only).

* Superposition over several tiles whose radii are batched upward
  (complete_superposition.cu:197-255 radius per tile, 465-490 batching, 264-335 footprint):
  the expected BEV is built from Gaussian cell integrals (scipy's normal CDF, not erf) with
  hand-derived launch radii, so a slip in the radius rule, the batching or the footprint fails.
* Transfer with a finite source distance (transfer_param_struct_div3_impl.cu:14-34,
  transfer_param_struct_div3.cuh:35-45, prim_transf_kernel.cu:11-44): a BEV that is linear in
  its index reproduces exactly under trilinear interpolation, so every interior voxel must hold
  the closed form at its divergent fan index x d / (d - z).
* Transfer launch slack (raytracedicom_wrapper.cu:5343-5404): whole 32-wide thread blocks run
  from the box start, so columns past the box up to the block edge are sampled and the next
  column is not.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.special import erf, ndtr

import oracle as O

SQ2 = math.sqrt(2.0)


def _radius_f32(r):
    """min(int(3 / (sqrt(2) r) + 0.5), 33) in float32 (complete_superposition.cu:242)."""
    q = np.float32(3.0) / (np.float32(SQ2) * np.float32(r)) + np.float32(0.5)
    return min(int(q), 33)


def _cell_masses(r, R):
    """Mass of a unit Gaussian of sigma = 1 / (sqrt(2) r) over the cells [i - 1/2, i + 1/2],
    i = 0..R, from the normal CDF."""
    s = 1.0 / (SQ2 * r)
    i = np.arange(R + 1, dtype=np.float64)
    return ndtr((i + 0.5) / s) - ndtr((i - 0.5) / s)


# Tile minima of rSigmaEff and the radius each gives on its own (checked below in float32).
OWN_RADIUS = {0.1: 21, 0.15: 14, 0.2: 11, 0.3: 7, 1.0: 2}


def _multi_tile_field(seed=11):
    """Steps 1..5 of a 64 x 16 grid (2 x 2 tiles per step). Tile minima: 2 tiles at r = 0.1,
    3 at 0.15, 5 at 0.2, 6 at 0.3, 4 at 1.0; one further step holds only dead tiles."""
    rng = np.random.default_rng(seed)
    K, sy, sx = 8, 16, 64
    idd = np.zeros((K, sy, sx))
    rs = np.full((K, sy, sx), np.inf)
    minima = [0.1, 0.1, 0.15, 0.15, 0.15, 0.2, 0.2, 0.2, 0.2, 0.2,
              0.3, 0.3, 0.3, 0.3, 0.3, 0.3, 1.0, 1.0, 1.0, 1.0]
    rng.shuffle(minima)
    tiles = {}
    n = 0
    for k in range(1, 6):
        for ty in range(2):
            for tx in range(2):
                m = minima[n]
                n += 1
                tiles[(k, ty, tx)] = m
                ys, xs = slice(ty * 8, ty * 8 + 8), slice(tx * 32, tx * 32 + 32)
                live = rng.random((8, 32)) < 0.15
                live[rng.integers(8), rng.integers(32)] = True
                r = m * (1.0 + rng.random((8, 32)))
                yy, xx = np.nonzero(live)
                j = rng.integers(yy.size)
                r[yy[j], xx[j]] = m  # the tile minimum sits on a live pixel
                blk_r = np.where(live, r, np.inf)
                blk_d = np.where(live, 0.2 + rng.random((8, 32)), 0.0)
                rs[k, ys, xs] = blk_r
                idd[k, ys, xs] = blk_d
    return idd, rs, tiles


def _superpose_plan(sx, sy, k0, k1, bevz):
    p = O.Plan()
    p.superpDims[0], p.superpDims[1] = sx, sy
    p.beamFirstInside = k0
    p.layerPassive[0] = k1
    p.guaranteedPassive = bevz
    return p


def test_superpose_multi_tile_batched_radii():
    for m, R in OWN_RADIUS.items():
        assert _radius_f32(m) == R
    idd, rs, tiles = _multi_tile_field()
    K, sy, sx = idd.shape
    k0, k1 = 1, 7  # step 6 is all dead: its four tiles fall in bin 0
    p = _superpose_plan(sx, sy, k0, k1, K)
    bev = np.zeros((K, sy + 64, sx + 64))
    O.lib().orc_superpose(C.byref(p), 0, O._dp(idd), O._dp(rs), O._dp(bev))

    # Histogram: bins 21:2, 14:3, 11:5, 7:6, 2:4, 0:4 (the dead step).
    hist = np.array(p.radHist[0][:34])
    want_hist = np.zeros(34, int)
    want_hist[[21, 14, 11, 7, 2, 0]] = [2, 3, 5, 6, 4, 4]
    np.testing.assert_array_equal(hist, want_hist)
    # Batching from the largest bin down (complete_superposition.cu:482-490): the launch at
    # radius 21 collects 2 + 3 + 5 + 6 = 16 tiles and closes at bin 7, so the next launch is at
    # radius 6, which takes the four radius-2 tiles although no tile has radius 6 on its own.
    bt = np.array(p.batchRad[0][:33])
    want_bt = np.zeros(33, int)
    want_bt[[21, 6, 0]] = [16, 4, 4]
    np.testing.assert_array_equal(bt, want_bt)
    launch = {21: 21, 14: 21, 11: 21, 7: 21, 2: 6}

    def build(radius_of):
        out = np.zeros_like(bev)
        mass = 0.0
        for (k, ty, tx), m in tiles.items():
            R = radius_of(OWN_RADIUS[m])
            for y in range(ty * 8, ty * 8 + 8):
                for x in range(tx * 32, tx * 32 + 32):
                    d = idd[k, y, x]
                    if d == 0.0:
                        continue
                    e = _cell_masses(rs[k, y, x], R)
                    prof = np.concatenate([e[:0:-1], e])  # e_|i| for i = -R..R
                    out[k, y + 32 - R:y + 33 + R, x + 32 - R:x + 33 + R] += d * np.outer(prof, prof)
                    mass += d * erf(rs[k, y, x] * (R + 0.5)) ** 2
        return out, mass

    want, mass = build(lambda r: launch[r])
    np.testing.assert_allclose(bev, want, rtol=1e-11, atol=1e-14)
    assert bev.sum() == pytest.approx(mass, rel=1e-12)
    # The check discriminates: each tile at its own radius gives another field.
    own, own_mass = build(lambda r: r)
    assert np.abs(bev - own).max() > 1e-6
    assert abs(own_mass - mass) > 1e-9 * mass


def _set_affine(p, M, off):
    for i, v in enumerate(np.asarray(M, np.float32).ravel()):
        p.g2dose[i] = float(v)
    for i in range(3):
        p.g2doseOff[i] = float(off[i])


def test_transfer_finite_sad_closed_form():
    """Gantry -> dose index: (x, y, z) = (g.x + 20, g.y + 6, -g.z); fan steps of -1 along
    gantry z from 0, so fan step k sits at dose z = k; rays 1.0 x 0.75 wide from (-18, -5) (exact in float);
    source distances 300 / 250. The BEV is a + b i + c j + e k on every texel, which trilinear
    interpolation reproduces exactly between texel centres."""
    d = (300.0, 250.0)
    first, passive, bevz = 2, 12, 16
    sx, sy = 32, 8
    fc, fd = (-18.0, -5.0), (1.0, 0.75)
    p = O.Plan()
    p.superpDims[0], p.superpDims[1] = sx, sy
    p.beamFirstInside = first
    p.layerPassive[0] = passive
    p.guaranteedPassive = bevz
    _set_affine(p, [[1, 0, 0], [0, 1, 0], [0, 0, -1]], (20.0, 6.0, 0.0))
    p.fanCorner[0], p.fanCorner[1], p.fanCorner[2] = fc[0], fc[1], 0.0
    p.fanDelta[0], p.fanDelta[1], p.fanDelta[2] = fd[0], fd[1], -1.0
    p.sourceDist[0], p.sourceDist[1] = d
    a, b, c, e = 1.0, 0.013, 0.021, 0.057
    kk, jj, ii = np.meshgrid(np.arange(bevz), np.arange(sy + 64), np.arange(sx + 64), indexing="ij")
    bev = a + b * ii + c * jj + e * kk
    dims = np.array([48, 24, 20], np.int32)
    dose = np.zeros((20, 24, 48))
    O.lib().orc_transfer(C.byref(p), 0, O._dp(np.ascontiguousarray(bev)), O._dp(dose), O._ip(dims))

    z, y, x = np.meshgrid(np.arange(20), np.arange(24), np.arange(48), indexing="ij")
    gx, gy, gz = x - 20.0, y - 6.0, -z.astype(np.float64)
    # divergent fan: a point at gantry z, lateral g, lies on the ray through g d / (d - z)
    fi = (gx * d[0] / (d[0] - gz) - fc[0]) / fd[0] + 32
    fj = (gy * d[1] / (d[1] - gz) - fc[1]) / fd[1] + 32
    fk = (gz - 0.0) / -1.0 - first
    want = a + b * fi + c * fj + e * fk
    # interior samples (all eight texels inside the BEV) of voxels in the launched box
    interior = ((fi >= 0) & (fi <= sx + 63) & (fj >= 0) & (fj <= sy + 63) & (fk >= 0) & (fk <= bevz - 1))
    # box from the eight fan corners of steps [first, passive - 1] over the padded ray grid
    cx, cy, cz = [], [], []
    for k in (first, passive - 1):
        for jc in (-32, sy + 31):
            for ic in (-32, sx + 31):
                fzg = 0.0 + k * -1.0
                fx = (ic * fd[0] + fc[0]) * (1 - fzg / d[0])
                fy = (jc * fd[1] + fc[1]) * (1 - fzg / d[1])
                cx.append(fx + 20.0); cy.append(fy + 6.0); cz.append(-fzg)
    x0 = int(math.floor(min(cx))) // 32 * 32 if min(cx) >= 0 else 0
    y0, z0 = max(int(math.floor(min(cy))), 0), max(int(math.floor(min(cz))), 0)
    x1, y1, z1 = min(math.ceil(max(cx)), 47), min(math.ceil(max(cy)), 23), min(math.ceil(max(cz)), 19)
    xe = min(x0 + (x1 - x0 + 32) // 32 * 32, 48)
    ye = min(y0 + (y1 - y0 + 8) // 8 * 8, 24)
    inbox = (x >= x0) & (x < xe) & (y >= y0) & (y < ye) & (z >= z0) & (z <= z1)
    sel = interior & inbox
    assert sel.sum() > 2000
    np.testing.assert_allclose(dose[sel], want[sel], rtol=1e-12, atol=1e-12)
    assert not dose[~inbox].any()
    # a wrong divergence sign (x d / (d + z)) moves the far-lateral voxels by over a texel
    wrong = a + b * ((gx * d[0] / (d[0] + gz) - fc[0]) / fd[0] + 32) + c * fj + e * fk
    assert np.abs(wrong[sel] - dose[sel]).max() > 0.5 * b


def test_transfer_launch_slack():
    """Beam along dose +x: fan step k sits at dose x = k + 3. Steps [first, passive - 1] =
    [8, 13] give a box x in [11, 16], whose start floors to 0, so one 32-wide block covers
    x = 0..31. Each column x samples BEV plane x - 3 - first (the BEV plane index is the fan
    step minus the first inside step): the columns past the box, 17..31, read planes 6..20,
    and column 32 is never launched."""
    first, passive, bevz = 8, 14, 40
    sx, sy = 32, 8
    p = O.Plan()
    p.superpDims[0], p.superpDims[1] = sx, sy
    p.beamFirstInside = first
    p.layerPassive[0] = passive
    p.guaranteedPassive = bevz
    # dose x = -g.z + 3, dose y = g.y + 4, dose z = g.x + 16 (a rotation: det = 1)
    _set_affine(p, [[0, 0, -1], [0, 1, 0], [1, 0, 0]], (3.0, 4.0, 16.0))
    p.fanCorner[0], p.fanCorner[1], p.fanCorner[2] = -16.0, -4.0, 0.0
    p.fanDelta[0], p.fanDelta[1], p.fanDelta[2] = 1.0, 1.0, -1.0
    p.sourceDist[0] = p.sourceDist[1] = 1.0e30
    rng = np.random.default_rng(7)
    bev = 0.5 + rng.random((bevz, sy + 64, sx + 64))
    dims = np.array([64, 8, 32], np.int32)
    dose = np.zeros((32, 8, 64))
    O.lib().orc_transfer(C.byref(p), 0, O._dp(bev), O._dp(dose), O._ip(dims))
    # voxel (x, y, z): fan i = z - 16 + 16 + 32, j = y - 4 + 4 + 32, plane x - 3 - first
    for x in range(64):
        col = dose[:, :, x]
        plane = x - 3 - first
        if x >= 32 or plane < 0:
            assert not col.any(), x
            continue
        want = bev[plane, 32:40, 32:64].T  # [z][y] = bev[plane][y + 32][z + 32]
        np.testing.assert_allclose(col, want, rtol=0, atol=1e-12, err_msg=f"column {x}")
    # the slack columns past the box end (x = 17..31) do carry dose
    assert dose[:, :, 17:32].min() > 0.5
