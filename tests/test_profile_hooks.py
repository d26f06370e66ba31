"""CarbonPBS profile-sigma and overall-weight hooks (raytracedicom_wrapper.cu:2080-2188,
2816-2862, 4577-4832; switched off in the reference at 2812-2814, switchable here), on the CPU.

* The oracle's hooks (orc_profile_hooks) against closed forms: a table with one or two Gaussians,
  constant or linear in depth and row, and beam parameters a + 2 b d + c d^2 at the physical depth
  along each divergent ray, give rSigmaEff = (w/2) / (sqrt(2) (sqrt((a + 2 b d + c d^2 +
  sum w s^2 / sum w) / 2) + 0.21)) with w the mean divergent ray width; the overall channel
  scales the IDD; dead steps get +inf; steps outside [first inside, after last) stay untouched.
* The host planner's profile plan (pbs_plan_debug) against the oracle's and against hand-written
  table shapes and rows for each of the reference's shape rules.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import _abi
from subsecondpbs_amd import fixtures as F


def profile_beam(beam, L, depthN=40, step=4.0, rows=None, hooks=3, para_rows=None, energies=None, channels=5):
    """The beam with a synthetic two-Gaussian profile table [rows][depthN][channels] (weights
    0.7 - 0.002 d and 0.3 + 0.002 d, sigmas 1.5 + 0.01 d and 3.0 + 0.005 d mm, overall weight
    1 + 0.002 d + 0.05 row) and beam parameters (8 + 0.5 row, 0.01, 1e-4) per row."""
    rows = L if rows is None else rows
    d = np.arange(depthN, dtype=np.float64)
    t = np.zeros((rows, depthN, channels), np.float32)
    nG = (channels - 1) // 2
    for r in range(rows):
        if nG >= 1:
            t[r, :, 0] = 0.7 - 0.002 * d
            t[r, :, nG] = 1.5 + 0.01 * d
        if nG >= 2:
            t[r, :, 1] = 0.3 + 0.002 * d
            t[r, :, nG + 1] = 3.0 + 0.005 * d
        t[r, :, channels - 1] = 1.0 + 0.002 * d + 0.05 * r
    pr = rows if para_rows is None else para_rows
    para = np.stack([8.0 + 0.5 * np.arange(pr), np.full(pr, 0.01), np.full(pr, 1e-4)], 1).astype(np.float32)
    b = dict(beam)
    b["profile_data"] = t
    b["profile_setting"] = np.array([0.0, step, depthN], np.float32)
    b["beam_para_data"] = para
    b["beam_para_pos"] = 0.0
    if energies is not None:
        b["profile_energies"] = np.asarray(energies, np.float32)
    b["carbon_profile_hooks"] = hooks
    return b


def _hook_plan(K=12, rx=32, ry=8, first=2, after=9, rows=1, depthN=6, channels=3, hooks=3, dist=(2000.0, 1800.0)):
    p = O.Plan()
    p.rayDims[0], p.rayDims[1] = rx, ry
    p.superpDims[0], p.superpDims[1] = rx, ry
    p.steps = K
    p.beamFirstInside = first
    p.afterLast[0] = after
    p.fanCorner[0], p.fanCorner[1], p.fanCorner[2] = -16.0, -4.0, 30.0
    p.fanDelta[0], p.fanDelta[1], p.fanDelta[2] = 1.0, 0.75, -2.0
    p.sourceDist[0], p.sourceDist[1] = dist
    p.profileHooks = hooks
    p.profileRows, p.profileDepthN, p.profileChannels = rows, depthN, channels
    p.profileDepth0, p.profileDepthStep = 10.0, 5.0
    p.beamParaPos = 40.0
    return p


def _geometry(p, K, rx, ry):
    k, y, x = np.meshgrid(np.arange(K), np.arange(ry), np.arange(rx), indexing="ij")
    fx = -16.0 + x * 1.0
    fy = -4.0 + y * 0.75
    fz = 30.0 + k * -2.0
    dx, dy = p.sourceDist[0], p.sourceDist[1]
    sad = 0.5 * (dx + dy)
    # distance from the source along the ray through (fx, fy) at axial distance sad - fz
    slant = np.sqrt(1.0 + (fx / dx) ** 2 + (fy / dy) ** 2)
    phy = (sad - fz) * slant - (sad - 40.0)
    width = 0.5 * (1.0 * (1.0 - fz / dx) + 0.75 * (1.0 - fz / dy))
    return k, phy, width


def test_hooks_one_gaussian_closed_form():
    K, rx, ry, first, after = 12, 32, 8, 2, 9
    p = _hook_plan(K, rx, ry, first, after)
    s0 = 2.5
    table = np.zeros((1, 6, 3), np.float32)
    table[0, :, 0] = 3.0                      # weight (the sigma is a weighted mean: one Gaussian)
    table[0, :, 1] = s0                       # sigma
    table[0, :, 2] = 0.5 + 0.125 * np.arange(6)  # overall weight, linear in the depth index (exact in float)
    p.profileTable = table.ctypes.data_as(O._f)
    a, b, c = 6.0, 0.015625, 2.0 ** -12  # exact in float
    p.profileLayer[0][0], p.profileLayer[0][1], p.profileLayer[0][2], p.profileLayer[0][3] = 0.0, a, b, c
    rng = np.random.default_rng(0)
    wepl = rng.uniform(0.0, 50.0, (K, ry, rx))
    idd = rng.uniform(0.1, 1.0, (K, ry, rx))
    idd[5, 3, 7] = 0.0
    rs = np.full((K, ry, rx), 0.123)
    idd0 = idd.copy()
    O.lib().orc_profile_hooks(C.byref(p), 0, O._dp(wepl), O._dp(idd), O._dp(rs))
    k, phy, width = _geometry(p, K, rx, ry)
    r2 = a + 2 * b * phy + c * phy * phy
    want_rs = 0.5 * width / (math.sqrt(2.0) * (np.sqrt(0.5 * (r2 + s0 * s0)) + 0.21))
    dix = np.clip((wepl - 10.0) / 5.0, 0.0, 5.0)
    want_idd = idd0 * (0.5 + 0.125 * dix)
    act = (k >= first) & (k < after)
    live = act & (idd0 > 0)
    np.testing.assert_allclose(rs[live], want_rs[live], rtol=1e-12)
    np.testing.assert_allclose(idd[act], want_idd[act], rtol=1e-12)
    assert np.isinf(rs[5, 3, 7]) and idd[5, 3, 7] == 0.0
    assert (rs[~act] == 0.123).all() and np.array_equal(idd[~act], idd0[~act])
    # the physical depth matters: a ray far off axis sits deeper than the axial one at one step
    assert abs(phy[4, 0, 0] - phy[4, 4, 16]) > 0.05


def test_hooks_two_gaussians_rows_and_depth():
    """Two Gaussians whose weights and sigmas change with depth, and a table row halfway between
    two rows: the weighted sigma^2 mean of the bilinearly read table."""
    K, rx, ry = 10, 32, 8
    p = _hook_plan(K, rx, ry, 0, 10, rows=2, depthN=6, channels=5, hooks=1)
    d = np.arange(6, dtype=np.float64)
    table = np.zeros((2, 6, 5), np.float32)
    for r in range(2):
        table[r, :, 0] = 0.6 + 0.05 * d + 0.1 * r
        table[r, :, 1] = 0.4 - 0.02 * d
        table[r, :, 2] = 1.2 + 0.1 * d + 0.3 * r
        table[r, :, 3] = 2.8 - 0.05 * d
        table[r, :, 4] = 7.0
    p.profileTable = table.ctypes.data_as(O._f)
    p.profileLayer[0][0], p.profileLayer[0][1], p.profileLayer[0][2], p.profileLayer[0][3] = 0.5, 5.0, 0.0, 0.0
    rng = np.random.default_rng(1)
    wepl = rng.uniform(10.0, 35.0, (K, ry, rx))  # depth index in [0, 5]: no clamping
    idd = rng.uniform(0.1, 1.0, (K, ry, rx))
    idd0 = idd.copy()
    rs = np.zeros((K, ry, rx))
    O.lib().orc_profile_hooks(C.byref(p), 0, O._dp(wepl), O._dp(idd), O._dp(rs))
    di = (wepl - 10.0) / 5.0
    w1, w2 = 0.6 + 0.05 * di + 0.05, 0.4 - 0.02 * di
    s1, s2 = 1.2 + 0.1 * di + 0.15, 2.8 - 0.05 * di
    m2 = (w1 * s1 * s1 + w2 * s2 * s2) / (w1 + w2)
    _, phy, width = _geometry(p, K, rx, ry)
    want = 0.5 * width / (math.sqrt(2.0) * (np.sqrt(0.5 * (5.0 + m2)) + 0.21))
    np.testing.assert_allclose(rs, want, rtol=1e-6)
    assert np.array_equal(idd, idd0)  # the weight hook is off


def _plan_pair(case, beam):
    got = _abi.plan_debug(case, beam)
    c = dict(case)
    c["beams"] = [beam]
    p = O.plan_beam(c)
    assert p.ok == 1, p.error
    return got, p


@pytest.mark.parametrize("name", ["rows_are_layers", "rows_from_beam_para", "rows_from_energies",
                                  "rows_are_table_energies", "eleven_channel_export", "hooks_off", "sigma_needs_para"])
def test_profile_plan_matches_oracle(name):
    case = F.make_config1()
    beam = case["beams"][0]
    L = len(beam["energies"])  # 100..140 MeV in 10 MeV steps
    E = case["energy"]
    if name == "rows_are_layers":
        b = profile_beam(beam, L)
        want = dict(hooks=3, rows=5, channels=5, row=[0, 1, 2, 3, 4])
    elif name == "rows_from_beam_para":
        b = profile_beam(beam, L, rows=7, para_rows=7)
        want = dict(hooks=3, rows=7, channels=5, row=[0, 1, 2, 3, 4])
    elif name == "rows_from_energies":
        # 9 rows on their own energy axis 80..160 MeV; 5 beam-parameter rows do not divide the table
        b = profile_beam(beam, L, rows=9, para_rows=5, channels=7, energies=np.linspace(80, 160, 9))
        want = dict(hooks=3, rows=9, channels=7, row=[2, 3, 4, 5, 6])
    elif name == "rows_are_table_energies":
        n = int(E["n_energies"])
        b = profile_beam(beam, L, rows=n, para_rows=n, depthN=8)
        want = dict(hooks=3, rows=n, channels=5, row=None)
    elif name == "eleven_channel_export":
        b = profile_beam(beam, L, rows=3, para_rows=0, channels=11, hooks=2)
        b.pop("beam_para_data")
        want = dict(hooks=2, rows=3, channels=11, row=[0, 1, 2, 2, 2])
    elif name == "hooks_off":
        b = profile_beam(beam, L, hooks=0)
        want = dict(hooks=0, rows=0, channels=0, row=None)
    else:  # the sigma hook needs beam parameters for every layer
        b = profile_beam(beam, L, para_rows=2, hooks=1)
        want = dict(hooks=0, rows=0, channels=0, row=None)
    got, p = _plan_pair(case, b)
    assert got["profileHooks"] == p.profileHooks == want["hooks"]
    assert got["profileRows"] == p.profileRows == want["rows"]
    assert got["profileChannels"] == p.profileChannels == want["channels"]
    if want["hooks"] == 0:
        return
    ref = np.array([list(p.profileLayer[l]) for l in range(L)], np.float32)
    np.testing.assert_allclose(got["profileLayer"], ref, rtol=1e-6, atol=1e-6)
    if want["row"] is not None:
        np.testing.assert_allclose(ref[:, 0], want["row"], atol=1e-6)
    if name == "rows_are_table_energies":
        # the row is the layer's fractional index in the energy table
        e = np.asarray(E["energies_per_u"], np.float64)
        np.testing.assert_allclose(ref[:, 0], np.interp(np.asarray(beam["energies"], np.float64), e, np.arange(e.size)),
                                   atol=1e-3)
    # beam parameters interpolated at the row: 8 + 0.5 row, 0.01, 1e-4 (the table rows' own values)
    if b.get("beam_para_data") is not None and want["rows"] == np.asarray(b["beam_para_data"]).reshape(-1, 3).shape[0]:
        np.testing.assert_allclose(ref[:, 1], 8.0 + 0.5 * ref[:, 0], rtol=1e-6)
        np.testing.assert_allclose(ref[:, 2:], [[0.01, 1e-4]] * L, rtol=1e-6)


def test_hooks_change_the_oracle_dose():
    """With the hooks on, the oracle's dose differs from the plain method; with them off (the
    reference's setting) the profile arrays change nothing."""
    case = F.make_config1()
    beam = case["beams"][0]
    L = len(beam["energies"])
    ref, _ = O.beam_dose(case)
    for hooks, same in ((0, True), (1, False), (2, False), (3, False)):
        c = dict(case)
        c["beams"] = [profile_beam(beam, L, hooks=hooks)]
        d, p = O.beam_dose(c)
        assert p.profileHooks == hooks
        assert np.array_equal(d, ref) == same, hooks


@pytest.mark.parametrize("bad", ["setting", "para"])
def test_profile_shapes_checked_at_entry(bad):
    """The entry validation's profile-model shapes (raytracedicom_wrapper.cu:1143-1148): a
    profile table needs the three-value depth setting, and beam parameters come in rows of 3."""
    case = F.make_config1()
    b = dict(case["beams"][0])
    L = len(b["energies"])
    b = profile_beam(b, L, hooks=0)
    if bad == "setting":
        b["profile_setting"] = np.asarray(b["profile_setting"], np.float32)[:2]
        needle = "profileSetting"
    else:
        b["beam_para_data"] = np.asarray(b["beam_para_data"], np.float32).ravel()[:-1]
        needle = "multiple of 3"
    with pytest.raises(RuntimeError, match=needle):
        _abi.plan_debug(case, b)
