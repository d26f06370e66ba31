"""Unit heuristics, the reference's environment overrides and the CT adapter, on the CPU.

* The centimetre heuristic (raytracedicom_wrapper.cu:2728-2729: ctResolution.z < 0.3 means cm)
  and the LUT depth unit (2742-2790: a cm CT with peak depths in the hundreds reads the tables
  as mm), with the RTD_ENERGY_DEPTH_SCALE / RTD_ENERGY_DEPTH_UNIT overrides: the host planner
  (pbs_plan_debug) against hand-written units and against the oracle's plan.
* RTD_CPB_COVERAGE_CM (3193-3202) on the legacy CPB grid.
* The CT adapter (2288-2441): HU, HU+1000, density and SPR volumes, detected or forced by
  RTD_CT_INPUT_TYPE. The oracle's restatement is pinned by the table round trip (the table
  interpolated at the returned index gives back the input) and, on the density table's
  non-monotonic stretch, by the crossing the reference's bisection visits, found here by an
  independent walk; the engine's adapter must equal it bitwise.
"""
import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import _abi
from subsecondpbs_amd import fixtures as F
from subsecondpbs_amd import lut

E = lut.load_energy_tables()
ENV = ("RTD_ENERGY_DEPTH_SCALE", "RTD_ENERGY_DEPTH_UNIT", "RTD_CPB_COVERAGE_CM", "RTD_CT_INPUT_TYPE")


@pytest.fixture(autouse=True)
def _clean_env(monkeypatch):
    for k in ENV:
        monkeypatch.delenv(k, raising=False)


def _to_cm(case):
    """The case with every length in cm (what CarbonPBS callers pass)."""
    t = np.float32(10)
    c = dict(case)
    for k in ("ct_resolution", "ct_corner", "dose_resolution", "dose_corner"):
        c[k] = tuple(float(np.float32(v) / t) for v in case[k])
    b = dict(case["beams"][0])
    for k in ("ray_spacing", "source_dist", "source_position", "spot_delta", "spot_offset", "spot_sigmas",
              "spot_positions"):
        if k in b:
            b[k] = np.asarray(b[k], np.float32) / t
    b["sad"] = float(np.float32(b["sad"]) / t)
    ss = np.array(b["subspot_data"], np.float32)
    ss[..., [0, 1, 3, 4]] /= t
    b["subspot_data"] = ss
    c["beams"] = [b]
    return c


def _oracle_plan(case, mode=0, to_mm=0.0, coverage=0.0):
    c = dict(case)
    b = dict(case["beams"][0])
    b["energy_depth_mode"], b["energy_depth_to_mm"], b["cpb_coverage"] = mode, to_mm, coverage
    c["beams"] = [b]
    p = O.plan_beam(c)
    assert p.ok == 1, p.error
    return p


# (environment, energy depth mode and value the oracle is given, expected LUT depth unit in mm)
CM_CASES = [
    ({}, 0, 0.0, 1.0),                                    # heuristic: peak depths > 200 read as mm
    ({"RTD_ENERGY_DEPTH_UNIT": "cm"}, 1, 10.0, 10.0),
    ({"RTD_ENERGY_DEPTH_UNIT": "MM"}, 1, 1.0, 1.0),
    ({"RTD_ENERGY_DEPTH_UNIT": "furlong"}, 2, 0.0, 10.0),  # unknown unit: the CT's, no heuristic
    ({"RTD_ENERGY_DEPTH_SCALE": "2.5"}, 1, 2.5, 2.5),
    ({"RTD_ENERGY_DEPTH_SCALE": "0"}, 2, 0.0, 10.0),      # zero scale: the CT's unit
    ({"RTD_ENERGY_DEPTH_SCALE": "-3"}, 2, 0.0, 10.0),
    ({"RTD_ENERGY_DEPTH_SCALE": "4", "RTD_ENERGY_DEPTH_UNIT": "cm"}, 1, 4.0, 4.0),  # the scale wins
]


@pytest.mark.parametrize("env,mode,to_mm,want", CM_CASES, ids=[str(c[0]) for c in CM_CASES])
def test_energy_depth_unit_on_a_cm_ct(monkeypatch, env, mode, to_mm, want):
    mm = F.make_config1()
    cm = _to_cm(mm)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    got = _abi.plan_debug(cm, cm["beams"][0])
    assert got["lenToMm"] == 10.0  # ctResolution.z = 0.2 < 0.3
    assert got["energyDepthToMm"] == pytest.approx(want, rel=1e-7)
    p = _oracle_plan(cm, mode, to_mm)
    assert p.energyDepthToMm == pytest.approx(want, rel=1e-7)
    L = got["numLayers"]
    # peak depths are the table's, times the depth unit (KAT: 100 MeV -> 74.209 mm in the mm table)
    assert got["peakDepth"][0] == pytest.approx(74.209 * want, rel=2e-5)
    keep = O._Keep()
    O.lib().orc_layer_params(O.C.byref(p), O.C.byref(O.make_beam(cm["beams"][0], keep)),
                             O.C.byref(O.make_energy(cm["energy"], keep)), O._dp(np.zeros(max(p.steps, 1))))
    np.testing.assert_allclose(got["peakDepth"], np.asarray(p.peakDepth[:L]), rtol=1e-6)
    # the geometry is the mm case's: the same grids, steps and (mm) fan
    ref = _abi.plan_debug(mm, mm["beams"][0])
    for k in ("rayDims", "superpDims", "steps", "spotDims"):
        assert tuple(np.atleast_1d(got[k])) == tuple(np.atleast_1d(ref[k])), k
    np.testing.assert_allclose(got["fanCorner"], ref["fanCorner"], rtol=1e-6, atol=1e-5)
    np.testing.assert_allclose(got["fanDelta"], ref["fanDelta"], rtol=1e-6)
    np.testing.assert_allclose(got["sourceDist"], ref["sourceDist"], rtol=1e-6)
    assert tuple(p.rayDims) == tuple(got["rayDims"]) and p.steps == got["steps"]


def test_energy_depth_unit_on_a_mm_ct(monkeypatch):
    mm = F.make_config1()
    got = _abi.plan_debug(mm, mm["beams"][0])
    assert got["lenToMm"] == 1.0 and got["energyDepthToMm"] == 1.0
    monkeypatch.setenv("RTD_ENERGY_DEPTH_UNIT", "cm")
    got = _abi.plan_debug(mm, mm["beams"][0])
    assert got["energyDepthToMm"] == 10.0
    assert got["peakDepth"][0] == pytest.approx(742.09, rel=2e-5)
    assert _oracle_plan(mm, 1, 10.0).energyDepthToMm == 10.0


def _legacy_case():
    case = F.make_config1()
    b = dict(case["beams"][0])
    L = len(b["energies"])
    counts = np.asarray(b["layer_spot_counts"])
    pos = np.asarray(b["spot_positions"], np.float32).reshape(-1, 2)
    w = np.asarray(b["spot_weights"], np.float32)
    sub = np.zeros((L, int(counts.max()), 5), np.float32)
    o = 0
    for l in range(L):
        n = int(counts[l])
        sub[l, :n, 0:2] = pos[o:o + n]
        sub[l, :n, 2] = w[o:o + n]
        sub[l, :n, 3:5] = np.asarray(b["spot_sigmas"], np.float32).reshape(L, 2)[l]
        o += n
    for k in ("layer_spot_counts", "spot_positions", "spot_weights"):
        b.pop(k)
    b["subspot_data"] = sub
    case["beams"] = [b]
    return case


@pytest.mark.parametrize("value,applies", [("30", True), ("12.25", True), ("0", False), ("-5", False),
                                           ("nan", False), ("abc", False)])
def test_cpb_coverage_override(monkeypatch, value, applies):
    case = _legacy_case()
    base = _abi.plan_debug(case, case["beams"][0])
    assert base["legacyPath"] == 1
    monkeypatch.setenv("RTD_CPB_COVERAGE_CM", value)
    got = _abi.plan_debug(case, case["beams"][0])
    rs = case["beams"][0]["ray_spacing"]
    if applies:
        c = float(value)
        want = (int(np.ceil(np.float32(c) / np.float32(rs[0]))), int(np.ceil(np.float32(c) / np.float32(rs[1]))))
        assert tuple(got["rayDims"]) == want
        p = _oracle_plan(case, coverage=c)
        assert tuple(p.rayDims) == want
        assert tuple(got["cpbCorner"]) == tuple(base["cpbCorner"])  # the corner stays the projected one
    else:
        assert tuple(got["rayDims"]) == tuple(base["rayDims"])


# ---------------------------------------------------------------- CT adapter ---------------
def _bisection_crossing(v, value):
    """The reference's bisection restated as a walk over (lo, hi) pairs: the crossing it ends on."""
    lo, hi = 0, len(v) - 1
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if v[mid] <= value:
            lo = mid
        else:
            hi = mid
    return lo


def test_ct_adapter_oracle_pins():
    dens = np.asarray(E["density_vector"], np.float32)
    sp = np.asarray(E["sp_vector"], np.float32)
    en = E
    # HU (auto: a value below -200) -> HU + 1000; HU+1000 passes unchanged
    hu = np.array([-1000.0, -250.0, 0.0, 40.0, 1200.0], np.float32)
    np.testing.assert_array_equal(O.ct_as_hu_plus_1000(hu, en), hu + np.float32(1000))
    hp = np.array([0.0, 1000.0, 2500.0], np.float32)
    np.testing.assert_array_equal(O.ct_as_hu_plus_1000(hp, en), hp)
    # SPR (forced; the table rises strictly): texel values come back as their index, values
    # between texels as the fraction, and the table reproduces the input
    idx = np.array([1.0, 17.0, 1000.0, 2000.0, 3000.0], np.float32)
    np.testing.assert_array_equal(O.ct_as_hu_plus_1000(sp[idx.astype(int)], en, forced_type=3), idx)
    rng = np.random.default_rng(0)
    x = (rng.random(4000) * 3070 + 0.5).astype(np.float32)
    vals = np.interp(x, np.arange(sp.size), sp).astype(np.float32)
    got = O.ct_as_hu_plus_1000(vals, en, forced_type=3)
    np.testing.assert_allclose(np.interp(got, np.arange(sp.size), sp), vals, rtol=2e-6, atol=2e-7)
    np.testing.assert_allclose(got, x, atol=2e-2)
    # clamps at the table's ends
    np.testing.assert_array_equal(O.ct_as_hu_plus_1000(np.array([0.0, 9.0], np.float32), en, forced_type=3),
                                  np.array([0.0, sp.size - 1], np.float32))
    # density (auto: every value in [0, 20]): round trip everywhere; on the stretch where the
    # table falls (1.12 -> 1.08 at texel 1100) several texel pairs bracket a value, and the
    # result lies on the pair the bisection ends on
    vals = np.concatenate([rng.random(3000).astype(np.float32) * np.float32(2.2) + np.float32(0.01),
                           np.linspace(1.081, 1.119, 200, dtype=np.float32)])
    got = O.ct_as_hu_plus_1000(vals, en)
    np.testing.assert_allclose(np.interp(got, np.arange(dens.size), dens), vals, rtol=3e-6, atol=3e-7)
    for v, g in zip(vals[-200:], got[-200:]):
        lo = _bisection_crossing(dens, v)
        crossings = [i for i in range(dens.size - 1) if min(dens[i], dens[i + 1]) <= v <= max(dens[i], dens[i + 1])]
        assert len(crossings) >= 2 and lo in crossings
        assert lo <= g <= lo + 1


@pytest.mark.parametrize("kind,forced", [("hu", None), ("hu+1000", None), ("density", None), ("spr", "spr"),
                                         ("density", "DENSITY"), ("hu+1000", "hu"), ("density", "huplus1000"),
                                         ("hu", "bogus")])
def test_ct_adapter_engine_matches_oracle(monkeypatch, kind, forced):
    rng = np.random.default_rng(4)
    base = {"hu": rng.uniform(-1000, 1500, 5000), "hu+1000": rng.uniform(0, 2500, 5000),
            "density": rng.uniform(0.0005, 2.3, 5000), "spr": rng.uniform(0.0005, 1.8, 5000)}[kind]
    ct = base.astype(np.float32).reshape(10, 20, 25)
    if forced:
        monkeypatch.setenv("RTD_CT_INPUT_TYPE", forced)
    # auto detection sees SPR-range values as density (both lie in [0, 20])
    auto = {"hu": 0, "hu+1000": 1, "density": 2, "spr": 2}[kind]
    want_type = {"hu": 0, "hu+1000": 1, "huplus1000": 1, "density": 2, "spr": 3}.get(str(forced).lower(), auto)
    got, t = _abi.ct_to_hu_plus_1000(ct, E)
    assert t == want_type
    ref = O.ct_as_hu_plus_1000(ct, E, forced_type=want_type)
    np.testing.assert_array_equal(got, ref)
