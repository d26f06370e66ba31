"""The C ABI library loads and exports every entry point include/pbs_abi.h declares (CPU)."""
import ctypes as C
import os
import re

import pytest

from subsecondpbs_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pbs_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-zA-Z_][\w\s\*]*?\b([a-zA-Z_]\w*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "for", "return", "sizeof")))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_abi.LIB_PATH), "build the engine first"
    lib = C.CDLL(_abi.LIB_PATH)
    declared = _declared()
    assert set(_abi.EXPORTED) == set(declared), (declared, _abi.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name


def test_bindings_load_without_gpu():
    L = _abi.lib()
    assert L.pbs_device_count() >= 0
    assert isinstance(_abi.last_error(), str)


def test_invalid_arguments_fail_cleanly():
    """Validation errors return a code and a message instead of touching the GPU."""
    L = _abi.lib()
    d, dp = _abi.i3((4, 4, 4))
    r, rp = _abi.f3((1, 1, 1))
    rc = L.pbs_subsecond_wrapper(None, dp, rp, rp, None, dp, rp, rp, None, 1, None, 0, 0, 0)
    assert rc == -1 and "ctData" in _abi.last_error()


def test_final_dose_roi_integers():
    """ROI bounding box + margins (5,5,2), intersection sign and the 0.5 mm CPB grid."""
    L = _abi.lib()
    roi = (_abi.Vec3i * 3)(_abi.Vec3i(10, 12, 5), _abi.Vec3i(20, 14, 9), _abi.Vec3i(15, 30, 7))
    res, corner = _abi.Vec3f(2.0, 2.0, 3.0), _abi.Vec3f(-50.0, -60.0, -20.0)
    mn, mx, dims, cpb = _abi.Vec3f(), _abi.Vec3f(), _abi.Vec3i(), _abi.CPBGrid()
    hit = L.pbs_final_dose_roi(roi, 3, res, corner, _abi.Vec3f(0, 0, 1), _abi.Vec3f(0, 0, -1000), 4,
                               C.byref(mn), C.byref(mx), C.byref(dims), C.byref(cpb))
    assert hit == 1
    assert (mn.x, mn.y, mn.z) == (-50 + 20 - 5, -60 + 24 - 5, -20 + 15 - 2)
    assert (mx.x, mx.y, mx.z) == (-50 + 42 + 5, -60 + 62 + 5, -20 + 30 + 2)
    assert (dims.x, dims.y, dims.z) == (11, 19, 5)
    assert (cpb.dims.x, cpb.dims.y, cpb.dims.z) == (int((mx.x - mn.x) / 0.5), int((mx.y - mn.y) / 0.5), 4)
    assert (cpb.corner.x, cpb.corner.y, cpb.corner.z) == (mn.x, mn.y, 0.0)
    # beam pointing away from the ROI centre
    hit = L.pbs_final_dose_roi(roi, 3, res, corner, _abi.Vec3f(0, 0, -1), _abi.Vec3f(0, 0, -1000), 4,
                               None, None, None, None)
    assert hit == 0


def test_cxx_dropin_symbols_exported():
    """include/raytracedicom_integration.h: the reference's original C++ entry points."""
    lib = C.CDLL(_abi.LIB_PATH)
    for name in ("subsecondWrapper", "createRTDBeamSettings", "destroyRTDBeamSettings",
                 "createRTDEnergyStruct", "destroyRTDEnergyStruct"):
        assert hasattr(lib, name), name


def test_struct_mirrors_match_the_library():
    """The ctypes mirrors of pbs_beam_settings / pbs_energy_struct have the compiled sizes."""
    import ctypes as C

    from subsecondpbs_amd import _abi

    L = _abi.lib()
    assert L.pbs_abi_sizeof(0) == C.sizeof(_abi.BeamSettings)
    assert L.pbs_abi_sizeof(1) == C.sizeof(_abi.EnergyStruct)
    assert L.pbs_abi_sizeof(2) == C.sizeof(_abi.PlanInfo)
    assert L.pbs_abi_sizeof(7) == -1
