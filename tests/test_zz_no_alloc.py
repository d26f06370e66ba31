"""pbs_engine_beam_dose_noalloc through the C ABI: the same dose as pbs_engine_beam_dose, bitwise,
on two calls in a row (the second with a smaller beam) on an engine that has run no beam, with no
growth of the device memory in use; a beam too large for the scratch fails cleanly."""
import ctypes as C

import numpy as np
import pytest

from subsecondpbs_amd import _abi, fixtures as F

gpu = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _engine(case):
    from subsecondpbs_amd.engine import Engine
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    return eng


def _smaller(beam):
    """The beam's first three layers (fewer layers and spots)."""
    b = dict(beam)
    n = 3
    counts = np.asarray(beam["layer_spot_counts"])
    k = int(counts[:n].sum())
    b["energies"] = np.asarray(beam["energies"])[:n]
    b["spot_sigmas"] = np.asarray(beam["spot_sigmas"]).reshape(-1, 2)[:n]
    b["subspot_data"] = np.asarray(beam["subspot_data"])[:n]
    b["layer_spot_counts"] = counts[:n]
    b["spot_positions"] = np.asarray(beam["spot_positions"]).reshape(-1, 2)[:k]
    b["spot_weights"] = np.asarray(beam["spot_weights"])[:k]
    return b


def _call(lib, fn, eng, beam, dose):
    keep = _abi.Keep()
    bs = _abi.beam_settings(beam, keep)
    return getattr(lib, fn)(C.c_void_p(eng.handle), C.byref(bs), C.c_void_p(dose.data_ptr()))


@gpu
def test_noalloc_matches_beam_dose_twice():
    torch = _torch()
    case = F.make_config1()
    nx, ny, nz = case["dose_dims"]
    lib = _abi.lib()
    beams = [case["beams"][0], _smaller(case["beams"][0])]
    ref_eng = _engine(case)
    eng = _engine(case)  # has run no beam: every work buffer must come from the scratch
    refs, gots = [], []
    for i, beam in enumerate(beams):
        print(f"[no-alloc] call {i}: {len(beam['energies'])} layers, {len(beam['spot_weights'])} spots", flush=True)
        ref = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
        got = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _abi.check(_call(lib, "pbs_engine_beam_dose", ref_eng, beam, ref), "pbs_engine_beam_dose")
        ref_eng.sync()
        free0 = torch.cuda.mem_get_info()[0]
        rc = _call(lib, "pbs_engine_beam_dose_noalloc", eng, beam, got)
        _abi.check(rc, "pbs_engine_beam_dose_noalloc")
        eng.sync()
        free1 = torch.cuda.mem_get_info()[0]
        refs.append(ref.cpu().numpy())
        gots.append(got.cpu().numpy())
        assert refs[-1].max() > 0
        assert np.array_equal(gots[-1], refs[-1]), f"call {i}: max |d| {np.abs(gots[-1] - refs[-1]).max()}"
        # the engine's normal call allocates ~230 MB of work buffers for this beam; this one none
        assert free0 - free1 < (8 << 20), f"device memory in use grew by {(free0 - free1) >> 20} MiB"
    assert not np.array_equal(refs[0], refs[1])
    ref_eng.close()
    eng.close()


@gpu
def test_noalloc_rejects_a_beam_beyond_the_scratch():
    torch = _torch()
    case = F.make_config2()  # head-size beam: its work buffers exceed the 256 MiB scratch
    nx, ny, nz = case["dose_dims"]
    eng = _engine(case)
    dose = torch.full((nz, ny, nx), 3.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = _call(_abi.lib(), "pbs_engine_beam_dose_noalloc", eng, case["beams"][0], dose)
    assert rc == -4  # PBS_E_NOMEM
    assert "no-allocation scratch" in _abi.last_error()
    eng.sync()
    eng.close()
    assert bool((dose == 3.0).all())
