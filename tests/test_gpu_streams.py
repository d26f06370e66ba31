"""Stream order between torch and the engine (regression test).

A dose tensor that torch fills and zeroes behind a few milliseconds of queued torch work is
handed to ``Engine.beam_dose`` straight away. The engine's stream must wait for torch's before
it accumulates; before the wait was added the engine's writes could land first and torch's
zero then wiped them. Every repetition must equal the dose computed with torch idle, bitwise.
"""
import numpy as np
import pytest

from subsecondpbs_amd import fixtures as F

torch = pytest.importorskip("torch")
gpu = pytest.mark.gpu


@gpu
def test_engine_waits_for_queued_torch_work():
    from subsecondpbs_amd.engine import Engine

    case = F.make_config1()
    beam = case["beams"][0]
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    nx, ny, nz = case["dose_dims"]
    ref = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    eng.beam_dose(beam, ref)
    eng.sync()
    ref = ref.cpu().numpy()
    assert ref.max() > 0.0
    print("\n[gpu-test] engine after queued torch work", flush=True)
    g = torch.randn((4096, 4096), dtype=torch.float32, device="cuda") * 1e-2
    for rep in range(4):
        for _ in range(4):
            g = (g @ g) * 1e-2  # milliseconds of work queued on torch's stream
        d = torch.full((nz, ny, nx), 1.0e30, dtype=torch.float32, device="cuda")
        d.zero_()
        eng.beam_dose(beam, d)
        eng.sync()
        got = d.cpu().numpy()
        assert np.array_equal(got, ref), f"repetition {rep}: max |diff| {np.abs(got - ref).max():.3e}"
    eng.close()
