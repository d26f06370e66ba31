"""Launch shapes never change a result: the head-size beam's dose is bitwise the same under
every superposition grid, transfer brick and stream mode the autotuner may pick, and the tuner
leaves the engine at its best setting."""
import numpy as np
import pytest

from subsecondpbs_amd import autotune, fixtures as F

gpu = pytest.mark.gpu


@gpu
def test_every_launch_shape_gives_the_same_dose():
    import torch
    from subsecondpbs_amd.engine import Engine

    case = F.make_config1()
    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    nx, ny, nz = case["dose_dims"]
    beam = case["beams"][0]

    def dose():
        d = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
        eng.beam_dose(beam, d)
        eng.sync()
        return d.cpu().numpy()

    ref = dose()
    print("\n[gpu-test] launch shapes on config1", flush=True)
    for knob, values in autotune.CANDIDATES.items():
        for v in values:
            eng.set_tuning(**{knob: v})
            assert np.array_equal(dose(), ref), (knob, v)
        eng.set_tuning(**{knob: {"superpose_grid": 4096, "transfer_shape": 0, "one_stream": 0}[knob]})
    d = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    res = autotune.tune(eng, beam, d, reps=3)
    assert eng.tuning() == res["best"]
    assert np.array_equal(dose(), ref)
    print(f"tuned {res['best']}: {res['default_ms']:.3f} -> {res['best_ms']:.3f} ms")
    with pytest.raises(RuntimeError):
        eng.set_tuning(transfer_shape=9)
    eng.close()
