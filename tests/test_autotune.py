"""Autotuner logic on the CPU with a stand-in engine: coordinate descent over the launch knobs
picks the fastest value of each and leaves it applied; engines without side streams drop the
stream knob; the JSON cache round-trips. The GPU side (results identical under every launch
shape) is in tests/test_gpu_tuning.py."""
import pytest

from subsecondpbs_amd import autotune


class FakeEngine:
    def __init__(self, side_streams=True):
        self.t = {"superpose_grid": 4096, "transfer_shape": 0, "one_stream": 0 if side_streams else 1}
        self.side = side_streams

    def tuning(self):
        return dict(self.t)

    def set_tuning(self, **kw):
        for k, v in kw.items():
            if k == "one_stream" and v == 0 and not self.side:
                raise RuntimeError("pbs_engine_set_tuning: this engine was created without side streams")
            self.t[k] = v


def _cost(eng):
    t = eng.t
    return 1.0 + abs(t["superpose_grid"] - 2048) / 1e4 + (0.0 if t["transfer_shape"] == 3 else 0.1) \
        + (0.05 if t["one_stream"] else 0.0)


def test_tune_picks_the_fastest_values():
    eng = FakeEngine()
    res = autotune.tune(eng, {}, None, measure=lambda: _cost(eng))
    assert res["best"] == {"superpose_grid": 2048, "transfer_shape": 3, "one_stream": 0}
    assert eng.tuning() == res["best"]
    assert res["best_ms"] == pytest.approx(1.0)
    assert res["default_ms"] > res["best_ms"]
    assert set(res["ms"]) == {"superpose_grid", "transfer_shape", "one_stream"}


def test_engine_without_side_streams_keeps_one_stream():
    eng = FakeEngine(side_streams=False)
    res = autotune.tune(eng, {}, None, measure=lambda: _cost(eng))
    assert "one_stream" not in res["ms"] and eng.tuning()["one_stream"] == 1


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    case = {"ct_dims": (64, 64, 64), "dose_dims": (64, 64, 64)}
    key = autotune.signature(case, {"energies": [1, 2], "spot_weights": [1.0] * 10}, "gfx950")
    assert autotune.load(path, key) is None
    autotune.save(path, key, {"superpose_grid": 2048})
    autotune.save(path, "other", {"superpose_grid": 1024})
    assert autotune.load(path, key) == {"superpose_grid": 2048}


def test_small_gains_keep_the_current_value():
    eng = FakeEngine()
    # transfer shape 3 is only 1 % faster than the default 0: below the 2 % threshold
    cost = lambda: 1.0 + (0.0 if eng.t["transfer_shape"] == 3 else 0.01)
    res = autotune.tune(eng, {}, None, measure=cost, candidates={"transfer_shape": (0, 1, 2, 3, 4)})
    assert res["best"]["transfer_shape"] == 0
