"""The multi-beam batch on the GPU: every pass of the chain runs once for all beams of a batch
(one launch per pass and radius class), and the transfers add the beams voxel by voxel in beam
order, so the batch's dose equals the one-beam-at-a-time engine's sum bitwise. The passes of a
batch run as one HIP graph, captured on the first call and replayed while the batch's shape
repeats. Cases: the config-3 proton beams alone, the mixed proton + carbon batch (two table
sets, the carbon halo branch on) captured once and replayed twice, and a batch holding a beam
the passes do not cover (legacy subspot -> CPB path), which runs on the lanes."""
import numpy as np
import pytest

from subsecondpbs_amd import fixtures as F

gpu = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def config3():
    case = F.make_config3()
    print(f"\n[gpu-test] config3: {len(case['beams'])} beams", flush=True)
    return case


def _sequential(case, order):
    """One beam at a time on plain engines (one per table set), in the given beam order."""
    torch = _torch()
    from subsecondpbs_amd.engine import Engine
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    engines = {}
    for kind, energy, model in (("proton", case["energy"], 0),
                                ("carbon", case["carbon_energy"], case["carbon_nuclear"])):
        e = Engine(energy)
        e.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
        e.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
        e.set_nuclear(model)
        engines[kind] = e
    for i in order:
        e = engines[case["kinds"][i]]
        e.beam_dose(case["beams"][i], dose)
        e.sync()
    for e in engines.values():
        e.close()
    return dose.cpu().numpy()


def _batch(case, order, repeats=1):
    torch = _torch()
    from subsecondpbs_amd.batch import Batch
    nx, ny, nz = case["dose_dims"]
    bt = Batch(case["energy"], lanes=4)
    carbon = bt.add_table(case["carbon_energy"], case["carbon_nuclear"])
    bt.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    bt.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    beams = bt.prepare([case["beams"][i] for i in order])
    tables = [carbon if case["kinds"][i] == "carbon" else 0 for i in order]
    out = []
    for _ in range(repeats):
        dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
        bt.dose(beams, dose, tables=tables)
        bt.sync()
        out.append(dose.cpu().numpy())
    launches = bt.last_launches
    captures = bt.graph_captures
    bt.close()
    assert captures == 1  # the first call captures the passes' graph, the repeats replay it
    return out, launches


@gpu
def test_proton_batch_equals_sequential_bitwise(config3):
    order = [i for i, k in enumerate(config3["kinds"]) if k == "proton"]
    (got,), launches = _batch(config3, order)
    ref = _sequential(config3, order)
    print(f"12 proton beams: {launches} launches, max |d| = {np.abs(got - ref).max():.3e}")
    assert launches <= 30
    assert np.array_equal(got, ref)


@gpu
def test_mixed_batch_equals_sequential_bitwise_and_repeats(config3):
    # carbons interleaved with protons: the table sets share the passes
    order = [0, 12, 1, 2, 13, 3, 4, 5, 14, 6, 7, 8, 15, 9, 10, 11]
    outs, launches = _batch(config3, order, repeats=3)
    ref = _sequential(config3, order)
    print(f"16 mixed beams: {launches} launches, max |d| = {np.abs(outs[0] - ref).max():.3e}")
    assert launches <= 40
    for o in outs:  # the captured graph and its replays, each against the eager one-beam sum
        assert np.array_equal(o, ref)


@gpu
def test_batch_with_a_legacy_beam_runs_on_the_lanes(config3):
    """A subspot -> CPB beam is outside the multi-beam passes: the batch falls back to the lanes
    (no multi-beam launch), and its dose still matches the sequential sum to rounding."""
    torch = _torch()
    from subsecondpbs_amd.batch import Batch
    case = config3
    b = dict(case["beams"][0])
    L = len(b["energies"])
    counts = np.asarray(b["layer_spot_counts"])
    pos = np.asarray(b["spot_positions"], np.float32).reshape(-1, 2)
    w = np.asarray(b["spot_weights"], np.float32)
    S = int(counts.max())
    sub = np.zeros((L, S, 5), np.float32)
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
    beams = [b, case["beams"][1]]
    nx, ny, nz = case["dose_dims"]
    bt = Batch(case["energy"], lanes=2)
    bt.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    bt.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    bt.dose(beams, dose)
    bt.sync()
    assert bt.last_launches == 0
    bt.close()
    from subsecondpbs_amd.engine import Engine
    e = Engine(case["energy"])
    e.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    e.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    ref = torch.zeros_like(dose)
    for x in beams:
        e.beam_dose(x, ref)
        e.sync()
    e.close()
    got, want = dose.cpu().numpy(), ref.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


@gpu
def test_submits_into_one_buffer_accumulate_in_order(config3):
    """Three served requests into the same dose buffer, submitted back to back: they share one
    lane (the transfer adds without atomics), so the buffer holds the three beams' sum in
    submission order, bitwise the plain engine's sequential sum; a request into another buffer
    meanwhile takes the next lane and still gets its own beam's dose."""
    torch = _torch()
    from subsecondpbs_amd.batch import Batch
    case = config3
    order = [0, 1, 2]
    nx, ny, nz = case["dose_dims"]
    bt = Batch(case["energy"], lanes=4)
    bt.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    bt.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    shared = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    other = torch.zeros_like(shared)
    tickets = [bt.submit(case["beams"][order[0]], shared), bt.submit(case["beams"][3], other)]
    tickets += [bt.submit(case["beams"][i], shared) for i in order[1:]]
    for t in tickets:
        bt.wait(t)
    got, got_other = shared.cpu().numpy(), other.cpu().numpy()
    bt.close()
    ref = _sequential(case, order)
    ref_other = _sequential(case, [3])
    print(f"3 requests into one buffer: max |d| = {np.abs(got - ref).max():.3e}")
    assert np.array_equal(got, ref)
    assert np.array_equal(got_other, ref_other)
