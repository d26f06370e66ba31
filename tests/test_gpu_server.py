"""Served form: the request server (resident CTs, engine slots, admission by device memory,
resident beams with weight-only updates).

Checks: a full request equals the plain engine's dose bitwise; a resident beam's first result
and its weight-only updates agree with the plain engine at the new weights (one rounding of
IDD = W x unit IDD apart) and with the FP64 oracle within 0.5 % of Dmax; every result is the same
alone and in a mix of overlapping requests (bitwise); a request larger than the budget fails
with an out-of-memory error; the halo branch is refused for resident beams; requests of the
serving mix on each CT size agree with the FP64 oracle within 0.5 % of Dmax (the oracle takes
the float trace's entry step on rays whose cumulative HU+1000 passes within 1 of the 150 entry
threshold, where the two traces may round to different steps).
"""
import copy

import numpy as np
import pytest

import oracle as O
from subsecondpbs_amd import fixtures as F

from .oracle_stages import dose_following_entry

torch = pytest.importorskip("torch")
gpu = pytest.mark.gpu


def _plain(case, beam, entry=False):
    """The plain engine's dose of ``beam``; with ``entry``, also its per-ray entry steps."""
    from subsecondpbs_amd.engine import Engine

    eng = Engine(case["energy"])
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    nx, ny, nz = case["dose_dims"]
    d = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    eng.beam_dose(beam, d)
    eng.sync()
    out = d.cpu().numpy()
    fi = eng.debug_fetch(("fi",))["fi"] if entry else None
    eng.close()
    return (out, fi) if entry else out


def _grid(case):
    return case["dose_dims"], case["dose_resolution"], case["dose_corner"]


def _reweighted(beam, seed):
    b = dict(beam)
    rng = np.random.default_rng(seed)
    w = np.asarray(beam["spot_weights"], np.float32)
    b["spot_weights"] = (w * rng.uniform(0.5, 1.5, w.size)).astype(np.float32)
    return b


@pytest.fixture(scope="module")
def served():
    from subsecondpbs_amd.server import Server

    case = F.make_config1()
    sv = Server(case["energy"], slots=4)
    cid = sv.add_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    print("\n[gpu-test] server with 4 slots, config1 CT resident", flush=True)
    yield case, sv, cid
    sv.close()


@gpu
def test_full_request_equals_plain_engine(served):
    case, sv, cid = served
    beam = case["beams"][0]
    t = sv.submit(cid, *_grid(case), beam)
    got = sv.result(t)
    ref = _plain(case, beam)
    assert np.array_equal(got, ref)
    q, total = sv.timing(t)
    assert q >= 0.0 and total >= q


@gpu
def test_resident_beam_weight_updates(served):
    case, sv, cid = served
    beam = case["beams"][0]
    h = sv.submit(cid, *_grid(case), beam, keep=True)
    first = sv.result(h)
    ref = _plain(case, beam)
    assert np.abs(first - ref).max() <= 1e-5 * ref.max()
    for seed in (1, 2):
        b2 = _reweighted(beam, seed)
        t = sv.submit_weights(h, b2["spot_weights"])
        got = sv.result(t)
        plain = _plain(case, b2)
        err_plain = np.abs(got - plain).max() / plain.max()
        c2 = copy.deepcopy(case)
        c2["beams"] = [b2]
        orc, _ = O.beam_dose(c2)
        err_orc = np.abs(got.astype(np.float64) - orc).max() / orc.max()
        print(f"weights seed {seed}: |served - plain|/Dmax = {err_plain:.2e}, vs FP64 = {err_orc:.2e}")
        assert err_plain <= 1e-5
        assert err_orc <= 5e-3
    sv.release(h)


@gpu
def test_same_result_alone_and_in_the_mix(served):
    case, sv, cid = served
    beam = case["beams"][0]
    alone = sv.result(sv.submit(cid, *_grid(case), beam))
    h = sv.submit(cid, *_grid(case), beam, keep=True)
    sv.result(h)
    w2 = _reweighted(beam, 7)["spot_weights"]
    upd_alone = sv.result(sv.submit_weights(h, w2))
    # a mix of overlapping requests: full beams at other weights, updates, the probes in between
    tickets, probes, upd_probes = [], [], []
    for i in range(12):
        if i % 4 == 1:
            probes.append(sv.submit(cid, *_grid(case), beam))
        elif i % 4 == 3:
            upd_probes.append(sv.submit_weights(h, w2))
        elif i % 2 == 0:
            tickets.append(sv.submit(cid, *_grid(case), _reweighted(beam, 100 + i)))
        else:
            tickets.append(sv.submit_weights(h, _reweighted(beam, 200 + i)["spot_weights"]))
    for t in probes:
        assert np.array_equal(sv.result(t), alone)
    for t in upd_probes:
        assert np.array_equal(sv.result(t), upd_alone)
    for t in tickets:
        sv.result(t)
    sv.release(h)


@gpu
def test_budget_too_small_fails_cleanly():
    from subsecondpbs_amd.server import Server

    case = F.make_config1()
    sv = Server(case["energy"], slots=1, memory_budget_bytes=2.0e6)
    cid = sv.add_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    t = sv.submit(cid, *_grid(case), case["beams"][0])
    with pytest.raises(RuntimeError, match="budget"):
        sv.result(t)
    sv.close()


@gpu
@pytest.mark.parametrize("ri", [0, 4, 7])
def test_serving_mix_requests_against_oracle(ri):
    """Requests of the serving mix on each CT size (512x512x200 with a 256x256x100 dose grid,
    128^3, 256x256x160) through the server, against the FP64 oracle within 0.5 % of Dmax. Request
    7 has rays whose cumulative HU+1000 passes within 2e-4 of the entry threshold."""
    from subsecondpbs_amd.server import Server

    cfg = F.make_config5()
    r = cfg["requests"][ri]
    c = cfg["cts"][r["ct"]]
    print(f"\n[gpu-test] serving-mix request {ri}: CT {c['ct_dims']}, {len(r['beam']['energies'])} layers, "
          f"{np.asarray(r['beam']['spot_weights']).size} spots", flush=True)
    sv = Server(cfg["energy"], slots=2)
    cid = sv.add_ct(c["ct"], c["ct_dims"], c["ct_resolution"], c["ct_corner"])
    got = sv.result(sv.submit(cid, c["dose_dims"], c["dose_resolution"], c["dose_corner"], r["beam"]))
    sv.close()
    case = dict(c)
    case["beams"] = [r["beam"]]
    case["energy"] = cfg["energy"]
    # the same beam through a plain engine: bitwise the server's dose, and its entry steps
    plain, fi = _plain(case, r["beam"], entry=True)
    assert np.array_equal(plain, got)
    ref, n_followed = dose_following_entry(case, fi)
    err = float(np.abs(got.astype(np.float64) - ref).max() / ref.max())
    print(f"request {ri}: max|dD|/Dmax = {err:.3e} ({n_followed} rays on the entry threshold followed)")
    assert err <= 5e-3


@gpu
def test_serving_mix_device_results_every_pass():
    """The whole serving mix (48 requests on three CT sizes, 9 resident beams, 11 weight-only
    updates) through 4 slots, three times, with every result handed over on the device at once:
    each equals the plain engine's dose (full requests bitwise; updates within one rounding of
    IDD = W x unit IDD). Regression test for dose buffers recycled while a result copy of the
    same memory could still be pending."""
    from subsecondpbs_amd.engine import Engine
    from subsecondpbs_amd.server import Server

    cfg = F.make_config5(n_requests=48)
    reqs, cts = cfg["requests"], cfg["cts"]
    plain = []
    engines = []
    for c in cts:
        e = Engine(cfg["energy"])
        e.set_ct(c["ct"], c["ct_dims"], c["ct_resolution"], c["ct_corner"])
        e.set_dose_grid(c["dose_dims"], c["dose_resolution"], c["dose_corner"])
        engines.append(e)
    for r in reqs:
        nx, ny, nz = cts[r["ct"]]["dose_dims"]
        d = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
        engines[r["ct"]].beam_dose(r["beam"], d)
        engines[r["ct"]].sync()
        plain.append(d)
    for e in engines:
        e.close()
    kept = {r["update_of"] for r in reqs if r["update_of"] >= 0}
    sv = Server(cfg["energy"], slots=4)
    cids = [sv.add_ct(c["ct"], c["ct_dims"], c["ct_resolution"], c["ct_corner"]) for c in cts]
    for rep in range(3):
        handles, tickets = {}, []
        for i, r in enumerate(reqs):
            c = cts[r["ct"]]
            if r["update_of"] >= 0:
                tickets.append(sv.submit_weights(handles[r["update_of"]], r["beam"]["spot_weights"]))
            else:
                t = sv.submit(cids[r["ct"]], *_grid(c), r["beam"], keep=i in kept)
                tickets.append(t)
                if i in kept:
                    handles[i] = t
        res = [sv.result(t, device=True) for t in tickets]
        for h in handles.values():
            sv.release(h)
        for i, (a, b) in enumerate(zip(res, plain)):
            err = float((a - b).abs().max().item()) / float(b.abs().max().item())
            if reqs[i]["update_of"] < 0 and i not in kept:
                assert err == 0.0, f"pass {rep} request {i}: {err:.3e}"
            else:
                assert err < 1e-5, f"pass {rep} request {i} (update of {reqs[i]['update_of']}): {err:.3e}"
    sv.close()
